/* scan_trunc_host.cpp -- host side of the truncated GPU scanners
 * RefTrCYKScan (src/cm_dpsearch_trunc.c:79) and RefITrInsideScan (:741). */
#include <hip/hip_runtime.h>
#include <stdlib.h>
#include "cm_internal.h"
#include "scan_host.h"

extern "C" void infx_trdev_release(void *trdev) { (void) trdev; }

extern "C" int RefTrCYKScan(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                            int64_t i0, int64_t j0, float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                            int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, char *ret_mode, float *ret_sc)
{
  return infx_fail(eslEINVAL, errbuf, "RefTrCYKScan: not built yet");
}

extern "C" int RefITrInsideScan(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                                int64_t i0, int64_t j0, float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                                int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, char *ret_mode, float *ret_sc)
{
  return infx_fail(eslEINVAL, errbuf, "RefITrInsideScan: not built yet");
}
