/* oracle_trunc.c -- CPU restatement of the truncated QDB scanners
 * RefTrCYKScan (src/cm_dpsearch_trunc.c:79-700) and RefITrInsideScan
 * (:741-1338).  TEST INFRASTRUCTURE ONLY (see oracle.h). */
#include <stdio.h>
#include "oracle.h"

int oracle_trcyk(const CM_t *cm, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0, float cutoff,
                 int do_null3, int greedy, float env_cutoff, int want_rows, ORACLE_RESULT *res, char *errbuf)
{
  if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "truncated oracle not built yet");
  return eslEINVAL;
}
int oracle_triinside(const CM_t *cm, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0, float cutoff,
                     int do_null3, int greedy, float env_cutoff, int want_rows, ORACLE_RESULT *res, char *errbuf)
{
  if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "truncated oracle not built yet");
  return eslEINVAL;
}
