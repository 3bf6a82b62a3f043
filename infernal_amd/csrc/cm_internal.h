/* cm_internal.h -- private helpers shared by the host-side sources of the
 * library (model I/O and configuration, hit lists, GPU scanners).
 */
#ifndef INFX_CM_INTERNAL_H
#define INFX_CM_INTERNAL_H

#include <stdarg.h>
#include <stdio.h>
#include "../../include/infernal_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ESL_MAX(a,b) (((a)>(b))?(a):(b))
#define ESL_MIN(a,b) (((a)<(b))?(a):(b))

/* Fill errbuf (if non-NULL) and return status, as Easel's ESL_FAIL. */
int infx_fail(int status, char *errbuf, const char *fmt, ...);

/* Alphabet (Easel RNA layout). */
extern const int   infx_ndegen[INFX_KP];
extern const char  infx_degen[INFX_KP][INFX_K];
extern const char  infx_sym[INFX_KP + 1];

/* State helpers (src/cm.c:1177-1270, :1775). */
int StateDelta(int sttype);
int StateLeftDelta(int sttype);
int StateRightDelta(int sttype);
int Emitmode(int sttype);
int StateIsDetached(const CM_t *cm, int v);

/* Easel vector helpers restated from Easel's documented behaviour. */
float infx_vec_FSum(const float *vec, int n);
void  infx_vec_FNorm(float *vec, int n);

/* Log-sum lookups (tables built by FLogsumInit/init_ilogsum). */
const float *infx_flogsum_table(void);
const int   *infx_ilogsum_table(void);

/* Init scores (src/cm.c:4100-4183): [M][W+1] blocks, caller frees [0] and row array. */
float **FCalcInitDPScores(const CM_t *cm);
int   **ICalcInitDPScores(const CM_t *cm);
float   Scorify(int sc);

/* NULL3 (src/cm_parsetree.c:2352) and act-vector count (Easel esl_abc_DCount). */
void ScoreCorrectionNull3(const float *null0, float *comp, int len, float omega, float *ret_sc);
void infx_abc_DCount(double *ct, ESL_DSQ x, double wt);

/* Hit-list internals shared with the scanners. */
int  cm_tophits_Grow(CM_TOPHITS *th);

/* Standard greedy/gamma hit resolution from a candidate set, used by the GPU
 * scanners after the device pass (see cm_hits.cpp). */
typedef struct infx_cand_s {
  int32_t j;        /* end position (absolute) */
  int32_t d;        /* length */
  float   sc;       /* bestsc[d] at j */
  int32_t r;        /* bestr[d] */
  int32_t mode;     /* TRMODE_* (TRMODE_J for standard scanners) */
} INFX_CAND;

int infx_resolve_hits(CM_t *cm, char *errbuf, int pass_idx, int truncated, INFX_CAND *cand, int64_t ncand,
                      ESL_DSQ *dsq, int64_t i0, int64_t j0, int W, float cutoff, int do_null3,
                      int greedy, CM_TOPHITS *hitlist);

#ifdef __cplusplus
}
#endif
#endif
