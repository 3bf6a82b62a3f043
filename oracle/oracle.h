/* oracle.h -- CPU restatement of Infernal's QDB scanners.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the GPU scanners are
 * compared against (tests/, __graft_entry__.smoke(), and bench.py's CPU
 * baseline).  The product library never links, loads or calls it.
 *
 * It restates, on the reference's own data layout (full alpha[2][M][W+1]
 * rows, alpha_begl[W+1][nbegl][W+1] decks) and in the reference's operation
 * order, the scanners of src/cm_dpsearch.c and src/cm_dpsearch_trunc.c, the
 * scan-matrix set-up of src/cm_mx.c, the hit reporters of src/cm_mx.c and the
 * overlap removal of src/cm_tophits.c.  It reads a configured model through
 * the CM_t mirror declared in include/infernal_gpu.h.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stdint.h>
#include "../include/infernal_gpu.h"

#define ORACLE_CYK      0
#define ORACLE_FINSIDE  1
#define ORACLE_IINSIDE  2
#define ORACLE_TRCYK    3
#define ORACLE_TRIINSIDE 4

typedef struct oracle_hit_s {
  int64_t start, stop;
  int     root;
  int     mode;
  float   score;
} ORACLE_HIT;

typedef struct oracle_result_s {
  int         nhits;
  ORACLE_HIT *hits;          /* kept hits, in the order they are appended to the caller's list */
  int64_t     envi, envj;    /* -1 if no cell reached env_cutoff */
  float       sc;            /* best score over all (j, d) */
  int         mode;          /* truncated scanners: mode of the best hit */
  float      *vsc;           /* [M] best score per state */
  /* optional per-j dump of the root row: for each j in i0..j0 and each
   * d = 0..W, bestsc and bestr; allocated only if want_rows was set */
  int         W;
  float      *rows_sc;       /* [(j-i0)*(W+1) + d] */
  int        *rows_r;
  char       *rows_mode;
  /* truncated scanners: per (j, d) and mode (TRMODE_T/R/L/J = 0..3) the
   * best truncated begin max_y X_mode[y][d] + penalty[y], before the root's
   * record rule (IMPOSSIBLE where no begin reaches the cell) */
  float      *rows_msc;      /* [((j-i0)*(W+1) + d)*4 + mode] */
} ORACLE_RESULT;

/* Run one scan.  algo: ORACLE_*; greedy: 1 greedy overlap removal, 0 the
 * gamma semi-HMM; pass_idx is used by the truncated scanners.  Returns eslOK
 * or an Easel-style error code (message in errbuf). */
int  oracle_scan(const CM_t *cm, int algo, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0,
                 float cutoff, int do_null3, int greedy, float env_cutoff, int want_rows,
                 ORACLE_RESULT *res, char *errbuf);
void oracle_result_free(ORACLE_RESULT *res);

/* Hit accumulation driven directly (the reporters the scanners call per j):
 * begin, one update per end position j with that row's bestsc/bestr[/bestmode]
 * over d = dmin..dmax, finish (kept hits into res; res NULL: discard). */
int  oracle_trunc_hits_begin(void **acc, int L, int64_t i0, int64_t j0, int W, int greedy);
int  oracle_trunc_hits_update(void *acc, const CM_t *cm, int pass_idx, int j, int dmin, int dmax, const float *bestsc,
                              const int *bestr, const char *bestmode, const double *act, float cutoff, int do_null3);
int  oracle_trunc_hits_finish(void *acc, ORACLE_RESULT *res, char *errbuf);
/* Single rules, for the tests that pin them: cm_hit_AllowTruncation, and the
 * NULL3 correction of hit i..j computed as a reporter computes it. */
int   oracle_allow_truncation(const CM_t *cm, int pass_idx, int64_t start, int64_t stop, int64_t i0, int64_t j0, int mode, int b);
float oracle_null3_hit(const CM_t *cm, const ESL_DSQ *dsq, int64_t i, int64_t j);

/* The oracle's own log-sum tables (src/logsum.c:74-148), for table checks. */
const float *oracle_flogsum_table(void);
const int   *oracle_ilogsum_table(void);

#endif
