/* oracle.c -- CPU restatement of Infernal's QDB scanners.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the GPU scanners, used by tests/,
 * __graft_entry__.smoke() and bench.py's CPU baseline.  Never linked into the
 * product library.
 *
 * Built with gcc -O3 -ffp-contract=off and no -march flag (x86-64 baseline,
 * no FMA), as the reference's default CFLAGS="-O3" build (configure.ac:306).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ESL_MAX(a,b) (((a)>(b))?(a):(b))
#define ESL_MIN(a,b) (((a)<(b))?(a):(b))
/* src/infernal.h:153 */
#define sreLOG2(x)  ((x) > 0 ? log(x) * 1.44269504 : IMPOSSIBLE)
#define sreEXP2(x)  (exp((x) * 0.69314718 ))
#define IADD(a,b)   ((int) ((unsigned) (a) + (unsigned) (b)))

/* ------------------------------------------------------------------------ */
/* log-sum tables: src/logsum.c:72-169                                        */

static float flogsum_lookup[LOGSUM_TBL];
static int   ilogsum_lookup[LOGSUM_TBL];
static int   tables_ready = 0;

static void oracle_tables_init(void)
{
  if (tables_ready) return;
  for (int i = 0; i < LOGSUM_TBL; i++) {
    flogsum_lookup[i] = sreLOG2(1. + sreEXP2((double) -i / INTSCALE));
    ilogsum_lookup[i] = rint(INTSCALE * (sreLOG2(1. + sreEXP2((double) -i / INTSCALE))));
  }
  tables_ready = 1;
}
const float *oracle_flogsum_table(void) { oracle_tables_init(); return flogsum_lookup; }
const int   *oracle_ilogsum_table(void) { oracle_tables_init(); return ilogsum_lookup; }

static inline float oracle_FLogsum(float s1, float s2)
{
  const float max = ESL_MAX(s1, s2);
  const float min = ESL_MIN(s1, s2);
  return (min == -INFINITY || (max - min) >= 23.f) ? max : max + flogsum_lookup[(int) ((max - min) * INTSCALE)];
}
static inline int oracle_ILogsum(int s1, int s2)
{
  const int max = ESL_MAX(-INFTY, ESL_MAX(s1, s2));
  const int min = ESL_MIN(s1, s2);
  return (min <= -INFTY || (max - min) >= LOGSUM_TBL) ? max : IADD(max, ilogsum_lookup[max - min]);
}
static inline float oracle_Scorify(int sc) { return (sc == -INFTY) ? IMPOSSIBLE : ((float) sc / INTSCALE); }

/* ------------------------------------------------------------------------ */
/* state helpers: src/cm.c:1177-1270                                          */

static int oracle_sdelta(int t) { return t == MP_st ? 2 : (t == ML_st || t == MR_st || t == IL_st || t == IR_st) ? 1 : 0; }
static int oracle_rdelta(int t) { return (t == MP_st || t == MR_st || t == IR_st) ? 1 : 0; }
static int oracle_emitmode(int t)
{
  switch (t) { case IL_st: case ML_st: return EMITLEFT; case IR_st: case MR_st: return EMITRIGHT; case MP_st: return EMITPAIR; }
  return EMITNONE;
}

/* Easel esl_abc_DCount for the RNA alphabet: a degenerate code spreads its
 * count evenly over the residues it stands for. */
static void oracle_DCount(double *ct, ESL_DSQ x, double wt)
{
  static const char deg[16][4] = {
    {1,0,0,0},{0,1,0,0},{0,0,1,0},{0,0,0,1},{0,0,0,0},{1,0,1,0},{0,1,0,1},{1,1,0,0},
    {0,0,1,1},{0,1,1,0},{1,0,0,1},{1,1,0,1},{0,1,1,1},{1,1,1,0},{1,0,1,1},{1,1,1,1}};
  if (x < 4) { ct[x] += wt; return; }
  if (x >= 16) return;
  int n = deg[x][0] + deg[x][1] + deg[x][2] + deg[x][3];
  for (int y = 0; y < 4; y++) if (deg[x][y]) ct[y] += wt / (double) n;
}

/* Easel esl_vec_FNorm (Kahan-summed esl_vec_FSum) */
static void oracle_FNorm(float *vec, int n)
{
  float sum = 0., c = 0., yy, t;
  for (int i = 0; i < n; i++) { yy = vec[i] - c; t = sum + yy; c = (t - sum) - yy; sum = t; }
  if (sum != 0.0) for (int x = 0; x < n; x++) vec[x] /= sum;
  else            for (int x = 0; x < n; x++) vec[x] = 1. / (float) n;
}

/* src/cm_parsetree.c:2352 ScoreCorrectionNull3 (+ LogSum2, src/logsum.c:151) */
static float oracle_null3(const float *null0, float *comp, int len, float omega)
{
  float score = 0.;
  for (int a = 0; a < INFX_K; a++) score += sreLOG2(comp[a] / null0[a]) * comp[a] * len;
  score += sreLOG2(omega);
  const float s1 = 0.f, s2 = score;
  const float max = ESL_MAX(s1, s2), min = ESL_MIN(s1, s2);
  return (min == -INFINITY || (max - min) >= 23.f) ? max : max + flogsum_lookup[(int) ((max - min) * INTSCALE)];
}

/* ------------------------------------------------------------------------ */
/* hit accumulation: ReportHitsGreedily src/cm_mx.c:7547, UpdateGammaHitMx
 * :7440, TBackGammaHitMx :7640; overlap removal src/cm_tophits.c:220-245,
 * :400-420, :1097-1150 and the clone loop of src/cm_dpsearch.c:651-661.     */

typedef struct { int64_t start, stop; int root, mode; float score; int removed; } HREC;
typedef struct {
  int greedy, L, W;
  float cutoff;
  int64_t i0, j0;
  HREC *h; int n, nalloc;
  float *mx, *savesc; int64_t *gback; int *saver, *savemode;
} HITACC;

static void hitacc_init(HITACC *a, int L, int64_t i0, int64_t j0, int W, int greedy)
{
  memset(a, 0, sizeof(*a));
  a->greedy = greedy; a->L = L; a->W = W; a->i0 = i0; a->j0 = j0;
  if (!greedy) {
    a->mx = (float *) malloc(sizeof(float) * (L + 1));
    a->savesc = (float *) malloc(sizeof(float) * (L + 1));
    a->gback = (int64_t *) malloc(sizeof(int64_t) * (L + 1));
    a->saver = (int *) malloc(sizeof(int) * (L + 1));
    a->savemode = (int *) malloc(sizeof(int) * (L + 1));
    a->mx[0] = 0; a->gback[0] = -1;
  }
}
static void hitacc_free(HITACC *a)
{
  free(a->h); free(a->mx); free(a->savesc); free(a->gback); free(a->saver); free(a->savemode);
  memset(a, 0, sizeof(*a));
}
static void hitacc_push(HITACC *a, int64_t start, int64_t stop, int root, int mode, float score)
{
  if (a->n == a->nalloc) { a->nalloc = a->nalloc ? 2 * a->nalloc : 256; a->h = (HREC *) realloc(a->h, sizeof(HREC) * a->nalloc); }
  HREC *r = &a->h[a->n++];
  r->start = start; r->stop = stop; r->root = root; r->mode = mode; r->score = score; r->removed = 0;
}

static int oracle_allow_trunc(const CM_t *cm, int pass_idx, int64_t start, int64_t stop, int64_t i0, int64_t j0, int mode, int b);

static float null3_corr(const CM_t *cm, const double *act, int W, int64_t ip, int64_t jp, int d)
{
  float comp[INFX_K];
  for (int a = 0; a < INFX_K; a++) comp[a] = act[(size_t) (jp % (W + 1)) * INFX_K + a] - act[(size_t) ((ip - 1) % (W + 1)) * INFX_K + a];
  oracle_FNorm(comp, INFX_K);
  return oracle_null3(cm->null, comp, d, cm->null3_omega);
}

static int hitacc_update(HITACC *a, const CM_t *cm, int pass_idx, int j, int dmin, int dmax, const float *bestsc, const int *bestr,
                         const char *bestmode, const double *act, float cutoff, int do_null3)
{
  int64_t jp = j - a->i0 + 1;
  a->cutoff = cutoff;
  if (a->greedy) {
    if (dmin > dmax) return eslOK;
    float max_sc_reported = IMPOSSIBLE;
    for (int d = ESL_MAX(1, dmin); d <= dmax; d++) {
      int64_t i = j - d + 1;
      int mode = (bestmode == NULL) ? TRMODE_J : bestmode[d];
      float hit_sc = bestsc[d];
      if (hit_sc > max_sc_reported && hit_sc >= cutoff && NOT_IMPOSSIBLE(hit_sc)) {
        int do_report = 1;
        if (bestmode != NULL) do_report = oracle_allow_trunc(cm, pass_idx, i, j, a->i0, a->j0, mode, bestr[d]);
        if (do_report && do_null3) {
          int64_t ip = i - a->i0 + 1;
          hit_sc -= null3_corr(cm, act, a->W, ip, jp, d);
          do_report = (hit_sc > max_sc_reported && hit_sc >= cutoff) ? 1 : 0;
        }
        if (do_report) { hitacc_push(a, i, j, bestr[d], mode, hit_sc); max_sc_reported = hit_sc; }
      }
    }
    return eslOK;
  }
  a->mx[jp] = a->mx[jp - 1] + 0;
  a->gback[jp] = -1; a->savesc[jp] = IMPOSSIBLE; a->saver[jp] = -1; a->savemode[jp] = -1;
  if (dmin > dmax) return eslOK;
  for (int d = ESL_MAX(1, dmin); d <= dmax; d++) {
    int64_t i = j - d + 1, ip = jp - d + 1;
    int mode = (bestmode == NULL) ? TRMODE_J : bestmode[d];
    float hit_sc = bestsc[d];
    float cumulative_sc = a->mx[ip - 1] + hit_sc;
    if (cumulative_sc > a->mx[jp] && NOT_IMPOSSIBLE(hit_sc)) {
      int do_report = 1;
      if (bestmode != NULL) do_report = oracle_allow_trunc(cm, pass_idx, i, j, a->i0, a->j0, mode, bestr[d]);
      if (do_report && do_null3) {
        float corr = null3_corr(cm, act, a->W, ip, jp, d);
        hit_sc -= corr;
        cumulative_sc -= corr;
        do_report = (cumulative_sc > a->mx[jp]) ? 1 : 0;
      }
      if (do_report) { a->mx[jp] = cumulative_sc; a->gback[jp] = i; a->savesc[jp] = hit_sc; a->saver[jp] = bestr[d]; a->savemode[jp] = mode; }
    }
  }
  return eslOK;
}

static int hit_cmp(const void *p1, const void *p2)
{
  const HREC *h1 = *(const HREC * const *) p1, *h2 = *(const HREC * const *) p2;
  if (h1->score < h2->score) return 1;
  if (h1->score > h2->score) return -1;
  if (h1->start > h2->start) return 1;
  if (h1->start < h2->start) return -1;
  return 0;
}

static int hitacc_finish_gamma(HITACC *a, float cutoff, ORACLE_RESULT *res)
{
  int64_t j = a->j0;
  res->hits = (ORACLE_HIT *) malloc(sizeof(ORACLE_HIT) * ((size_t) a->L + 1));
  res->nhits = 0;
  while (j >= a->i0) {
    int64_t jp = j - a->i0 + 1;
    if (a->gback[jp] == -1) j--;
    else {
      if (a->savesc[jp] >= cutoff) {
        ORACLE_HIT *o = &res->hits[res->nhits++];
        o->start = a->gback[jp]; o->stop = j; o->root = a->saver[jp]; o->mode = a->savemode[jp]; o->score = a->savesc[jp];
      }
      j = a->gback[jp] - 1;
    }
  }
  return eslOK;
}

static int hitacc_finish(HITACC *a, ORACLE_RESULT *res, char *errbuf)
{
  (void) errbuf;
  res->nhits = 0; res->hits = NULL;
  if (!a->greedy) return hitacc_finish_gamma(a, a->cutoff, res);
  const size_t ncap = a->n > 0 ? (size_t) a->n : 1;
  HREC **p = (HREC **) malloc(sizeof(HREC *) * ncap);
  for (int h = 0; h < a->n; h++) p[h] = &a->h[h];
  if (a->n > 1) qsort(p, a->n, sizeof(HREC *), hit_cmp);
  for (int i = 0; i < a->n; i++) {
    if (p[i]->removed) continue;
    for (int k = i + 1; k < a->n; k++) {
      if (p[k]->removed) continue;
      if (!(p[k]->stop < p[i]->start) && !(p[i]->stop < p[k]->start)) p[k]->removed = 1;
    }
  }
  res->hits = (ORACLE_HIT *) malloc(sizeof(ORACLE_HIT) * ncap);
  for (int i = 0; i < a->n; i++) {
    if (p[i]->removed) continue;
    ORACLE_HIT *o = &res->hits[res->nhits++];
    o->start = p[i]->start; o->stop = p[i]->stop; o->root = p[i]->root; o->mode = p[i]->mode; o->score = p[i]->score;
  }
  free(p);
  return eslOK;
}

/* src/cm_tophits.c:3048 cm_hit_AllowTruncation */
static int oracle_allow_trunc(const CM_t *cm, int pass_idx, int64_t start, int64_t stop, int64_t i0, int64_t j0, int mode, int b)
{
  int in_local_mode = (cm->flags & CMH_LOCAL_BEGIN) ? 1 : 0;
  int nd = cm->ndidx[b];
  int lpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd) ? cm->emap->lpos[nd] : cm->emap->lpos[nd] + 1;
  int rpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATR_nd) ? cm->emap->rpos[nd] : cm->emap->rpos[nd] - 1;
  if (pass_idx == PLI_PASS_STD_ANY || pass_idx == PLI_PASS_5P_AND_3P_ANY || pass_idx == PLI_PASS_HMM_ONLY_ANY) return 1;
  if (start == i0 && stop == j0) return 1;
  if (in_local_mode) {
    switch (mode) { case TRMODE_J: return 1; case TRMODE_L: return stop == j0; case TRMODE_R: return start == i0; default: return 0; }
  }
  switch (mode) {
  case TRMODE_J: return (lpos == 1 && rpos == cm->clen);
  case TRMODE_L: return (stop == j0 && lpos == 1);
  case TRMODE_R: return (start == i0 && rpos == cm->clen);
  default: return 0;
  }
}

/* ------------------------------------------------------------------------ */
/* init scores and d = 0 cells                                                */

/* src/cm.c:4100 FCalcInitDPScores */
static float **oracle_FInitSc(const CM_t *cm)
{
  float *el = (float *) malloc(sizeof(float) * (cm->W + 1));
  for (int d = 0; d <= cm->W; d++) el[d] = cm->el_selfsc * d;
  float **a = (float **) malloc(sizeof(float *) * cm->M);
  a[0] = (float *) malloc(sizeof(float) * (size_t) cm->M * (cm->W + 1));
  for (int v = 0; v < cm->M; v++) {
    a[v] = a[0] + (size_t) v * (cm->W + 1);
    if (NOT_IMPOSSIBLE(cm->endsc[v])) for (int d = 0; d <= cm->W; d++) a[v][d] = el[d] + cm->endsc[v];
    else                              for (int d = 0; d <= cm->W; d++) a[v][d] = IMPOSSIBLE;
  }
  free(el);
  return a;
}
/* src/cm.c:4146 ICalcInitDPScores */
static int **oracle_IInitSc(const CM_t *cm)
{
  int *el = (int *) malloc(sizeof(int) * (cm->W + 1));
  for (int d = 0; d <= cm->W; d++) el[d] = (int) ((unsigned) cm->iel_selfsc * (unsigned) d);
  int **a = (int **) malloc(sizeof(int *) * cm->M);
  a[0] = (int *) malloc(sizeof(int) * (size_t) cm->M * (cm->W + 1));
  for (int v = 0; v < cm->M; v++) {
    a[v] = a[0] + (size_t) v * (cm->W + 1);
    if (cm->iendsc[v] != -INFTY) for (int d = 0; d <= cm->W; d++) a[v][d] = IADD(el[d], cm->iendsc[v]);
    else                         for (int d = 0; d <= cm->W; d++) a[v][d] = -INFTY;
  }
  free(el);
  return a;
}

/* src/cm_mx.c:6039 cm_scan_mx_InitializeFloats */
static void oracle_InitD0F(const CM_t *cm, float *alpha, float *begl, const int *begl_idx, int nbegl, int M, int Wsmx)
{
  const size_t W1 = (size_t) Wsmx + 1;
#define A(jj, vv, dd)  alpha[((size_t) (jj) * M + (vv)) * W1 + (dd)]
#define BG(jj, vv, dd) begl[((size_t) (jj) * nbegl + begl_idx[vv]) * W1 + (dd)]
  for (size_t x = 0; x < 2 * (size_t) M * W1; x++) alpha[x] = IMPOSSIBLE;
  for (size_t x = 0; x < (size_t) nbegl * W1 * W1; x++) begl[x] = IMPOSSIBLE;
  for (int v = M - 1; v >= 0; v--) {
    if (cm->stid[v] != BEGL_S) {
      if (cm->sttype[v] == E_st) { A(0, v, 0) = A(1, v, 0) = 0.; }
      else if (cm->sttype[v] == S_st || cm->sttype[v] == D_st) {
        int y = cm->cfirst[v];
        A(0, v, 0) = cm->endsc[v];
        for (int k = 0; k < cm->cnum[v]; k++) A(0, v, 0) = ESL_MAX(A(0, v, 0), (A(0, y + k, 0) + cm->tsc[v][k]));
        A(0, v, 0) = ESL_MAX(A(0, v, 0), IMPOSSIBLE);
      }
      else if (cm->sttype[v] == B_st) {
        int w = cm->cfirst[v], y = cm->cnum[v];
        A(0, v, 0) = BG(0, w, 0) + A(0, y, 0);
      }
      A(1, v, 0) = A(0, v, 0);
    } else {
      int y = cm->cfirst[v];
      BG(0, v, 0) = cm->endsc[v];
      for (int k = 0; k < cm->cnum[v]; k++) BG(0, v, 0) = ESL_MAX(BG(0, v, 0), (A(0, y + k, 0) + cm->tsc[v][k]));
      BG(0, v, 0) = ESL_MAX(BG(0, v, 0), IMPOSSIBLE);
      for (int j = 1; j <= Wsmx; j++) BG(j, v, 0) = BG(0, v, 0);
    }
  }
#undef A
#undef BG
}
/* src/cm_mx.c:6095 cm_scan_mx_InitializeIntegers */
static void oracle_InitD0I(const CM_t *cm, int *alpha, int *begl, const int *begl_idx, int nbegl, int M, int Wsmx)
{
  const size_t W1 = (size_t) Wsmx + 1;
#define A(jj, vv, dd)  alpha[((size_t) (jj) * M + (vv)) * W1 + (dd)]
#define BG(jj, vv, dd) begl[((size_t) (jj) * nbegl + begl_idx[vv]) * W1 + (dd)]
  for (size_t x = 0; x < 2 * (size_t) M * W1; x++) alpha[x] = -INFTY;
  for (size_t x = 0; x < (size_t) nbegl * W1 * W1; x++) begl[x] = -INFTY;
  for (int v = M - 1; v >= 0; v--) {
    if (cm->stid[v] != BEGL_S) {
      if (cm->sttype[v] == E_st) { A(0, v, 0) = A(1, v, 0) = 0; }
      else if (cm->sttype[v] == S_st || cm->sttype[v] == D_st) {
        int y = cm->cfirst[v];
        A(0, v, 0) = cm->iendsc[v];
        for (int k = 0; k < cm->cnum[v]; k++) A(0, v, 0) = ESL_MAX(A(0, v, 0), IADD(A(0, y + k, 0), cm->itsc[v][k]));
        A(0, v, 0) = ESL_MAX(A(0, v, 0), -INFTY);
      }
      else if (cm->sttype[v] == B_st) {
        int w = cm->cfirst[v], y = cm->cnum[v];
        A(0, v, 0) = IADD(BG(0, w, 0), A(0, y, 0));
      }
      A(1, v, 0) = A(0, v, 0);
    } else {
      int y = cm->cfirst[v];
      BG(0, v, 0) = cm->iendsc[v];
      for (int k = 0; k < cm->cnum[v]; k++) BG(0, v, 0) = ESL_MAX(BG(0, v, 0), IADD(A(0, y + k, 0), cm->itsc[v][k]));
      BG(0, v, 0) = ESL_MAX(BG(0, v, 0), -INFTY);
      for (int j = 1; j <= Wsmx; j++) BG(j, v, 0) = BG(0, v, 0);
    }
  }
#undef A
#undef BG
}

/* ------------------------------------------------------------------------ */
/* the three standard scanners                                                */

/* FastCYKScan */
#define FN           oracle_cyk
#define T            float
#define ESCT         float
#define TSCT         float
#define ESC_IMP      IMPOSSIBLE
#define OP(a,b)      ESL_MAX((a),(b))
#define ADD(a,b)     ((a) + (b))
#define IMP          IMPOSSIBLE
#define ROOTMAX(a)   ESL_MAX(IMPOSSIBLE, (a))
#define TSC(v)       ((const float *) cm->tsc[v])
#define ESC(v)       ((const float *) cm->oesc[v])
#define BEGINSC(y)   cm->beginsc[y]
#define NOTIMP_BEGIN(y) NOT_IMPOSSIBLE(cm->beginsc[y])
#define SCORE(x)     (x)
#define INITSC       oracle_FInitSc
#define INIT_D0      oracle_InitD0F
#include "oracle_std_body.h"
#undef FN
#undef OP

/* FastFInsideScan */
#define FN           oracle_finside
#define OP(a,b)      oracle_FLogsum((a),(b))
#include "oracle_std_body.h"
#undef FN
#undef T
#undef ESCT
#undef TSCT
#undef ESC_IMP
#undef OP
#undef ADD
#undef IMP
#undef ROOTMAX
#undef TSC
#undef ESC
#undef BEGINSC
#undef NOTIMP_BEGIN
#undef SCORE
#undef INITSC
#undef INIT_D0

/* FastIInsideScan */
#define FN           oracle_iinside
#define T            int
#define ESCT         int
#define TSCT         int
#define ESC_IMP      (-INFTY)
#define OP(a,b)      oracle_ILogsum((a),(b))
#define ADD(a,b)     IADD((a),(b))
#define IMP          (-INFTY)
#define ROOTMAX(a)   ESL_MAX(-INFTY, (a))
#define TSC(v)       ((const int *) cm->itsc[v])
#define ESC(v)       ((const int *) cm->ioesc[v])
#define BEGINSC(y)   cm->ibeginsc[y]
#define NOTIMP_BEGIN(y) (cm->ibeginsc[y] != -INFTY)
#define SCORE(x)     oracle_Scorify(x)
#define INITSC       oracle_IInitSc
#define INIT_D0      oracle_InitD0I
#include "oracle_std_body.h"

int oracle_trcyk(const CM_t *cm, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0, float cutoff,
                 int do_null3, int greedy, float env_cutoff, int want_rows, ORACLE_RESULT *res, char *errbuf);
int oracle_triinside(const CM_t *cm, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0, float cutoff,
                     int do_null3, int greedy, float env_cutoff, int want_rows, ORACLE_RESULT *res, char *errbuf);

int oracle_scan(const CM_t *cm, int algo, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0,
                float cutoff, int do_null3, int greedy, float env_cutoff, int want_rows,
                ORACLE_RESULT *res, char *errbuf)
{
  memset(res, 0, sizeof(*res));
  if (errbuf) errbuf[0] = '\0';
  oracle_tables_init();
  if (!(cm->flags & CMH_BITS)) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "CMH_BITS flag is not raised"); return eslEINCOMPAT; }
  if (j0 < i0)                 { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "i0: %lld j0: %lld", (long long) i0, (long long) j0); return eslEINCOMPAT; }
  if (qdbidx < 0 || qdbidx > 2) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "qdbidx is invalid"); return eslEINCOMPAT; }
  int status;
  switch (algo) {
  case ORACLE_CYK:     status = oracle_cyk(cm, qdbidx, dsq, i0, j0, cutoff, do_null3, greedy, env_cutoff, want_rows, res, errbuf); break;
  case ORACLE_FINSIDE: status = oracle_finside(cm, qdbidx, dsq, i0, j0, cutoff, do_null3, greedy, env_cutoff, want_rows, res, errbuf); break;
  case ORACLE_IINSIDE: status = oracle_iinside(cm, qdbidx, dsq, i0, j0, cutoff, do_null3, greedy, env_cutoff, want_rows, res, errbuf); break;
  case ORACLE_TRCYK:   status = oracle_trcyk(cm, qdbidx, pass_idx, dsq, i0, j0, cutoff, do_null3, greedy, env_cutoff, want_rows, res, errbuf); break;
  case ORACLE_TRIINSIDE: status = oracle_triinside(cm, qdbidx, pass_idx, dsq, i0, j0, cutoff, do_null3, greedy, env_cutoff, want_rows, res, errbuf); break;
  default: if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "unknown algorithm %d", algo); return eslEINVAL;
  }
  return status;
}

void oracle_result_free(ORACLE_RESULT *res)
{
  free(res->hits); free(res->vsc); free(res->rows_sc); free(res->rows_r); free(res->rows_mode); free(res->rows_msc);
  memset(res, 0, sizeof(*res));
}

/* ------------------------------------------------------------------------ */
/* helpers for the truncated scanners (oracle_trunc.c)                        */

int oracle_trunc_hits_begin(void **acc, int L, int64_t i0, int64_t j0, int W, int greedy)
{
  HITACC *a = (HITACC *) malloc(sizeof(HITACC));
  hitacc_init(a, L, i0, j0, W, greedy);
  *acc = a;
  return eslOK;
}
int oracle_trunc_hits_update(void *acc, const CM_t *cm, int pass_idx, int j, int dmin, int dmax, const float *bestsc,
                             const int *bestr, const char *bestmode, const double *act, float cutoff, int do_null3)
{
  return hitacc_update((HITACC *) acc, cm, pass_idx, j, dmin, dmax, bestsc, bestr, bestmode, act, cutoff, do_null3);
}
int oracle_trunc_hits_finish(void *acc, ORACLE_RESULT *res, char *errbuf)
{
  HITACC *a = (HITACC *) acc;
  int st = res ? hitacc_finish(a, res, errbuf) : eslOK;
  hitacc_free(a);
  free(a);
  return st;
}
/* act[jp % (W+1)] = act[(jp-1) % (W+1)] + count(x) (src/cm_dpsearch.c:223-230) */
void oracle_act_update(double *act, int W, int jp, ESL_DSQ x)
{
  double *dst = &act[(size_t) (jp % (W + 1)) * INFX_K];
  const double *src = &act[(size_t) ((jp - 1) % (W + 1)) * INFX_K];
  for (int a = 0; a < INFX_K; a++) dst[a] = src[a];
  oracle_DCount(dst, x, 1.);
}

/* ------------------------------------------------------------------------ */
/* direct access to single rules, for the tests that pin them                 */

/* cm_hit_AllowTruncation as the hit reporters apply it */
int oracle_allow_truncation(const CM_t *cm, int pass_idx, int64_t start, int64_t stop, int64_t i0, int64_t j0, int mode, int b)
{
  return oracle_allow_trunc(cm, pass_idx, start, stop, i0, j0, mode, b);
}

/* the NULL3 correction a reporter subtracts from a hit i..j of dsq: its
 * composition from the running residue counts, as during a scan */
float oracle_null3_hit(const CM_t *cm, const ESL_DSQ *dsq, int64_t i, int64_t j)
{
  oracle_tables_init();
  const int W = (int) (j - i + 1);
  double *act = (double *) calloc((size_t) (W + 1) * INFX_K, sizeof(double));
  for (int jp = 1; jp <= W; jp++) oracle_act_update(act, W, jp, dsq[i + jp - 1]);
  float c = null3_corr(cm, act, W, 1, W, W);
  free(act);
  return c;
}
