/* oracle_trunc.c -- CPU restatement of the truncated QDB scanners
 *   RefTrCYKScan      src/cm_dpsearch_trunc.c:79-700
 *   RefITrInsideScan  src/cm_dpsearch_trunc.c:741-1338
 * with the matrices of cm_tr_scan_mx_Create (src/cm_mx.c:6391-6640) and the
 * initialisation of cm_tr_scan_mx_InitializeFloats/Integers (:6853-6980).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Behaviours of the reference kept on purpose:
 *  - the root cells X[cur][0][d] are never reset per j, so a truncated begin
 *    at j is only taken if it beats the best value from j-2 (same parity);
 *  - RefITrInsideScan keeps J/L/R/T sums in float temporaries, so every int
 *    result passes through float (x86 conversion: out-of-range -> INT_MIN);
 *  - its d = 0 J cells of S/D/BEGL states start from the float endsc
 *    converted to int, not from iendsc;
 *  - its vsc[v] is the value at the last j, not a maximum over j;
 *  - d = 1 MATP cells compute J terms that are thrown away (the reference
 *    reads index d-2 = -1 for them; this restatement skips the read).
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define ESL_MAX(a,b) (((a)>(b))?(a):(b))
#define ESL_MIN(a,b) (((a)<(b))?(a):(b))
#define IADD(a,b)   ((int) ((unsigned) (a) + (unsigned) (b)))

/* helpers shared with oracle.c */
extern const int *oracle_ilogsum_table(void);
int  oracle_trunc_hits_begin(void **acc, int L, int64_t i0, int64_t j0, int W, int greedy);
int  oracle_trunc_hits_update(void *acc, const CM_t *cm, int pass_idx, int j, int dmin, int dmax, const float *bestsc,
                              const int *bestr, const char *bestmode, const double *act, float cutoff, int do_null3);
int  oracle_trunc_hits_finish(void *acc, ORACLE_RESULT *res, char *errbuf);
void oracle_act_update(double *act, int W, int jp, ESL_DSQ x);

/* x86 cvttss2si semantics for float -> int (out of range -> INT_MIN) */
static inline int f2i(float f) { return (f >= -2147483648.0f && f < 2147483648.0f) ? (int) f : INT_MIN; }

static inline int oracle_ILogsum_t(int s1, int s2)
{
  const int *tbl = oracle_ilogsum_table();
  const int max = ESL_MAX(-INFTY, ESL_MAX(s1, s2));
  const int min = ESL_MIN(s1, s2);
  return (min <= -INFTY || (max - min) >= LOGSUM_TBL) ? max : IADD(max, tbl[max - min]);
}
static inline float Scorify_t(int sc) { return (sc == -INFTY) ? IMPOSSIBLE : ((float) sc / INTSCALE); }

static int sdelta(int t) { return t == MP_st ? 2 : (t == ML_st || t == MR_st || t == IL_st || t == IR_st) ? 1 : 0; }
static int rdelta(int t) { return (t == MP_st || t == MR_st || t == IR_st) ? 1 : 0; }
static int emitmode(int t)
{
  switch (t) { case IL_st: case ML_st: return EMITLEFT; case IR_st: case MR_st: return EMITRIGHT; case MP_st: return EMITPAIR; }
  return EMITNONE;
}
static int detached(const CM_t *cm, int v) { return (v + 1 < cm->M && cm->stid[v + 1] == END_E); }

static int fill_from_pass(int pass_idx, int *L, int *R, int *T)
{
  *L = *R = *T = 0;
  switch (pass_idx) {
  case PLI_PASS_5P_AND_3P_FORCE: case PLI_PASS_5P_AND_3P_ANY: *L = *R = *T = 1; return eslOK;
  case PLI_PASS_5P_ONLY_FORCE: *R = 1; return eslOK;
  case PLI_PASS_3P_ONLY_FORCE: *L = 1; return eslOK;
  }
  return eslEINVAL;
}
static int pty_from_pass(int pass_idx)
{
  switch (pass_idx) {
  case PLI_PASS_5P_ONLY_FORCE: return TRPENALTY_5P_ONLY;
  case PLI_PASS_3P_ONLY_FORCE: return TRPENALTY_3P_ONLY;
  case PLI_PASS_5P_AND_3P_FORCE: case PLI_PASS_5P_AND_3P_ANY: return TRPENALTY_5P_AND_3P;
  }
  return -1;
}

/* ======================================================================== */
/* RefTrCYKScan                                                              */

int oracle_trcyk(const CM_t *cm, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0, float cutoff,
                 int do_null3, int greedy, float env_cutoff, int want_rows, ORACLE_RESULT *res, char *errbuf)
{
  int fill_L, fill_R, fill_T, pty_idx;
  if (fill_from_pass(pass_idx, &fill_L, &fill_R, &fill_T) != eslOK) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "RefTrCYKScan(), unexpected pass idx: %d", pass_idx); return eslEINVAL; }
  if ((pty_idx = pty_from_pass(pass_idx)) == -1) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "RefTrCYKScan(), unexpected pass idx: %d", pass_idx); return eslEINCOMPAT; }
  if (cm->trp == NULL) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "no truncation penalties"); return eslEINCOMPAT; }
  const int M = cm->M, Wt = cm->W;
  const int L = (int) (j0 - i0 + 1);
  const int W = (Wt > L) ? L : Wt;
  const int *dmax = (qdbidx == SMX_QDB1_TIGHT) ? cm->qdbinfo->dmax1 : (qdbidx == SMX_QDB2_LOOSE) ? cm->qdbinfo->dmax2 : NULL;
  const int do_banded = (dmax != NULL);
  const size_t W1 = (size_t) Wt + 1;
  int v, d, k, j, y, w, yoffset, i;

  /* bands (src/cm_mx.c:6428-6441): dn = 1, dx = min(j, dmax, W) */
  int *dxAA = (int *) malloc(sizeof(int) * (size_t) (Wt + 1) * M);
  for (int jr = 1; jr <= Wt; jr++)
    for (v = 0; v < M; v++) dxAA[(size_t) jr * M + v] = do_banded ? ESL_MIN(jr, ESL_MIN(dmax[v], Wt)) : ESL_MIN(jr, Wt);

  int nbegl = 0;
  int *bi = (int *) malloc(sizeof(int) * M);
  for (v = 0; v < M; v++) bi[v] = (cm->stid[v] == BEGL_S) ? nbegl++ : -1;
  float *J = (float *) malloc(sizeof(float) * 2 * M * W1), *Lm = (float *) malloc(sizeof(float) * 2 * M * W1);
  float *R = (float *) malloc(sizeof(float) * 2 * M * W1), *T = (float *) malloc(sizeof(float) * 2 * M * W1);
  size_t nb = (size_t) (nbegl ? nbegl : 1) * W1 * W1;
  float *JB = (float *) malloc(sizeof(float) * nb), *LB = (float *) malloc(sizeof(float) * nb), *RB = (float *) malloc(sizeof(float) * nb);
#define X(A, jj, vv, dd)  A[((size_t) (jj) * M + (vv)) * W1 + (dd)]
#define XB(A, jj, vv, dd) A[((size_t) (jj) * nbegl + bi[vv]) * W1 + (dd)]
  for (size_t x = 0; x < 2 * (size_t) M * W1; x++) J[x] = Lm[x] = R[x] = T[x] = IMPOSSIBLE;
  for (size_t x = 0; x < nb; x++) JB[x] = LB[x] = RB[x] = IMPOSSIBLE;
  /* cm_tr_scan_mx_InitializeFloats */
  for (v = M - 1; v >= 0; v--) {
    if (cm->stid[v] != BEGL_S) {
      if (cm->sttype[v] == E_st) {
        X(J, 0, v, 0) = X(J, 1, v, 0) = 0.;
        X(Lm, 0, v, 0) = X(Lm, 1, v, 0) = 0.;
        X(R, 0, v, 0) = X(R, 1, v, 0) = 0.;
      } else if (cm->sttype[v] == S_st || cm->sttype[v] == D_st) {
        y = cm->cfirst[v];
        X(J, 0, v, 0) = cm->endsc[v];
        for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) X(J, 0, v, 0) = ESL_MAX(X(J, 0, v, 0), (X(J, 0, y + yoffset, 0) + cm->tsc[v][yoffset]));
        X(J, 0, v, 0) = ESL_MAX(X(J, 0, v, 0), IMPOSSIBLE);
      } else if (cm->sttype[v] == B_st) {
        w = cm->cfirst[v]; y = cm->cnum[v];
        X(J, 0, v, 0) = XB(JB, 0, w, 0) + X(J, 0, y, 0);
      }
      X(J, 1, v, 0) = X(J, 0, v, 0);
    } else {
      y = cm->cfirst[v];
      XB(JB, 0, v, 0) = cm->endsc[v];
      for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) XB(JB, 0, v, 0) = ESL_MAX(XB(JB, 0, v, 0), (X(J, 0, y + yoffset, 0) + cm->tsc[v][yoffset]));
      XB(JB, 0, v, 0) = ESL_MAX(XB(JB, 0, v, 0), IMPOSSIBLE);
      for (j = 1; j <= Wt; j++) XB(JB, j, v, 0) = XB(JB, 0, v, 0);
    }
  }

  /* FCalcInitDPScores */
  float *el = (float *) malloc(sizeof(float) * W1);
  for (d = 0; d <= Wt; d++) el[d] = cm->el_selfsc * d;
  float *init = (float *) malloc(sizeof(float) * M * W1);
  for (v = 0; v < M; v++)
    for (d = 0; d <= Wt; d++) init[(size_t) v * W1 + d] = NOT_IMPOSSIBLE(cm->endsc[v]) ? el[d] + cm->endsc[v] : IMPOSSIBLE;
#define INIT(vv, dd) init[(size_t) (vv) * W1 + (dd)]

  int *jp_wA = (int *) malloc(sizeof(int) * W1);
  float *bestsc = (float *) malloc(sizeof(float) * W1);
  int *bestr = (int *) malloc(sizeof(int) * W1);
  char *bestmode = (char *) malloc(W1);
  float *vsc = (float *) malloc(sizeof(float) * M);
  for (v = 0; v < M; v++) vsc[v] = IMPOSSIBLE;
  float vsc_root = IMPOSSIBLE;
  int vmode_root = TRMODE_UNKNOWN;
  int64_t envi = j0 + 1, envj = i0 - 1;
  double *act = do_null3 ? (double *) calloc((size_t) (W + 1) * INFX_K, sizeof(double)) : NULL;
  void *hacc = NULL;
  oracle_trunc_hits_begin(&hacc, L, i0, j0, W, greedy);
  if (want_rows) {
    res->W = W;
    res->rows_sc = (float *) malloc(sizeof(float) * (size_t) L * (W + 1));
    res->rows_r = (int *) malloc(sizeof(int) * (size_t) L * (W + 1));
    res->rows_mode = (char *) malloc((size_t) L * (W + 1));
    res->rows_msc = (float *) malloc(sizeof(float) * (size_t) L * (W + 1) * 4);
    for (size_t x = 0; x < (size_t) L * (W + 1) * 4; x++) res->rows_msc[x] = IMPOSSIBLE;
  }
  int status = eslOK;

  for (j = (int) i0; j <= (int) j0; j++) {
    float Jsc, Lsc, Rsc, Tsc;
    int jp_g = j - (int) i0 + 1;
    int cur = j % 2, prv = (j - 1) % 2;
    const int *dxA = (jp_g >= W) ? &dxAA[(size_t) W * M] : &dxAA[(size_t) jp_g * M];
    const int dn_all = 1;
    for (d = 0; d <= W; d++) jp_wA[d] = (j - d) % (W + 1);
    if (act) oracle_act_update(act, W, jp_g, dsq[j]);

    for (v = M - 1; v > 0; v--) {
      int st = cm->sttype[v];
      if (st == E_st) continue;
      const float *esc_v = cm->oesc[v], *tsc_v = cm->tsc[v], *lmesc_v = cm->lmesc[v], *rmesc_v = cm->rmesc[v];
      int em = emitmode(st);
      int jp_v = (cm->stid[v] == BEGL_S) ? (j % (W + 1)) : cur;
      int jp_y = rdelta(st) ? prv : cur;
      int jq_y = rdelta(st) ? cur : prv;
      int sd = sdelta(st);
      float esc_j = IMPOSSIBLE, rmesc_j = IMPOSSIBLE;
      if (st == IR_st || st == MR_st) { esc_j = esc_v[dsq[j]]; rmesc_j = rmesc_v[dsq[j]]; }
      if (st == MP_st) rmesc_j = rmesc_v[dsq[j]];
      int dn = dn_all, dx = dxA[v];

      if (st == B_st) {
        w = cm->cfirst[v]; y = cm->cnum[v];
        for (d = dn; d <= dx; d++) {
          int kmin, kmax;
          if (do_banded) {
            int dx_y = ESL_MIN(dmax[y], Wt), dx_w = ESL_MIN(dmax[w], Wt);
            kmin = ESL_MAX(0, d - dx_w);
            kmax = ESL_MIN(dx_y, d);
          } else { kmin = 0; kmax = d; }
          Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - sd);
          if (fill_R) Rsc = INIT(v, d - sd);
          if (fill_T) Tsc = INIT(v, d - sd);
          for (k = kmin; k <= kmax; k++) {
            Jsc = ESL_MAX(Jsc, (XB(JB, jp_wA[k], w, d - k) + X(J, jp_y, y, k)));
            if (fill_L) Lsc = ESL_MAX(Lsc, (XB(JB, jp_wA[k], w, d - k) + X(Lm, jp_y, y, k)));
            if (fill_R) Rsc = ESL_MAX(Rsc, (XB(RB, jp_wA[k], w, d - k) + X(J, jp_y, y, k)));
          }
          if (fill_T) {
            int kn = ESL_MAX(1, kmin), kx = ESL_MIN(d - 1, kmax);
            for (k = kn; k <= kx; k++) Tsc = ESL_MAX(Tsc, (XB(RB, jp_wA[k], w, d - k) + X(Lm, jp_y, y, k)));
          }
          X(J, jp_v, v, d) = Jsc;
          if (fill_T) X(T, jp_v, v, d) = Tsc;
          if (fill_L) X(Lm, jp_v, v, d) = (kmin == 0) ? ESL_MAX(Lsc, ESL_MAX(XB(JB, jp_wA[0], w, d), XB(LB, jp_wA[0], w, d))) : Lsc;
          if (fill_R) X(R, jp_v, v, d) = (kmax == d) ? ESL_MAX(Rsc, ESL_MAX(X(J, jp_y, y, d), X(R, jp_y, y, d))) : Rsc;
        }
      } else if (cm->stid[v] == BEGL_S) {
        y = cm->cfirst[v];
        for (d = dn; d <= dx; d++) {
          Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - sd);
          if (fill_R) Rsc = INIT(v, d - sd);
          for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
            Jsc = ESL_MAX(Jsc, X(J, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
            if (fill_L) Lsc = ESL_MAX(Lsc, X(Lm, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
            if (fill_R) Rsc = ESL_MAX(Rsc, X(R, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
          }
          XB(JB, jp_v, v, d) = Jsc;
          if (fill_L) XB(LB, jp_v, v, d) = Lsc;
          if (fill_R) XB(RB, jp_v, v, d) = Rsc;
        }
      } else if (em == EMITLEFT) {
        if (!detached(cm, v)) {
          y = cm->cfirst[v];
          i = j - dn + 1;
          int Ry0 = (st == IL_st) ? 1 : 0;
          for (d = dn; d <= dx; d++) {
            Jsc = INIT(v, d - sd);
            if (fill_L) Lsc = INIT(v, d - sd);
            if (fill_R) { Rsc = INIT(v, d); X(R, jp_v, v, d) = Rsc; }
            for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
              Jsc = ESL_MAX(Jsc, X(J, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
              if (fill_L) Lsc = ESL_MAX(Lsc, X(Lm, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
            }
            X(J, jp_v, v, d) = Jsc + esc_v[dsq[i]];
            if (fill_L) X(Lm, jp_v, v, d) = (d >= 2) ? Lsc + esc_v[dsq[i]] : esc_v[dsq[i]];
            if (fill_R) {
              for (yoffset = Ry0; yoffset < cm->cnum[v]; yoffset++)
                Rsc = ESL_MAX(Rsc, ESL_MAX(X(J, jp_y, y + yoffset, d) + tsc_v[yoffset], X(R, jp_y, y + yoffset, d) + tsc_v[yoffset]));
              X(R, jp_v, v, d) = Rsc;
            }
            i--;
          }
        }
      } else if (em == EMITRIGHT) {
        if (!detached(cm, v)) {
          y = cm->cfirst[v];
          int Ly0 = (st == IR_st) ? 1 : 0;
          for (d = dn; d <= dx; d++) {
            Jsc = INIT(v, d - sd);
            if (fill_R) Rsc = INIT(v, d - sd);
            if (fill_L) { Lsc = INIT(v, d); X(Lm, jp_v, v, d) = Lsc; }
            for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
              Jsc = ESL_MAX(Jsc, X(J, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
              if (fill_R) Rsc = ESL_MAX(Rsc, X(R, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
            }
            X(J, jp_v, v, d) = Jsc + esc_j;
            if (fill_R) X(R, jp_v, v, d) = (d >= 2) ? Rsc + esc_j : esc_j;
            if (fill_L) {
              for (yoffset = Ly0; yoffset < cm->cnum[v]; yoffset++)
                Lsc = ESL_MAX(Lsc, ESL_MAX(X(J, jq_y, y + yoffset, d) + tsc_v[yoffset], X(Lm, jq_y, y + yoffset, d) + tsc_v[yoffset]));
              X(Lm, jp_v, v, d) = Lsc;
            }
          }
        }
      } else if (em == EMITPAIR) {
        y = cm->cfirst[v];
        i = j - dn + 1;
        for (d = dn; d <= dx; d++) {
          if (d >= 2) Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - 1);
          if (fill_R) Rsc = INIT(v, d - 1);
          for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
            if (d >= 2) Jsc = ESL_MAX(Jsc, X(J, jp_y, y + yoffset, d - 2) + tsc_v[yoffset]);
            if (fill_L) Lsc = ESL_MAX(Lsc, ESL_MAX(X(J, jq_y, y + yoffset, d - 1) + tsc_v[yoffset], X(Lm, jq_y, y + yoffset, d - 1) + tsc_v[yoffset]));
            if (fill_R) Rsc = ESL_MAX(Rsc, ESL_MAX(X(J, jp_y, y + yoffset, d - 1) + tsc_v[yoffset], X(R, jp_y, y + yoffset, d - 1) + tsc_v[yoffset]));
          }
          X(J, jp_v, v, d) = (d >= 2) ? Jsc + esc_v[dsq[i] * INFX_KP + dsq[j]] : IMPOSSIBLE;
          if (fill_L) X(Lm, jp_v, v, d) = (d >= 2) ? Lsc + lmesc_v[dsq[i]] : lmesc_v[dsq[i]];
          if (fill_R) X(R, jp_v, v, d) = (d >= 2) ? Rsc + rmesc_j : rmesc_j;
          i--;
        }
      } else {
        y = cm->cfirst[v];
        for (d = dn; d <= dx; d++) {
          Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - sd);
          if (fill_R) Rsc = INIT(v, d - sd);
          for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
            Jsc = ESL_MAX(Jsc, X(J, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
            if (fill_L) Lsc = ESL_MAX(Lsc, X(Lm, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
            if (fill_R) Rsc = ESL_MAX(Rsc, X(R, jp_y, y + yoffset, d - sd) + tsc_v[yoffset]);
          }
          X(J, jp_v, v, d) = Jsc;
          if (fill_L) X(Lm, jp_v, v, d) = Lsc;
          if (fill_R) X(R, jp_v, v, d) = Rsc;
        }
      }

      if (cm->stid[v] == BEGL_S) {
        for (d = dn; d <= dx; d++) {
          vsc[v] = ESL_MAX(vsc[v], XB(JB, jp_v, v, d));
          if (fill_L) vsc[v] = ESL_MAX(vsc[v], XB(LB, jp_v, v, d));
          if (fill_R) vsc[v] = ESL_MAX(vsc[v], XB(RB, jp_v, v, d));
        }
      } else {
        for (d = dn; d <= dx; d++) {
          vsc[v] = ESL_MAX(vsc[v], X(J, jp_v, v, d));
          if (fill_L) vsc[v] = ESL_MAX(vsc[v], X(Lm, jp_v, v, d));
          if (fill_R) vsc[v] = ESL_MAX(vsc[v], X(R, jp_v, v, d));
          if (cm->stid[v] == BIF_B && fill_T) vsc[v] = ESL_MAX(vsc[v], X(T, jp_v, v, d));
        }
      }
    }

    /* truncated begins from ROOT_S (src/cm_dpsearch_trunc.c:525-605) */
    {
      const int jp_v = cur;   /* left over from v = 1, as in the reference */
      for (d = 0; d <= W; d++) { bestr[d] = 0; bestsc[d] = IMPOSSIBLE; bestmode[d] = TRMODE_UNKNOWN; }
      for (y = 1; y < M; y++) {
        float trpenalty = (cm->flags & CMH_LOCAL_BEGIN) ? cm->trp->l_ptyAA[pty_idx][y] : cm->trp->g_ptyAA[pty_idx][y];
        if (!NOT_IMPOSSIBLE(trpenalty)) continue;
        int dn = ESL_MAX(dn_all, dn_all);
        int dx = ESL_MIN(dxA[0], dxA[y]);
        int jp_y = cur;
        float sc;
        if (want_rows)
          for (d = dn; d <= dx; d++) {
            float *m = &res->rows_msc[((size_t) (jp_g - 1) * (W + 1) + d) * 4];
            m[TRMODE_J] = ESL_MAX(m[TRMODE_J], X(J, jp_y, y, d) + trpenalty);
            if (fill_L) m[TRMODE_L] = ESL_MAX(m[TRMODE_L], X(Lm, jp_y, y, d) + trpenalty);
            if (fill_R) m[TRMODE_R] = ESL_MAX(m[TRMODE_R], X(R, jp_y, y, d) + trpenalty);
            if (fill_T && cm->sttype[y] == B_st) m[TRMODE_T] = ESL_MAX(m[TRMODE_T], X(T, jp_y, y, d) + trpenalty);
          }
        for (d = dn; d <= dx; d++) {
          sc = X(J, jp_y, y, d) + trpenalty;
          if (sc > X(J, jp_v, 0, d)) { X(J, jp_v, 0, d) = sc; if (sc > bestsc[d]) { bestsc[d] = sc; bestmode[d] = TRMODE_J; bestr[d] = y; } }
        }
        if (fill_L) for (d = dn; d <= dx; d++) {
          sc = X(Lm, jp_y, y, d) + trpenalty;
          if (sc > X(Lm, jp_v, 0, d)) { X(Lm, jp_v, 0, d) = sc; if (sc > bestsc[d]) { bestsc[d] = sc; bestmode[d] = TRMODE_L; bestr[d] = y; } }
        }
        if (fill_R) for (d = dn; d <= dx; d++) {
          sc = X(R, jp_y, y, d) + trpenalty;
          if (sc > X(R, jp_v, 0, d)) { X(R, jp_v, 0, d) = sc; if (sc > bestsc[d]) { bestsc[d] = sc; bestmode[d] = TRMODE_R; bestr[d] = y; } }
        }
        if (fill_T && cm->sttype[y] == B_st) for (d = dn; d <= dx; d++) {
          sc = X(T, jp_y, y, d) + trpenalty;
          if (sc > X(T, jp_v, 0, d)) { X(T, jp_v, 0, d) = sc; if (sc > bestsc[d]) { bestsc[d] = sc; bestmode[d] = TRMODE_T; bestr[d] = y; } }
        }
      }
      for (d = dn_all; d <= dxA[0]; d++)
        if (bestsc[d] > vsc_root) { vsc_root = bestsc[d]; vmode_root = bestmode[d]; }
      for (d = dn_all; d <= dxA[0]; d++)
        if (bestsc[d] >= env_cutoff) { envi = ESL_MIN(envi, j - d + 1); envj = ESL_MAX(envj, j); }
      if (want_rows)
        for (d = 0; d <= W; d++) {
          res->rows_sc[(size_t) (jp_g - 1) * (W + 1) + d] = bestsc[d];
          res->rows_r[(size_t) (jp_g - 1) * (W + 1) + d] = bestr[d];
          res->rows_mode[(size_t) (jp_g - 1) * (W + 1) + d] = bestmode[d];
        }
      if ((status = oracle_trunc_hits_update(hacc, cm, pass_idx, j, dn_all, dxA[0], bestsc, bestr, bestmode, act, cutoff, do_null3)) != eslOK) goto DONE;
    }
  }
  vsc[0] = vsc_root;
  res->envi = (envi == j0 + 1) ? -1 : envi;
  res->envj = (envj == i0 - 1) ? -1 : envj;
  res->sc = vsc_root;
  res->mode = vmode_root;
  res->vsc = vsc; vsc = NULL;
  status = oracle_trunc_hits_finish(hacc, res, errbuf);
  hacc = NULL;
DONE:
#undef X
#undef XB
#undef INIT
  if (hacc) oracle_trunc_hits_finish(hacc, NULL, NULL);
  free(dxAA); free(bi); free(J); free(Lm); free(R); free(T); free(JB); free(LB); free(RB);
  free(el); free(init); free(jp_wA); free(bestsc); free(bestr); free(bestmode); free(vsc); free(act);
  return status;
}

/* ======================================================================== */
/* RefITrInsideScan                                                          */

int oracle_triinside(const CM_t *cm, int qdbidx, int pass_idx, const ESL_DSQ *dsq, int64_t i0, int64_t j0, float cutoff,
                     int do_null3, int greedy, float env_cutoff, int want_rows, ORACLE_RESULT *res, char *errbuf)
{
  int fill_L, fill_R, fill_T, pty_idx;
  if (fill_from_pass(pass_idx, &fill_L, &fill_R, &fill_T) != eslOK) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "RefITrInsideScan(), unexpected pass idx: %d", pass_idx); return eslEINVAL; }
  if ((pty_idx = pty_from_pass(pass_idx)) == -1) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "RefITrInsideScan(), unexpected pass idx: %d", pass_idx); return eslEINCOMPAT; }
  if (cm->trp == NULL) { if (errbuf) snprintf(errbuf, eslERRBUFSIZE, "no truncation penalties"); return eslEINCOMPAT; }
  const int M = cm->M, Wt = cm->W;
  const int L = (int) (j0 - i0 + 1);
  const int W = (Wt > L) ? L : Wt;
  const int *dmax = (qdbidx == SMX_QDB1_TIGHT) ? cm->qdbinfo->dmax1 : (qdbidx == SMX_QDB2_LOOSE) ? cm->qdbinfo->dmax2 : NULL;
  const int do_banded = (dmax != NULL);
  const size_t W1 = (size_t) Wt + 1;
  int v, d, k, j, y, w, yoffset, i;
#define LS(a, b) oracle_ILogsum_t((a), (b))

  int *dxAA = (int *) malloc(sizeof(int) * (size_t) (Wt + 1) * M);
  for (int jr = 1; jr <= Wt; jr++)
    for (v = 0; v < M; v++) dxAA[(size_t) jr * M + v] = do_banded ? ESL_MIN(jr, ESL_MIN(dmax[v], Wt)) : ESL_MIN(jr, Wt);
  int nbegl = 0;
  int *bi = (int *) malloc(sizeof(int) * M);
  for (v = 0; v < M; v++) bi[v] = (cm->stid[v] == BEGL_S) ? nbegl++ : -1;
  int *J = (int *) malloc(sizeof(int) * 2 * M * W1), *Lm = (int *) malloc(sizeof(int) * 2 * M * W1);
  int *R = (int *) malloc(sizeof(int) * 2 * M * W1), *T = (int *) malloc(sizeof(int) * 2 * M * W1);
  size_t nb = (size_t) (nbegl ? nbegl : 1) * W1 * W1;
  int *JB = (int *) malloc(sizeof(int) * nb), *LB = (int *) malloc(sizeof(int) * nb), *RB = (int *) malloc(sizeof(int) * nb);
#define X(A, jj, vv, dd)  A[((size_t) (jj) * M + (vv)) * W1 + (dd)]
#define XB(A, jj, vv, dd) A[((size_t) (jj) * nbegl + bi[vv]) * W1 + (dd)]
  for (size_t x = 0; x < 2 * (size_t) M * W1; x++) J[x] = Lm[x] = R[x] = T[x] = -INFTY;
  for (size_t x = 0; x < nb; x++) JB[x] = LB[x] = RB[x] = -INFTY;
  /* cm_tr_scan_mx_InitializeIntegers: S/D/BEGL cells start from the float
   * endsc converted to int (src/cm_mx.c:6941, :6955) */
  for (v = M - 1; v >= 0; v--) {
    if (cm->stid[v] != BEGL_S) {
      if (cm->sttype[v] == E_st) {
        X(J, 0, v, 0) = X(J, 1, v, 0) = 0; X(Lm, 0, v, 0) = X(Lm, 1, v, 0) = 0; X(R, 0, v, 0) = X(R, 1, v, 0) = 0;
      } else if (cm->sttype[v] == S_st || cm->sttype[v] == D_st) {
        y = cm->cfirst[v];
        X(J, 0, v, 0) = f2i(cm->endsc[v]);
        for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) X(J, 0, v, 0) = ESL_MAX(X(J, 0, v, 0), IADD(X(J, 0, y + yoffset, 0), cm->itsc[v][yoffset]));
        X(J, 0, v, 0) = ESL_MAX(X(J, 0, v, 0), -INFTY);
      } else if (cm->sttype[v] == B_st) {
        w = cm->cfirst[v]; y = cm->cnum[v];
        X(J, 0, v, 0) = IADD(XB(JB, 0, w, 0), X(J, 0, y, 0));
      }
      X(J, 1, v, 0) = X(J, 0, v, 0);
    } else {
      y = cm->cfirst[v];
      XB(JB, 0, v, 0) = f2i(cm->endsc[v]);
      for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) XB(JB, 0, v, 0) = ESL_MAX(XB(JB, 0, v, 0), IADD(X(J, 0, y + yoffset, 0), cm->itsc[v][yoffset]));
      XB(JB, 0, v, 0) = ESL_MAX(XB(JB, 0, v, 0), -INFTY);
      for (j = 1; j <= Wt; j++) XB(JB, j, v, 0) = XB(JB, 0, v, 0);
    }
  }
  /* ICalcInitDPScores */
  int *init = (int *) malloc(sizeof(int) * M * W1);
  for (v = 0; v < M; v++)
    for (d = 0; d <= Wt; d++) init[(size_t) v * W1 + d] = (cm->iendsc[v] != -INFTY) ? IADD((int) ((unsigned) cm->iel_selfsc * (unsigned) d), cm->iendsc[v]) : -INFTY;
#define INIT(vv, dd) init[(size_t) (vv) * W1 + (dd)]

  int *jp_wA = (int *) malloc(sizeof(int) * W1);
  float *bestsc = (float *) malloc(sizeof(float) * W1);
  int *bestr = (int *) malloc(sizeof(int) * W1);
  char *bestmode = (char *) malloc(W1);
  float *vsc = (float *) malloc(sizeof(float) * M);
  for (v = 0; v < M; v++) vsc[v] = IMPOSSIBLE;
  float vsc_root = IMPOSSIBLE;
  int vmode_root = TRMODE_UNKNOWN;
  int64_t envi = j0 + 1, envj = i0 - 1;
  double *act = do_null3 ? (double *) calloc((size_t) (W + 1) * INFX_K, sizeof(double)) : NULL;
  void *hacc = NULL;
  oracle_trunc_hits_begin(&hacc, L, i0, j0, W, greedy);
  if (want_rows) {
    res->W = W;
    res->rows_sc = (float *) malloc(sizeof(float) * (size_t) L * (W + 1));
    res->rows_r = (int *) malloc(sizeof(int) * (size_t) L * (W + 1));
    res->rows_mode = (char *) malloc((size_t) L * (W + 1));
    res->rows_msc = (float *) malloc(sizeof(float) * (size_t) L * (W + 1) * 4);
    for (size_t x = 0; x < (size_t) L * (W + 1) * 4; x++) res->rows_msc[x] = IMPOSSIBLE;
  }
  int status = eslOK;

  for (j = (int) i0; j <= (int) j0; j++) {
    float Jsc, Lsc, Rsc, Tsc;   /* float temporaries, as the reference */
    int jp_g = j - (int) i0 + 1;
    int cur = j % 2, prv = (j - 1) % 2;
    const int *dxA = (jp_g >= W) ? &dxAA[(size_t) W * M] : &dxAA[(size_t) jp_g * M];
    const int dn_all = 1;
    for (d = 0; d <= W; d++) jp_wA[d] = (j - d) % (W + 1);
    if (act) oracle_act_update(act, W, jp_g, dsq[j]);

    for (v = M - 1; v > 0; v--) {
      int st = cm->sttype[v];
      if (st == E_st) continue;
      const int *esc_v = cm->ioesc[v], *tsc_v = cm->itsc[v], *lmesc_v = cm->ilmesc[v], *rmesc_v = cm->irmesc[v];
      int em = emitmode(st);
      int jp_v = (cm->stid[v] == BEGL_S) ? (j % (W + 1)) : cur;
      int jp_y = rdelta(st) ? prv : cur;
      int jq_y = rdelta(st) ? cur : prv;
      int sd = sdelta(st);
      int esc_j = -INFTY, rmesc_j = -INFTY;
      if (st == IR_st || st == MR_st) { esc_j = esc_v[dsq[j]]; rmesc_j = rmesc_v[dsq[j]]; }
      if (st == MP_st) rmesc_j = rmesc_v[dsq[j]];
      int dn = dn_all, dx = dxA[v];

      if (st == B_st) {
        w = cm->cfirst[v]; y = cm->cnum[v];
        for (d = dn; d <= dx; d++) {
          int kmin, kmax;
          if (do_banded) {
            int dx_y = ESL_MIN(dmax[y], Wt), dx_w = ESL_MIN(dmax[w], Wt);
            kmin = ESL_MAX(0, d - dx_w);
            kmax = ESL_MIN(dx_y, d);
          } else { kmin = 0; kmax = d; }
          Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - sd);
          if (fill_R) Rsc = INIT(v, d - sd);
          if (fill_T) Tsc = INIT(v, d - sd);
          for (k = kmin; k <= kmax; k++) {
            Jsc = LS(f2i(Jsc), IADD(XB(JB, jp_wA[k], w, d - k), X(J, jp_y, y, k)));
            if (fill_L) Lsc = LS(f2i(Lsc), IADD(XB(JB, jp_wA[k], w, d - k), X(Lm, jp_y, y, k)));
            if (fill_R) Rsc = LS(f2i(Rsc), IADD(XB(RB, jp_wA[k], w, d - k), X(J, jp_y, y, k)));
          }
          if (fill_T) {
            int kn = ESL_MAX(1, kmin), kx = ESL_MIN(d - 1, kmax);
            for (k = kn; k <= kx; k++) Tsc = LS(f2i(Tsc), IADD(XB(RB, jp_wA[k], w, d - k), X(Lm, jp_y, y, k)));
          }
          X(J, jp_v, v, d) = f2i(Jsc);
          if (fill_T) X(T, jp_v, v, d) = f2i(Tsc);
          if (fill_L) X(Lm, jp_v, v, d) = (kmin == 0) ? LS(f2i(Lsc), ESL_MAX(XB(JB, jp_wA[0], w, d), XB(LB, jp_wA[0], w, d))) : f2i(Lsc);
          if (fill_R) X(R, jp_v, v, d) = (kmax == d) ? LS(f2i(Rsc), ESL_MAX(X(J, jp_y, y, d), X(R, jp_y, y, d))) : f2i(Rsc);
        }
      } else if (cm->stid[v] == BEGL_S) {
        y = cm->cfirst[v];
        for (d = dn; d <= dx; d++) {
          Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - sd);
          if (fill_R) Rsc = INIT(v, d - sd);
          for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
            Jsc = LS(f2i(Jsc), IADD(X(J, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
            if (fill_L) Lsc = LS(f2i(Lsc), IADD(X(Lm, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
            if (fill_R) Rsc = LS(f2i(Rsc), IADD(X(R, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
          }
          XB(JB, jp_v, v, d) = f2i(Jsc);
          if (fill_L) XB(LB, jp_v, v, d) = f2i(Lsc);
          if (fill_R) XB(RB, jp_v, v, d) = f2i(Rsc);
        }
      } else if (em == EMITLEFT) {
        if (!detached(cm, v)) {
          y = cm->cfirst[v];
          i = j - dn + 1;
          int Ry0 = (st == IL_st) ? 1 : 0;
          for (d = dn; d <= dx; d++) {
            Jsc = INIT(v, d - sd);
            if (fill_L) Lsc = INIT(v, d - sd);
            if (fill_R) { Rsc = INIT(v, d); X(R, jp_v, v, d) = f2i(Rsc); }
            for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
              Jsc = LS(f2i(Jsc), IADD(X(J, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
              if (fill_L) Lsc = LS(f2i(Lsc), IADD(X(Lm, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
            }
            X(J, jp_v, v, d) = f2i(Jsc + (float) esc_v[dsq[i]]);
            if (fill_L) X(Lm, jp_v, v, d) = f2i((d >= 2) ? Lsc + (float) esc_v[dsq[i]] : (float) esc_v[dsq[i]]);
            if (fill_R) {
              for (yoffset = Ry0; yoffset < cm->cnum[v]; yoffset++)
                Rsc = LS(f2i(Rsc), LS(IADD(X(J, jp_y, y + yoffset, d), tsc_v[yoffset]), IADD(X(R, jp_y, y + yoffset, d), tsc_v[yoffset])));
              X(R, jp_v, v, d) = f2i(Rsc);
            }
            i--;
          }
        }
      } else if (em == EMITRIGHT) {
        if (!detached(cm, v)) {
          y = cm->cfirst[v];
          int Ly0 = (st == IR_st) ? 1 : 0;
          for (d = dn; d <= dx; d++) {
            Jsc = INIT(v, d - sd);
            if (fill_R) Rsc = INIT(v, d - sd);
            if (fill_L) { Lsc = INIT(v, d); X(Lm, jp_v, v, d) = f2i(Lsc); }
            for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
              Jsc = LS(f2i(Jsc), IADD(X(J, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
              if (fill_R) Rsc = LS(f2i(Rsc), IADD(X(R, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
            }
            X(J, jp_v, v, d) = f2i(Jsc + (float) esc_j);
            if (fill_R) X(R, jp_v, v, d) = f2i((d >= 2) ? Rsc + (float) esc_j : (float) esc_j);
            if (fill_L) {
              for (yoffset = Ly0; yoffset < cm->cnum[v]; yoffset++)
                Lsc = LS(f2i(Lsc), LS(IADD(X(J, jq_y, y + yoffset, d), tsc_v[yoffset]), IADD(X(Lm, jq_y, y + yoffset, d), tsc_v[yoffset])));
              X(Lm, jp_v, v, d) = f2i(Lsc);
            }
          }
        }
      } else if (em == EMITPAIR) {
        y = cm->cfirst[v];
        i = j - dn + 1;
        for (d = dn; d <= dx; d++) {
          if (d >= 2) Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - 1);
          if (fill_R) Rsc = INIT(v, d - 1);
          for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
            if (d >= 2) Jsc = LS(f2i(Jsc), IADD(X(J, jp_y, y + yoffset, d - 2), tsc_v[yoffset]));
            if (fill_L) Lsc = LS(f2i(Lsc), ESL_MAX(IADD(X(J, jq_y, y + yoffset, d - 1), tsc_v[yoffset]), IADD(X(Lm, jq_y, y + yoffset, d - 1), tsc_v[yoffset])));
            if (fill_R) Rsc = LS(f2i(Rsc), ESL_MAX(IADD(X(J, jp_y, y + yoffset, d - 1), tsc_v[yoffset]), IADD(X(R, jp_y, y + yoffset, d - 1), tsc_v[yoffset])));
          }
          X(J, jp_v, v, d) = f2i((d >= 2) ? Jsc + (float) esc_v[dsq[i] * INFX_KP + dsq[j]] : (float) -INFTY);
          if (fill_L) X(Lm, jp_v, v, d) = f2i((d >= 2) ? Lsc + (float) lmesc_v[dsq[i]] : (float) lmesc_v[dsq[i]]);
          if (fill_R) X(R, jp_v, v, d) = f2i((d >= 2) ? Rsc + (float) rmesc_j : (float) rmesc_j);
          i--;
        }
      } else {
        y = cm->cfirst[v];
        for (d = dn; d <= dx; d++) {
          Jsc = INIT(v, d - sd);
          if (fill_L) Lsc = INIT(v, d - sd);
          if (fill_R) Rsc = INIT(v, d - sd);
          for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) {
            Jsc = LS(f2i(Jsc), IADD(X(J, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
            if (fill_L) Lsc = LS(f2i(Lsc), IADD(X(Lm, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
            if (fill_R) Rsc = LS(f2i(Rsc), IADD(X(R, jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
          }
          X(J, jp_v, v, d) = f2i(Jsc);
          if (fill_L) X(Lm, jp_v, v, d) = f2i(Lsc);
          if (fill_R) X(R, jp_v, v, d) = f2i(Rsc);
        }
      }

      {
        int ivsc = -INFTY;
        if (cm->stid[v] == BEGL_S) {
          for (d = dn; d <= dx; d++) {
            ivsc = ESL_MAX(ivsc, XB(JB, jp_v, v, d));
            if (fill_L) ivsc = ESL_MAX(ivsc, XB(LB, jp_v, v, d));
            if (fill_R) ivsc = ESL_MAX(ivsc, XB(RB, jp_v, v, d));
          }
        } else {
          for (d = dn; d <= dx; d++) {
            ivsc = ESL_MAX(ivsc, X(J, jp_v, v, d));
            if (fill_L) ivsc = ESL_MAX(ivsc, X(Lm, jp_v, v, d));
            if (fill_R) ivsc = ESL_MAX(ivsc, X(R, jp_v, v, d));
            if (cm->stid[v] == BIF_B && fill_T) ivsc = ESL_MAX(ivsc, X(T, jp_v, v, d));
          }
        }
        vsc[v] = Scorify_t(ivsc);
      }
    }

    {
      const int jp_v = cur;
      for (d = 0; d <= W; d++) { bestr[d] = 0; bestsc[d] = IMPOSSIBLE; bestmode[d] = TRMODE_UNKNOWN; }
      for (y = 1; y < M; y++) {
        int itr = (cm->flags & CMH_LOCAL_BEGIN) ? cm->trp->il_ptyAA[pty_idx][y] : cm->trp->ig_ptyAA[pty_idx][y];
        if (itr == -INFTY) continue;
        int dn = dn_all;
        int dx = ESL_MIN(dxA[0], dxA[y]);
        int jp_y = cur, sc;
        float fsc;
        if (want_rows)
          for (d = dn; d <= dx; d++) {
            float *m = &res->rows_msc[((size_t) (jp_g - 1) * (W + 1) + d) * 4];
            m[TRMODE_J] = ESL_MAX(m[TRMODE_J], Scorify_t(IADD(X(J, jp_y, y, d), itr)));
            if (fill_L) m[TRMODE_L] = ESL_MAX(m[TRMODE_L], Scorify_t(IADD(X(Lm, jp_y, y, d), itr)));
            if (fill_R) m[TRMODE_R] = ESL_MAX(m[TRMODE_R], Scorify_t(IADD(X(R, jp_y, y, d), itr)));
            if (fill_T && cm->sttype[y] == B_st) m[TRMODE_T] = ESL_MAX(m[TRMODE_T], Scorify_t(IADD(X(T, jp_y, y, d), itr)));
          }
        for (d = dn; d <= dx; d++) {
          sc = IADD(X(J, jp_y, y, d), itr);
          if (sc > X(J, jp_v, 0, d)) { X(J, jp_v, 0, d) = sc; fsc = Scorify_t(sc); if (fsc > bestsc[d]) { bestsc[d] = fsc; bestmode[d] = TRMODE_J; bestr[d] = y; } }
        }
        if (fill_L) for (d = dn; d <= dx; d++) {
          sc = IADD(X(Lm, jp_y, y, d), itr);
          if (sc > X(Lm, jp_v, 0, d)) { X(Lm, jp_v, 0, d) = sc; fsc = Scorify_t(sc); if (fsc > bestsc[d]) { bestsc[d] = fsc; bestmode[d] = TRMODE_L; bestr[d] = y; } }
        }
        if (fill_R) for (d = dn; d <= dx; d++) {
          sc = IADD(X(R, jp_y, y, d), itr);
          if (sc > X(R, jp_v, 0, d)) { X(R, jp_v, 0, d) = sc; fsc = Scorify_t(sc); if (fsc > bestsc[d]) { bestsc[d] = fsc; bestmode[d] = TRMODE_R; bestr[d] = y; } }
        }
        if (fill_T && cm->sttype[y] == B_st) for (d = dn; d <= dx; d++) {
          sc = IADD(X(T, jp_y, y, d), itr);
          if (sc > X(T, jp_v, 0, d)) { X(T, jp_v, 0, d) = sc; fsc = Scorify_t(sc); if (fsc > bestsc[d]) { bestsc[d] = fsc; bestmode[d] = TRMODE_T; bestr[d] = y; } }
        }
      }
      for (d = dn_all; d <= dxA[0]; d++)
        if (bestsc[d] > vsc_root) { vsc_root = bestsc[d]; vmode_root = bestmode[d]; }
      for (d = dn_all; d <= dxA[0]; d++)
        if (bestsc[d] >= env_cutoff) { envi = ESL_MIN(envi, j - d + 1); envj = ESL_MAX(envj, j); }
      if (want_rows)
        for (d = 0; d <= W; d++) {
          res->rows_sc[(size_t) (jp_g - 1) * (W + 1) + d] = bestsc[d];
          res->rows_r[(size_t) (jp_g - 1) * (W + 1) + d] = bestr[d];
          res->rows_mode[(size_t) (jp_g - 1) * (W + 1) + d] = bestmode[d];
        }
      if ((status = oracle_trunc_hits_update(hacc, cm, pass_idx, j, dn_all, dxA[0], bestsc, bestr, bestmode, act, cutoff, do_null3)) != eslOK) goto DONE;
    }
  }
  vsc[0] = vsc_root;
  res->envi = (envi == j0 + 1) ? -1 : envi;
  res->envj = (envj == i0 - 1) ? -1 : envj;
  res->sc = vsc_root;
  res->mode = vmode_root;
  res->vsc = vsc; vsc = NULL;
  status = oracle_trunc_hits_finish(hacc, res, errbuf);
  hacc = NULL;
DONE:
#undef X
#undef XB
#undef INIT
#undef LS
  if (hacc) oracle_trunc_hits_finish(hacc, NULL, NULL);
  free(dxAA); free(bi); free(J); free(Lm); free(R); free(T); free(JB); free(LB); free(RB);
  free(init); free(jp_wA); free(bestsc); free(bestr); free(bestmode); free(vsc); free(act);
  return status;
}
