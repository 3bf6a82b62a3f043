/* cm_qdband.cpp -- query-dependent bands (QDBs) and the window length W.
 *
 *   CalculateQueryDependentBands  src/cm_qdband.c:45-63
 *   BandCalculationEngine         src/cm_qdband.c:163-430
 *   BandTruncationNegligible      src/cm_qdband.c:503-534
 *   cm_CalculateLocalBeginProbs   src/cm_modelconfig.c:397-426
 *
 * gamma_v(n) is the probability that the parse subtree rooted at state v
 * emits n residues, computed bottom-up (v = M-1..0) in double precision from
 * the model's transition probabilities, with local ends off and standard
 * local begins on (ROOT_S's density is the begin-weighted mixture of the
 * states it can begin in).  The bands keep each tail's mass <= beta:
 * dmin is the first n whose cumulative mass exceeds beta (left to right),
 * dmax the first n whose cumulative mass exceeds beta right to left; W is
 * dmax of ROOT_S at beta_W.  Z (the largest length tracked) starts at
 * 4 * clen and doubles until the truncated tail is negligible.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "cm_internal.h"

/* Easel's esl_vec_DSum (Kahan-compensated) and esl_vec_DNorm */
static double dsum(const double *v, int n)
{
  double sum = 0., c = 0.;
  for (int x = 0; x < n; x++) {
    const double y = v[x] - c;
    const double t = sum + y;
    c = (t - sum) - y;
    sum = t;
  }
  return sum;
}
static void dnorm(double *v, int n)
{
  const double sum = dsum(v, n);
  if (sum != 0.0) for (int x = 0; x < n; x++) v[x] /= sum;
  else            for (int x = 0; x < n; x++) v[x] = 1. / (double) n;
}

/* src/cm_qdband.c:503 */
static int truncation_negligible(const double *density, int b, int Z)
{
  double C = 0.;
  for (int i = b + 1; i <= Z; i++) C += density[i];
  const double logbeta = (log(density[Z]) - log(density[b + 1])) / (Z - b - 1);
  const double beta = exp(logbeta);
  const double D = (beta / (1. - beta)) * density[Z];
  return (D < C * DBL_EPSILON) ? 1 : 0;
}

/* src/cm_modelconfig.c:397: begin[node 1] = 1 - p, p shared by MATP/MATL/MATR/BIF nodes 2.. */
static void local_begin_probs(const CM_t *cm, float p_internal_start, std::vector<float> &begin)
{
  int nstarts = 0;
  for (int nd = 2; nd < cm->nodes; nd++)
    if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd || cm->ndtype[nd] == BIF_nd) nstarts++;
  begin.assign(cm->M, 0.f);
  begin[cm->nodemap[1]] = (float) (1. - (double) p_internal_start);   /* double, then rounded: as the reference */
  const float p = p_internal_start / (float) nstarts;
  for (int nd = 2; nd < cm->nodes; nd++)
    if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd || cm->ndtype[nd] == BIF_nd)
      begin[cm->nodemap[nd]] = p;
}

/* src/cm_qdband.c:163.  Returns eslOK, eslERANGE (Z too small), eslEINVAL.
 * ret_gamma (optional, caller-allocated [M][Z+1]): each state's density as its
 * parents used it (after the renormalisation); E states hold the end density. */
extern "C" int infx_BandCalculationEngine(CM_t *cm, int Z, CM_QDBINFO *qdbinfo, double beta_W, int *ret_W, double *ret_gamma)
{
  if (qdbinfo != NULL && ((qdbinfo->beta2 - qdbinfo->beta1) > 1E-20)) return eslEINVAL;
  const int M = cm->M;
  double max_beta = beta_W;
  if (qdbinfo != NULL) max_beta = ESL_MAX(max_beta, ESL_MAX(qdbinfo->beta1, qdbinfo->beta2));

  /* transition copies with local ends off; standard local begins on */
  std::vector<float> t((size_t) M * MAXCONNECT);
  for (int v = 0; v < M; v++) for (int k = 0; k < MAXCONNECT; k++) t[(size_t) v * MAXCONNECT + k] = cm->t[v][k];
  if (cm->flags & CMH_LOCAL_END)
    for (int v = 0; v < M; v++)
      if (cm->end[v] != 0.f) infx_vec_FNorm(&t[(size_t) v * MAXCONNECT], cm->cnum[v]);
  std::vector<float> begin;
  local_begin_probs(cm, cm->pbegin, begin);

  std::vector<std::vector<double>> gamma(M);
  gamma[0].assign((size_t) Z + 1, 0.);
  std::vector<double> gend((size_t) Z + 1, 0.);
  gend[0] = 1.0;
  int W = 0;

  for (int v = M - 1; v >= 0; v--) {
    const int st = cm->sttype[v];
    if (st == E_st) {
      if (qdbinfo) { qdbinfo->dmin1[v] = qdbinfo->dmax1[v] = 0; qdbinfo->dmin2[v] = qdbinfo->dmax2[v] = 0; }
      if (ret_gamma) memcpy(ret_gamma + (size_t) v * (Z + 1), gend.data(), sizeof(double) * (Z + 1));
      continue;
    }
    std::vector<double> &g = gamma[v];
    if (v != 0) g.assign((size_t) Z + 1, 0.);
    auto child = [&](int y) -> const std::vector<double> & { return (cm->sttype[y] == E_st) ? gend : gamma[y]; };
    double pdf = 0.;
    if (st == B_st) {
      const std::vector<double> &gl = child(cm->cfirst[v]), &gr = child(cm->cnum[v]);
      for (int n = 0; n <= Z; n++) {
        for (int leftn = 0; leftn <= n; leftn++) g[n] += gl[leftn] * gr[n - leftn];
        pdf += g[n];
      }
    } else if (v != 0) {
      const int dv = StateDelta(st);
      for (int n = dv; n <= Z; n++) {
        for (int yo = 0; yo < cm->cnum[v]; yo++) {
          const int y = cm->cfirst[v] + yo;
          g[n] += t[(size_t) v * MAXCONNECT + yo] * child(y)[n - dv];
        }
        pdf += g[n];
      }
    }
    if (begin[v] > 0.)
      for (int n = 0; n <= Z; n++) gamma[0][n] += begin[v] * g[n];

    if (v == 0) pdf = dsum(gamma[0].data(), Z + 1);
    if (pdf <= 0.999 || g[Z] > max_beta * DBL_EPSILON) return eslERANGE;
    if (pdf > 1.0) dnorm(g.data(), Z + 1);
    if (ret_gamma) memcpy(ret_gamma + (size_t) v * (Z + 1), g.data(), sizeof(double) * (Z + 1));

    if (qdbinfo != NULL) {
      pdf = 0.;
      int set2 = 0;
      for (int n = 0; n <= Z; n++) {
        pdf += g[n];
        if (!set2 && pdf > qdbinfo->beta2) { qdbinfo->dmin2[v] = n; set2 = 1; }
        if (set2 && pdf > qdbinfo->beta1) { qdbinfo->dmin1[v] = n; break; }
      }
      pdf = 0.;
      set2 = 0;
      for (int n = Z; n >= 0; n--) {
        pdf += g[n];
        if (!set2 && pdf > qdbinfo->beta2) { qdbinfo->dmax2[v] = n; set2 = 1; }
        if (set2 && pdf > qdbinfo->beta1) { qdbinfo->dmax1[v] = n; break; }
      }
    }
    if (v == 0) {
      pdf = 0.;
      for (int n = Z; n >= 0; n--) {
        pdf += gamma[0][n];
        if (pdf > beta_W) { W = n; break; }
      }
    }
    /* children no higher state needs any more can be dropped (memory only) */
    if (st == B_st) { std::vector<double>().swap(gamma[cm->cfirst[v]]); std::vector<double>().swap(gamma[cm->cnum[v]]); }
  }

  if (!truncation_negligible(gamma[0].data(), W, Z) ||
      (qdbinfo != NULL && !truncation_negligible(gamma[0].data(), qdbinfo->dmax1[0], Z)) ||
      (qdbinfo != NULL && !truncation_negligible(gamma[0].data(), qdbinfo->dmax2[0], Z)))
    return eslERANGE;
  if (ret_W) *ret_W = W;
  return eslOK;
}

/* src/cm_qdband.c:45 */
extern "C" int infx_CalculateQueryDependentBands(CM_t *cm, char *errbuf, CM_QDBINFO *qdbinfo, double beta_W, int *ret_W, int *ret_Z)
{
  int status;
  if (qdbinfo != NULL && ((qdbinfo->beta2 - qdbinfo->beta1) > 1E-20))
    return infx_fail(eslEINVAL, errbuf, "Calculating QDBs, qdbinfo->beta1 < qdbinfo->beta2");
  int Z = cm->clen * 4;
  while ((status = infx_BandCalculationEngine(cm, Z, qdbinfo, beta_W, ret_W, NULL)) != eslOK) {
    if (status != eslERANGE) return infx_fail(status, errbuf, "Calculating QDBs, unexpected error");
    Z *= 2;
    if (Z > cm->clen * 1000) return infx_fail(eslEINCONCEIVABLE, errbuf, "Calculating QDBs, Z got insanely large (> 1000*clen)");
  }
  if (ret_Z) *ret_Z = Z;
  return eslOK;
}
