/* cm_trunc.cpp -- truncated-begin penalties (host side).
 *
 * Restates cm_tr_penalties_Create (src/cm_trunc.c:51-450) and what it
 * depends on: cm_CalculateLocalBeginProbs (src/cm_modelconfig.c:404),
 * cm_ExpectedStateOccupancy (src/cm.c:4250), cm_ExpectedPositionOccupancy
 * (src/cm.c:4405) and InsertsGivenNodeIndex (src/cm.c:4897).  Called by
 * cm_Configure before the model is localized, as the reference does
 * (src/cm_modelconfig.c:274 precedes :312).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

#include "cm_internal.h"

static inline double sreLOG2(double x) { return (x > 0) ? log(x) * 1.44269504 : IMPOSSIBLE; }

static int Prob2Score(float p, float null)
{
  if (p == 0.0) return -INFTY;
  return (int) floor(0.5 + INTSCALE * sreLOG2(p / null));
}

/* psi[v]: expected number of times state v is entered in the global model.
 * The transition x -> v is t[x][v - cfirst[x]]: a state's children are the
 * contiguous block cfirst..cfirst+cnum-1 in transition order, which is what
 * the reference's transition map (cm_CreateTransitionMap) encodes. */
static int expected_state_occupancy(const CM_t *cm, std::vector<double> &psi, char *errbuf)
{
  const int M = cm->M;
  psi.assign(M, 0.0);
  std::vector<float> t0(cm->cnum[0] > 0 ? cm->cnum[0] : 1);
  for (int k = 0; k < cm->cnum[0]; k++) t0[k] = cm->root_trans ? cm->root_trans[k] : cm->t[0][k];
  auto tr = [&](int x, int k) -> float { return x == 0 ? t0[k] : cm->t[x][k]; };
  for (int v = 0; v < M; v++) {
    int is_insert = (cm->sttype[v] == IL_st || cm->sttype[v] == IR_st);
    if (cm->sttype[v] == S_st) { psi[v] = 1.0; continue; }
    int final_y = is_insert ? 1 : 0;
    for (int y = cm->pnum[v] - 1; y >= final_y; y--) {
      int x = cm->plast[v] - y;
      if (x < 0 || x >= M || cm->sttype[x] == B_st) continue;
      int k = v - cm->cfirst[x];
      if (k < 0 || k >= cm->cnum[x]) return infx_fail(eslEINVAL, errbuf, "state %d is not a child of its parent %d", v, x);
      psi[v] += psi[x] * tr(x, k);
    }
    if (is_insert) psi[v] += psi[v] * (cm->t[v][0] / (1. - cm->t[v][0]));
  }
  return eslOK;
}

static void inserts_given_node(const CM_t *cm, int nd, int *i1, int *i2)
{
  *i1 = *i2 = -1;
  if (nd < 0) return;
  int v = cm->nodemap[nd];
  switch (cm->ndtype[nd]) {
  case MATP_nd: *i1 = v + 4; *i2 = v + 5; break;
  case MATL_nd: *i1 = v + 2; break;
  case MATR_nd: *i1 = v + 2; break;
  case BEGR_nd: *i1 = v + 1; break;
  case ROOT_nd: *i1 = v + 1; *i2 = v + 2; break;
  }
}

template <typename T> static T **alloc2(int r, int c, T val)
{
  T **a = (T **) malloc(sizeof(T *) * r);
  T *b = (T *) malloc(sizeof(T) * (size_t) r * c);
  for (int i = 0; i < r; i++) { a[i] = b + (size_t) i * c; for (int j = 0; j < c; j++) a[i][j] = val; }
  return a;
}

static void tr_penalties_free(CM_TR_PENALTIES *trp)
{
  if (!trp) return;
  CM_TR_PENALTIES *t = trp;
  if (t->g_ptyAA)  { free(t->g_ptyAA[0]);  free(t->g_ptyAA); }
  if (t->l_ptyAA)  { free(t->l_ptyAA[0]);  free(t->l_ptyAA); }
  if (t->ig_ptyAA) { free(t->ig_ptyAA[0]); free(t->ig_ptyAA); }
  if (t->il_ptyAA) { free(t->il_ptyAA[0]); free(t->il_ptyAA); }
  free(t);
}

/* ignore_inserts: the reference's flag (src/cm_trunc.c:51): the whole
 * probability of a node goes to its match or bifurcation state, none to
 * the inserts before it.  cm_Configure builds the table without it. */
static int tr_penalties_build(const CM_t *cm, int ignore_inserts, CM_TR_PENALTIES **ret_trp, char *errbuf)
{
  int status;
  *ret_trp = NULL;
  if (cm->emap == NULL) return infx_fail(eslEINVAL, errbuf, "truncation penalties need the emit map");
  const int M = cm->M;
  CM_TR_PENALTIES *trp = (CM_TR_PENALTIES *) calloc(1, sizeof(CM_TR_PENALTIES));
  *ret_trp = trp;
  trp->M = M;
  trp->ignored_inserts = ignore_inserts ? 1 : 0;
  trp->g_ptyAA  = alloc2<float>(NTRPENALTY, M, (float) IMPOSSIBLE);
  trp->l_ptyAA  = alloc2<float>(NTRPENALTY, M, (float) IMPOSSIBLE);
  trp->ig_ptyAA = alloc2<int>(NTRPENALTY, M, -INFTY);
  trp->il_ptyAA = alloc2<int>(NTRPENALTY, M, -INFTY);

  /* local begin probabilities of the global model (cm_CalculateLocalBeginProbs) */
  std::vector<float> begin(M, 0.f);
  {
    int nstarts = 0;
    for (int nd = 2; nd < cm->nodes; nd++)
      if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd || cm->ndtype[nd] == BIF_nd) nstarts++;
    begin[cm->nodemap[1]] = 1. - cm->pbegin;
    float p = cm->pbegin / (float) nstarts;
    for (int nd = 2; nd < cm->nodes; nd++)
      if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd || cm->ndtype[nd] == BIF_nd)
        begin[cm->nodemap[nd]] = p;
  }
  std::vector<double> psi;
  if ((status = expected_state_occupancy(cm, psi, errbuf)) != eslOK) return status;

  float g_5and3 = 2. / (cm->clen * (cm->clen + 1));
  float g_5or3  = 1. / cm->clen;
  float prv5 = 0., prv3 = 0., prv53 = 0.;
  for (int nd = 0; nd < cm->nodes; nd++) {
    int lpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd) ? cm->emap->lpos[nd] : cm->emap->lpos[nd] + 1;
    int rpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATR_nd) ? cm->emap->rpos[nd] : cm->emap->rpos[nd] - 1;
    if (cm->ndtype[nd] == END_nd) {
      prv5 = prv3 = prv53 = 0.;
    } else if (cm->ndtype[nd] == BEGL_nd || cm->ndtype[nd] == BEGR_nd) {
      int b = cm->plast[cm->nodemap[nd]];
      prv5  = (cm->ndtype[nd] == BEGL_nd) ? 0. : trp->l_ptyAA[TRPENALTY_5P_ONLY][b];
      prv3  = (cm->ndtype[nd] == BEGR_nd) ? 0. : trp->l_ptyAA[TRPENALTY_3P_ONLY][b];
      prv53 = trp->l_ptyAA[TRPENALTY_5P_AND_3P][b];
    } else if (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd || cm->ndtype[nd] == MATR_nd || cm->ndtype[nd] == BIF_nd) {
      int m = cm->nodemap[nd], i1, i2;
      inserts_given_node(cm, nd - 1, &i1, &i2);
      float m_psi = psi[m];
      /* The reference compares the node type with the state id MATP_MP here
       * (src/cm_trunc.c:290), which never matches: ML/MR occupancy is not
       * added to MP's. */
      float i1_psi = (i1 == -1) ? 0. : psi[i1];
      float i2_psi = (i2 == -1) ? 0. : psi[i2];
      float summed_psi = m_psi + i1_psi + i2_psi;
      if (ignore_inserts) {
        i1_psi = i2_psi = 0.;
        summed_psi = m_psi;
      }

      trp->g_ptyAA[TRPENALTY_5P_AND_3P][m] = (m_psi / summed_psi) * g_5and3;
      if (i1 != -1) trp->g_ptyAA[TRPENALTY_5P_AND_3P][i1] = (i1_psi / summed_psi) * g_5and3;
      if (i2 != -1) trp->g_ptyAA[TRPENALTY_5P_AND_3P][i2] = (i2_psi / summed_psi) * g_5and3;
      if (rpos == cm->clen) {
        trp->g_ptyAA[TRPENALTY_5P_ONLY][m] = (m_psi / summed_psi) * g_5or3;
        if (i1 != -1) trp->g_ptyAA[TRPENALTY_5P_ONLY][i1] = (i1_psi / summed_psi) * g_5or3;
        if (i2 != -1) trp->g_ptyAA[TRPENALTY_5P_ONLY][i2] = (i2_psi / summed_psi) * g_5or3;
      }
      if (lpos == 1) {
        trp->g_ptyAA[TRPENALTY_3P_ONLY][m] = (m_psi / summed_psi) * g_5or3;
        if (i1 != -1) trp->g_ptyAA[TRPENALTY_3P_ONLY][i1] = (i1_psi / summed_psi) * g_5or3;
        if (i2 != -1) trp->g_ptyAA[TRPENALTY_3P_ONLY][i2] = (i2_psi / summed_psi) * g_5or3;
      }
      int subtree_clen = rpos - lpos + 1;
      int nfrag5 = subtree_clen, nfrag3 = subtree_clen, nfrag53 = (subtree_clen * (subtree_clen + 1)) / 2;
      float cur5  = begin[m] / (float) nfrag5 + prv5;
      float cur3  = begin[m] / (float) nfrag3 + prv3;
      float cur53 = begin[m] / (float) nfrag53 + prv53;
      trp->l_ptyAA[TRPENALTY_5P_AND_3P][m] = (m_psi / summed_psi) * cur53;
      if (i1 != -1) trp->l_ptyAA[TRPENALTY_5P_AND_3P][i1] = (i1_psi / summed_psi) * cur53;
      if (i2 != -1) trp->l_ptyAA[TRPENALTY_5P_AND_3P][i2] = (i2_psi / summed_psi) * cur53;
      trp->l_ptyAA[TRPENALTY_5P_ONLY][m] = (m_psi / summed_psi) * cur5;
      if (i1 != -1) trp->l_ptyAA[TRPENALTY_5P_ONLY][i1] = (i1_psi / summed_psi) * cur5;
      if (i2 != -1) trp->l_ptyAA[TRPENALTY_5P_ONLY][i2] = (i2_psi / summed_psi) * cur5;
      trp->l_ptyAA[TRPENALTY_3P_ONLY][m] = (m_psi / summed_psi) * cur3;
      if (i1 != -1) trp->l_ptyAA[TRPENALTY_3P_ONLY][i1] = (i1_psi / summed_psi) * cur3;
      if (i2 != -1) trp->l_ptyAA[TRPENALTY_3P_ONLY][i2] = (i2_psi / summed_psi) * cur3;
      prv5  = (cm->ndtype[nd] == MATL_nd) ? cur5 : 0.;
      prv3  = (cm->ndtype[nd] == MATR_nd) ? cur3 : 0.;
      prv53 = cur53;
    }
  }

  /* probabilities -> bit scores and x1000 integer scores */
  for (int v = 0; v < M; v++) {
    bool emitter_begin = (cm->stid[v] == MATP_MP || cm->stid[v] == MATL_ML || cm->stid[v] == MATR_MR || cm->stid[v] == BIF_B) ||
                         ((cm->sttype[v] == IL_st || cm->sttype[v] == IR_st) && !StateIsDetached(cm, v));
    if (!emitter_begin) continue;
    if (cm->stid[v] == MATP_IL && cm->ndtype[cm->ndidx[v] + 1] == END_nd) continue;
    if (!NOT_IMPOSSIBLE(trp->g_ptyAA[TRPENALTY_5P_AND_3P][v]))
      return infx_fail(eslEINVAL, errbuf, "couldn't create truncation penalties for the CM (state %d)", v);
    trp->ig_ptyAA[TRPENALTY_5P_AND_3P][v] = Prob2Score(trp->g_ptyAA[TRPENALTY_5P_AND_3P][v], 1.0);
    trp->g_ptyAA[TRPENALTY_5P_AND_3P][v]  = sreLOG2(trp->g_ptyAA[TRPENALTY_5P_AND_3P][v]);
    if (NOT_IMPOSSIBLE(trp->g_ptyAA[TRPENALTY_5P_ONLY][v])) {
      trp->ig_ptyAA[TRPENALTY_5P_ONLY][v] = Prob2Score(trp->g_ptyAA[TRPENALTY_5P_ONLY][v], 1.0);
      trp->g_ptyAA[TRPENALTY_5P_ONLY][v]  = sreLOG2(trp->g_ptyAA[TRPENALTY_5P_ONLY][v]);
    }
    if (NOT_IMPOSSIBLE(trp->g_ptyAA[TRPENALTY_3P_ONLY][v])) {
      trp->ig_ptyAA[TRPENALTY_3P_ONLY][v] = Prob2Score(trp->g_ptyAA[TRPENALTY_3P_ONLY][v], 1.0);
      trp->g_ptyAA[TRPENALTY_3P_ONLY][v]  = sreLOG2(trp->g_ptyAA[TRPENALTY_3P_ONLY][v]);
    }
    trp->il_ptyAA[TRPENALTY_5P_AND_3P][v] = Prob2Score(trp->l_ptyAA[TRPENALTY_5P_AND_3P][v], 1.0);
    trp->il_ptyAA[TRPENALTY_5P_ONLY][v]   = Prob2Score(trp->l_ptyAA[TRPENALTY_5P_ONLY][v], 1.0);
    trp->il_ptyAA[TRPENALTY_3P_ONLY][v]   = Prob2Score(trp->l_ptyAA[TRPENALTY_3P_ONLY][v], 1.0);
    trp->l_ptyAA[TRPENALTY_5P_AND_3P][v]  = sreLOG2(trp->l_ptyAA[TRPENALTY_5P_AND_3P][v]);
    trp->l_ptyAA[TRPENALTY_5P_ONLY][v]    = sreLOG2(trp->l_ptyAA[TRPENALTY_5P_ONLY][v]);
    trp->l_ptyAA[TRPENALTY_3P_ONLY][v]    = sreLOG2(trp->l_ptyAA[TRPENALTY_3P_ONLY][v]);
  }
  return eslOK;
}

/* cm->trp for cm_Configure (on a failure the partial table is attached, and
 * released with the model) */
extern "C" int infx_tr_penalties_create(CM_t *cm, char *errbuf)
{
  CM_TR_PENALTIES *trp = NULL;
  const int status = tr_penalties_build(cm, 0, &trp, errbuf);
  cm->trp = trp;
  return status;
}

/* The float penalty tables of an unconfigured (global) model, built with or
 * without the inserts' shares, into out[2][NTRPENALTY][M]: glocal, then
 * local; the model is not changed. */
extern "C" int infx_debug_tr_penalties(const CM_t *cm, int ignore_inserts, float *out, char *errbuf)
{
  CM_TR_PENALTIES *trp = NULL;
  const int status = tr_penalties_build(cm, ignore_inserts, &trp, errbuf);
  if (status == eslOK) {
    const size_t M = (size_t) cm->M;
    for (int i = 0; i < NTRPENALTY; i++) {
      memcpy(out + (size_t) i * M, trp->g_ptyAA[i], M * sizeof(float));
      memcpy(out + (size_t) (NTRPENALTY + i) * M, trp->l_ptyAA[i], M * sizeof(float));
    }
  }
  tr_penalties_free(trp);
  return status;
}
