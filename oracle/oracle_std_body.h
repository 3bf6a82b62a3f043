/* oracle_std_body.h -- body of the oracle's standard scanners, included once
 * per score type.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The includer defines:
 *   FN        function name
 *   T         cell type (float or int)
 *   OP(a,b)   child combination: ESL_MAX (CYK), FLogsum, ILogsum
 *   ADD(a,b)  cell + score addition
 *   IMP       the impossible value (IMPOSSIBLE or -INFTY)
 *   ROOTMAX(a) ESL_MAX(IMP, a)
 *   TSC, ESC, BEGINSC, NOTIMP_BEGIN(v)
 *   SCORE(x)  bit score of a cell (identity or Scorify)
 *   INITSC    init score table builder
 *   INIT_D0   d = 0 matrix initializer
 * and follows src/cm_dpsearch.c:90-681 (FastCYKScan), :1106-1690
 * (FastIInsideScan) and :1731-2320 (FastFInsideScan).
 */
static int FN(const CM_t *cm, const int qdbidx, const ESL_DSQ *dsq, int64_t i0, int64_t j0, float cutoff,
              int do_null3, int greedy, float env_cutoff, int want_rows, ORACLE_RESULT *res, char *errbuf)
{
  int M = cm->M;
  int L = (int) (j0 - i0 + 1);
  int Wsmx = cm->W;
  int W = (Wsmx > L) ? L : Wsmx;
  int v, d, k, j, y, w, yoffset, i;
  int status = eslOK;
  const int *dmin = NULL, *dmax = NULL;
  int do_banded = 0;
  if (qdbidx == SMX_QDB1_TIGHT)      { do_banded = 1; dmin = cm->qdbinfo->dmin1; dmax = cm->qdbinfo->dmax1; }
  else if (qdbidx == SMX_QDB2_LOOSE) { do_banded = 1; dmin = cm->qdbinfo->dmin2; dmax = cm->qdbinfo->dmax2; }

  /* band rows for jrow = 1..Wsmx (src/cm_mx.c:5725-5739) */
  int *dnAA = (int *) malloc(sizeof(int) * (size_t) (Wsmx + 1) * M);
  int *dxAA = (int *) malloc(sizeof(int) * (size_t) (Wsmx + 1) * M);
  for (int jr = 1; jr <= Wsmx; jr++)
    for (v = 0; v < M; v++) {
      int mp = (cm->sttype[v] == MP_st);
      int dn = mp ? 2 : 1, dx = jr;
      if (do_banded) { dn = mp ? ESL_MAX(dmin[v], 2) : ESL_MAX(dmin[v], 1); dx = ESL_MIN(jr, ESL_MIN(dmax[v], Wsmx)); }
      dn = ESL_MIN(dn, Wsmx);
      dnAA[(size_t) jr * M + v] = dn;
      dxAA[(size_t) jr * M + v] = dx;
    }

  /* matrices: alpha[2][M][Wsmx+1]; BEGL decks [Wsmx+1][nbegl][Wsmx+1] */
  int nbegl = 0;
  int *begl_idx = (int *) malloc(sizeof(int) * M);
  for (v = 0; v < M; v++) begl_idx[v] = (cm->stid[v] == BEGL_S) ? nbegl++ : -1;
  const size_t W1 = (size_t) Wsmx + 1;
  T *alpha = (T *) malloc(sizeof(T) * 2 * M * W1);
  T *begl  = (T *) malloc(sizeof(T) * (nbegl ? nbegl : 1) * W1 * W1);
#define A(jj, vv, dd)  alpha[((size_t) (jj) * M + (vv)) * W1 + (dd)]
#define BG(jj, vv, dd) begl[((size_t) (jj) * nbegl + begl_idx[vv]) * W1 + (dd)]
  INIT_D0(cm, alpha, begl, begl_idx, nbegl, M, Wsmx);

  T **init_scAA = INITSC(cm);
  T *sc_v = (T *) malloc(sizeof(T) * W1);
  for (d = 0; d <= (int) Wsmx; d++) sc_v[d] = IMP;
  int *jp_wA = (int *) malloc(sizeof(int) * W1);
  float *bestsc = (float *) malloc(sizeof(float) * W1);
  int   *bestr  = (int *) malloc(sizeof(int) * W1);
  float *vsc = (float *) malloc(sizeof(float) * M);
  for (v = 0; v < M; v++) vsc[v] = IMPOSSIBLE;
  float vsc_root = IMPOSSIBLE;
  int64_t envi = j0 + 1, envj = i0 - 1;
  double *act = NULL;
  if (do_null3) act = (double *) calloc((size_t) (W + 1) * INFX_K, sizeof(double));
  HITACC hacc;
  hitacc_init(&hacc, L, i0, j0, W, greedy);
  if (want_rows) {
    res->W = W;
    res->rows_sc = (float *) malloc(sizeof(float) * (size_t) L * (W + 1));
    res->rows_r  = (int *) malloc(sizeof(int) * (size_t) L * (W + 1));
    res->rows_mode = NULL;
  }
  const float **esc_vAA = (const float **) NULL; (void) esc_vAA;

  for (j = (int) i0; j <= (int) j0; j++) {
    int jp_g = j - (int) i0 + 1;
    int cur = j % 2, prv = (j - 1) % 2;
    const int *dnA = (jp_g >= W) ? &dnAA[(size_t) W * M] : &dnAA[(size_t) jp_g * M];
    const int *dxA = (jp_g >= W) ? &dxAA[(size_t) W * M] : &dxAA[(size_t) jp_g * M];
    for (d = 0; d <= W; d++) jp_wA[d] = (j - d) % (W + 1);
    if (act) {
      double *dst = &act[(size_t) (jp_g % (W + 1)) * INFX_K];
      const double *src = &act[(size_t) ((jp_g - 1) % (W + 1)) * INFX_K];
      for (int x = 0; x < INFX_K; x++) dst[x] = src[x];
      oracle_DCount(dst, dsq[j], 1.);
    }

    for (v = M - 1; v > 0; v--) {
      int st = cm->sttype[v];
      if (st == E_st) continue;
      const ESCT *esc_v = ESC(v);
      const TSCT *tsc_v = TSC(v);
      int emitmode = oracle_emitmode(st);
      int jp_v = (cm->stid[v] == BEGL_S) ? (j % (W + 1)) : cur;
      int jp_y = (oracle_rdelta(st) > 0) ? prv : cur;
      int sd = oracle_sdelta(st);
      int cnum = cm->cnum[v];
      int dn = dnA[v], dx = dxA[v];
      ESCT esc_j = ESC_IMP;
      if (st == IR_st || st == MR_st) esc_j = esc_v[dsq[j]];
      T sc;

      if (st == B_st) {
        w = cm->cfirst[v]; y = cm->cnum[v];
        for (d = dn; d <= dx; d++) {
          int kmin, kmax;
          if (do_banded) {
            int dn_y = ESL_MIN(dmin[y], Wsmx), dx_y = ESL_MIN(dmax[y], Wsmx);
            int dn_w = ESL_MIN(dmin[w], Wsmx), dx_w = ESL_MIN(dmax[w], Wsmx);
            kmin = ESL_MAX(0, ESL_MAX(dn_y, d - dx_w));
            kmax = ESL_MIN(dx_y, d - dn_w);
          } else { kmin = 0; kmax = d; }
          sc = init_scAA[v][d - sd];
          for (k = kmin; k <= kmax; k++) sc = OP(sc, ADD(BG(jp_wA[k], w, d - k), A(jp_y, y, k)));
          A(jp_v, v, d) = sc;
        }
      }
      else if (cm->stid[v] == BEGL_S) {
        y = cm->cfirst[v];
        for (d = dn; d <= dx; d++) {
          sc = init_scAA[v][d - sd];
          for (yoffset = 0; yoffset < cm->cnum[v]; yoffset++) sc = OP(sc, ADD(A(jp_y, y + yoffset, d - sd), tsc_v[yoffset]));
          BG(jp_v, v, d) = sc;
        }
      }
      else if (st == IL_st || st == IR_st) {
        y = cm->cfirst[v];
        int dp_y = dn - sd;
        i = j - dn + 1;
        for (d = dn; d <= dx; d++, dp_y++) {
#define C(kk) ADD(A(jp_y, y + (kk), dp_y), tsc_v[kk])
          switch (cnum) {
          case 2: sc = OP(C(1), init_scAA[v][dp_y]); sc = OP(sc, C(0)); break;
          case 3: sc = OP(C(2), C(1)); sc = OP(sc, init_scAA[v][dp_y]); sc = OP(sc, C(0)); break;
          case 4: sc = OP(C(3), C(2)); sc = OP(sc, init_scAA[v][dp_y]); sc = OP(sc, C(1)); sc = OP(sc, C(0)); break;
          case 5: sc = OP(C(4), C(3)); sc = OP(sc, init_scAA[v][dp_y]); sc = OP(sc, C(1)); sc = OP(sc, C(2)); sc = OP(sc, C(0)); break;
          case 6: sc = OP(C(5), init_scAA[v][dp_y]); sc = OP(sc, C(4)); sc = OP(sc, C(3)); sc = OP(sc, C(2)); sc = OP(sc, C(1)); sc = OP(sc, C(0)); break;
          default: status = eslEINVAL; goto DONE;
          }
          if (emitmode == EMITLEFT)       sc = ADD(sc, esc_v[dsq[i--]]);
          else if (emitmode == EMITRIGHT) sc = ADD(sc, esc_j);
          A(jp_v, v, d) = sc;
        }
      }
      else {
        y = cm->cfirst[v];
        int dp_y = dn - sd;
        i = j - dn + 1;
        for (d = dn; d <= dx; d++, dp_y++) {
          switch (cnum) {
          case 2: sc_v[d] = OP(C(1), init_scAA[v][dp_y]); sc_v[d] = OP(sc_v[d], C(0)); break;
          case 3: sc_v[d] = OP(C(2), C(1)); sc_v[d] = OP(sc_v[d], init_scAA[v][dp_y]); sc_v[d] = OP(sc_v[d], C(0)); break;
          case 4: sc_v[d] = OP(C(3), C(2)); sc_v[d] = OP(sc_v[d], init_scAA[v][dp_y]); sc_v[d] = OP(sc_v[d], C(1)); sc_v[d] = OP(sc_v[d], C(0)); break;
          case 5: sc_v[d] = OP(C(4), C(3)); sc_v[d] = OP(sc_v[d], init_scAA[v][dp_y]); sc_v[d] = OP(sc_v[d], C(1)); sc_v[d] = OP(sc_v[d], C(2)); sc_v[d] = OP(sc_v[d], C(0)); break;
          case 6: sc_v[d] = OP(C(5), init_scAA[v][dp_y]); sc_v[d] = OP(sc_v[d], C(4)); sc_v[d] = OP(sc_v[d], C(3)); sc_v[d] = OP(sc_v[d], C(2)); sc_v[d] = OP(sc_v[d], C(1)); sc_v[d] = OP(sc_v[d], C(0)); break;
          default: status = eslEINVAL; goto DONE;
          }
#undef C
        }
        switch (emitmode) {
        case EMITLEFT:  for (d = dn; d <= dx; d++) A(jp_v, v, d) = ADD(sc_v[d], esc_v[dsq[i--]]); break;
        case EMITNONE:  for (d = dn; d <= dx; d++) A(jp_v, v, d) = sc_v[d]; break;
        case EMITRIGHT: for (d = dn; d <= dx; d++) A(jp_v, v, d) = ADD(sc_v[d], esc_j); break;
        case EMITPAIR:  for (d = dn; d <= dx; d++) A(jp_v, v, d) = ADD(sc_v[d], esc_v[dsq[i--] * INFX_KP + dsq[j]]); break;
        }
      }
      if (cm->stid[v] != BEGL_S) { for (d = dn; d <= dx; d++) vsc[v] = ESL_MAX(vsc[v], SCORE(A(jp_v, v, d))); }
      else                       { for (d = dn; d <= dx; d++) vsc[v] = ESL_MAX(vsc[v], SCORE(BG(jp_v, v, d))); }
    }

    /* ROOT_S and local begins */
    {
      const TSCT *tsc_v = TSC(0);
      for (d = 0; d <= W; d++) { bestr[d] = -1; bestsc[d] = IMPOSSIBLE; }
      int jp_v = cur;
      for (d = dnA[0]; d <= dxA[0]; d++) {
        bestr[d] = 0;
        y = cm->cfirst[0];
        A(jp_v, 0, d) = ROOTMAX(ADD(A(cur, y, d), tsc_v[0]));
        for (yoffset = 1; yoffset < cm->cnum[0]; yoffset++)
          A(jp_v, 0, d) = OP(A(jp_v, 0, d), ADD(A(cur, y + yoffset, d), tsc_v[yoffset]));
      }
      if (cm->flags & CMH_LOCAL_BEGIN) {
        for (y = 1; y < M; y++) {
          if (NOTIMP_BEGIN(y)) {
            int dn = ESL_MAX(dnA[0], dnA[y]);
            int dx = ESL_MIN(dxA[0], dxA[y]);
            if (cm->stid[y] == BEGL_S) {
              int jp_y = j % (W + 1);
              for (d = dn; d <= dx; d++)
                if (A(jp_v, 0, d) < ADD(BG(jp_y, y, d), BEGINSC(y))) { A(jp_v, 0, d) = ADD(BG(jp_y, y, d), BEGINSC(y)); bestr[d] = y; }
            } else {
              int jp_y = cur;
              for (d = dn; d <= dx; d++)
                if (A(jp_v, 0, d) < ADD(A(jp_y, y, d), BEGINSC(y))) { A(jp_v, 0, d) = ADD(A(jp_y, y, d), BEGINSC(y)); bestr[d] = y; }
            }
          }
        }
      }
      for (d = dnA[0]; d <= dxA[0]; d++) {
        bestsc[d] = SCORE(A(jp_v, 0, d));
        vsc_root = ESL_MAX(vsc_root, SCORE(A(jp_v, 0, d)));
      }
      for (d = dnA[0]; d <= dxA[0]; d++)
        if (A(jp_v, 0, d) >= env_cutoff) { envi = ESL_MIN(envi, j - d + 1); envj = ESL_MAX(envj, j); }
      if (want_rows) {
        for (d = 0; d <= W; d++) {
          res->rows_sc[(size_t) (jp_g - 1) * (W + 1) + d] = bestsc[d];
          res->rows_r[(size_t) (jp_g - 1) * (W + 1) + d] = bestr[d];
        }
      }
      if ((status = hitacc_update(&hacc, cm, PLI_PASS_STD_ANY, j, dnA[0], dxA[0], bestsc, bestr, NULL, act, cutoff, do_null3)) != eslOK) goto DONE;
    }
  }
  vsc[0] = vsc_root;
  res->envi = (envi == j0 + 1) ? -1 : envi;
  res->envj = (envj == i0 - 1) ? -1 : envj;
  res->sc = vsc_root;
  res->mode = TRMODE_J;
  res->vsc = vsc; vsc = NULL;
  status = hitacc_finish(&hacc, res, errbuf);

DONE:
#undef A
#undef BG
  hitacc_free(&hacc);
  free(dnAA); free(dxAA); free(begl_idx); free(alpha); free(begl);
  free(init_scAA[0]); free(init_scAA);
  free(sc_v); free(jp_wA); free(bestsc); free(bestr); free(vsc); free(act);
  if (status != eslOK && errbuf && !errbuf[0]) snprintf(errbuf, eslERRBUFSIZE, "oracle scan failed (%d)", status);
  return status;
}
