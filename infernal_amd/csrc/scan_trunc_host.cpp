/* scan_trunc_host.cpp -- host side of the truncated GPU scanners
 * RefTrCYKScan (src/cm_dpsearch_trunc.c:79) and RefITrInsideScan (:741):
 * the device profile of a configured CM, the windowed launch and the C-ABI
 * entry points.  infx_TrScanWindows runs many windows of one model in a single
 * launch (one workgroup per window).
 */
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <algorithm>
#include <vector>

#include "cm_internal.h"
#include "scan_kernels.h"
#include "scan_host.h"

extern "C" hipError_t infx_launch_tr(int semi, int want_vsc, dim3 grid, dim3 block, hipStream_t stream, const TrScanArgs *a);
extern "C" int infx_tr_occupancy(int semi, int want_vsc, int nw);

static uint32_t fb(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static int f2i_host(float f) { return (f >= -2147483648.0f && f < 2147483648.0f) ? (int) f : INT_MIN; }

struct TrProfile {
  int qdbidx = -1, semi = -1, pty = -1, local = -1;
  std::vector<TState> st;
  std::vector<uint32_t> esc, rtmpl, dtmpl;
  std::vector<TrPen> pen;
  std::vector<int> pblk;     /* per root block: offsets, then begin indices */
  int R = 0, RB = 0, RT = 0, root_dx = 0, nlev = 0;
  uint32_t el = 0;
  DevBuf d_st, d_esc, d_rtmpl, d_dtmpl, d_pen, d_pblk;
  bool uploaded = false;
  /* the row schedule for sched_nw wavefronts (per-state items if sched_vsc) */
  int sched_nw = 0, sched_vsc = -1;
  std::vector<int> sched;
  std::vector<TrItem> items;
  DevBuf d_sched, d_items;
};

struct TrDev {
  std::vector<TrProfile *> prof;
  DevBuf d_dsq, d_scratch, d_cand, d_ncand, d_win, d_wout, d_lsum;
  ~TrDev() {
    for (auto *p : prof) {
      p->d_st.release(); p->d_esc.release(); p->d_rtmpl.release(); p->d_dtmpl.release(); p->d_pen.release(); p->d_pblk.release();
      p->d_sched.release(); p->d_items.release();
      delete p;
    }
    d_dsq.release(); d_scratch.release(); d_cand.release(); d_ncand.release(); d_win.release(); d_wout.release(); d_lsum.release();
  }
};

extern "C" void infx_trdev_release(void *trdev) { if (trdev) delete (TrDev *) trdev; }

/* d = 0 J cells (cm_tr_scan_mx_InitializeFloats/Integers, src/cm_mx.c:6853-6980).
 * The integer version starts S/D/BEGL cells from the float endsc converted to
 * int, as the reference does. */
static void tr_d0_init(const CM_t *cm, int semi, std::vector<uint32_t> &a0)
{
  const int M = cm->M;
  std::vector<float> f(M, (float) IMPOSSIBLE);
  std::vector<int> n(M, -INFTY);
  for (int v = M - 1; v >= 0; v--) {
    int st = cm->sttype[v];
    bool begl = (cm->stid[v] == BEGL_S);
    if (!begl && st == E_st) { f[v] = 0.; n[v] = 0; }
    else if (begl || st == S_st || st == D_st) {
      int y = cm->cfirst[v];
      float x = cm->endsc[v];
      int xi = f2i_host(cm->endsc[v]);
      for (int k = 0; k < cm->cnum[v]; k++) {
        x = ESL_MAX(x, (f[y + k] + cm->tsc[v][k]));
        xi = ESL_MAX(xi, (int) ((unsigned) n[y + k] + (unsigned) cm->itsc[v][k]));
      }
      x = ESL_MAX(x, IMPOSSIBLE);
      xi = ESL_MAX(xi, -INFTY);
      f[v] = x; n[v] = xi;
    } else if (st == B_st) {
      f[v] = f[cm->cfirst[v]] + f[cm->cnum[v]];
      n[v] = (int) ((unsigned) n[cm->cfirst[v]] + (unsigned) n[cm->cnum[v]]);
    }
  }
  a0.resize(M);
  for (int v = 0; v < M; v++) a0[v] = (semi == TSEMI_IINSIDE) ? (uint32_t) n[v] : fb(f[v]);
}

static int build_tr_profile(CM_t *cm, CM_TR_SCAN_MX *trsmx, int qdbidx, int semi, int pty, TrProfile &P, char *errbuf,
                            std::vector<int> *ret_lo = nullptr, std::vector<int> *ret_hi = nullptr)
{
  const int M = cm->M, W = trsmx->W;
  const int *dmax = qdbidx == SMX_QDB1_TIGHT ? cm->qdbinfo->dmax1 : qdbidx == SMX_QDB2_LOOSE ? cm->qdbinfo->dmax2 : NULL;
  const int banded = (dmax != NULL);
  std::vector<int> dx(M), lo(M, INT32_MAX), hi(M, INT32_MIN);
  for (int v = 0; v < M; v++) dx[v] = trsmx->dxAAA[qdbidx][W][v];
  auto extend = [&](int v, int a, int b) { if (a > b) return; lo[v] = ESL_MIN(lo[v], a); hi[v] = ESL_MAX(hi[v], b); };
  for (int v = 1; v < M; v++) {
    int st = cm->sttype[v];
    if (st == E_st || dx[v] < 1) continue;
    extend(v, 1, dx[v]);
    if (st == B_st) {
      int w = cm->cfirst[v], y = cm->cnum[v];
      for (int d = 1; d <= dx[v]; d++) {
        int kmin, kmax;
        if (banded) { kmin = ESL_MAX(0, d - ESL_MIN(dmax[w], W)); kmax = ESL_MIN(ESL_MIN(dmax[y], W), d); }
        else { kmin = 0; kmax = d; }
        if (kmin == 0) extend(w, d, d);    /* L mode: BEGL w alone, this row */
        if (kmax == d) extend(y, d, d);    /* R mode: BEGR y alone */
        extend(w, 0, 0);                   /* idle lanes of a block read cell 0 */
        if (kmin > kmax) continue;
        extend(y, kmin, kmax);
        extend(w, d - kmax, d - kmin);
      }
    } else {
      int em = Emitmode(st);
      int a = (em == EMITNONE) ? 1 : 0;
      int b = (em == EMITPAIR) ? dx[v] - 1 : dx[v];
      for (int k = 0; k < cm->cnum[v]; k++) extend(cm->cfirst[v] + k, a, b);
    }
  }
  /* deck ring rows of each BEGL state: its parent reads rows j-k for k up to
   * the BEGR band (kdx_y), so a power of two above kdx_y rows suffices */
  std::vector<int> ring(M, 0), parent_b(M, -1);
  for (int v = 1; v < M; v++)
    if (cm->sttype[v] == B_st) {
      const int w = cm->cfirst[v], y = cm->cnum[v];
      parent_b[w] = v;
      const int kdy = banded ? ESL_MIN(dmax[y], W) : W;
      int r = 1;
      while (r < kdy + 1) r <<= 1;
      ring[w] = r;
    }
  std::vector<uint32_t> a0;
  tr_d0_init(cm, semi, a0);
  const uint32_t imp = (semi == TSEMI_IINSIDE) ? (uint32_t) (-INFTY) : fb((float) IMPOSSIBLE);
  const uint32_t zero = (semi == TSEMI_IINSIDE) ? 0u : fb(0.f);
  std::vector<int> base(M, 0), tbase(M, 0), stride(M, 0);
  std::vector<uint32_t> rJ, rL, rR, dJ, dL, dR;
  for (int v = 1; v < M; v++) {
    if (lo[v] > hi[v]) continue;
    lo[v] = ESL_MAX(lo[v], 0); hi[v] = ESL_MIN(hi[v], W);
    if (lo[v] > hi[v]) continue;
    bool begl = (cm->stid[v] == BEGL_S);
    if (begl && ring[v] == 0) return infx_fail(eslEINVAL, errbuf, "BEGL state %d has no bifurcation parent", v);
    std::vector<uint32_t> &tj = begl ? dJ : rJ, &tl = begl ? dL : rL, &tr = begl ? dR : rR;
    base[v] = (int) tj.size() - lo[v];
    stride[v] = hi[v] - lo[v] + 1;
    for (int r = 0; r < (begl ? ring[v] : 1); r++)     /* a BEGL state: every ring row starts as the template */
      for (int d = lo[v]; d <= hi[v]; d++) {
        tj.push_back(d == 0 ? a0[v] : imp);
        tl.push_back((d == 0 && cm->sttype[v] == E_st) ? zero : imp);
        tr.push_back((d == 0 && cm->sttype[v] == E_st) ? zero : imp);
      }
    if ((int64_t) tj.size() > INT32_MAX / 4) return infx_fail(eslERANGE, errbuf, "truncated scan profile too large");
  }
  if (ret_lo) *ret_lo = lo;
  if (ret_hi) *ret_hi = hi;
  while (rJ.size() % 64) { rJ.push_back(imp); rL.push_back(imp); rR.push_back(imp); }
  while (dJ.size() % 64) { dJ.push_back(imp); dL.push_back(imp); dR.push_back(imp); }
  P.R = (int) rJ.size(); P.RB = (int) dJ.size();
  P.rtmpl = rJ; P.rtmpl.insert(P.rtmpl.end(), rL.begin(), rL.end()); P.rtmpl.insert(P.rtmpl.end(), rR.begin(), rR.end());
  P.dtmpl = dJ; P.dtmpl.insert(P.dtmpl.end(), dL.begin(), dL.end()); P.dtmpl.insert(P.dtmpl.end(), dR.begin(), dR.end());
  P.RT = 0;
  for (int v = 1; v < M; v++)
    if (cm->sttype[v] == B_st) { tbase[v] = P.RT; P.RT += dx[v] + 1; }
  P.RT = (P.RT + 63) / 64 * 64;

  P.esc.clear();
  P.st.assign(M, TState());
  for (int v = 0; v < M; v++) {
    TState &s = P.st[v];
    memset(&s, 0, sizeof(s));
    int st = cm->sttype[v];
    s.kind = (st == E_st) ? KIND_SKIP : (st == B_st) ? KIND_B : (cm->stid[v] == BEGL_S) ? KIND_BEGL : (st == IL_st) ? KIND_IL : KIND_GENERIC;
    if (s.kind == KIND_IL && cm->cfirst[v] != v) return infx_fail(eslEINVAL, errbuf, "IL state %d does not loop on itself", v);
    s.emitmode = Emitmode(st);
    s.cnum = (st == B_st) ? 0 : cm->cnum[v];
    s.dx = dx[v];
    s.base = base[v];
    s.tbase = tbase[v];
    s.detached = StateIsDetached(cm, v);
    s.first_r = (st == IL_st || st == IR_st) ? 1 : 0;
    s.esc_off = s.lesc_off = s.resc_off = 0;
    if (st == ML_st || st == MR_st || st == IL_st || st == IR_st || st == MP_st) {
      int n = (st == MP_st) ? INFX_KP * INFX_KP : INFX_KP;
      s.esc_off = (int) P.esc.size();
      for (int x = 0; x < n; x++) P.esc.push_back(semi == TSEMI_IINSIDE ? (uint32_t) cm->ioesc[v][x] : fb(cm->oesc[v][x]));
      if (st == MP_st) {
        s.lesc_off = (int) P.esc.size();
        for (int x = 0; x < INFX_KP; x++) P.esc.push_back(semi == TSEMI_IINSIDE ? (uint32_t) cm->ilmesc[v][x] : fb(cm->lmesc[v][x]));
        s.resc_off = (int) P.esc.size();
        for (int x = 0; x < INFX_KP; x++) P.esc.push_back(semi == TSEMI_IINSIDE ? (uint32_t) cm->irmesc[v][x] : fb(cm->rmesc[v][x]));
      }
    }
    if (semi == TSEMI_IINSIDE) {
      s.init_ok = cm->iendsc[v] != -INFTY;
      s.endsc = (uint32_t) cm->iendsc[v];
      if (st != B_st && st != E_st) for (int k = 0; k < cm->cnum[v] && k < 6; k++) s.tsc[k] = (uint32_t) cm->itsc[v][k];
    } else {
      s.init_ok = NOT_IMPOSSIBLE(cm->endsc[v]);
      s.endsc = fb(cm->endsc[v]);
      if (st != B_st && st != E_st) for (int k = 0; k < cm->cnum[v] && k < 6; k++) s.tsc[k] = fb(cm->tsc[v][k]);
    }
    if (cm->stid[v] == BEGL_S) { s.ring_mask = ring[v] - 1; s.ring_stride = stride[v]; }
    if (st == B_st) {
      int w = cm->cfirst[v], y = cm->cnum[v];
      s.child_base[0] = base[w]; s.child_base[1] = base[y];
      s.child_base[2] = ring[w] - 1; s.child_base[3] = stride[w];
      s.banded = banded;
      if (banded) { s.kdx_y = ESL_MIN(dmax[y], W); s.kdx_w = ESL_MIN(dmax[w], W); }
    } else if (st != E_st) {
      for (int k = 0; k < cm->cnum[v] && k < 6; k++) s.child_base[k] = base[cm->cfirst[v] + k];
    }
  }
  /* same-row dependency levels: a state reads cells of the same row j from
   * its children (B: BEGL w and BEGR y), so it comes after all of them */
  P.nlev = 0;
  for (int v = M - 1; v >= 1; v--) {
    TState &s = P.st[v];
    if (s.kind == KIND_SKIP || s.dx < 1) { s.level = -1; continue; }
    int lv = 0;
    auto dep = [&](int c) { if (c != v && c < M && P.st[c].level >= 0 && P.st[c].kind != KIND_SKIP && P.st[c].dx >= 1) lv = ESL_MAX(lv, P.st[c].level + 1); };
    if (s.kind == KIND_B) { dep(cm->cfirst[v]); dep(cm->cnum[v]); }
    else for (int k = 0; k < cm->cnum[v]; k++) dep(cm->cfirst[v] + k);
    s.level = lv;
    P.nlev = ESL_MAX(P.nlev, lv + 1);
  }
  P.pen.clear();
  const int local = (cm->flags & CMH_LOCAL_BEGIN) ? 1 : 0;
  for (int y = 1; y < M; y++) {
    TrPen p;
    bool ok;
    if (semi == TSEMI_IINSIDE) {
      int ip = local ? cm->trp->il_ptyAA[pty][y] : cm->trp->ig_ptyAA[pty][y];
      ok = (ip != -INFTY); p.pen = (uint32_t) ip;
    } else {
      float fp = local ? cm->trp->l_ptyAA[pty][y] : cm->trp->g_ptyAA[pty][y];
      ok = NOT_IMPOSSIBLE(fp); p.pen = fb(fp);
    }
    if (!ok) continue;
    if (cm->stid[y] == BEGL_S || cm->sttype[y] == E_st) return infx_fail(eslEINVAL, errbuf, "truncated begin into state %d", y);
    p.y = y; p.base = base[y]; p.tbase = tbase[y]; p.dx = dx[y];
    if (!cm->emap) return infx_fail(eslEINVAL, errbuf, "truncated scan needs the model's emit map");
    {
      /* src/cm_tophits.c:3048: the consensus span of the begin state's node */
      const int nd = cm->ndidx[y];
      const int lpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATL_nd) ? cm->emap->lpos[nd] : cm->emap->lpos[nd] + 1;
      const int rpos = (cm->ndtype[nd] == MATP_nd || cm->ndtype[nd] == MATR_nd) ? cm->emap->rpos[nd] : cm->emap->rpos[nd] - 1;
      p.flags = (cm->sttype[y] == B_st ? 1 : 0) | (lpos == 1 ? 2 : 0) | (rpos == cm->clen ? 4 : 0);
    }
    P.pen.push_back(p);
  }
  P.root_dx = dx[0];
  /* the root's truncated begins per 64-cell block of d: a begin whose band
   * ends before the block changes none of its cells, so the block visits
   * only the others, still in ascending state order */
  {
    const int nblk = (W + 63) / 64 + 1;
    P.pblk.assign(nblk + 1, 0);
    std::vector<int> idx;
    for (int b = 0; b < nblk; b++) {
      P.pblk[b] = (int) idx.size();
      for (int k = 0; k < (int) P.pen.size(); k++)
        if (P.pen[k].dx >= 64 * b + 1) idx.push_back(k);
    }
    P.pblk[nblk] = (int) idx.size();
    P.pblk.insert(P.pblk.end(), idx.begin(), idx.end());
  }
  P.el = (semi == TSEMI_IINSIDE) ? (uint32_t) cm->iel_selfsc : fb(cm->el_selfsc);
  if (P.esc.empty()) P.esc.push_back(0);
  if (P.pen.empty()) { TrPen z; memset(&z, 0, sizeof(z)); P.pen.push_back(z); P.pen.pop_back(); }
  P.qdbidx = qdbidx; P.semi = semi; P.pty = pty; P.local = local;
  return eslOK;
}

/* A row's work for nw wavefronts: the states of each dependency level cut
 * into items (a bifurcation or ordinary state: one item per 64-cell block of
 * its steady-state band; an IL state, whose blocks chain, or any state when
 * per-state maxima are wanted: one item), dealt largest first to the least
 * loaded wavefront.  Costs are estimates in lane-steps: cells x child terms
 * x modes; a bifurcation block its k range (plus the lanes' spread) x modes;
 * an IL chain a few fixed-point sweeps per block. */
static void build_tr_schedule(const CM_t *cm, TrProfile &P, int nw, int vsc, int fill_L, int fill_R, int fill_T)
{
  const int M = cm->M;
  const int modes = 1 + fill_L + fill_R;
  std::vector<std::vector<std::pair<double, TrItem>>> lev(P.nlev);
  for (int v = M - 1; v >= 1; v--) {
    const TState &s = P.st[v];
    if (s.level < 0) continue;
    const int nb = (s.dx + 63) / 64;
    auto add = [&](int b0, int b1, double cost) { TrItem it; it.v = v; it.b0 = b0; it.b1 = b1; lev[s.level].push_back({cost, it}); };
    if (s.kind == KIND_B) {
      std::vector<double> c(nb);
      for (int b = 0; b < nb; b++) {
        const int d0 = 64 * b + 1, dl = ESL_MIN(d0 + 63, s.dx);
        const int kb0 = s.banded ? ESL_MAX(0, d0 - s.kdx_w) : 0, kb1 = s.banded ? ESL_MIN(s.kdx_y, dl) : dl;
        c[b] = 64.0 * (ESL_MAX(0, kb1 - kb0 + 1) + 8) * (modes + fill_T);
      }
      if (vsc) { double t = 0; for (double x : c) t += x; add(0, nb, t); }
      else for (int b = 0; b < nb; b++) add(b, b + 1, c[b]);
    } else {
      const double per = 64.0 * (ESL_MAX(1, s.cnum) + 2) * modes * (s.kind == KIND_IL ? 3 : 1);
      if (vsc || s.kind == KIND_IL) add(0, nb, per * nb);
      else for (int b = 0; b < nb; b++) add(b, b + 1, per);
    }
  }
  P.sched.assign((size_t) P.nlev * nw + 1, 0);
  P.items.clear();
  for (int l = 0; l < P.nlev; l++) {
    auto &L = lev[l];
    std::stable_sort(L.begin(), L.end(), [](const std::pair<double, TrItem> &a, const std::pair<double, TrItem> &b) { return a.first > b.first; });
    std::vector<double> load(nw, 0.0);
    std::vector<std::vector<TrItem>> per(nw);
    for (auto &x : L) {
      int k = 0;
      for (int q = 1; q < nw; q++) if (load[q] < load[k]) k = q;
      load[k] += x.first;
      per[k].push_back(x.second);
    }
    for (int q = 0; q < nw; q++) {
      P.sched[(size_t) l * nw + q] = (int) P.items.size();
      P.items.insert(P.items.end(), per[q].begin(), per[q].end());
    }
  }
  P.sched[(size_t) P.nlev * nw] = (int) P.items.size();
  P.sched_nw = nw; P.sched_vsc = vsc;
}

static int upload_buf(DevBuf &b, const void *src, size_t n, hipStream_t s)
{
  if (b.reserve(n + 64) != eslOK) return eslEMEM;
  if (n == 0) return eslOK;
  return hipMemcpyAsync(b.ptr, src, n, hipMemcpyHostToDevice, s) == hipSuccess ? eslOK : eslESYS;
}

/* Window results returned to the caller of infx_TrScanWindows. */
struct TrOut { float best; int mode; int64_t envi, envj; };

static int tr_scan_windows(int semi, const char *fname, CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx,
                           ESL_DSQ *dsq, int nwin, const int64_t *wi0, const int64_t *wj0,
                           float cutoff, int do_null3, float env_cutoff, int want_hits, int want_vsc,
                           std::vector<TrCand> &cands, std::vector<TrOut> &outs, std::vector<float> &vsc_all)
{
  int status;
  int fill_L, fill_R, fill_T, pty;
  if (!(cm->flags & CMH_BITS)) return infx_fail(eslEINCOMPAT, errbuf, "%s, CMH_BITS flag is not raised.\n", fname);
  if (dsq == NULL)              return infx_fail(eslEINCOMPAT, errbuf, "%s, dsq is NULL\n", fname);
  if (semi == TSEMI_CYK && (cm->search_opts & CM_SEARCH_INSIDE))   return infx_fail(eslEINCOMPAT, errbuf, "%s, CM_SEARCH_INSIDE flag raised", fname);
  if (semi != TSEMI_CYK && !(cm->search_opts & CM_SEARCH_INSIDE))  return infx_fail(eslEINCOMPAT, errbuf, "%s, CM_SEARCH_INSIDE flag not raised", fname);
  if (trsmx == NULL)            return infx_fail(eslEINCOMPAT, errbuf, "%s, trsmx == NULL\n", fname);
  if (semi == TSEMI_CYK && !trsmx->floats_valid) return infx_fail(eslEINCOMPAT, errbuf, "%s, trsmx->floats_valid if FALSE", fname);
  if (semi != TSEMI_CYK && !trsmx->ints_valid)   return infx_fail(eslEINCOMPAT, errbuf, "%s, trsmx->ints_valid if FALSE", fname);
  if (qdbidx < 0 || qdbidx > 2) return infx_fail(eslEINCOMPAT, errbuf, "%s, qdbidx is invalid", fname);
  if (cm_TrFillFromPassIdx(pass_idx, &fill_L, &fill_R, &fill_T) != eslOK) return infx_fail(eslEINVAL, errbuf, "%s(), unexpected pass idx: %d", fname, pass_idx);
  if ((pty = cm_tr_penalties_IdxForPass(pass_idx)) == -1) return infx_fail(eslEINCOMPAT, errbuf, "%s(), unexpected pass idx: %d", fname, pass_idx);
  if (cm->trp == NULL) return infx_fail(eslEINCOMPAT, errbuf, "%s, the CM has no truncation penalties (configure with CM_CONFIG_TRUNC)", fname);
  int64_t maxlen = 0, total = 0;
  for (int w = 0; w < nwin; w++) {
    if (wj0[w] < wi0[w]) return infx_fail(eslEINCOMPAT, errbuf, "%s, i0: %lld j0: %lldd\n", fname, (long long) wi0[w], (long long) wj0[w]);
    if (wi0[w] < 1) return infx_fail(eslEINCOMPAT, errbuf, "%s, i0 must be >= 1", fname);
    if (wj0[w] - wi0[w] + 1 > INT32_MAX / 2) return infx_fail(eslERANGE, errbuf, "%s, window too long", fname);
    if (wj0[w] > (int64_t) INT32_MAX - 1)   /* hit candidates carry 32-bit positions */
      return infx_fail(eslERANGE, errbuf, "%s, j0 = %lld: positions past %d do not fit the hit candidates", fname, (long long) wj0[w], INT32_MAX - 1);
    for (int64_t x = wi0[w]; x <= wj0[w]; x++)
      if (dsq[x] > 15) return infx_fail(eslEINVAL, errbuf, "%s, residue code %d at %lld is not a residue", fname, (int) dsq[x], (long long) x);
    maxlen = ESL_MAX(maxlen, wj0[w] - wi0[w] + 1);
    total += wj0[w] - wi0[w] + 1;
  }
  /* profile (host only: its errors are reported with or without a device) */
  TrProfile fresh;
  if ((status = build_tr_profile(cm, trsmx, qdbidx, semi, pty, fresh, errbuf)) != eslOK) return status;

  if ((status = infx_dev_init(errbuf)) != eslOK) return status;
  hipStream_t stream = infx_stream();
  if (!trsmx->dev) trsmx->dev = new TrDev();
  TrDev *D = (TrDev *) trsmx->dev;
  TrProfile *P = NULL;
  for (auto *p : D->prof) if (p->qdbidx == qdbidx && p->semi == semi && p->pty == pty) P = p;
  if (!P) { P = new TrProfile(); D->prof.push_back(P); }
  bool same = P->uploaded && P->st.size() == fresh.st.size() &&
              memcmp(P->st.data(), fresh.st.data(), fresh.st.size() * sizeof(TState)) == 0 && P->esc == fresh.esc &&
              P->rtmpl == fresh.rtmpl && P->dtmpl == fresh.dtmpl && P->pen.size() == fresh.pen.size() &&
              (fresh.pen.empty() || memcmp(P->pen.data(), fresh.pen.data(), fresh.pen.size() * sizeof(TrPen)) == 0) &&
              P->root_dx == fresh.root_dx && P->el == fresh.el && P->local == fresh.local;
  if (!same) {
    P->st = fresh.st; P->esc = fresh.esc; P->rtmpl = fresh.rtmpl; P->dtmpl = fresh.dtmpl; P->pen = fresh.pen; P->pblk = fresh.pblk;
    P->R = fresh.R; P->RB = fresh.RB; P->RT = fresh.RT; P->root_dx = fresh.root_dx; P->el = fresh.el;
    P->qdbidx = qdbidx; P->semi = semi; P->pty = pty; P->local = fresh.local;
    P->nlev = fresh.nlev; P->sched_nw = 0; P->sched_vsc = -1;   /* the schedule follows the profile */
    if (upload_buf(P->d_st, P->st.data(), P->st.size() * sizeof(TState), stream) != eslOK ||
        upload_buf(P->d_esc, P->esc.data(), P->esc.size() * 4, stream) != eslOK ||
        upload_buf(P->d_rtmpl, P->rtmpl.data(), P->rtmpl.size() * 4, stream) != eslOK ||
        upload_buf(P->d_dtmpl, P->dtmpl.data(), P->dtmpl.size() * 4, stream) != eslOK ||
        upload_buf(P->d_pen, P->pen.data(), P->pen.size() * sizeof(TrPen), stream) != eslOK ||
        upload_buf(P->d_pblk, P->pblk.data(), P->pblk.size() * sizeof(int), stream) != eslOK)
      return infx_fail(eslESYS, errbuf, "%s: device upload failed", fname);
    P->uploaded = true;
  }
  if (semi == TSEMI_IINSIDE && D->d_lsum.cap == 0)
    if (upload_buf(D->d_lsum, infx_ilogsum_table(), LOGSUM_TBL * 4, stream) != eslOK) return infx_fail(eslESYS, errbuf, "upload failed");

  /* windows: residues copied back to back, the longest first (each
   * wavefront scans one window at a time, in this order) */
  std::vector<int> order(nwin);
  for (int w = 0; w < nwin; w++) order[w] = w;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return wj0[a] - wi0[a] > wj0[b] - wi0[b]; });
  std::vector<TrWindow> hw(nwin);
  std::vector<uint8_t> packed;
  packed.reserve((size_t) total);
  for (int k = 0; k < nwin; k++) {
    const int w = order[k];
    int64_t off = (int64_t) packed.size();
    packed.insert(packed.end(), dsq + wi0[w], dsq + wj0[w] + 1);
    hw[k].i0 = wi0[w]; hw[k].j0 = wj0[w]; hw[k].dsq_off = off - wi0[w]; hw[k].id = w; hw[k].pad = 0;
  }
  const int W = trsmx->W;
  /* wavefronts per window: eight when the windows cannot fill the chip
   * with four, else four (fewer barriers per row) */
  int nw = (nwin * 4 < infx_num_cu() * 8) ? 8 : 4;
  if (infx_wpb_override()) nw = ESL_MIN(infx_wpb_override(), 8);
  if (P->sched_nw != nw || P->sched_vsc != (want_vsc ? 1 : 0)) {
    build_tr_schedule(cm, *P, nw, want_vsc ? 1 : 0, fill_L, fill_R, fill_T);
    if (upload_buf(P->d_sched, P->sched.data(), P->sched.size() * sizeof(int), stream) != eslOK ||
        upload_buf(P->d_items, P->items.data(), ESL_MAX((size_t) 1, P->items.size()) * sizeof(TrItem), stream) != eslOK)
      return infx_fail(eslESYS, errbuf, "%s: schedule upload failed", fname);
  }
  /* slot: J, L, R rows [2][R] each | T rows [2][RT] | J, L, R deck rings
   * [RB] each | root history [4][2][W+1] | vsc [M]   (4-byte cells) */
  const size_t slot_bytes = (((size_t) 6 * P->R + 2 * (size_t) P->RT + 3 * (size_t) P->RB + 8 * (size_t) (W + 1) + cm->M) * 4 + 255) / 256 * 256;
  size_t mem_free = 0, mem_total = 0;
  if (hipMemGetInfo(&mem_free, &mem_total) != hipSuccess) { (void) hipGetLastError(); mem_free = (size_t) 16 << 30; }
  const size_t budget = ESL_MIN((size_t) 96 << 30, mem_free > ((size_t) 8 << 30) ? mem_free - ((size_t) 8 << 30) : mem_free / 2)
                        + D->d_scratch.cap;
  int nslots = (int) ESL_MIN((int64_t) nwin, (int64_t) (budget / slot_bytes));
  const int occ = ESL_MAX(1, infx_tr_occupancy(semi, want_vsc, nw));
  nslots = ESL_MIN(nslots, infx_num_cu() * occ);
  if (nslots < 1) return infx_fail(eslEMEM, errbuf, "%s: one DP slot needs %zu bytes", fname, slot_bytes);
  size_t wout_bytes = (size_t) nwin * (4 + 4 + 8 + 8) + (want_vsc ? (size_t) nwin * cm->M * 4 : 0) + 1024;
  if (upload_buf(D->d_dsq, packed.data(), packed.size(), stream) != eslOK ||
      upload_buf(D->d_win, hw.data(), hw.size() * sizeof(TrWindow), stream) != eslOK ||
      D->d_scratch.reserve(slot_bytes * nslots) != eslOK || D->d_wout.reserve(wout_bytes) != eslOK ||
      D->d_ncand.reserve(128) != eslOK)
    return infx_fail(eslEMEM, errbuf, "%s: device allocation failed", fname);
  /* candidate buffer: eight per residue (a cutoff near 0 keeps many cells of
   * each row), at most 8 GB; an overflow re-runs the launch once with the
   * exact size */
  int64_t cand_cap = ESL_MAX((int64_t) 1 << 16, ESL_MIN((int64_t) 8 * (int64_t) total, (int64_t) (((size_t) 8 << 30) / sizeof(TrCand))));
  cand_cap = ESL_MAX(cand_cap, (int64_t) (D->d_cand.cap / sizeof(TrCand)));
  if (D->d_cand.reserve((size_t) cand_cap * sizeof(TrCand)) != eslOK) return infx_fail(eslEMEM, errbuf, "%s: device allocation failed", fname);
  cand_cap = (int64_t) (D->d_cand.cap / sizeof(TrCand));

  TrScanArgs A;
  memset(&A, 0, sizeof(A));
  A.st = (const TState *) P->d_st.ptr; A.esc = P->d_esc.ptr; A.dsq = (const uint8_t *) D->d_dsq.ptr;
  A.row_tmpl = P->d_rtmpl.ptr; A.deck_tmpl = P->d_dtmpl.ptr; A.lsum = D->d_lsum.ptr;
  A.pen = (const TrPen *) P->d_pen.ptr; A.npen = (int) P->pen.size();
  A.pen_blk = (const int *) P->d_pblk.ptr; A.pen_nblk = (W + 63) / 64 + 1;
  A.items = (const TrItem *) P->d_items.ptr; A.sched = (const int *) P->d_sched.ptr; A.nlev = P->nlev; A.nw = nw;
  A.win = (const TrWindow *) D->d_win.ptr; A.nwin = nwin;
  A.scratch = (char *) D->d_scratch.ptr; A.slot_bytes = (int64_t) slot_bytes; A.nslots = nslots;
  A.M = cm->M; A.W = W; A.R = P->R; A.RB = P->RB; A.RT = P->RT; A.root_dx = P->root_dx;
  A.fill_L = fill_L; A.fill_R = fill_R; A.fill_T = fill_T;
  A.el_selfsc = P->el;
  const int greedy = !(cm->search_opts & CM_SEARCH_CMNOTGREEDY);
  /* greedy without NULL3: the kernel drops the candidates the greedy
   * reporter can never report (a shorter admitted d of the same 64-cell
   * block scores at least as high); NULL3 changes the scores the reporter
   * compares, so every candidate goes to the host */
  A.cand_mode = !want_hits ? 0 : (!greedy ? 2 : (do_null3 ? 0 : 1));
  A.pass_any = (pass_idx == PLI_PASS_STD_ANY || pass_idx == PLI_PASS_5P_AND_3P_ANY || pass_idx == PLI_PASS_HMM_ONLY_ANY) ? 1 : 0;
  A.local = (cm->flags & CMH_LOCAL_BEGIN) ? 1 : 0;
  A.cutoff = want_hits ? cutoff : INFINITY;
  A.env_cutoff = env_cutoff;
  A.cand = (TrCand *) D->d_cand.ptr; A.ncand = (unsigned long long *) D->d_ncand.ptr; A.cand_cap = cand_cap;
  A.next_win = (unsigned *) ((char *) D->d_ncand.ptr + 64);
  char *wo = (char *) D->d_wout.ptr;
  A.win_best = (float *) wo;
  A.win_mode = (int *) (wo + (size_t) nwin * 4);
  A.win_envi = (int64_t *) (wo + (((size_t) nwin * 8 + 7) / 8) * 8);
  A.win_envj = A.win_envi + nwin;
  A.win_vsc = (float *) (A.win_envj + nwin);

  dim3 block(64 * nw), grid(nslots);
  unsigned long long ncand = 0;
  float kms = 0; int launches = 0;
  struct Events {   /* kernel timing; destroyed on every return */
    hipEvent_t a = NULL, b = NULL;
    ~Events() { if (a) (void) hipEventDestroy(a); if (b) (void) hipEventDestroy(b); }
  } ev;
  if (hipEventCreate(&ev.a) != hipSuccess || hipEventCreate(&ev.b) != hipSuccess)
    return infx_fail(eslESYS, errbuf, "%s: event creation failed", fname);
  hipEvent_t e0 = ev.a, e1 = ev.b;
  for (int attempt = 0; attempt < 3; attempt++) {
    if (hipMemsetAsync(A.ncand, 0, 128, stream) != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: memset failed", fname);
    (void) hipEventRecord(e0, stream);   /* timing only */
    hipError_t e = infx_launch_tr(semi, want_vsc, grid, block, stream, &A);
    (void) hipEventRecord(e1, stream);
    if (e != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: kernel launch failed: %s", fname, hipGetErrorString(e));
    if (hipMemcpyAsync(&ncand, A.ncand, 8, hipMemcpyDeviceToHost, stream) != hipSuccess || hipStreamSynchronize(stream) != hipSuccess)
      return infx_fail(eslESYS, errbuf, "%s: kernel execution failed", fname);
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { (void) hipGetLastError(); ms = 0; }
    kms += ms; launches++;
    if ((int64_t) ncand <= cand_cap) break;
    if (D->d_cand.reserve((size_t) ncand * sizeof(TrCand) + 4096) != eslOK) return infx_fail(eslEMEM, errbuf, "%s: candidate buffer", fname);
    cand_cap = (int64_t) (D->d_cand.cap / sizeof(TrCand));
    A.cand = (TrCand *) D->d_cand.ptr; A.cand_cap = cand_cap;
  }
  if ((int64_t) ncand > cand_cap) return infx_fail(eslEMEM, errbuf, "%s: candidate buffer overflow", fname);
  cands.resize(ncand);
  std::vector<float> wb(nwin); std::vector<int> wm(nwin); std::vector<int64_t> wi(nwin), wj(nwin);
  if ((ncand && hipMemcpyAsync(cands.data(), A.cand, ncand * sizeof(TrCand), hipMemcpyDeviceToHost, stream) != hipSuccess) ||
      hipMemcpyAsync(wb.data(), A.win_best, nwin * 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipMemcpyAsync(wm.data(), A.win_mode, nwin * 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipMemcpyAsync(wi.data(), A.win_envi, nwin * 8, hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipMemcpyAsync(wj.data(), A.win_envj, nwin * 8, hipMemcpyDeviceToHost, stream) != hipSuccess)
    return infx_fail(eslESYS, errbuf, "%s: result download failed", fname);
  if (want_vsc) {
    vsc_all.resize((size_t) nwin * cm->M);
    if (hipMemcpyAsync(vsc_all.data(), A.win_vsc, (size_t) nwin * cm->M * 4, hipMemcpyDeviceToHost, stream) != hipSuccess)
      return infx_fail(eslESYS, errbuf, "%s: vsc download failed", fname);
  }
  if (hipStreamSynchronize(stream) != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: download failed", fname);
  outs.resize(nwin);
  for (int w = 0; w < nwin; w++) { outs[w].best = wb[w]; outs[w].mode = wm[w]; outs[w].envi = wi[w]; outs[w].envj = wj[w]; }
  double cells = 0;
  {
    CellRows cr;
    infx_cell_rows(cm, trsmx, qdbidx, 1, &cr);
    for (int w = 0; w < nwin; w++) cells += cr.cells(wj0[w] - wi0[w] + 1);
  }
  infx_timing_add(0, kms, (int64_t) cells, launches);
  return eslOK;
}

static int tr_scan_one(int semi, const char *fname, CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                       int64_t i0, int64_t j0, float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                       int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, char *ret_mode, float *ret_sc)
{
  auto t0 = std::chrono::steady_clock::now();
  infx_timing_reset();
  std::vector<TrCand> cands; std::vector<TrOut> outs; std::vector<float> vsc_all;
  int status = tr_scan_windows(semi, fname, cm, errbuf, trsmx, qdbidx, pass_idx, dsq, 1, &i0, &j0, cutoff, do_null3, env_cutoff,
                               hitlist != NULL, ret_vsc != NULL, cands, outs, vsc_all);
  if (status != eslOK) return status;
  if (ret_envi) *ret_envi = (outs[0].envi == INT64_MAX) ? -1 : outs[0].envi;
  if (ret_envj) *ret_envj = (outs[0].envj == INT64_MIN) ? -1 : outs[0].envj;
  if (ret_sc) *ret_sc = outs[0].best;
  if (ret_mode) *ret_mode = (char) outs[0].mode;
  if (ret_vsc) {
    float *v = (float *) malloc(sizeof(float) * cm->M);
    memcpy(v, vsc_all.data(), sizeof(float) * cm->M);
    v[0] = outs[0].best;
    *ret_vsc = v;
  }
  if (hitlist) {
    std::vector<INFX_CAND> c(cands.size());
    for (size_t k = 0; k < cands.size(); k++) { c[k].j = cands[k].j; c[k].d = cands[k].d; c[k].sc = cands[k].sc; c[k].r = cands[k].r; c[k].mode = cands[k].mode; }
    const int64_t L = j0 - i0 + 1;
    const int W = (trsmx->W > L) ? (int) L : trsmx->W;
    int greedy = !(cm->search_opts & CM_SEARCH_CMNOTGREEDY);
    if ((status = infx_resolve_hits(cm, errbuf, pass_idx, 1, c.data(), (int64_t) c.size(), dsq, i0, j0, W, cutoff, do_null3, greedy, hitlist)) != eslOK)
      return status;
  }
  double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  infx_timing_add(wall, 0, 0, 0);
  return eslOK;
}

/* The device profile RefTrCYKScan (semi 0) / RefITrInsideScan (semi 1)
 * would upload for pass pass_idx, for checks of the kernel's address
 * arithmetic on the host.  dims[0..6] = R, RB, RT, number of emission
 * entries, root band max, number of truncated begins, length of the
 * per-block begin lists; lohi[2v], lohi[2v+1] = the band of d cells state v
 * holds in its rows (lo > hi: none).  Returns the number of descriptors, or
 * -status on error. */
extern "C" int infx_debug_tr_profile(CM_t *cm, CM_TR_SCAN_MX *trsmx, int qdbidx, int semi, int pass_idx, void *out_st, int st_cap,
                                     void *out_pen, int pen_cap, int *out_pblk, int pblk_cap, int *lohi, int *dims, char *errbuf)
{
  int pty;
  if (!trsmx || qdbidx < 0 || qdbidx > 2 || (semi != TSEMI_CYK && semi != TSEMI_IINSIDE)) return -eslEINVAL;
  if ((pty = cm_tr_penalties_IdxForPass(pass_idx)) == -1) return -eslEINCOMPAT;
  if (cm->trp == NULL) return -eslEINCOMPAT;
  TrProfile P;
  std::vector<int> lo, hi;
  int status = build_tr_profile(cm, trsmx, qdbidx, semi, pty, P, errbuf, &lo, &hi);
  if (status != eslOK) return -status;
  for (int i = 0; i < (int) P.st.size() && i < st_cap; i++) memcpy((char *) out_st + (size_t) i * sizeof(TState), &P.st[i], sizeof(TState));
  for (int i = 0; i < (int) P.pen.size() && i < pen_cap; i++) memcpy((char *) out_pen + (size_t) i * sizeof(TrPen), &P.pen[i], sizeof(TrPen));
  for (int i = 0; i < (int) P.pblk.size() && i < pblk_cap; i++) out_pblk[i] = P.pblk[i];
  for (int v = 0; v < cm->M; v++) { lohi[2 * v] = lo[v]; lohi[2 * v + 1] = hi[v]; }
  dims[0] = P.R; dims[1] = P.RB; dims[2] = P.RT; dims[3] = (int) P.esc.size(); dims[4] = P.root_dx;
  dims[5] = (int) P.pen.size(); dims[6] = (int) P.pblk.size();
  return (int) P.st.size();
}

/* The row schedule the truncated kernel runs for nw wavefronts per window
 * (per-state items if want_vsc): sched[nlev * nw + 1] item offsets and
 * items[] as (v, b0, b1) triples.  Returns the item count (or -status) and
 * the level count in *ret_nlev. */
extern "C" int infx_debug_tr_schedule(CM_t *cm, CM_TR_SCAN_MX *trsmx, int qdbidx, int semi, int pass_idx, int nw, int want_vsc,
                                      int *out_sched, int sched_cap, int *out_items, int items_cap, int *ret_nlev, char *errbuf)
{
  int pty, fL, fR, fT;
  if (!trsmx || qdbidx < 0 || qdbidx > 2 || (semi != TSEMI_CYK && semi != TSEMI_IINSIDE) || nw < 1 || nw > 8) return -eslEINVAL;
  if ((pty = cm_tr_penalties_IdxForPass(pass_idx)) == -1 || cm->trp == NULL) return -eslEINCOMPAT;
  if (cm_TrFillFromPassIdx(pass_idx, &fL, &fR, &fT) != eslOK) return -eslEINVAL;
  TrProfile P;
  int status = build_tr_profile(cm, trsmx, qdbidx, semi, pty, P, errbuf);
  if (status != eslOK) return -status;
  build_tr_schedule(cm, P, nw, want_vsc ? 1 : 0, fL, fR, fT);
  for (int i = 0; i < (int) P.sched.size() && i < sched_cap; i++) out_sched[i] = P.sched[i];
  for (int i = 0; i < (int) P.items.size() && 3 * i + 2 < items_cap; i++) {
    out_items[3 * i] = P.items[i].v; out_items[3 * i + 1] = P.items[i].b0; out_items[3 * i + 2] = P.items[i].b1;
  }
  *ret_nlev = P.nlev;
  return (int) P.items.size();
}

/* Attach truncated-scanner resources to trsmx without touching a device: a
 * TrDev holding nprof host-side profiles (descriptors and tables sized as for
 * the model, device buffers empty), as a scan on the GPU leaves behind.
 * Lets a CPU test free a CM whose truncated matrix carries them. */
extern "C" int infx_debug_attach_trdev(CM_t *cm, CM_TR_SCAN_MX *trsmx, int nprof)
{
  if (!cm || !trsmx || trsmx->dev || nprof < 0 || nprof > 8) return eslEINVAL;
  TrDev *D = new TrDev();
  for (int k = 0; k < nprof; k++) {
    TrProfile *P = new TrProfile();
    P->qdbidx = k % 3; P->semi = k % 2;
    P->st.resize(cm->M); P->esc.resize(16 * cm->M); P->rtmpl.resize(cm->M); P->dtmpl.resize(cm->M);
    P->pen.resize(cm->M); P->pblk.resize(cm->M + 1);
    D->prof.push_back(P);
  }
  trsmx->dev = D;
  return eslOK;
}

extern "C" int RefTrCYKScan(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                            int64_t i0, int64_t j0, float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                            int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, char *ret_mode, float *ret_sc)
{
  return tr_scan_one(TSEMI_CYK, "RefTrCYKScan", cm, errbuf, trsmx, qdbidx, pass_idx, dsq, i0, j0, cutoff, hitlist, do_null3,
                     env_cutoff, ret_envi, ret_envj, ret_vsc, ret_mode, ret_sc);
}

extern "C" int RefITrInsideScan(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                                int64_t i0, int64_t j0, float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                                int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, char *ret_mode, float *ret_sc)
{
  return tr_scan_one(TSEMI_IINSIDE, "RefITrInsideScan", cm, errbuf, trsmx, qdbidx, pass_idx, dsq, i0, j0, cutoff, hitlist, do_null3,
                     env_cutoff, ret_envi, ret_envj, ret_vsc, ret_mode, ret_sc);
}

/* Batched truncated scans: nwin windows [wi0[w], wj0[w]] of one target dsq
 * (windows may overlap), each scanned as an independent call of the scanner
 * named by <inside> (0 RefTrCYKScan, 1 RefITrInsideScan).  For window w,
 * hits are appended to hitlists[w] (if hitlists is non-NULL) and the best
 * score / mode / envelope are written to ret_sc[w], ret_mode[w], ret_envi[w],
 * ret_envj[w].  No counterpart in the reference, which makes one call per
 * window. */
extern "C" int infx_TrScanWindows(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int inside, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                                  int nwin, const int64_t *wi0, const int64_t *wj0, float cutoff, CM_TOPHITS **hitlists,
                                  int do_null3, float env_cutoff, float *ret_sc, char *ret_mode, int64_t *ret_envi, int64_t *ret_envj)
{
  auto t0 = std::chrono::steady_clock::now();
  infx_timing_reset();
  if (nwin < 1) return infx_fail(eslEINVAL, errbuf, "infx_TrScanWindows: no windows");
  int semi = inside ? TSEMI_IINSIDE : TSEMI_CYK;
  std::vector<TrCand> cands; std::vector<TrOut> outs; std::vector<float> vsc_all;
  int status = tr_scan_windows(semi, inside ? "RefITrInsideScan" : "RefTrCYKScan", cm, errbuf, trsmx, qdbidx, pass_idx, dsq, nwin, wi0, wj0,
                               cutoff, do_null3, env_cutoff, hitlists != NULL, 0, cands, outs, vsc_all);
  if (status != eslOK) return status;
  for (int w = 0; w < nwin; w++) {
    if (ret_sc) ret_sc[w] = outs[w].best;
    if (ret_mode) ret_mode[w] = (char) outs[w].mode;
    if (ret_envi) ret_envi[w] = (outs[w].envi == INT64_MAX) ? -1 : outs[w].envi;
    if (ret_envj) ret_envj[w] = (outs[w].envj == INT64_MIN) ? -1 : outs[w].envj;
  }
  if (hitlists) {
    std::vector<std::vector<INFX_CAND>> per(nwin);
    for (auto &c : cands) { INFX_CAND x; x.j = c.j; x.d = c.d; x.sc = c.sc; x.r = c.r; x.mode = c.mode; per[c.win].push_back(x); }
    int greedy = !(cm->search_opts & CM_SEARCH_CMNOTGREEDY);
    for (int w = 0; w < nwin; w++) {
      if (!hitlists[w]) continue;
      const int64_t L = wj0[w] - wi0[w] + 1;
      const int W = (trsmx->W > L) ? (int) L : trsmx->W;
      if ((status = infx_resolve_hits(cm, errbuf, pass_idx, 1, per[w].data(), (int64_t) per[w].size(), dsq, wi0[w], wj0[w], W, cutoff,
                                      do_null3, greedy, hitlists[w])) != eslOK)
        return status;
    }
  }
  double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  infx_timing_add(wall, 0, 0, 0);
  return eslOK;
}
