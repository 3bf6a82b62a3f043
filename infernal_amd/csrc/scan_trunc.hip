/* scan_trunc.hip -- gfx950 kernels of the truncated QDB scanners:
 *   RefTrCYKScan      src/cm_dpsearch_trunc.c:79-700    (max-plus, FP32)
 *   RefITrInsideScan  src/cm_dpsearch_trunc.c:741-1338  (ILogsum, int32 cells, FP32 temporaries)
 *
 * One workgroup owns one scan call (a target window i0..j0) and sweeps
 * j = i0..j0.  Within a row the states are visited level by level: a state's
 * level is one more than the highest level among the children whose cells of
 * the same row it reads, so the states of one level depend only on earlier
 * levels of this row and on the previous row.  The host splits each level's
 * work -- 64-cell blocks of d, lanes over d -- among the workgroup's
 * wavefronts (largest first onto the least loaded), and a barrier ends the
 * level.  A state's cells come out exactly as in the reference's v = M-1..1
 * sweep: the values a cell is made of are final when it is computed, and the
 * operations on them are the reference's, in its order.  The J, L, R (and T
 * for B states) cells of a block are computed together.
 *
 * BEGL decks are rings of per-state rows: a bifurcation reads BEGL cells of
 * rows j-k for k up to its BEGR band, so a ring of a power of two above that
 * many rows holds every row it can read (the reference keeps W+1 full decks).
 * A bifurcation block walks k over the union of its lanes' k ranges: for one
 * k every lane reads the same deck row and BEGR cell, and each lane folds in
 * only the k of its own range, in ascending k, the reference's order.
 *
 * The scan is not split into chunks: the reference never resets the root
 * cells X[cur][0][d] between positions, so whether a truncated begin at j is
 * taken depends on every earlier j of the same parity (src/cm_dpsearch_trunc.c
 * :538-605).  Truncated scans run on short terminal windows (the pipeline's
 * truncated passes use windows of length maxW), and many windows are batched
 * into one launch, one workgroup each.
 *
 * IL states are the only same-j chains (J and L modes read the state's own
 * d-1 cell); they are solved by the exact fixed-point sweep described in
 * scan_std.hip.  RefITrInsideScan keeps its sums in float temporaries; every
 * int result here passes through float exactly as there, with the x86
 * float -> int conversion (out of range -> INT_MIN).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "scan_kernels.h"

#define IMPF        (-1e36f)
#define NOTIMP_F    (-9.999e35f)
#define INFTY_I     987654321
#define LOGSUM_N    23000
#define KP          18

__device__ __forceinline__ void wave_sync_t()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
/* read-only tables through the constant address space: uniform indices
 * become scalar loads (one s_load per field group, cached in the K$) */
typedef const __attribute__((address_space(4))) TState CTState;
typedef const __attribute__((address_space(4))) TrPen CTrPen;
typedef const __attribute__((address_space(4))) int CInt;
__device__ __forceinline__ TrPen load_pen(const TrPen *base, int i)
{
  CTrPen *q = (CTrPen *) base + i;
  TrPen p;
  p.y = q->y; p.base = q->base; p.tbase = q->tbase; p.dx = q->dx; p.flags = q->flags; p.pen = q->pen;
  return p;
}
/* lane i <- x of lane i-1, lane 0 <- lane 0 of `first` (DPP wave_shr:1) */
__device__ __forceinline__ int shr1_t(int first, int x) { return __builtin_amdgcn_update_dpp(first, x, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int wadd(int a, int b) { return (int) ((unsigned) a + (unsigned) b); }
__device__ __forceinline__ int wmul(int a, int b) { return (int) ((unsigned) a * (unsigned) b); }
/* x86 cvttss2si: truncation toward zero, out of range or NaN -> INT_MIN */
__device__ __forceinline__ int f2i(float f) { return (f >= -2147483648.0f && f < 2147483648.0f) ? (int) f : INT_MIN; }

__device__ __forceinline__ int ilogsum(int s1, int s2, const int *tbl)
{
  const int mx = max(-INFTY_I, max(s1, s2));
  const int mn = min(s1, s2);
  const int diff = (int) ((unsigned) mx - (unsigned) mn);
  const bool use_max = (mn <= -INFTY_I) || (diff >= LOGSUM_N);
  const int t = tbl[use_max ? 0 : diff];
  return use_max ? mx : wadd(mx, t);
}

/* C: cell type; A: accumulator (temporary) type. */
template <int S> struct TR;

template <> struct TR<TSEMI_CYK> {
  typedef float C;
  typedef float A;
  static __device__ __forceinline__ C imp() { return IMPF; }
  static __device__ __forceinline__ A a_init(C x) { return x; }
  static __device__ __forceinline__ A acc(A a, C term, const int *) { return fmaxf(a, term); }
  static __device__ __forceinline__ C acc_direct(A a, C term, const int *) { return fmaxf(a, term); }
  static __device__ __forceinline__ C store(A a) { return a; }
  static __device__ __forceinline__ C store_emit(A a, C e) { return __fadd_rn(a, e); }
  static __device__ __forceinline__ C emit_only(C e) { return e; }
  static __device__ __forceinline__ C term(C cell, uint32_t t) { return __fadd_rn(cell, __uint_as_float(t)); }
  static __device__ __forceinline__ C add(C a, C b) { return __fadd_rn(a, b); }
  static __device__ __forceinline__ C inner_ls(C a, C b, const int *) { return fmaxf(a, b); }  /* ESL_MAX in TrCYK */
  static __device__ __forceinline__ C cmax(C a, C b) { return fmaxf(a, b); }
  static __device__ __forceinline__ C init(int ok, uint32_t el, uint32_t endsc, int x) {
    return ok ? __fadd_rn(__fmul_rn(__uint_as_float(el), (float) x), __uint_as_float(endsc)) : IMPF;
  }
  static __device__ __forceinline__ C bits(uint32_t u) { return __uint_as_float(u); }
  static __device__ __forceinline__ float score(C x) { return x; }
  static __device__ __forceinline__ bool same(C a, C b) { return __float_as_uint(a) == __float_as_uint(b); }
  static __device__ __forceinline__ int tob(C x) { return __float_as_int(x); }
  static __device__ __forceinline__ C fromb(int x) { return __int_as_float(x); }
};

template <> struct TR<TSEMI_IINSIDE> {
  typedef int   C;
  typedef float A;
  static __device__ __forceinline__ C imp() { return -INFTY_I; }
  static __device__ __forceinline__ A a_init(C x) { return (float) x; }
  static __device__ __forceinline__ A acc(A a, C term, const int *tbl) { return (float) ilogsum(f2i(a), term, tbl); }
  /* ILogsum(float temporary, int term) assigned straight to an int cell */
  static __device__ __forceinline__ C acc_direct(A a, C term, const int *tbl) { return ilogsum(f2i(a), term, tbl); }
  static __device__ __forceinline__ C store(A a) { return f2i(a); }
  static __device__ __forceinline__ C store_emit(A a, C e) { return f2i(__fadd_rn(a, (float) e)); }
  static __device__ __forceinline__ C emit_only(C e) { return f2i((float) e); }
  static __device__ __forceinline__ C term(C cell, uint32_t t) { return wadd(cell, (int) t); }
  static __device__ __forceinline__ C add(C a, C b) { return wadd(a, b); }
  static __device__ __forceinline__ C inner_ls(C a, C b, const int *tbl) { return ilogsum(a, b, tbl); }
  static __device__ __forceinline__ C cmax(C a, C b) { return max(a, b); }
  static __device__ __forceinline__ C init(int ok, uint32_t el, uint32_t endsc, int x) {
    return ok ? wadd(wmul((int) el, x), (int) endsc) : -INFTY_I;
  }
  static __device__ __forceinline__ C bits(uint32_t u) { return (int) u; }
  static __device__ __forceinline__ float score(C x) { return (x == -INFTY_I) ? IMPF : ((float) x / 1000.0f); }
  static __device__ __forceinline__ bool same(C a, C b) { return a == b; }
  static __device__ __forceinline__ int tob(C x) { return x; }
  static __device__ __forceinline__ C fromb(int x) { return x; }
};

/* Materialise a loaded value here (issued loads go out together and are
 * waited for once, not one round trip per use). */
template <typename T> __device__ __forceinline__ void hold_t(T &x) { asm volatile("" : "+v"(x)); }
#define TR_KU 8
#define TR_RU 8

/* the row being computed: buffers of the window's slot and the row's indices */
template <int S> struct TrRow {
  typedef typename TR<S>::C Cc;
  Cc *Jc, *Jp, *Lc, *Lp, *Rc, *Rp, *Tc;
  Cc *JB, *LB, *RB;          /* deck rings (all rows) */
  const uint8_t *dsq;
  int64_t j;
  int jp, We, xj;
};

/* blocks b0..b1-1 of state v at the row described by w (one wavefront) */
template <int S, bool VSC>
__device__ __forceinline__ void tr_state_blocks(const TrScanArgs &a, const TrRow<S> &w, int v, int b0, int b1, float *vscs,
                                                const int *tbl, int lane)
{
  typedef TR<S> R;
  typedef typename R::C Cc;
  typedef typename R::A Aa;
  const Cc *esc = (const Cc *) a.esc;
  const int fL = a.fill_L, fR = a.fill_R, fT = a.fill_T;
  Cc *Jc = w.Jc, *Jp = w.Jp, *Lc = w.Lc, *Rc = w.Rc, *Rp = w.Rp, *Tc = w.Tc;
  const uint8_t *dsq = w.dsq;
  const int64_t j = w.j;
  const int jp = w.jp, xj = w.xj;

  CTState *s = (CTState *) (a.st + v);   /* scalar loads: the descriptor is read-only */
  const int kind = s->kind;
  if (kind == KIND_SKIP) return;
  const int dx = min(min(jp, w.We), s->dx);
  if (1 > dx) return;
  const int cnum = s->cnum;
  const int em = s->emitmode, eo = s->esc_off;
  const uint32_t el = a.el_selfsc;
  const int iok = s->init_ok;
  Cc vmax = R::imp();
  const int bend = min(b1, (dx + 63) >> 6);   /* blocks that hold cells at this row */

  if (kind == KIND_B) {
    const int bw = s->child_base[0], by = s->child_base[1];
    const int wmask = s->child_base[2], wstride = s->child_base[3];
    const int banded = s->banded, kdy = s->kdx_y, kdw = s->kdx_w;
    const int64_t rowc = (int64_t) (jp & wmask) * wstride;   /* this row's deck ring row */
    for (int b = b0; b < bend; b++) {
      const int d0 = 64 * b + 1;
      const int d = d0 + lane;
      const bool act = d <= dx;
      const int kmin = banded ? max(0, d - kdw) : 0;
      const int kmax = banded ? min(kdy, d) : d;
      const int kn = max(1, kmin), kx = min(d - 1, kmax);
      const Cc ini = R::init(iok, el, s->endsc, d);
      Aa Js = R::a_init(ini), Ls = R::a_init(ini), Rs = R::a_init(ini), Ts = R::a_init(ini);
      /* the union of the block's k ranges: kmin grows and kmax grows with d */
      const int dl = min(d0 + 63, dx);
      const int kb0 = banded ? max(0, d0 - kdw) : 0;
      const int kb1 = banded ? min(kdy, dl) : dl;
      /* TR_KU values of k at a time: every load of the group is issued
       * before the first is used (the k steps depend on each other only
       * through the accumulators), then the group is folded in ascending k.
       * k past the union's end loads the last k's cells and is not used. */
      for (int kc = kb0; kc <= kb1; kc += TR_KU) {
        Cc jy[TR_KU], ly[TR_KU], jb[TR_KU], rb[TR_KU];
#pragma unroll
        for (int u = 0; u < TR_KU; u++) {
          const int k = min(kc + u, kb1);
          const bool use = (act && k >= kmin && k <= kmax) || (fT && act && k >= kn && k <= kx);
          const int64_t bo = (int64_t) ((jp - k) & wmask) * wstride + bw + (use ? d - k : 0);
          jy[u] = Jc[by + k];
          ly[u] = (fL || fT) ? Lc[by + k] : R::imp();
          jb[u] = w.JB[bo];
          rb[u] = (fR || fT) ? w.RB[bo] : R::imp();
        }
#pragma unroll
        for (int u = 0; u < TR_KU; u++) { hold_t(jy[u]); hold_t(ly[u]); hold_t(jb[u]); hold_t(rb[u]); }
#pragma unroll
        for (int u = 0; u < TR_KU; u++) {
          const int k = kc + u;
          if (k > kb1) break;
          const bool in = act && k >= kmin && k <= kmax;
          const bool inT = fT && act && k >= kn && k <= kx;
          if (in) {
            Js = R::acc(Js, R::add(jb[u], jy[u]), tbl);
            if (fL) Ls = R::acc(Ls, R::add(jb[u], ly[u]), tbl);
            if (fR) Rs = R::acc(Rs, R::add(rb[u], jy[u]), tbl);
          }
          if (inT) Ts = R::acc(Ts, R::add(rb[u], ly[u]), tbl);
        }
      }
      if (act) {
        const Cc jv = R::store(Js);
        Jc[s->base + d] = jv;
        if (VSC) vmax = R::cmax(vmax, jv);
        if (fT) { const Cc tv = R::store(Ts); Tc[s->tbase + d] = tv; if (VSC) vmax = R::cmax(vmax, tv); }
        if (fL) {
          const Cc lv = (kmin == 0) ? R::acc_direct(Ls, R::cmax(w.JB[rowc + bw + d], w.LB[rowc + bw + d]), tbl) : R::store(Ls);
          Lc[s->base + d] = lv;
          if (VSC) vmax = R::cmax(vmax, lv);
        }
        if (fR) {
          const Cc rv = (kmax == d) ? R::acc_direct(Rs, R::cmax(Jc[by + d], Rc[by + d]), tbl) : R::store(Rs);
          Rc[s->base + d] = rv;
          if (VSC) vmax = R::cmax(vmax, rv);
        }
      }
    }
  } else if (kind == KIND_BEGL) {
    const int64_t rowo = (int64_t) (jp & s->ring_mask) * s->ring_stride;
    for (int b = b0; b < bend; b++) {
      const int d = 64 * b + 1 + lane;
      if (d <= dx) {
        const Cc ini = R::init(iok, el, s->endsc, d);
        Aa Js = R::a_init(ini), Ls = R::a_init(ini), Rs = R::a_init(ini);
        for (int k = 0; k < cnum; k++) {
          const int cb = s->child_base[k];
          Js = R::acc(Js, R::term(Jc[cb + d], s->tsc[k]), tbl);
          if (fL) Ls = R::acc(Ls, R::term(Lc[cb + d], s->tsc[k]), tbl);
          if (fR) Rs = R::acc(Rs, R::term(Rc[cb + d], s->tsc[k]), tbl);
        }
        const Cc jv = R::store(Js);
        w.JB[rowo + s->base + d] = jv;
        if (VSC) vmax = R::cmax(vmax, jv);
        if (fL) { const Cc lv = R::store(Ls); w.LB[rowo + s->base + d] = lv; if (VSC) vmax = R::cmax(vmax, lv); }
        if (fR) { const Cc rv = R::store(Rs); w.RB[rowo + s->base + d] = rv; if (VSC) vmax = R::cmax(vmax, rv); }
      }
    }
  } else if (em == EMITLEFT) {
    if (!s->detached) {
      const int Ry0 = s->first_r;     /* 1 for IL (no self transition in R mode) */
      const bool chain = (kind == KIND_IL);
      /* d-1 = 0 cells: initial values of the IL chain (only IL bands hold a
       * d = 0 cell; other left emitters never read one).  An IL state is one
       * work item, all its blocks in order. */
      Cc carryJ = R::imp(), carryL = R::imp();
      if (chain) { carryJ = Jc[s->base + 0]; carryL = Lc[s->base + 0]; }
      for (int b = b0; b < bend; b++) {
        const int d = 64 * b + 1 + lane;
        const bool act = d <= dx;
        const int dd = act ? d : dx;
        const Cc e = esc[eo + dsq[j - dd + 1]];
        const Cc iniJ = R::init(iok, el, s->endsc, dd - 1);
        Cc jv, lv = R::imp();
        if (!chain) {
          Aa Js = R::a_init(iniJ), Ls = R::a_init(iniJ);
          for (int k = 0; k < cnum; k++) {
            const int cb = s->child_base[k];
            Js = R::acc(Js, R::term(Jc[cb + dd - 1], s->tsc[k]), tbl);
            if (fL) Ls = R::acc(Ls, R::term(Lc[cb + dd - 1], s->tsc[k]), tbl);
          }
          jv = R::store_emit(Js, e);
          if (fL) lv = (dd >= 2) ? R::store_emit(Ls, e) : R::emit_only(e);
        } else {
          /* J_d = f(J_{d-1}): the self term (k = 0) comes first, then k = 1.. */
          Cc ct[6];
#pragma unroll
          for (int k = 1; k < 6; k++) if (k < cnum) ct[k] = R::term(Jc[s->child_base[k] + dd - 1], s->tsc[k]);
          auto gJ = [&](Cc prevJ) -> Cc {
            Aa A = R::acc(R::a_init(iniJ), R::term(prevJ, s->tsc[0]), tbl);
            for (int k = 1; k < cnum; k++) A = R::acc(A, ct[k], tbl);
            return R::store_emit(A, e);
          };
          const int cJ = R::tob(carryJ);
          Cc x = gJ((lane == 0) ? carryJ : R::imp());
          for (int it = 0; it < 65; it++) {
            const Cc up = R::fromb(shr1_t(cJ, R::tob(x)));   /* lane 0: the carry */
            const Cc xn = gJ(up);
            const bool ch = act && !R::same(xn, x);
            x = xn;
            if (!__any(ch)) break;
          }
          jv = x;
          carryJ = R::fromb(__builtin_amdgcn_readlane(R::tob(x), 63));
          if (fL) {
            Cc cl[6];
#pragma unroll
            for (int k = 1; k < 6; k++) if (k < cnum) cl[k] = R::term(Lc[s->child_base[k] + dd - 1], s->tsc[k]);
            auto gL = [&](Cc prevL) -> Cc {
              Aa A = R::acc(R::a_init(iniJ), R::term(prevL, s->tsc[0]), tbl);
              for (int k = 1; k < cnum; k++) A = R::acc(A, cl[k], tbl);
              return (dd >= 2) ? R::store_emit(A, e) : R::emit_only(e);
            };
            const int cL = R::tob(carryL);
            Cc y = gL((lane == 0) ? carryL : R::imp());
            for (int it = 0; it < 65; it++) {
              const Cc up = R::fromb(shr1_t(cL, R::tob(y)));
              const Cc yn = gL(up);
              const bool ch = act && !R::same(yn, y);
              y = yn;
              if (!__any(ch)) break;
            }
            lv = y;
            carryL = R::fromb(__builtin_amdgcn_readlane(R::tob(y), 63));
          }
        }
        if (act) {
          Jc[s->base + d] = jv;
          if (VSC) vmax = R::cmax(vmax, jv);
          if (fL) { Lc[s->base + d] = lv; if (VSC) vmax = R::cmax(vmax, lv); }
          if (fR) {
            Aa Rs = R::a_init(R::init(iok, el, s->endsc, d));
            for (int k = Ry0; k < cnum; k++) {
              const int cb = s->child_base[k];
              Rs = R::acc(Rs, R::inner_ls(R::term(Jc[cb + d], s->tsc[k]), R::term(Rc[cb + d], s->tsc[k]), tbl), tbl);
            }
            const Cc rv = R::store(Rs);
            Rc[s->base + d] = rv;
            if (VSC) vmax = R::cmax(vmax, rv);
          }
        }
      }
    }
  } else if (em == EMITRIGHT) {
    if (!s->detached) {
      const int Ly0 = s->first_r;     /* 1 for IR (no self transition in L mode) */
      const Cc e = esc[eo + xj];
      for (int b = b0; b < bend; b++) {
        const int d = 64 * b + 1 + lane;
        if (d <= dx) {
          const Cc iniJ = R::init(iok, el, s->endsc, d - 1);
          Aa Js = R::a_init(iniJ), Rs = R::a_init(iniJ);
          for (int k = 0; k < cnum; k++) {
            const int cb = s->child_base[k];
            Js = R::acc(Js, R::term(Jp[cb + d - 1], s->tsc[k]), tbl);
            if (fR) Rs = R::acc(Rs, R::term(Rp[cb + d - 1], s->tsc[k]), tbl);
          }
          const Cc jv = R::store_emit(Js, e);
          Jc[s->base + d] = jv;
          if (VSC) vmax = R::cmax(vmax, jv);
          if (fR) { const Cc rv = (d >= 2) ? R::store_emit(Rs, e) : R::emit_only(e); Rc[s->base + d] = rv; if (VSC) vmax = R::cmax(vmax, rv); }
          if (fL) {
            Aa Ls = R::a_init(R::init(iok, el, s->endsc, d));
            for (int k = Ly0; k < cnum; k++) {
              const int cb = s->child_base[k];
              Ls = R::acc(Ls, R::inner_ls(R::term(Jc[cb + d], s->tsc[k]), R::term(Lc[cb + d], s->tsc[k]), tbl), tbl);
            }
            const Cc lv = R::store(Ls);
            Lc[s->base + d] = lv;
            if (VSC) vmax = R::cmax(vmax, lv);
          }
        }
      }
    }
  } else if (em == EMITPAIR) {
    const Cc re = esc[s->resc_off + xj];
    for (int b = b0; b < bend; b++) {
      const int d = 64 * b + 1 + lane;
      if (d <= dx) {
        const int xi = dsq[j - d + 1];
        Cc jv = R::imp();
        if (d >= 2) {
          Aa Js = R::a_init(R::init(iok, el, s->endsc, d - 2));
          for (int k = 0; k < cnum; k++) Js = R::acc(Js, R::term(Jp[s->child_base[k] + d - 2], s->tsc[k]), tbl);
          jv = R::store_emit(Js, esc[eo + xi * KP + xj]);
        } else if (S == TSEMI_IINSIDE) {
          jv = R::emit_only(R::imp());
        }
        Jc[s->base + d] = jv;
        if (VSC) vmax = R::cmax(vmax, jv);
        const Cc ini1 = R::init(iok, el, s->endsc, d - 1);
        if (fL) {
          Aa Ls = R::a_init(ini1);
          for (int k = 0; k < cnum; k++) {
            const int cb = s->child_base[k];
            Ls = R::acc(Ls, R::cmax(R::term(Jc[cb + d - 1], s->tsc[k]), R::term(Lc[cb + d - 1], s->tsc[k])), tbl);
          }
          const Cc le = esc[s->lesc_off + xi];
          const Cc lv = (d >= 2) ? R::store_emit(Ls, le) : R::emit_only(le);
          Lc[s->base + d] = lv;
          if (VSC) vmax = R::cmax(vmax, lv);
        }
        if (fR) {
          Aa Rs = R::a_init(ini1);
          for (int k = 0; k < cnum; k++) {
            const int cb = s->child_base[k];
            Rs = R::acc(Rs, R::cmax(R::term(Jp[cb + d - 1], s->tsc[k]), R::term(Rp[cb + d - 1], s->tsc[k])), tbl);
          }
          const Cc rv = (d >= 2) ? R::store_emit(Rs, re) : R::emit_only(re);
          Rc[s->base + d] = rv;
          if (VSC) vmax = R::cmax(vmax, rv);
        }
      }
    }
  } else { /* EMITNONE: D, S */
    for (int b = b0; b < bend; b++) {
      const int d = 64 * b + 1 + lane;
      if (d <= dx) {
        const Cc ini = R::init(iok, el, s->endsc, d);
        Aa Js = R::a_init(ini), Ls = R::a_init(ini), Rs = R::a_init(ini);
        for (int k = 0; k < cnum; k++) {
          const int cb = s->child_base[k];
          Js = R::acc(Js, R::term(Jc[cb + d], s->tsc[k]), tbl);
          if (fL) Ls = R::acc(Ls, R::term(Lc[cb + d], s->tsc[k]), tbl);
          if (fR) Rs = R::acc(Rs, R::term(Rc[cb + d], s->tsc[k]), tbl);
        }
        const Cc jv = R::store(Js);
        Jc[s->base + d] = jv;
        if (VSC) vmax = R::cmax(vmax, jv);
        if (fL) { const Cc lv = R::store(Ls); Lc[s->base + d] = lv; if (VSC) vmax = R::cmax(vmax, lv); }
        if (fR) { const Cc rv = R::store(Rs); Rc[s->base + d] = rv; if (VSC) vmax = R::cmax(vmax, rv); }
      }
    }
  }
  if (VSC) {
    /* with per-state maxima requested, a state is one work item (all its
     * blocks on one wavefront) */
    for (int o = 32; o >= 1; o >>= 1) vmax = R::cmax(vmax, __shfl_xor(vmax, o, 64));
    /* RefTrCYKScan keeps a running max; RefITrInsideScan stores
     * Scorify(max over the cells of this j), overwriting the previous
     * j (src/cm_dpsearch_trunc.c:1117-1135). */
    if (lane == 0) vscs[v] = (S == TSEMI_CYK) ? fmaxf(vscs[v], R::score(vmax)) : R::score(vmax);
  }
}

#define TR_MAX_WAVES 8

template <int S, bool VSC>
__global__ void __launch_bounds__(64 * TR_MAX_WAVES) tr_scan_kernel(TrScanArgs a)
{
  typedef TR<S> R;
  typedef typename R::C Cc;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const int nw = a.nw;                         /* == blockDim.x / 64 (the schedule's wave count) */
  const int slot = blockIdx.x;
  if (slot >= a.nslots) return;
  const int *tbl = (const int *) a.lsum;
  const Cc *rt = (const Cc *) a.row_tmpl;    /* [3][R]: J, L, R templates */
  const Cc *dt = (const Cc *) a.deck_tmpl;   /* [3][RB]: J, L, R deck rings */
  const int fL = a.fill_L, fR = a.fill_R, fT = a.fill_T;
  const int nthr = blockDim.x;

  __shared__ int s_win;
  __shared__ float s_best[TR_MAX_WAVES];
  __shared__ int s_mode[TR_MAX_WAVES], s_d[TR_MAX_WAVES];
  __shared__ long long s_bj[TR_MAX_WAVES], s_ei[TR_MAX_WAVES], s_ej[TR_MAX_WAVES];

  char *sbase = a.scratch + (size_t) slot * a.slot_bytes;
  const size_t RR = a.R;
  Cc *J0 = (Cc *) sbase;                      /* J rows [2][R] */
  Cc *L0 = J0 + 2 * RR;                       /* L rows */
  Cc *R0 = L0 + 2 * RR;                       /* R rows */
  Cc *T0 = R0 + 2 * RR;                       /* T rows [2][RT] */
  const int W = a.W, W1 = a.W + 1;
  Cc *JB = T0 + 2 * (size_t) a.RT;            /* deck rings [RB] each */
  Cc *LB = JB + (size_t) a.RB;
  Cc *RB = LB + (size_t) a.RB;
  Cc *H  = RB + (size_t) a.RB;                /* root history [4 modes][2][W+1] */
  float *vscs = (float *) (H + 8 * (size_t) W1);

  for (;;) {
    /* the next window, longest first; one thread takes it, all read it
     * after the barrier (the next write comes after this window's last
     * barrier, when every thread has read it) */
    if (threadIdx.x == 0) s_win = (int) atomicAdd(a.next_win, 1u);
    __syncthreads();
    const int nxt = s_win;
    if (nxt >= a.nwin) break;
    const TrWindow win = a.win[nxt];
    const int wdx = win.id;
    const int64_t i0 = win.i0, j0 = win.j0;
    const int Lw = (int) (j0 - i0 + 1);
    const int We = min(W, Lw);                  /* effective W of this window */

    for (int x = threadIdx.x; x < a.R; x += nthr) {
      const Cc tj = rt[x], tl = rt[a.R + x], tr = rt[2 * a.R + x];
      J0[x] = tj; J0[RR + x] = tj; L0[x] = tl; L0[RR + x] = tl; R0[x] = tr; R0[RR + x] = tr;
    }
    for (int x = threadIdx.x; x < 2 * a.RT; x += nthr) T0[x] = R::imp();
    for (int x = threadIdx.x; x < a.RB; x += nthr) { JB[x] = dt[x]; LB[x] = dt[a.RB + x]; RB[x] = dt[2 * a.RB + x]; }
    for (int x = threadIdx.x; x < 8 * W1; x += nthr) H[x] = R::imp();
    if (VSC) for (int v = threadIdx.x; v < a.M; v += nthr) vscs[v] = IMPF;
    __syncthreads();

    float best = IMPF; int best_mode = TRMODE_UNKNOWN_K; int64_t best_j = INT64_MAX; int best_d = INT32_MAX;
    int64_t envi = INT64_MAX, envj = INT64_MIN;
    TrRow<S> rw;
    rw.JB = JB; rw.LB = LB; rw.RB = RB;
    rw.dsq = a.dsq + win.dsq_off;   /* dsq[j] for j in i0..j0 */
    rw.We = We;

    for (int64_t j = i0; j <= j0; j++) {
      const int jp = (int) (j - i0 + 1);
      const int cur = (int) (j & 1), prv = cur ^ 1;
      rw.Jc = J0 + (size_t) cur * RR; rw.Jp = J0 + (size_t) prv * RR;
      rw.Lc = L0 + (size_t) cur * RR; rw.Lp = L0 + (size_t) prv * RR;
      rw.Rc = R0 + (size_t) cur * RR; rw.Rp = R0 + (size_t) prv * RR;
      rw.Tc = T0 + (size_t) cur * a.RT;
      rw.j = j; rw.jp = jp; rw.xj = rw.dsq[j];

      const CInt *sc = (const CInt *) a.sched;
      const TrItem *items = a.items;
      for (int lvl = 0; lvl < a.nlev; lvl++) {
        const int t0 = sc[lvl * nw + wid], t1 = sc[lvl * nw + wid + 1];
        for (int t = t0; t < t1; t++) {
          const CInt *ip = (const CInt *) (items + t);
          tr_state_blocks<S, VSC>(a, rw, ip[0], ip[1], ip[2], vscs, tbl, lane);
        }
        __syncthreads();
      } /* levels */

      /* truncated begins out of ROOT_S, y ascending, modes J, L, R, T; the
       * root's 64-cell blocks dealt round-robin to the wavefronts (a block's
       * history cells stay with one wavefront) */
      const int rdx = min(min(jp, We), a.root_dx);
      Cc *Jc = rw.Jc, *Lc = rw.Lc, *Rc = rw.Rc, *Tc = rw.Tc;
      Cc *HJ = H + (size_t) (0 * 2 + cur) * W1, *HL = H + (size_t) (1 * 2 + cur) * W1;
      Cc *HR = H + (size_t) (2 * 2 + cur) * W1, *HT = H + (size_t) (3 * 2 + cur) * W1;
      for (int blk = wid; 64 * blk + 1 <= rdx; blk += nw) {
        const int d = 64 * blk + 1 + lane;
        const bool act = d <= rdx;
        float bsc = IMPF; int bmode = TRMODE_UNKNOWN_K, br = 0, bfl = 0;
        if (act) {
          Cc hj = HJ[d], hl = HL[d], hr = HR[d], ht = HT[d];
          /* the begins whose band reaches this block, ascending */
          const CInt *pb = (const CInt *) a.pen_blk;
          const int t0 = pb[blk], t1 = pb[blk + 1];
          /* TR_RU begins at a time: their descriptors and cells are loaded
           * before the first is folded in, then they are folded in
           * ascending order (the reference's y order) */
          for (int tg = t0; tg < t1; tg += TR_RU) {
            TrPen pg[TR_RU];
            Cc cj[TR_RU], cl[TR_RU], cr[TR_RU], ct[TR_RU];
#pragma unroll
            for (int u = 0; u < TR_RU; u++) {
              pg[u] = load_pen(a.pen, pb[a.pen_nblk + 1 + min(tg + u, t1 - 1)]);
              const bool okd = d <= min(rdx, min(min(jp, We), pg[u].dx));
              const int dq = okd ? d : 1;        /* a lane past this begin's band reads its cell 1 */
              cj[u] = Jc[pg[u].base + dq];
              cl[u] = fL ? Lc[pg[u].base + dq] : R::imp();
              cr[u] = fR ? Rc[pg[u].base + dq] : R::imp();
              ct[u] = (fT && (pg[u].flags & 1)) ? Tc[pg[u].tbase + dq] : R::imp();
            }
#pragma unroll
            for (int u = 0; u < TR_RU; u++) { hold_t(cj[u]); hold_t(cl[u]); hold_t(cr[u]); hold_t(ct[u]); }
#pragma unroll
            for (int u = 0; u < TR_RU; u++) {
              if (tg + u >= t1) break;
              const TrPen &p = pg[u];
              if (d > min(rdx, min(min(jp, We), p.dx))) continue;
              const Cc pen = R::bits(p.pen);
              Cc sc2 = R::add(cj[u], pen);
              if (sc2 > hj) { hj = sc2; const float f = R::score(sc2); if (f > bsc) { bsc = f; bmode = TRMODE_J_K; br = p.y; bfl = p.flags; } }
              if (fL) { sc2 = R::add(cl[u], pen); if (sc2 > hl) { hl = sc2; const float f = R::score(sc2); if (f > bsc) { bsc = f; bmode = TRMODE_L_K; br = p.y; bfl = p.flags; } } }
              if (fR) { sc2 = R::add(cr[u], pen); if (sc2 > hr) { hr = sc2; const float f = R::score(sc2); if (f > bsc) { bsc = f; bmode = TRMODE_R_K; br = p.y; bfl = p.flags; } } }
              if (fT && (p.flags & 1)) { sc2 = R::add(ct[u], pen); if (sc2 > ht) { ht = sc2; const float f = R::score(sc2); if (f > bsc) { bsc = f; bmode = TRMODE_T_K; br = p.y; bfl = p.flags; } } }
            }
          }
          HJ[d] = hj; HL[d] = hl; HR[d] = hr; HT[d] = ht;
          if (bsc > best) { best = bsc; best_mode = bmode; best_j = j; best_d = d; }   /* a lane's first maximum in scan order */
          if (bsc >= a.env_cutoff) { envi = min(envi, j - d + 1); envj = max(envj, j); }
        }
        bool rep;
        if (a.cand_mode == 2) {
          rep = act && bsc > 0.f && bsc > NOTIMP_F;
        } else {
          const bool q = act && bsc >= a.cutoff && bsc > NOTIMP_F;
          if (a.cand_mode == 1) {
            /* src/cm_mx.c:7547 ReportHitsGreedily reports d at this j only if
             * its score beats every admitted d' < d (cm_hit_AllowTruncation,
             * src/cm_tophits.c:3048), so a d that an admitted shorter d of
             * its block beats can never be reported; the host replay
             * applies the rule across blocks */
            const int64_t st = j - d + 1;
            bool adm = a.pass_any || (st == i0 && j == j0);
            if (!adm) {
              if (a.local) adm = bmode == TRMODE_J_K || (bmode == TRMODE_L_K && j == j0) || (bmode == TRMODE_R_K && st == i0);
              else adm = (bmode == TRMODE_J_K && (bfl & 6) == 6) || (bmode == TRMODE_L_K && j == j0 && (bfl & 2)) ||
                         (bmode == TRMODE_R_K && st == i0 && (bfl & 4));
            }
            const bool qa = q && adm;
            float inc = qa ? bsc : -INFINITY;
            for (int o = 1; o < 64; o <<= 1) {
              const float u = __shfl_up(inc, o, 64);
              if (lane >= o) inc = fmaxf(inc, u);
            }
            float ex = __shfl_up(inc, 1, 64);
            if (lane == 0) ex = -INFINITY;
            rep = qa && bsc > ex;
          } else {
            rep = q;
          }
        }
        const unsigned long long m = __ballot(rep);
        if (m) {
          const int n = __popcll(m);
          unsigned long long basei = 0;
          if (lane == 0) basei = atomicAdd(a.ncand, (unsigned long long) n);
          basei = __shfl(basei, 0, 64);
          if (rep) {
            const unsigned long long idx = basei + __popcll(m & ((1ull << lane) - 1ull));
            if ((int64_t) idx < a.cand_cap) {
              TrCand cd; cd.win = wdx; cd.j = (int32_t) j; cd.d = d; cd.sc = bsc; cd.r = br; cd.mode = bmode;
              a.cand[idx] = cd;
            }
          }
        }
      }
    } /* j */

    /* best over the window: the first (j, d) in scan order with the maximum
     * score, reduced over the lanes, then over the wavefronts */
    for (int o = 32; o >= 1; o >>= 1) {
      const float ob = __shfl_xor(best, o, 64);
      const int om = __shfl_xor(best_mode, o, 64);
      const long long oj = __shfl_xor((long long) best_j, o, 64);
      const int od = __shfl_xor(best_d, o, 64);
      const bool take = (ob > best) || (ob == best && (oj < best_j || (oj == best_j && od < best_d)));
      if (take) { best = ob; best_mode = om; best_j = oj; best_d = od; }
      envi = min(envi, (int64_t) __shfl_xor((long long) envi, o, 64));
      envj = max(envj, (int64_t) __shfl_xor((long long) envj, o, 64));
    }
    if (lane == 0) { s_best[wid] = best; s_mode[wid] = best_mode; s_bj[wid] = best_j; s_d[wid] = best_d; s_ei[wid] = envi; s_ej[wid] = envj; }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int k = 1; k < nw; k++) {
        const bool take = (s_best[k] > best) || (s_best[k] == best && (s_bj[k] < best_j || (s_bj[k] == best_j && s_d[k] < best_d)));
        if (take) { best = s_best[k]; best_mode = s_mode[k]; best_j = s_bj[k]; best_d = s_d[k]; }
        envi = min(envi, (int64_t) s_ei[k]); envj = max(envj, (int64_t) s_ej[k]);
      }
      a.win_best[wdx] = best; a.win_mode[wdx] = best_mode;
      a.win_envi[wdx] = envi; a.win_envj[wdx] = envj;
    }
    if (VSC) for (int v = threadIdx.x; v < a.M; v += nthr) a.win_vsc[(size_t) wdx * a.M + v] = vscs[v];
    __syncthreads();
  } /* windows */
}

extern "C" hipError_t infx_launch_tr(int semi, int want_vsc, dim3 grid, dim3 block, hipStream_t stream, const TrScanArgs *a)
{
  if (block.x % 64 || block.x / 64 > TR_MAX_WAVES || (int) (block.x / 64) != a->nw) return hipErrorInvalidValue;
  switch (semi * 2 + (want_vsc ? 1 : 0)) {
  case 0: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_CYK, false>), grid, block, 0, stream, *a); break;
  case 1: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_CYK, true>), grid, block, 0, stream, *a); break;
  case 2: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_IINSIDE, false>), grid, block, 0, stream, *a); break;
  case 3: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_IINSIDE, true>), grid, block, 0, stream, *a); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

/* resident workgroups per CU of the truncated kernel at nw wavefronts each */
extern "C" int infx_tr_occupancy(int semi, int want_vsc, int nw)
{
  const void *f = NULL;
  switch (semi * 2 + (want_vsc ? 1 : 0)) {
  case 0: f = (const void *) tr_scan_kernel<TSEMI_CYK, false>; break;
  case 1: f = (const void *) tr_scan_kernel<TSEMI_CYK, true>; break;
  case 2: f = (const void *) tr_scan_kernel<TSEMI_IINSIDE, false>; break;
  case 3: f = (const void *) tr_scan_kernel<TSEMI_IINSIDE, true>; break;
  default: return 0;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * nw, 0) != hipSuccess) { (void) hipGetLastError(); return 0; }
  return nb;
}
