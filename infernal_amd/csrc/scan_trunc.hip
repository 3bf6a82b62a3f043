/* scan_trunc.hip -- gfx950 kernels of the truncated QDB scanners:
 *   RefTrCYKScan      src/cm_dpsearch_trunc.c:79-700    (max-plus, FP32)
 *   RefITrInsideScan  src/cm_dpsearch_trunc.c:741-1338  (ILogsum, int32 cells, FP32 temporaries)
 *
 * One wavefront owns one scan call (a target window i0..j0) and sweeps
 * j = i0..j0 and v = M-1..1 in the reference's order, 64 lanes over d.  The
 * J, L, R (and T for B states) cells of a state are computed together.
 *
 * The scan is not split into chunks: the reference never resets the root
 * cells X[cur][0][d] between positions, so whether a truncated begin at j is
 * taken depends on every earlier j of the same parity (src/cm_dpsearch_trunc.c
 * :538-605).  Truncated scans run on short terminal windows (the pipeline's
 * truncated passes use windows of length maxW), and many windows are batched
 * into one launch, one wavefront each.
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
  p.y = q->y; p.base = q->base; p.tbase = q->tbase; p.dx = q->dx; p.is_b = q->is_b; p.pen = q->pen;
  return p;
}
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
};

template <int S, bool VSC>
__global__ void __launch_bounds__(256) tr_scan_kernel(TrScanArgs a)
{
  typedef TR<S> R;
  typedef typename R::C Cc;
  typedef typename R::A Aa;
  const int lane = threadIdx.x & 63;
  const int slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (slot >= a.nslots) return;
  const int *tbl = (const int *) a.lsum;
  const Cc *esc = (const Cc *) a.esc;
  const Cc *rt = (const Cc *) a.row_tmpl;    /* [3][R]: J, L, R templates */
  const Cc *dt = (const Cc *) a.deck_tmpl;   /* [3][RB] */
  const int fL = a.fill_L, fR = a.fill_R, fT = a.fill_T;

  char *sbase = a.scratch + (size_t) slot * a.slot_bytes;
  const size_t RR = a.R;
  Cc *J0 = (Cc *) sbase;                      /* J rows [2][R] */
  Cc *L0 = J0 + 2 * RR;                       /* L rows */
  Cc *R0 = L0 + 2 * RR;                       /* R rows */
  Cc *T0 = R0 + 2 * RR;                       /* T rows [2][RT] */
  const int W = a.W, W1 = a.W + 1;
  Cc *JB = T0 + 2 * (size_t) a.RT;            /* decks [W+1][RB] */
  Cc *LB = JB + (size_t) W1 * a.RB;
  Cc *RB = LB + (size_t) W1 * a.RB;
  Cc *H  = RB + (size_t) W1 * a.RB;           /* root history [4 modes][2][W+1] */
  float *vscs = (float *) (H + 8 * (size_t) W1);

  for (;;) {
    /* the next window, longest first; every wavefront leaves once the
     * counter has passed the last window.  All lanes take part in the add
     * (lane 0 adds 1, the others 0) and the index is lane 0's result read
     * into a scalar: with no lane-dependent branch around the atomic the
     * index, and the loop exit, are wave-uniform by construction. */
    const int nxt = __builtin_amdgcn_readlane((int) atomicAdd(a.next_win, lane == 0 ? 1u : 0u), 0);
    if (nxt >= a.nwin) break;
    const TrWindow win = a.win[nxt];
    const int wdx = win.id;
    const uint8_t *dsq = a.dsq + win.dsq_off;   /* dsq[j] for j in i0..j0 */
    const int64_t i0 = win.i0, j0 = win.j0;
    const int Lw = (int) (j0 - i0 + 1);
    const int We = min(W, Lw);                  /* effective W of this window */


    for (int x = lane; x < a.R; x += 64) {
      const Cc tj = rt[x], tl = rt[a.R + x], tr = rt[2 * a.R + x];
      J0[x] = tj; J0[RR + x] = tj; L0[x] = tl; L0[RR + x] = tl; R0[x] = tr; R0[RR + x] = tr;
    }
    for (int x = lane; x < 2 * a.RT; x += 64) T0[x] = R::imp();
    for (int r = 0; r < W1; r++)
      for (int x = lane; x < a.RB; x += 64) {
        JB[(size_t) r * a.RB + x] = dt[x];
        LB[(size_t) r * a.RB + x] = dt[a.RB + x];
        RB[(size_t) r * a.RB + x] = dt[2 * a.RB + x];
      }
    for (int x = lane; x < 8 * W1; x += 64) H[x] = R::imp();
    if (VSC) for (int v = lane; v < a.M; v += 64) vscs[v] = IMPF;
    wave_sync_t();

    float best = IMPF; int best_mode = TRMODE_UNKNOWN_K; int64_t best_j = INT64_MAX; int best_d = INT32_MAX;
    int64_t envi = INT64_MAX, envj = INT64_MIN;

    for (int64_t j = i0; j <= j0; j++) {
      const int jp = (int) (j - i0 + 1);
      const int jmod = (int) (j % (We + 1));
      const int cur = (int) (j & 1), prv = cur ^ 1;
      Cc *Jc = J0 + (size_t) cur * RR, *Jp = J0 + (size_t) prv * RR;
      Cc *Lc = L0 + (size_t) cur * RR, *Lp = L0 + (size_t) prv * RR;
      Cc *Rc = R0 + (size_t) cur * RR, *Rp = R0 + (size_t) prv * RR;
      Cc *Tc = T0 + (size_t) cur * a.RT;
      Cc *JBc = JB + (size_t) jmod * a.RB, *LBc = LB + (size_t) jmod * a.RB, *RBc = RB + (size_t) jmod * a.RB;
      const int xj = dsq[j];

      for (int v = a.M - 1; v >= 1; v--) {
        CTState *s = (CTState *) (a.st + v);   /* scalar loads: the descriptor is read-only */
        const int kind = s->kind;
        if (kind == KIND_SKIP) continue;
        const int dx = min(min(jp, We), s->dx);
        if (1 > dx) continue;
        const int cnum = s->cnum;
        const int em = s->emitmode, eo = s->esc_off;
        const uint32_t el = a.el_selfsc;
        const int iok = s->init_ok;
        Cc vmax = R::imp();

        if (kind == KIND_B) {
          const int bw = s->child_base[0], by = s->child_base[1];
          for (int d0 = 1; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
            if (d <= dx) {
              int kmin, kmax;
              if (s->banded) { kmin = max(0, d - s->kdx_w); kmax = min(s->kdx_y, d); }
              else { kmin = 0; kmax = d; }
              const Cc ini = R::init(iok, el, s->endsc, d);
              Aa Js = R::a_init(ini), Ls = R::a_init(ini), Rs = R::a_init(ini), Ts = R::a_init(ini);
              int r = jmod - kmin; if (r < 0) r += We + 1;
              for (int k = kmin; k <= kmax; k++) {
                const size_t bo = (size_t) r * a.RB + bw + d - k;
                const Cc jb = JB[bo], jy = Jc[by + k];
                Js = R::acc(Js, R::add(jb, jy), tbl);
                if (fL) Ls = R::acc(Ls, R::add(jb, Lc[by + k]), tbl);
                if (fR) Rs = R::acc(Rs, R::add(RB[bo], jy), tbl);
                r = (r == 0) ? We : r - 1;
              }
              if (fT) {
                const int kn = max(1, kmin), kx = min(d - 1, kmax);
                r = jmod - kn; if (r < 0) r += We + 1;
                for (int k = kn; k <= kx; k++) {
                  Ts = R::acc(Ts, R::add(RB[(size_t) r * a.RB + bw + d - k], Lc[by + k]), tbl);
                  r = (r == 0) ? We : r - 1;
                }
              }
              const Cc jv = R::store(Js);
              Jc[s->base + d] = jv;
              if (VSC) vmax = R::cmax(vmax, jv);
              if (fT) { const Cc tv = R::store(Ts); Tc[s->tbase + d] = tv; if (VSC) vmax = R::cmax(vmax, tv); }
              if (fL) {
                const Cc lv = (kmin == 0) ? R::acc_direct(Ls, R::cmax(JBc[bw + d], LBc[bw + d]), tbl) : R::store(Ls);
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
          for (int d0 = 1; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
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
              JBc[s->base + d] = jv;
              if (VSC) vmax = R::cmax(vmax, jv);
              if (fL) { const Cc lv = R::store(Ls); LBc[s->base + d] = lv; if (VSC) vmax = R::cmax(vmax, lv); }
              if (fR) { const Cc rv = R::store(Rs); RBc[s->base + d] = rv; if (VSC) vmax = R::cmax(vmax, rv); }
            }
          }
        } else if (em == EMITLEFT) {
          if (!s->detached) {
            const int Ry0 = s->first_r;     /* 1 for IL (no self transition in R mode) */
            const bool chain = (kind == KIND_IL);
            /* d-1 = 0 cells: initial values of the IL chain (only IL bands
             * hold a d = 0 cell; other left emitters never read one) */
            Cc carryJ = R::imp(), carryL = R::imp();
            if (chain) { carryJ = Jc[s->base + 0]; carryL = Lc[s->base + 0]; }
            for (int d0 = 1; d0 <= dx; d0 += 64) {
              const int d = d0 + lane;
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
                Cc x = gJ((lane == 0) ? carryJ : R::imp());
                for (int it = 0; it < 65; it++) {
                  Cc up = __shfl_up(x, 1, 64);
                  const Cc xn = gJ((lane == 0) ? carryJ : up);
                  const bool ch = act && !R::same(xn, x);
                  x = xn;
                  if (!__any(ch)) break;
                }
                jv = x;
                carryJ = __shfl(x, 63, 64);
                if (fL) {
                  Cc cl[6];
#pragma unroll
                  for (int k = 1; k < 6; k++) if (k < cnum) cl[k] = R::term(Lc[s->child_base[k] + dd - 1], s->tsc[k]);
                  auto gL = [&](Cc prevL) -> Cc {
                    Aa A = R::acc(R::a_init(iniJ), R::term(prevL, s->tsc[0]), tbl);
                    for (int k = 1; k < cnum; k++) A = R::acc(A, cl[k], tbl);
                    return (dd >= 2) ? R::store_emit(A, e) : R::emit_only(e);
                  };
                  Cc y = gL((lane == 0) ? carryL : R::imp());
                  for (int it = 0; it < 65; it++) {
                    Cc up = __shfl_up(y, 1, 64);
                    const Cc yn = gL((lane == 0) ? carryL : up);
                    const bool ch = act && !R::same(yn, y);
                    y = yn;
                    if (!__any(ch)) break;
                  }
                  lv = y;
                  carryL = __shfl(y, 63, 64);
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
            for (int d0 = 1; d0 <= dx; d0 += 64) {
              const int d = d0 + lane;
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
          for (int d0 = 1; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
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
          for (int d0 = 1; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
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
          for (int o = 32; o >= 1; o >>= 1) vmax = R::cmax(vmax, __shfl_xor(vmax, o, 64));
          /* RefTrCYKScan keeps a running max; RefITrInsideScan stores
           * Scorify(max over the cells of this j), overwriting the previous
           * j (src/cm_dpsearch_trunc.c:1117-1135). */
          if (lane == 0) vscs[v] = (S == TSEMI_CYK) ? fmaxf(vscs[v], R::score(vmax)) : R::score(vmax);
        }
        wave_sync_t();
      } /* states */

      /* truncated begins out of ROOT_S, y ascending, modes J, L, R, T */
      const int rdx = min(min(jp, We), a.root_dx);
      Cc *HJ = H + (size_t) (0 * 2 + cur) * W1, *HL = H + (size_t) (1 * 2 + cur) * W1;
      Cc *HR = H + (size_t) (2 * 2 + cur) * W1, *HT = H + (size_t) (3 * 2 + cur) * W1;
      for (int d0 = 1, blk = 0; d0 <= rdx; d0 += 64, blk++) {
        const int d = d0 + lane;
        const bool act = d <= rdx;
        float bsc = IMPF; int bmode = TRMODE_UNKNOWN_K, br = 0;
        if (act) {
          Cc hj = HJ[d], hl = HL[d], hr = HR[d], ht = HT[d];
          /* the begins whose band reaches this block, ascending */
          const CInt *pb = (const CInt *) a.pen_blk;
          const int t0 = pb[blk], t1 = pb[blk + 1];
          for (int t = t0; t < t1; t++) {
            const TrPen p = load_pen(a.pen, pb[a.pen_nblk + 1 + t]);
            if (d > min(rdx, min(min(jp, We), p.dx))) continue;
            const Cc pen = R::bits(p.pen);
            Cc sc = R::add(Jc[p.base + d], pen);
            if (sc > hj) { hj = sc; const float f = R::score(sc); if (f > bsc) { bsc = f; bmode = TRMODE_J_K; br = p.y; } }
            if (fL) { sc = R::add(Lc[p.base + d], pen); if (sc > hl) { hl = sc; const float f = R::score(sc); if (f > bsc) { bsc = f; bmode = TRMODE_L_K; br = p.y; } } }
            if (fR) { sc = R::add(Rc[p.base + d], pen); if (sc > hr) { hr = sc; const float f = R::score(sc); if (f > bsc) { bsc = f; bmode = TRMODE_R_K; br = p.y; } } }
            if (fT && p.is_b) { sc = R::add(Tc[p.tbase + d], pen); if (sc > ht) { ht = sc; const float f = R::score(sc); if (f > bsc) { bsc = f; bmode = TRMODE_T_K; br = p.y; } } }
          }
          HJ[d] = hj; HL[d] = hl; HR[d] = hr; HT[d] = ht;
          if (bsc > best) { best = bsc; best_mode = bmode; best_j = j; best_d = d; }
          if (bsc >= a.env_cutoff) { envi = min(envi, j - d + 1); envj = max(envj, j); }
        }
        const bool rep = act && ((a.cand_mode == 2) ? (bsc > 0.f && bsc > NOTIMP_F) : (bsc >= a.cutoff && bsc > NOTIMP_F));
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
      wave_sync_t();
    } /* j */

    /* best over the window: the first (j, d) in scan order with the maximum score */
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
    if (lane == 0) {
      a.win_best[wdx] = best; a.win_mode[wdx] = best_mode;
      a.win_envi[wdx] = envi; a.win_envj[wdx] = envj;
    }
    if (VSC) {
      wave_sync_t();
      for (int v = lane; v < a.M; v += 64) a.win_vsc[(size_t) wdx * a.M + v] = vscs[v];
    }
  } /* windows */
}

extern "C" hipError_t infx_launch_tr(int semi, int want_vsc, dim3 grid, dim3 block, hipStream_t stream, const TrScanArgs *a)
{
  switch (semi * 2 + (want_vsc ? 1 : 0)) {
  case 0: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_CYK, false>), grid, block, 0, stream, *a); break;
  case 1: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_CYK, true>), grid, block, 0, stream, *a); break;
  case 2: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_IINSIDE, false>), grid, block, 0, stream, *a); break;
  case 3: hipLaunchKernelGGL((tr_scan_kernel<TSEMI_IINSIDE, true>), grid, block, 0, stream, *a); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
