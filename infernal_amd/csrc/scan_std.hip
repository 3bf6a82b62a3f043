/* scan_std.hip -- gfx950 kernels of the standard QDB scanners:
 *   FastCYKScan      src/cm_dpsearch.c:90-681     (max-plus, FP32)
 *   FastFInsideScan  src/cm_dpsearch.c:1731-2320  (FLogsum, FP32)
 *   FastIInsideScan  src/cm_dpsearch.c:1106-1690  (ILogsum, int32)
 *
 * One wavefront (64 lanes) owns one sequence chunk of one scan: it sweeps
 * j = start..end sequentially, and for each j the states v = M-1..1 in the
 * reference's order; the 64 lanes cover 64 consecutive d of a state's QDB
 * band.  Chunks overlap by W-1 residues so that every cell (v, j, d) of a
 * reported j is computed from the same inputs, in the same order, as the
 * reference computes it (a cell only depends on residues j-d+1..j).
 *
 * Latency plan (each state of each j is one short dependent step of the
 * wavefront, so the step must cost about one memory round trip):
 *  - state descriptors are read with scalar loads (constant address space);
 *  - all child cells of a 64-cell block, the block's left residues and the
 *    state's emission row are loaded together, then combined in registers;
 *    emission scores are picked from the row with ds_bpermute (lane x of a
 *    register holds the score of residue code x);
 *  - the local-begin accumulators and the log-sum table live in LDS;
 *  - the IL self-loop sweep shifts with a DPP wave_shr:1;
 *  - the bifurcation k-loop runs all lanes on the same k (one BEGL deck row,
 *    consecutive columns), broadcasting BEGR(j,k) with v_readlane.
 *
 * Arithmetic follows the reference operation for operation:
 *  - every child term is alpha + tsc rounded to FP32 (no contraction), and
 *    the terms of a cell are combined in the reference's order for each
 *    child count (that order matters for FLogsum/ILogsum; for max it does not);
 *  - ESL_MAX(a,b) is (a > b ? a : b); v_max_f32 returns the same value
 *    because no operand is NaN and no DP value is -0.0 (a -0.0 needs
 *    (-0) + (-0) under round-to-nearest, and no table holds -0.0);
 *  - FLogsum/ILogsum read the host-built tables (src/logsum.c:74-169);
 *  - the IL self-loop (cfirst[v] == v, same j, d-1) is solved by an exact
 *    fixed-point sweep: all lanes evaluate the cell from their left
 *    neighbour's current value until no lane changes.  A fixed point satisfies
 *    the reference's forward recurrence, and the forward recurrence has one
 *    solution, so the result is bit-identical; lane k is final after k+1
 *    sweeps, so at most 65 sweeps are needed;
 *  - each lane of a bifurcation visits its own k = kmin(d)..kmax(d) in
 *    ascending order (src/cm_dpsearch.c:256-283); lanes outside their range
 *    keep their value.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "scan_kernels.h"

#define IMPF        (-1e36f)
#define NOTIMP_F    (-9.999e35f)
#define INFTY_I     987654321
#define LOGSUM_N    INFX_LOGSUM_N
#define KP          18
#define WPB_MAX     16

typedef const __attribute__((address_space(4))) uint32_t CU32;

/* state descriptor through the scalar cache (uniform address, read-only) */
__device__ __forceinline__ DState load_state(const DState *p)
{
  DState s;
  uint32_t *dst = (uint32_t *) &s;
  CU32 *src = (CU32 *) p;
#pragma unroll
  for (int i = 0; i < (int) (sizeof(DState) / 4); i++) dst[i] = src[i];
  return s;
}

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wrap_add(int a, int b) { return (int) ((unsigned) a + (unsigned) b); }
__device__ __forceinline__ int wrap_sub(int a, int b) { return (int) ((unsigned) a - (unsigned) b); }
__device__ __forceinline__ int wrap_mul(int a, int b) { return (int) ((unsigned) a * (unsigned) b); }

/* lane i <- x of lane i-1, lane 0 <- lane 0 of `first` (DPP wave_shr:1) */
__device__ __forceinline__ int shr1(int first, int x) { return __builtin_amdgcn_update_dpp(first, x, 0x138, 0xf, 0xf, false); }

/* ------------------------------------------------------------------------ */
/* semirings                                                                 */

template <int S> struct SR;

template <> struct SR<SEMI_CYK> {
  typedef float T;
  static __device__ __forceinline__ T imp() { return IMPF; }
  static __device__ __forceinline__ T neg_sentinel() { return -INFINITY; }
  static __device__ __forceinline__ T op(T a, T b, const T *) { return fmaxf(a, b); }
  static __device__ __forceinline__ T add(T a, T b) { return __fadd_rn(a, b); }
  static __device__ __forceinline__ T bits(uint32_t u) { return __uint_as_float(u); }
  static __device__ __forceinline__ int tob(T x) { return __float_as_int(x); }
  static __device__ __forceinline__ T fromb(int x) { return __int_as_float(x); }
  static __device__ __forceinline__ T clamp_root(T x) { return fmaxf(x, IMPF); }
  static __device__ __forceinline__ float score(T x) { return x; }
  static __device__ __forceinline__ float envval(T x) { return x; }
  static __device__ __forceinline__ T init(int ok, uint32_t el, uint32_t endsc, int x) {
    return ok ? __fadd_rn(__fmul_rn(__uint_as_float(el), (float) x), __uint_as_float(endsc)) : IMPF;
  }
  static __device__ __forceinline__ bool same(T a, T b) { return __float_as_uint(a) == __float_as_uint(b); }
};

template <> struct SR<SEMI_FINSIDE> {
  typedef float T;
  static __device__ __forceinline__ T imp() { return IMPF; }
  static __device__ __forceinline__ T neg_sentinel() { return -INFINITY; }
  /* src/logsum.c:160 FLogsum */
  static __device__ __forceinline__ T op(T a, T b, const T *tbl) {
    const float mx = fmaxf(a, b), mn = fminf(a, b);
    const float diff = __fsub_rn(mx, mn);
    const bool  big  = !(diff < 23.f);          /* (max-min) >= 23, or min == -inf */
    const int   idx  = big ? 0 : (int) __fmul_rn(diff, 1000.0f);
    const float t    = tbl[idx];
    return big ? mx : __fadd_rn(mx, t);
  }
  static __device__ __forceinline__ T add(T a, T b) { return __fadd_rn(a, b); }
  static __device__ __forceinline__ T bits(uint32_t u) { return __uint_as_float(u); }
  static __device__ __forceinline__ int tob(T x) { return __float_as_int(x); }
  static __device__ __forceinline__ T fromb(int x) { return __int_as_float(x); }
  static __device__ __forceinline__ T clamp_root(T x) { return fmaxf(x, IMPF); }
  static __device__ __forceinline__ float score(T x) { return x; }
  static __device__ __forceinline__ float envval(T x) { return x; }
  static __device__ __forceinline__ T init(int ok, uint32_t el, uint32_t endsc, int x) {
    return ok ? __fadd_rn(__fmul_rn(__uint_as_float(el), (float) x), __uint_as_float(endsc)) : IMPF;
  }
  static __device__ __forceinline__ bool same(T a, T b) { return __float_as_uint(a) == __float_as_uint(b); }
};

template <> struct SR<SEMI_IINSIDE> {
  typedef int T;
  static __device__ __forceinline__ T imp() { return -INFTY_I; }
  static __device__ __forceinline__ T neg_sentinel() { return INT_MIN; }
  /* src/logsum.c:87 ILogsum */
  static __device__ __forceinline__ T op(T s1, T s2, const T *tbl) {
    const int mx = max(-INFTY_I, max(s1, s2));
    const int mn = min(s1, s2);
    const int diff = wrap_sub(mx, mn);
    const bool use_max = (mn <= -INFTY_I) || (diff >= LOGSUM_N);
    const int t = tbl[use_max ? 0 : diff];
    return use_max ? mx : wrap_add(mx, t);
  }
  static __device__ __forceinline__ T add(T a, T b) { return wrap_add(a, b); }
  static __device__ __forceinline__ T bits(uint32_t u) { return (int) u; }
  static __device__ __forceinline__ int tob(T x) { return x; }
  static __device__ __forceinline__ T fromb(int x) { return x; }
  static __device__ __forceinline__ T clamp_root(T x) { return max(x, -INFTY_I); }
  /* src/cm.c:4225 Scorify */
  static __device__ __forceinline__ float score(T x) { return (x == -INFTY_I) ? IMPF : ((float) x / 1000.0f); }
  static __device__ __forceinline__ float envval(T x) { return (float) x; }
  static __device__ __forceinline__ T init(int ok, uint32_t el, uint32_t endsc, int x) {
    return ok ? wrap_add(wrap_mul((int) el, x), (int) endsc) : -INFTY_I;
  }
  static __device__ __forceinline__ bool same(T a, T b) { return a == b; }
};

/* Combine child terms c[0..CN-1] and the init term in the reference's
 * order for each child count (src/cm_dpsearch.c:289-427 and :446-520;
 * identical in the FInside and IInside twins).  `has_c0` false leaves the
 * c[0] term out (the IL self term, applied by the caller). */
template <int S, int CN>
__device__ __forceinline__ typename SR<S>::T combine_children(const typename SR<S>::T *c, typename SR<S>::T ini,
                                                              bool has_c0, const typename SR<S>::T *tbl)
{
  typedef SR<S> R;
  typename R::T sc;
  if (CN == 2) {
    sc = R::op(c[1], ini, tbl);
  } else if (CN == 3) {
    sc = R::op(c[2], c[1], tbl);
    sc = R::op(sc, ini, tbl);
  } else if (CN == 4) {
    sc = R::op(c[3], c[2], tbl);
    sc = R::op(sc, ini, tbl);
    sc = R::op(sc, c[1], tbl);
  } else if (CN == 5) {
    sc = R::op(c[4], c[3], tbl);
    sc = R::op(sc, ini, tbl);
    sc = R::op(sc, c[1], tbl);
    sc = R::op(sc, c[2], tbl);
  } else { /* 6 */
    sc = R::op(c[5], ini, tbl);
    sc = R::op(sc, c[4], tbl);
    sc = R::op(sc, c[3], tbl);
    sc = R::op(sc, c[2], tbl);
    sc = R::op(sc, c[1], tbl);
  }
  if (has_c0) sc = R::op(sc, c[0], tbl);
  return sc;
}

/* One 64-cell block of a non-B state with CN children: the combined child
 * terms of cell d (IL: without the self term).  BEGL decks combine
 * init, c0, c1, ... in child order (src/cm_dpsearch.c:430-444). */
template <int S, int CN, bool BEGL, bool IL>
__device__ __forceinline__ typename SR<S>::T block_cells(const typename SR<S>::T *crow, const int32_t *cb, const uint32_t *tsc,
                                                         int dy, typename SR<S>::T ini, const typename SR<S>::T *tbl)
{
  typedef SR<S> R;
  typename R::T c[6];
#pragma unroll
  for (int k = 0; k < CN; k++) c[k] = (IL && k == 0) ? R::imp() : R::add(crow[cb[k] + dy], R::bits(tsc[k]));
  if (BEGL) {
    typename R::T sc = ini;
#pragma unroll
    for (int k = 0; k < CN; k++) sc = R::op(sc, c[k], tbl);
    return sc;
  }
  return combine_children<S, CN>(c, ini, !IL, tbl);
}

template <int S, bool BEGL, bool IL>
__device__ __forceinline__ typename SR<S>::T block_dispatch(int cnum, const typename SR<S>::T *crow, const int32_t *cb, const uint32_t *tsc,
                                                            int dy, typename SR<S>::T ini, const typename SR<S>::T *tbl)
{
  switch (cnum) {
  case 1:  return block_cells<S, 1, true, false>(crow, cb, tsc, dy, ini, tbl);  /* BEGL only: other states have 2..6 children */
  case 2:  return block_cells<S, 2, BEGL, IL>(crow, cb, tsc, dy, ini, tbl);
  case 3:  return block_cells<S, 3, BEGL, IL>(crow, cb, tsc, dy, ini, tbl);
  case 4:  return block_cells<S, 4, BEGL, IL>(crow, cb, tsc, dy, ini, tbl);
  case 5:  return block_cells<S, 5, BEGL, IL>(crow, cb, tsc, dy, ini, tbl);
  default: return block_cells<S, 6, BEGL, IL>(crow, cb, tsc, dy, ini, tbl);
  }
}

/* emission score of residue code x from a register whose lane y holds the
 * score of code y (ds_bpermute crossbar, no LDS storage) */
template <int S>
__device__ __forceinline__ typename SR<S>::T pick_esc(int ev_bits, int x)
{
  return SR<S>::fromb(__builtin_amdgcn_ds_bpermute(x << 2, ev_bits));
}

/* ------------------------------------------------------------------------ */

template <int S, bool VSC>
__global__ void __launch_bounds__(64 * WPB_MAX) std_scan_kernel(ScanArgs a)
{
  typedef SR<S> R;
  typedef typename R::T T;
  extern __shared__ __align__(16) char lds_raw[];

  const int lane = threadIdx.x & 63;
  const int wib  = threadIdx.x >> 6;
  const int slot = blockIdx.x * (blockDim.x >> 6) + wib;

  /* log-sum table into LDS (Inside) */
  const T *tbl = (const T *) a.lsum;
  int lds_off = 0;
  if (S != SEMI_CYK) {
    T *lt = (T *) lds_raw;
    const T *gt = (const T *) a.lsum;
    for (int i = threadIdx.x; i < LOGSUM_N; i += blockDim.x) lt[i] = gt[i];
    __syncthreads();
    tbl = lt;
    lds_off = LOGSUM_N * 4;
  }
  if (slot >= a.nslots) return;

  T   *acc_sc = (T *) (lds_raw + lds_off + (size_t) wib * a.acc_n * 8);
  int *acc_r  = (int *) (acc_sc + a.acc_n);

  char *sbase = a.scratch + (size_t) slot * a.slot_bytes;
  T   *deck = (T *) (sbase + a.deck_off);
  T   *row0 = (T *) (sbase + a.row_off);
  T   *row1 = row0 + a.R;
  float *vscs = (float *) (sbase + a.vsc_off);
  const T *esc   = (const T *) a.esc;
  const T *rtmpl = (const T *) a.row_tmpl;
  const T *dtmpl = (const T *) a.deck_tmpl;
  const uint8_t *dsq = a.dsq;
  const int64_t dsq_off = a.dsq_off;
  const uint32_t el = a.el_selfsc;
  const int rdn = a.root_dn;

  if (VSC) for (int v = lane; v < a.M; v += 64) vscs[v] = IMPF;

  for (int c = slot; c < a.nchunks; c += a.nslots) {
    const int64_t ac = a.i0 + (int64_t) c * a.chunk_len;
    const int64_t bc = min(a.j0, ac + a.chunk_len - 1);
    const int64_t s0 = max(a.i0, ac - (int64_t) a.W + 1);

    for (int x = lane; x < a.R; x += 64) { T t = rtmpl[x]; row0[x] = t; row1[x] = t; }
    for (int x = lane; x < a.RB; x += 64) deck[x] = dtmpl[x];
    wave_sync();

    float   best = IMPF;
    int64_t envi = INT64_MAX, envj = INT64_MIN;

    for (int64_t j = s0; j <= bc; j++) {
      const int jp   = (int) (j - s0 + 1);
      T *cur = (j & 1) ? row1 : row0;
      T *prv = (j & 1) ? row0 : row1;
      const int xj = __builtin_amdgcn_readfirstlane((int) dsq[j - dsq_off]);
      const uint8_t *dsq_j = dsq + (j + 1 - dsq_off);   /* left residue of cell d: dsq_j[-d] */
      const int rdx = min(jp, a.root_dx);

      if (a.local_begin) {
        for (int d = rdn + lane; d <= rdx; d += 64) { acc_sc[d - rdn] = R::neg_sentinel(); acc_r[d - rdn] = 0; }
      }

      for (int v = a.M - 1; v >= 1; v--) {
        const DState s = load_state(a.st + v);
        const int kind = s.kind;
        if (kind == KIND_SKIP) continue;
        const int dn = s.dn;
        const int dx = min(jp, s.dx);
        if (dn > dx) continue;
        const int has_begin = a.local_begin && s.has_begin;
        const int bdn = max(rdn, dn), bdx = min(rdx, dx);
        const T bsc = R::bits(s.beginsc);
        float vmax = IMPF;

        if (kind == KIND_GENERIC || kind == KIND_BEGL) {
          const T *crow = s.rdelta ? prv : cur;
          T *out = (kind == KIND_BEGL) ? deck + (jp & s.ring_mask) * s.ring_stride : cur;
          const int sd = s.sd, em = s.emitmode;
          /* emission row: lane y <- score of residue code y (MP: left code y, right code xj) */
          int ev = 0;
          if (em != EMITNONE) {
            const int yy = lane < KP ? lane : 0;
            ev = R::tob(esc[s.esc_off + (em == EMITPAIR ? yy * KP + xj : yy)]);
          }
          const T eright = (em == EMITRIGHT) ? R::fromb(__builtin_amdgcn_readlane(ev, xj)) : R::imp();
          for (int d0 = dn; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
            const bool act = d <= dx;
            const int dd = act ? d : dx;
            const int dy = dd - sd;
            const int xl = (em == EMITLEFT || em == EMITPAIR) ? (int) dsq_j[-dd] : 0;
            const T ini = R::init(s.init_ok, el, s.endsc, dy);
            T sc;
            if (kind == KIND_BEGL) sc = block_dispatch<S, true, false>(s.cnum, crow, s.child_base, s.tsc, dy, ini, tbl);
            else                   sc = block_dispatch<S, false, false>(s.cnum, crow, s.child_base, s.tsc, dy, ini, tbl);
            if (em == EMITLEFT || em == EMITPAIR) sc = R::add(sc, pick_esc<S>(ev, xl));
            else if (em == EMITRIGHT)             sc = R::add(sc, eright);
            if (act) {
              out[s.base + d] = sc;
              if (has_begin && d >= bdn && d <= bdx) {
                const T cand = R::add(sc, bsc);
                if (cand >= acc_sc[d - rdn]) { acc_sc[d - rdn] = cand; acc_r[d - rdn] = v; }
              }
              if (VSC) vmax = fmaxf(vmax, R::score(sc));
            }
          }
        } else if (kind == KIND_IL) {
          /* x_d = op(pre_d, x_{d-1} + t0) + e_d, pre_d = the other terms */
          const T t0 = R::bits(s.tsc[0]);
          const int yy = lane < KP ? lane : 0;
          const int ev = R::tob(esc[s.esc_off + yy]);
          T carry = cur[s.base + dn - 1];
          for (int d0 = dn; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
            const bool act = d <= dx;
            const int dd = act ? d : dx;            /* inactive lanes mirror a valid cell */
            const int dy = dd - 1;
            const int xl = (int) dsq_j[-dd];
            const T ini = R::init(s.init_ok, el, s.endsc, dy);
            const T pre = block_dispatch<S, false, true>(s.cnum, cur, s.child_base, s.tsc, dy, ini, tbl);
            const T e   = pick_esc<S>(ev, xl);
            const int cb = R::tob(carry);
            T x = R::add(R::op(pre, R::add(lane == 0 ? carry : R::imp(), t0), tbl), e);
            for (int it = 0; it < 65; it++) {
              const T left = R::fromb(shr1(cb, R::tob(x)));
              const T xn = R::add(R::op(pre, R::add(left, t0), tbl), e);
              const bool changed = act && !R::same(xn, x);
              x = xn;
              if (!__any(changed)) break;
            }
            if (act) {
              cur[s.base + d] = x;
              if (has_begin && d >= bdn && d <= bdx) {
                const T cand = R::add(x, bsc);
                if (cand >= acc_sc[d - rdn]) { acc_sc[d - rdn] = cand; acc_r[d - rdn] = v; }
              }
              if (VSC) vmax = fmaxf(vmax, R::score(x));
            }
            carry = R::fromb(__builtin_amdgcn_readlane(R::tob(x), 63));
          }
        } else { /* KIND_B */
          const int by = s.child_base[1];
          const int rmask = s.ring_mask, rstride = s.ring_stride;
          const T *wbase = deck + s.child_base[0];
          for (int d0 = dn; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
            const bool act = d <= dx;
            const int dd = act ? d : dx;
            const int dl = min(d0 + 63, dx);
            int kmin, kmax, K0, K1;
            if (s.banded) {
              kmin = max(0, max(s.kdn_y, dd - s.kdx_w));
              kmax = min(s.kdx_y, dd - s.kdn_w);
              K0 = max(0, max(s.kdn_y, d0 - s.kdx_w));   /* both bounds grow with d */
              K1 = min(s.kdx_y, dl - s.kdn_w);
            } else { kmin = 0; kmax = dd; K0 = 0; K1 = dl; }
            T sc = R::init(s.init_ok, el, s.endsc, dd);
            for (int kb = K0; kb <= K1; kb += 64) {
              const int ke = min(K1, kb + 63);
              const int yb = R::tob(cur[by + min(kb + lane, ke)]);
              /* four k per step, loads first; k > ke is masked (its reads stay
               * inside the slot's guard cells) */
              for (int k = kb; k <= ke; k += 4) {
                T bl[4], yv[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                  const int kk = k + u;
                  yv[u] = R::fromb(__builtin_amdgcn_readlane(yb, min(kk - kb, 63)));
                  bl[u] = wbase[((jp - kk) & rmask) * rstride - kk + dd];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                  const int kk = k + u;
                  const T nsc = R::op(sc, R::add(bl[u], yv[u]), tbl);
                  sc = (kk <= ke && kk >= kmin && kk <= kmax) ? nsc : sc;
                }
              }
            }
            if (act) {
              cur[s.base + d] = sc;
              if (has_begin && d >= bdn && d <= bdx) {
                const T cand = R::add(sc, bsc);
                if (cand >= acc_sc[d - rdn]) { acc_sc[d - rdn] = cand; acc_r[d - rdn] = v; }
              }
              if (VSC) vmax = fmaxf(vmax, R::score(sc));
            }
          }
        }
        if (VSC) {
          for (int o = 32; o >= 1; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
          if (lane == 0 && j >= ac) vscs[v] = fmaxf(vscs[v], vmax);
        }
        wave_sync();
      } /* states */

      /* ROOT_S (src/cm_dpsearch.c:562-623) and hit candidates */
      if (j >= ac) {
        float carry_max = -INFINITY;   /* prefix max of reportable scores (record filter) */
        for (int d0 = rdn; d0 <= rdx; d0 += 64) {
          const int d = d0 + lane;
          const bool act = d <= rdx;
          float bs = IMPF;
          int   br = 0;
          if (act) {
            T al = R::clamp_root(R::add(cur[a.root_child_base[0] + d], R::bits(a.root_tsc[0])));
            for (int k = 1; k < a.root_cnum; k++)
              al = R::op(al, R::add(cur[a.root_child_base[k] + d], R::bits(a.root_tsc[k])), tbl);
            if (a.local_begin) {
              const T acv = acc_sc[d - rdn];
              if (acv > al) { al = acv; br = acc_r[d - rdn]; }
            }
            bs = R::score(al);
            best = fmaxf(best, bs);
            if (R::envval(al) >= a.env_cutoff) { envi = min(envi, j - d + 1); envj = max(envj, j); }
          }
          bool rep;
          const bool q = act && bs >= a.cutoff && bs > NOTIMP_F;
          if (a.cand_mode == 1) {
            float xq = q ? bs : -INFINITY;
            float inc = xq;
            for (int o = 1; o < 64; o <<= 1) {
              const float u = __shfl_up(inc, o, 64);
              if (lane >= o) inc = fmaxf(inc, u);
            }
            float ex = __shfl_up(inc, 1, 64);
            if (lane == 0) ex = -INFINITY;
            const float pmax = fmaxf(IMPF, fmaxf(carry_max, ex));
            rep = q && bs > pmax;
            carry_max = fmaxf(carry_max, __shfl(inc, 63, 64));
          } else if (a.cand_mode == 2) {
            rep = act && bs > 0.f && bs > NOTIMP_F;
          } else {
            rep = q;
          }
          const unsigned long long m = __ballot(rep);
          if (m) {
            const int n = __popcll(m);
            unsigned long long basei = 0;
            if (lane == 0) basei = atomicAdd(a.ncand, (unsigned long long) n);
            basei = __shfl(basei, 0, 64);
            if (rep) {
              const unsigned long long below = m & ((1ull << lane) - 1ull);
              const unsigned long long idx = basei + __popcll(below);
              if ((int64_t) idx < a.cand_cap) {
                Cand cd; cd.j = (int32_t) j; cd.d = d; cd.sc = bs; cd.r = br;
                a.cand[idx] = cd;
              }
            }
          }
        }
      }
      wave_sync();
    } /* j */

    for (int o = 32; o >= 1; o >>= 1) {
      best = fmaxf(best, __shfl_xor(best, o, 64));
      envi = min(envi, (int64_t) __shfl_xor((long long) envi, o, 64));
      envj = max(envj, (int64_t) __shfl_xor((long long) envj, o, 64));
    }
    if (lane == 0) { a.chunk_best[c] = best; a.chunk_envi[c] = envi; a.chunk_envj[c] = envj; }
  } /* chunks */
}

/* ------------------------------------------------------------------------ */
/* launch stubs for scan_host.cpp                                            */

static const void *std_kernel_ptr(int semi, int want_vsc)
{
  switch (semi * 2 + (want_vsc ? 1 : 0)) {
  case 0: return (const void *) std_scan_kernel<SEMI_CYK, false>;
  case 1: return (const void *) std_scan_kernel<SEMI_CYK, true>;
  case 2: return (const void *) std_scan_kernel<SEMI_FINSIDE, false>;
  case 3: return (const void *) std_scan_kernel<SEMI_FINSIDE, true>;
  case 4: return (const void *) std_scan_kernel<SEMI_IINSIDE, false>;
  case 5: return (const void *) std_scan_kernel<SEMI_IINSIDE, true>;
  default: return NULL;
  }
}

/* LDS bytes of a workgroup of `wpb` wavefronts */
extern "C" size_t infx_std_lds_bytes(int semi, int wpb, int acc_n)
{
  return (semi != SEMI_CYK ? (size_t) LOGSUM_N * 4 : 0) + (size_t) wpb * acc_n * 8;
}

/* resident workgroups per CU for this launch shape (0 if it cannot run) */
extern "C" int infx_std_occupancy(int semi, int want_vsc, int wpb, int acc_n)
{
  const void *f = std_kernel_ptr(semi, want_vsc);
  if (!f || wpb < 1 || wpb > WPB_MAX) return 0;
  const size_t lds = infx_std_lds_bytes(semi, wpb, acc_n);
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) != hipSuccess) { (void) hipGetLastError(); return 0; }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * wpb, lds) != hipSuccess) { (void) hipGetLastError(); return 0; }
  return nb;
}

extern "C" hipError_t infx_launch_std(int semi, int want_vsc, dim3 grid, dim3 block, hipStream_t stream, const ScanArgs *a)
{
  const void *f = std_kernel_ptr(semi, want_vsc);
  if (!f || block.x % 64 || block.x / 64 > WPB_MAX) return hipErrorInvalidValue;
  const size_t lds = infx_std_lds_bytes(semi, block.x / 64, a->acc_n);
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
  if (e != hipSuccess) return e;
  switch (semi * 2 + (want_vsc ? 1 : 0)) {
  case 0: hipLaunchKernelGGL((std_scan_kernel<SEMI_CYK, false>), grid, block, lds, stream, *a); break;
  case 1: hipLaunchKernelGGL((std_scan_kernel<SEMI_CYK, true>), grid, block, lds, stream, *a); break;
  case 2: hipLaunchKernelGGL((std_scan_kernel<SEMI_FINSIDE, false>), grid, block, lds, stream, *a); break;
  case 3: hipLaunchKernelGGL((std_scan_kernel<SEMI_FINSIDE, true>), grid, block, lds, stream, *a); break;
  case 4: hipLaunchKernelGGL((std_scan_kernel<SEMI_IINSIDE, false>), grid, block, lds, stream, *a); break;
  case 5: hipLaunchKernelGGL((std_scan_kernel<SEMI_IINSIDE, true>), grid, block, lds, stream, *a); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
