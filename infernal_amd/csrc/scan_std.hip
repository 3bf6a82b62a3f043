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
 *    sweeps, so at most 65 sweeps are needed.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "scan_kernels.h"

#define IMPF        (-1e36f)
#define NOTIMP_F    (-9.999e35f)
#define INFTY_I     987654321
#define LOGSUM_N    23000
#define KP          18

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int wrap_add(int a, int b) { return (int) ((unsigned) a + (unsigned) b); }
__device__ __forceinline__ int wrap_sub(int a, int b) { return (int) ((unsigned) a - (unsigned) b); }
__device__ __forceinline__ int wrap_mul(int a, int b) { return (int) ((unsigned) a * (unsigned) b); }

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
  static __device__ __forceinline__ T clamp_root(T x) { return max(x, -INFTY_I); }
  /* src/cm.c:4225 Scorify */
  static __device__ __forceinline__ float score(T x) { return (x == -INFTY_I) ? IMPF : ((float) x / 1000.0f); }
  static __device__ __forceinline__ float envval(T x) { return (float) x; }
  static __device__ __forceinline__ T init(int ok, uint32_t el, uint32_t endsc, int x) {
    return ok ? wrap_add(wrap_mul((int) el, x), (int) endsc) : -INFTY_I;
  }
  static __device__ __forceinline__ bool same(T a, T b) { return a == b; }
};

/* Combine child terms c[0..cnum-1] and the init term in the reference's
 * order for each child count (src/cm_dpsearch.c:289-427 and :446-520;
 * identical in the FInside and IInside twins).  `has_c0` false leaves the
 * c[0] term out (the IL self term, applied last by the caller). */
template <int S>
__device__ __forceinline__ typename SR<S>::T combine_children(int cnum, const typename SR<S>::T *c, typename SR<S>::T ini,
                                                              bool has_c0, const typename SR<S>::T *tbl)
{
  typedef SR<S> R;
  typename R::T sc;
  switch (cnum) {
  case 2:
    sc = R::op(c[1], ini, tbl);
    break;
  case 3:
    sc = R::op(c[2], c[1], tbl);
    sc = R::op(sc, ini, tbl);
    break;
  case 4:
    sc = R::op(c[3], c[2], tbl);
    sc = R::op(sc, ini, tbl);
    sc = R::op(sc, c[1], tbl);
    break;
  case 5:
    sc = R::op(c[4], c[3], tbl);
    sc = R::op(sc, ini, tbl);
    sc = R::op(sc, c[1], tbl);
    sc = R::op(sc, c[2], tbl);
    break;
  default: /* 6 */
    sc = R::op(c[5], ini, tbl);
    sc = R::op(sc, c[4], tbl);
    sc = R::op(sc, c[3], tbl);
    sc = R::op(sc, c[2], tbl);
    sc = R::op(sc, c[1], tbl);
    break;
  }
  if (has_c0) sc = R::op(sc, c[0], tbl);
  return sc;
}

/* ------------------------------------------------------------------------ */

template <int S, bool VSC>
__global__ void __launch_bounds__(256) std_scan_kernel(ScanArgs a)
{
  typedef SR<S> R;
  typedef typename R::T T;
  const int lane = threadIdx.x & 63;
  const int slot = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (slot >= a.nslots) return;

  const int W1 = a.W + 1;
  char *sbase = a.scratch + (size_t) slot * a.slot_bytes;
  T   *row0   = (T *) sbase;
  T   *row1   = row0 + a.R;
  T   *deck   = row1 + a.R;
  T   *acc_sc = deck + (size_t) W1 * a.RB;
  int *acc_r  = (int *) (acc_sc + W1);
  float *vscs = (float *) (acc_r + W1);
  const T *esc   = (const T *) a.esc;
  const T *tbl   = (const T *) a.lsum;
  const T *rtmpl = (const T *) a.row_tmpl;
  const T *dtmpl = (const T *) a.deck_tmpl;
  const uint8_t *dsq = a.dsq;
  const int64_t dsq_off = a.dsq_off;

  if (VSC) for (int v = lane; v < a.M; v += 64) vscs[v] = IMPF;

  for (int c = slot; c < a.nchunks; c += a.nslots) {
    const int64_t ac = a.i0 + (int64_t) c * a.chunk_len;
    const int64_t bc = min(a.j0, ac + a.chunk_len - 1);
    const int64_t s0 = max(a.i0, ac - (int64_t) a.W + 1);

    for (int x = lane; x < a.R; x += 64) { T t = rtmpl[x]; row0[x] = t; row1[x] = t; }
    for (int r = 0; r < W1; r++)
      for (int x = lane; x < a.RB; x += 64) deck[(size_t) r * a.RB + x] = dtmpl[x];
    wave_sync();

    float   best = IMPF;
    int64_t envi = INT64_MAX, envj = INT64_MIN;
    float   rec_carry_unused = 0.f; (void) rec_carry_unused;

    for (int64_t j = s0; j <= bc; j++) {
      const int jp   = (int) (j - s0 + 1);
      const int jmod = (int) (j % W1);
      T *cur = (j & 1) ? row1 : row0;
      T *prv = (j & 1) ? row0 : row1;
      T *dk  = deck + (size_t) jmod * a.RB;
      const int xj = dsq[j - dsq_off];
      const int rdn = a.root_dn, rdx = min(jp, a.root_dx);

      if (a.local_begin) {
        for (int d = rdn + lane; d <= rdx; d += 64) { acc_sc[d] = R::neg_sentinel(); acc_r[d] = 0; }
        wave_sync();
      }

      for (int v = a.M - 1; v >= 1; v--) {
        const DState *s = &a.st[v];
        const int kind = s->kind;
        if (kind == KIND_SKIP) continue;
        const int dn = s->dn;
        const int dx = min(jp, s->dx);
        if (dn > dx) continue;
        const int cnum = s->cnum;
        const uint32_t el = a.el_selfsc;
        const int init_ok = s->init_ok;
        const uint32_t endsc = s->endsc;
        const int has_begin = a.local_begin && s->has_begin;
        const int bdn = max(rdn, dn), bdx = min(rdx, dx);
        const T bsc = R::bits(s->beginsc);
        float vmax = IMPF;

        if (kind == KIND_GENERIC || kind == KIND_BEGL) {
          const T *crow = s->rdelta ? prv : cur;
          T *out = (kind == KIND_BEGL) ? dk : cur;
          const int sd = s->sd, em = s->emitmode, eo = s->esc_off;
          const T eright = (em == EMITRIGHT) ? esc[eo + xj] : R::imp();
          for (int d0 = dn; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
            if (d <= dx) {
              const int dy = d - sd;
              T cv[6];
#pragma unroll
              for (int k = 0; k < 6; k++)
                if (k < cnum) cv[k] = R::add(crow[s->child_base[k] + dy], R::bits(s->tsc[k]));
              const T ini = R::init(init_ok, el, endsc, dy);
              T sc;
              if (kind == KIND_BEGL) {
                sc = ini;
                for (int k = 0; k < cnum; k++) sc = R::op(sc, cv[k], tbl);
              } else {
                sc = combine_children<S>(cnum, cv, ini, true, tbl);
                if (em == EMITLEFT)       sc = R::add(sc, esc[eo + dsq[j - d + 1 - dsq_off]]);
                else if (em == EMITRIGHT) sc = R::add(sc, eright);
                else if (em == EMITPAIR)  sc = R::add(sc, esc[eo + dsq[j - d + 1 - dsq_off] * KP + xj]);
              }
              out[s->base + d] = sc;
              if (has_begin && d >= bdn && d <= bdx) {
                const T cand = R::add(sc, bsc);
                if (cand >= acc_sc[d]) { acc_sc[d] = cand; acc_r[d] = v; }
              }
              if (VSC) vmax = fmaxf(vmax, R::score(sc));
            }
          }
        } else if (kind == KIND_IL) {
          /* x_d = op(pre_d, x_{d-1} + t0) + e_d, pre_d = the other terms */
          const T t0 = R::bits(s->tsc[0]);
          const int eo = s->esc_off;
          T carry = cur[s->base + dn - 1];
          for (int d0 = dn; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
            const bool act = d <= dx;
            const int dd = act ? d : dx;            /* inactive lanes mirror a valid cell */
            const int dy = dd - 1;
            T cv[6];
#pragma unroll
            for (int k = 1; k < 6; k++)
              if (k < cnum) cv[k] = R::add(cur[s->child_base[k] + dy], R::bits(s->tsc[k]));
            cv[0] = R::imp();
            const T ini = R::init(init_ok, el, endsc, dy);
            const T pre = combine_children<S>(cnum, cv, ini, false, tbl);
            const T e   = esc[eo + dsq[j - dd + 1 - dsq_off]];
            T left = (lane == 0) ? carry : R::imp();
            T x = R::add(R::op(pre, R::add(left, t0), tbl), e);
            for (int it = 0; it < 65; it++) {
              T up = __shfl_up(x, 1, 64);
              left = (lane == 0) ? carry : up;
              const T xn = R::add(R::op(pre, R::add(left, t0), tbl), e);
              const bool changed = act && !R::same(xn, x);
              x = xn;
              if (!__any(changed)) break;
            }
            if (act) {
              cur[s->base + d] = x;
              if (has_begin && d >= bdn && d <= bdx) {
                const T cand = R::add(x, bsc);
                if (cand >= acc_sc[d]) { acc_sc[d] = cand; acc_r[d] = v; }
              }
              if (VSC) vmax = fmaxf(vmax, R::score(x));
            }
            carry = __shfl(x, 63, 64);
          }
        } else { /* KIND_B */
          const int bw = s->child_base[0], by = s->child_base[1];
          for (int d0 = dn; d0 <= dx; d0 += 64) {
            const int d = d0 + lane;
            if (d <= dx) {
              int kmin, kmax;
              if (s->banded) {
                kmin = max(0, max(s->kdn_y, d - s->kdx_w));
                kmax = min(s->kdx_y, d - s->kdn_w);
              } else { kmin = 0; kmax = d; }
              T sc = R::init(init_ok, el, endsc, d);
              int r = jmod - kmin; if (r < 0) r += W1;
              for (int k = kmin; k <= kmax; k++) {
                const T bl = deck[(size_t) r * a.RB + bw + d - k];
                const T yv = cur[by + k];
                sc = R::op(sc, R::add(bl, yv), tbl);
                r = (r == 0) ? a.W : r - 1;
              }
              cur[s->base + d] = sc;
              if (has_begin && d >= bdn && d <= bdx) {
                const T cand = R::add(sc, bsc);
                if (cand >= acc_sc[d]) { acc_sc[d] = cand; acc_r[d] = v; }
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
              const T acv = acc_sc[d];
              if (acv > al) { al = acv; br = acc_r[d]; }
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
        wave_sync();
      }
    } /* j */

    for (int o = 32; o >= 1; o >>= 1) {
      best = fmaxf(best, __shfl_xor(best, o, 64));
      envi = min(envi, (int64_t) __shfl_xor((long long) envi, o, 64));
      envj = max(envj, (int64_t) __shfl_xor((long long) envj, o, 64));
    }
    if (lane == 0) { a.chunk_best[c] = best; a.chunk_envi[c] = envi; a.chunk_envj[c] = envj; }
  } /* chunks */
}

/* explicit instantiations, launched from scan_host.cpp through these stubs */
extern "C" hipError_t infx_launch_std(int semi, int want_vsc, dim3 grid, dim3 block, hipStream_t stream, const ScanArgs *a)
{
  switch (semi * 2 + (want_vsc ? 1 : 0)) {
  case 0: hipLaunchKernelGGL((std_scan_kernel<SEMI_CYK, false>), grid, block, 0, stream, *a); break;
  case 1: hipLaunchKernelGGL((std_scan_kernel<SEMI_CYK, true>), grid, block, 0, stream, *a); break;
  case 2: hipLaunchKernelGGL((std_scan_kernel<SEMI_FINSIDE, false>), grid, block, 0, stream, *a); break;
  case 3: hipLaunchKernelGGL((std_scan_kernel<SEMI_FINSIDE, true>), grid, block, 0, stream, *a); break;
  case 4: hipLaunchKernelGGL((std_scan_kernel<SEMI_IINSIDE, false>), grid, block, 0, stream, *a); break;
  case 5: hipLaunchKernelGGL((std_scan_kernel<SEMI_IINSIDE, true>), grid, block, 0, stream, *a); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
