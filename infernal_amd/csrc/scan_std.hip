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
#include <stdio.h>
#include "scan_kernels.h"

#define IMPF        (-1e36f)
#define NOTIMP_F    (-9.999e35f)
#define INFTY_I     987654321
#define LOGSUM_N    INFX_LOGSUM_N
#define KP          18
#define WPB_MAX     16

typedef const __attribute__((address_space(4))) uint32_t CU32;
typedef const __attribute__((address_space(4))) ScanArgs CArgs;

typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
typedef const __attribute__((address_space(4))) u32x16 CU32x16;

/* state descriptor through the scalar cache: one s_load_dwordx16 */
__device__ __forceinline__ DState load_state(const DState *p)
{
  static_assert(sizeof(DState) == 64, "DState is one 64-byte scalar load");
  const u32x16 w = *(CU32x16 *) p;
  DState s;
  s.hdr = w[0]; s.dnx = w[1]; s.base = (int32_t) w[2]; s.aux = w[3];
  s.beginsc = w[4]; s.endsc = w[5];
  s.tsc[0] = w[6]; s.tsc[1] = w[7]; s.tsc[2] = w[8]; s.tsc[3] = w[9]; s.tsc[4] = w[10]; s.tsc[5] = w[11];
  s.cbase = (int32_t) w[12]; s.coff[0] = w[13]; s.coff[1] = w[14]; s.coff[2] = w[15];
  return s;
}

/* The launch parameters are read through a pointer the compiler cannot
 * hoist loads from: each use is a fresh scalar load (scalar-cache hit), so
 * the parameters of the root and hit-reporting steps hold no SGPRs during
 * the state sweep. */
__device__ __forceinline__ CArgs *fresh(CArgs *p) { asm volatile("" : "+s"(p)); return p; }

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Element i of a wave-uniform base pointer, addressed with a 32-bit byte
 * offset: the load/store takes the SGPR base + VGPR offset form, so no lane
 * holds a 64-bit address (fewer VGPRs, no 64-bit adds).  Offsets here are
 * below 2^31 bytes (a slot's rows and decks, the target, the score tables). */
template <typename T> __device__ __forceinline__ T ldu(const T *base, uint32_t i)
{ return *(const T *) ((const char *) base + (uint32_t) (i * (uint32_t) sizeof(T))); }
template <typename T> __device__ __forceinline__ void stu(T *base, uint32_t i, T v)
{ *(T *) ((char *) base + (uint32_t) (i * (uint32_t) sizeof(T))) = v; }

/* Materialise a loaded value here.  Called on every value of a block right
 * after all of the block's loads are issued: the loads go out back to back
 * and are waited for once, and the compiler cannot sink them into the
 * per-row store branches (which would serialise rows j and j+1 behind the
 * row-j store). */
template <typename T> __device__ __forceinline__ void hold(T &x) { asm volatile("" : "+v"(x)); }

__device__ __forceinline__ int wrap_add(int a, int b) { return (int) ((unsigned) a + (unsigned) b); }
__device__ __forceinline__ int wrap_sub(int a, int b) { return (int) ((unsigned) a - (unsigned) b); }
__device__ __forceinline__ int wrap_mul(int a, int b) { return (int) ((unsigned) a * (unsigned) b); }

/* lane i <- x of lane i-1, lane 0 <- lane 0 of `first` (DPP wave_shr:1) */
__device__ __forceinline__ int shr1(int first, int x) { return __builtin_amdgcn_update_dpp(first, x, 0x138, 0xf, 0xf, false); }

/* ------------------------------------------------------------------------ */
/* semirings                                                                 */

template <int S> struct SRB;

template <> struct SRB<SEMI_CYK> {
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
  /* order-preserving int key of a non-NaN float (an involution) */
  static __device__ __forceinline__ int key(T x) { const int b = __float_as_int(x); return b ^ ((b >> 31) & 0x7fffffff); }
  static __device__ __forceinline__ T unkey(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
};

template <> struct SRB<SEMI_FINSIDE> {
  typedef float T;
  static __device__ __forceinline__ T imp() { return IMPF; }
  static __device__ __forceinline__ T neg_sentinel() { return -INFINITY; }
  /* src/logsum.c:160 FLogsum */
  /* (max-min) >= 23 or min == -inf returns max: such a difference clamps to
   * index 23000, whose entry is 0 (max + 0 == max).  A difference below 23
   * gives (int)(diff*1000) <= 22999 (the float product of the largest float
   * below 23 and 1000 rounds below 23000).  -inf - -inf is NaN, and
   * fminf(NaN, 23000) is 23000.  max - min is |a - b| bit for bit (IEEE
   * subtraction rounds symmetrically; -inf - x is -inf), so the min is not
   * formed: the abs is a source modifier of the multiply. */
  static __device__ __forceinline__ T op(T a, T b, const T *tbl) {
    const float mx = fmaxf(a, b);
    const float diff = fabsf(__fsub_rn(a, b));
    const int   idx  = (int) fminf(__fmul_rn(diff, 1000.0f), 23000.0f);
    return __fadd_rn(mx, tbl[idx]);
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
  /* order-preserving int key of a non-NaN float (an involution) */
  static __device__ __forceinline__ int key(T x) { const int b = __float_as_int(x); return b ^ ((b >> 31) & 0x7fffffff); }
  static __device__ __forceinline__ T unkey(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
};

template <> struct SRB<SEMI_IINSIDE> {
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
  static __device__ __forceinline__ int key(T x) { return x; }
  static __device__ __forceinline__ T unkey(int k) { return k; }
};

/* S = semiring | SEMI_GACC: the local-begin accumulators live in the
 * wavefront's global scratch slot (integer-keyed atomic max at L2) instead
 * of LDS, for root bands too wide for LDS at full occupancy */
#define SEMI_GACC 8
template <int S> struct SR : SRB<S & 7> {
  static const bool GACC = (S & SEMI_GACC) != 0;
  /* loads of a visit's rows share their lane offsets (Inside), or each row
   * computes its own (CYK) */
  static const bool SHARED_OFF = (S & 7) != SEMI_CYK;
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

/* emission score of residue code x from a register whose lane y holds the
 * score of code y (ds_bpermute crossbar, no LDS storage) */
template <int S>
__device__ __forceinline__ typename SR<S>::T pick_esc(int ev_bits, int x)
{
  return SR<S>::fromb(__builtin_amdgcn_ds_bpermute(x << 2, ev_bits));
}

/* ------------------------------------------------------------------------ */
/* Multi-row sweeps.  A wavefront computes NR consecutive rows j..j+NR-1 of a
 * state before it moves to the next state.  Row j+t of a state depends only
 * on higher states (rows j+t and j+t-1), on its own row j+t (the IL
 * self-loop) and, for IR, on its own row j+t-1, so the order is the
 * reference's order for each row, while the state descriptor, emission rows
 * and loop overheads are paid once per NR rows.  The row storage is a ring of
 * 2*NR rows (j & (2NR-1)): rows j-1 .. j+NR-1 live.  Rows are addressed as
 * 32-bit cell offsets (ring row * R) from the slot's row base.  NR is 2 for
 * CYK (8 wavefronts per SIMD leave 64 VGPRs) and 4 for Inside with the LDS
 * table (one 16-wavefront workgroup per CU: 128 VGPRs). */

/* Per-sweep context passed as plain values (a struct would be selected by
 * field address and spilled to scratch).  xj[t]: right residue of row j+t;
 * the local-begin maxima of row j+t are at acc0 + t * acc_n. */
#define SW_PARAMS const typename SR<S>::T *tbl, const uint8_t *dsq_b0, int dsq_w, const int *xj, uint32_t el, int lane, \
                  typename SR<S>::T *acc0, int acc_n, int rdn
#define SW_ARGS   tbl, dsq_b0, dsq_w, xj, el, lane, acc0, acc_n, rdn

/* local-begin candidate of cell d (src/cm_dpsearch.c:599-612): the largest
 * alpha + beginsc over begin states; the state is found again only for
 * reported cells (begin_argmax) */
template <int S>
__device__ __forceinline__ void begin_update(typename SR<S>::T *acc, int rdn, bool on, int d, typename SR<S>::T sc, typename SR<S>::T bsc)
{
  typedef SR<S> R;
  /* max without return: no read to wait for; lanes hold distinct d, and no
   * value is NaN.  LDS: ds_max_f32 / ds_max_i32.  Global: global_atomic_smax
   * on order-preserving int keys (there is no global float max). */
  if (R::GACC) {
    int *a = (int *) ((char *) acc + (uint32_t) ((d - rdn) * 4));   /* SGPR base + 32-bit offset */
    if (on) __hip_atomic_fetch_max(a, R::key(R::add(sc, bsc)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    if (on) __hip_atomic_fetch_max(&acc[d - rdn], R::add(sc, bsc), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
  }
}

/* NB consecutive 64-cell blocks of a non-B, non-IL state, rows j+t for
 * t < NR, from d0: all loads first (NR*NB independent cells per child and
 * lane), then the combines, then the stores.  Children are read at
 * ibase[ioff[t] + ...] (row j+t, or j+t-1 for right emitters), results
 * stored at obase[ooff[t] + ...] (the row ring, or a BEGL deck ring).  Lanes
 * past a row's band (dx[t]; a row past the chunk has dx < dn) compute a copy
 * of a valid cell and store nothing.
 * IR (the state is its own first child on the j-1 row, at d-1): row j+t's
 * self term (t >= 1) is row j+t-1's cell d-1, which this visit computes.  It
 * is taken from that row's results in registers, shifted one lane (DPP
 * wave_shr:1; lane 0 from irc[t], row j+t-1's cell d0-1: the previous
 * block's lane 63, or the stored cell dn-1 that no visit writes).
 * combine_children applies the self term last, as the reference's order for
 * each child count does. */
template <int S, bool VSC, int CN, bool BEGL, int NB, bool IR, int NR>
__device__ __forceinline__ float gen_blocks(SW_PARAMS, const DState &s,
                                            const typename SR<S>::T *ibase, const int *ioff,
                                            typename SR<S>::T *obase, const int *ooff, const int *cb,
                                            int d0, int dn, const int *dx, bool hb, const int *bdx,
                                            const int *ev, const bool *vs, typename SR<S>::T *irc)
{
  typedef SR<S> R;
  typedef typename R::T T;
  const uint32_t hdr = s.hdr;
  const int em = (hdr >> 3) & 3, sd = (hdr >> 5) & 3;
  const int init_ok = (hdr >> 12) & 1;
  const T bsc = R::bits(s.beginsc);
  const int bdn = max(rdn, dn);
  float vmax = IMPF;
  int d[NB], xl[NR][NB];
  bool a[NR][NB];
  T c[NR][NB][CN];
  int dxm = dx[0];
#pragma unroll
  for (int t = 1; t < NR; t++) dxm = max(dxm, dx[t]);
#pragma unroll
  for (int b = 0; b < NB; b++) {
    d[b] = d0 + 64 * b + lane;
    /* Inside (SR<S>::SHARED_OFF): offsets shared by the rows, each row
     * adding its own wave-uniform base.  Lanes past the widest row's band
     * read at that band's edge (cell dxm - sd of each child, inside its
     * storage; residue j+t-dxm+1, at most NR-1 residues before the chunk's
     * first row: inside the 64 zero bytes before the target at its start),
     * so they repeat an active lane's addresses; lanes of a narrower row
     * read that row's stored out-of-band cells.  Such values are never
     * stored.  CYK: per-row offsets, idle lanes reading the stand-in cell dn
     * and residue j+t (fewer wave-uniform values live: CYK's few operations
     * per load do not hide the scalar spills of per-row bases). */
    const int dcl = min(d[b], dxm);
    const uint32_t dyb = (uint32_t) (dcl - sd);
    const uint32_t xo = (uint32_t) (dsq_w - dcl);
#pragma unroll
    for (int t = 0; t < NR; t++) {
      a[t][b] = d[b] <= dx[t];
      if (R::SHARED_OFF) {
        xl[t][b] = (int) ldu(dsq_b0 + t, xo);
        const typename R::T *rb = ibase + ioff[t];
#pragma unroll
        for (int k = 0; k < CN; k++)
          c[t][b][k] = (IR && t > 0 && k == 0) ? R::imp() : ldu(rb, (uint32_t) cb[k] + dyb);
      } else {
        const int dy = (a[t][b] ? d[b] : dn) - sd;
        xl[t][b] = (int) ldu(dsq_b0 + t, (uint32_t) (dsq_w - (a[t][b] ? d[b] : 1)));
#pragma unroll
        for (int k = 0; k < CN; k++)
          c[t][b][k] = (IR && t > 0 && k == 0) ? R::imp() : ldu(ibase, (uint32_t) (ioff[t] + cb[k] + dy));
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int t = 0; t < NR; t++) {
      hold(xl[t][b]);
#pragma unroll
      for (int k = 0; k < CN; k++) if (!(IR && t > 0 && k == 0)) hold(c[t][b][k]);
    }
  /* right residue's score (EMITRIGHT): read after the loads are issued */
  T er[NR];
#pragma unroll
  for (int t = 0; t < NR; t++) er[t] = R::fromb(__builtin_amdgcn_readlane(ev[t], xj[t]));
  T sc[NR][NB];
#pragma unroll
  for (int b = 0; b < NB; b++) {
    /* the init term of cell d: the same in every row (idle lanes get their
     * own d's, a value nothing reads) */
    const T ini = R::init(init_ok, el, s.endsc, d[b] - sd);
#pragma unroll
    for (int t = 0; t < NR; t++) {
#pragma unroll
      for (int k = 0; k < CN; k++) if (!(IR && t > 0 && k == 0)) c[t][b][k] = R::add(c[t][b][k], R::bits(s.tsc[k]));
      if (BEGL) {
        sc[t][b] = ini;
#pragma unroll
        for (int k = 0; k < CN; k++) sc[t][b] = R::op(sc[t][b], c[t][b][k], tbl);
      } else {
        if (IR && t > 0) {
          c[t][b][0] = R::add(R::fromb(shr1(R::tob(irc[t]), R::tob(sc[t - 1][b]))), R::bits(s.tsc[0]));
          irc[t] = R::fromb(__builtin_amdgcn_readlane(R::tob(sc[t - 1][b]), 63));
        }
        sc[t][b] = combine_children<S, CN>(c[t][b], ini, true, tbl);
        if (em != EMITNONE) sc[t][b] = R::add(sc[t][b], (em == EMITRIGHT) ? er[t] : pick_esc<S>(ev[t], xl[t][b]));
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; b++)
#pragma unroll
    for (int t = 0; t < NR; t++)
      if (a[t][b]) {
        if (R::SHARED_OFF) stu(obase + ooff[t], (uint32_t) (s.base + d[b]), sc[t][b]);
        else               stu(obase, (uint32_t) (ooff[t] + s.base + d[b]), sc[t][b]);
        begin_update<S>(acc0 + t * acc_n, rdn, hb && d[b] >= bdn && d[b] <= bdx[t], d[b], sc[t][b], bsc);
        if (VSC && vs[t]) vmax = fmaxf(vmax, R::score(sc[t][b]));
      }
  return vmax;
}

/* A non-B, non-IL state with CN children over rows j+t, t < NR (bands
 * [dn, dx[t]]).  BEGL decks combine init, c0, c1, ... in child order
 * (src/cm_dpsearch.c:430-444); other states use combine_children.  NBW > 1
 * takes up to NBW blocks at a time where the band has them (more loads in
 * flight per wavefront, more registers): the wide-band CYK kernel of models
 * whose targets cannot fill the chip with wavefronts.  Returns the max score
 * over the rows flagged in vs. */
template <int S, bool VSC, int CN, bool BEGL, int NBW, bool IR, int NR>
__device__ __forceinline__ float gen_state(SW_PARAMS, const DState &s,
                                           const typename SR<S>::T *ibase, const int *ioff,
                                           typename SR<S>::T *obase, const int *ooff,
                                           int dn, const int *dx, bool hb, const int *bdx, const int *ev, const bool *vs)
{
  typedef SR<S> R;
  typedef typename R::T T;
  int cb[CN];
#pragma unroll
  for (int k = 0; k < CN; k++) cb[k] = s.cbase + (int) (int16_t) (s.coff[k >> 1] >> (16 * (k & 1)));
  int dxm = dx[0];
#pragma unroll
  for (int t = 1; t < NR; t++) dxm = max(dxm, dx[t]);
  float vmax = IMPF;
  /* IR: row j+t-1's cell dn-1 (outside the band, never written by a visit) */
  T irc[NR];
#pragma unroll
  for (int t = 0; t < NR; t++) irc[t] = (IR && t > 0) ? ldu(ibase, (uint32_t) (ioff[t] + cb[0] + dn - 1)) : R::imp();
  int d0 = dn;
  if (NBW >= 4)
    for (; d0 + 192 <= dxm; d0 += 256)
      vmax = fmaxf(vmax, gen_blocks<S, VSC, CN, BEGL, 4, IR, NR>(SW_ARGS, s, ibase, ioff, obase, ooff, cb, d0, dn, dx, hb, bdx, ev, vs, irc));
  if (NBW >= 2)
    for (; d0 + 64 <= dxm; d0 += 128)
      vmax = fmaxf(vmax, gen_blocks<S, VSC, CN, BEGL, 2, IR, NR>(SW_ARGS, s, ibase, ioff, obase, ooff, cb, d0, dn, dx, hb, bdx, ev, vs, irc));
  for (; d0 <= dxm; d0 += 64)
    vmax = fmaxf(vmax, gen_blocks<S, VSC, CN, BEGL, 1, IR, NR>(SW_ARGS, s, ibase, ioff, obase, ooff, cb, d0, dn, dx, hb, bdx, ev, vs, irc));
  return vmax;
}

/* FInside IL chains: a starting point for the fixed-point sweep.  In the
 * linear domain the chain x_d = log2(2^pre_d + 2^(x_{d-1} + t0)) + e_d is the
 * affine recurrence X_d = 2^(t0+e_d) X_{d-1} + 2^(pre_d+e_d); its maps compose
 * associatively, (A2, B2) o (A1, B1) = (A2 + A1, lse(B2, A2 + B1)) in log2
 * terms, so a Kogge-Stone prefix over the 64 lanes (DPP row_shr 1, 2, 4, 8
 * within each 16-lane row, then row_bcast 15 and 31 across rows) gives every
 * lane its map from the carry x_{d0-1}.  lse is the FLogsum of the LDS table
 * (seven VALU operations and one LDS read; the hardware's exp2/log2 pair cost
 * more issue cycles and measured 3% slower over the whole kernel), but the
 * composition sums in a tree order, so the result is only close to the
 * reference's values (within a few table steps); the sweeps that follow make
 * every lane exact.  Lanes without a source take the identity (A 0, B far
 * below any DP value: the log-sum returns its max unchanged), which leaves
 * their map unchanged bit for bit. */
__device__ __forceinline__ float lse_tbl(float x, float y, const float *tbl)
{
  return SRB<SEMI_FINSIDE>::op(x, y, tbl);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ void il_scan_step(float &A, float &B, const float *tbl)
{
  const float oA = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(0.0f), __float_as_int(A), CTRL, ROWS, 0xf, false));
  const float oB = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(-1e37f), __float_as_int(B), CTRL, ROWS, 0xf, false));
  B = lse_tbl(B, __fadd_rn(A, oB), tbl);
  A = __fadd_rn(A, oA);
}

__device__ __forceinline__ float il_guess_finside(float pre, float e, float t0, float carry, const float *tbl)
{
  float A = __fadd_rn(t0, e), B = __fadd_rn(pre, e);
  il_scan_step<0x111, 0xf>(A, B, tbl);   /* row_shr:1 */
  il_scan_step<0x112, 0xf>(A, B, tbl);   /* row_shr:2 */
  il_scan_step<0x114, 0xf>(A, B, tbl);   /* row_shr:4 */
  il_scan_step<0x118, 0xf>(A, B, tbl);   /* row_shr:8 */
  il_scan_step<0x142, 0xa>(A, B, tbl);   /* row_bcast:15 into rows 1 and 3 */
  il_scan_step<0x143, 0xc>(A, B, tbl);   /* row_bcast:31 into rows 2 and 3 */
  return lse_tbl(B, __fadd_rn(A, carry), tbl);
}

/* IL state: x_d = op(pre_d, x_{d-1} + t0) + e_d with pre_d the other CN-1
 * child terms combined with init (the self term is applied last), solved
 * per 64-cell block by the exact fixed-point sweep, the NR rows together
 * (children on the same row: row j+t at rows[roff[t] + ...]). */
template <int S, bool VSC, int CN, int NR>
__device__ __forceinline__ float il_state(SW_PARAMS, const DState &s, typename SR<S>::T *rows, const int *roff,
                                          int dn, const int *dx, bool hb, const int *bdx, int ev, const bool *vs)
{
  typedef SR<S> R;
  typedef typename R::T T;
  const int init_ok = (s.hdr >> 12) & 1;
  const T bsc = R::bits(s.beginsc);
  const T t0 = R::bits(s.tsc[0]);
  const int bdn = max(rdn, dn);
  int cb[CN];
#pragma unroll
  for (int k = 0; k < CN; k++) cb[k] = s.cbase + (int) (int16_t) (s.coff[k >> 1] >> (16 * (k & 1)));
  T carry[NR];
  int dxm = dx[0];
#pragma unroll
  for (int t = 0; t < NR; t++) {
    carry[t] = ldu(rows, (uint32_t) (roff[t] + s.base + dn - 1));
    dxm = max(dxm, dx[t]);
  }
  float vmax = IMPF;
  for (int d0 = dn; d0 <= dxm; d0 += 64) {
    const int d = d0 + lane;
    bool a[NR];
    int xl[NR];
    T c[NR][CN];
    /* Inside: shared offsets, lanes past the widest band reading at its
     * edge, as in gen_blocks; CYK: idle lanes mirror a valid cell and read
     * residue j+t */
    const int dcl = min(d, dxm);
    const uint32_t xo = (uint32_t) (dsq_w - dcl);
#pragma unroll
    for (int t = 0; t < NR; t++) {
      a[t] = d <= dx[t];
      c[t][0] = R::imp();
      if (R::SHARED_OFF) {
        xl[t] = (int) ldu(dsq_b0 + t, xo);
        const T *rb = rows + roff[t];
#pragma unroll
        for (int k = 1; k < CN; k++) c[t][k] = ldu(rb, (uint32_t) (cb[k] + dcl - 1));
      } else {
        const int dd = a[t] ? d : max(dn, dx[t]);
        xl[t] = (int) ldu(dsq_b0 + t, (uint32_t) (dsq_w - (a[t] ? d : 1)));
#pragma unroll
        for (int k = 1; k < CN; k++) c[t][k] = ldu(rows, (uint32_t) (roff[t] + cb[k] + dd - 1));
      }
    }
#pragma unroll
    for (int t = 0; t < NR; t++) {
      hold(xl[t]);
#pragma unroll
      for (int k = 1; k < CN; k++) hold(c[t][k]);
    }
    T pre[NR], e[NR], x[NR];
    const T ini = R::init(init_ok, el, s.endsc, d - 1);   /* every row's (idle lanes: a value nothing reads) */
#pragma unroll
    for (int t = 0; t < NR; t++) {
#pragma unroll
      for (int k = 1; k < CN; k++) c[t][k] = R::add(c[t][k], R::bits(s.tsc[k]));
      pre[t] = combine_children<S, CN>(c[t], ini, false, tbl);
      e[t] = pick_esc<S>(ev, xl[t]);
      /* the start is only a guess: the sweeps run until every lane is a
       * fixed point of the exact update, which the sequential recursion
       * defines uniquely (lane k is exact after k + 1 sweeps from any
       * start).  CYK and IInside start from pre + e (the first sweep brings
       * in the carry); FInside from the prefix-scan estimate of the chain. */
      if ((S & 7) == SEMI_FINSIDE) x[t] = il_guess_finside(pre[t], e[t], t0, carry[t], (const float *) tbl);
      else                         x[t] = R::add(pre[t], e[t]);
    }
    for (int it = 0; it < 66; it++) {
      bool changed = false;
#pragma unroll
      for (int t = 0; t < NR; t++) {
        const T l = R::fromb(shr1(R::tob(carry[t]), R::tob(x[t])));
        const T n = R::add(R::op(pre[t], R::add(l, t0), tbl), e[t]);
        changed = changed || (a[t] && !R::same(n, x[t]));
        x[t] = n;
      }
      if (!__any(changed)) break;
    }
#pragma unroll
    for (int t = 0; t < NR; t++) {
      if (a[t]) {
        if (R::SHARED_OFF) stu(rows + roff[t], (uint32_t) (s.base + d), x[t]);
        else               stu(rows, (uint32_t) (roff[t] + s.base + d), x[t]);
        begin_update<S>(acc0 + t * acc_n, rdn, hb && d >= bdn && d <= bdx[t], d, x[t], bsc);
        if (VSC && vs[t]) vmax = fmaxf(vmax, R::score(x[t]));
      }
      carry[t] = R::fromb(__builtin_amdgcn_readlane(R::tob(x[t]), 63));
    }
  }
  return vmax;
}

/* FInside IL states with bands wider than one block: 128-cell blocks, two
 * consecutive cells per lane (d0 + 2l and d0 + 2l + 1).  A sweep takes the
 * left neighbour's second cell (DPP wave_shr:1), computes the lane's first
 * cell from it and its second from the first, so each sweep carries the
 * chain two cells on; lane l is exact after l + 1 sweeps from any start, and
 * a sweep that changes no lane is the unique fixed point, as in il_state.
 * The starting point is il_guess_finside's prefix over the lanes, applied to
 * the composition of each lane's two maps. */
__device__ __forceinline__ void il_guess_finside2(float preA, float eA, float preB, float eB, float t0, float carry,
                                                  float &xA, float &xB, const float *tbl)
{
  const float a0 = __fadd_rn(t0, eA), b0 = __fadd_rn(preA, eA);
  const float a1 = __fadd_rn(t0, eB), b1 = __fadd_rn(preB, eB);
  float A = __fadd_rn(a1, a0), B = lse_tbl(b1, __fadd_rn(a1, b0), tbl);   /* second cell's map from the lane's input */
  il_scan_step<0x111, 0xf>(A, B, tbl);   /* row_shr:1 */
  il_scan_step<0x112, 0xf>(A, B, tbl);   /* row_shr:2 */
  il_scan_step<0x114, 0xf>(A, B, tbl);   /* row_shr:4 */
  il_scan_step<0x118, 0xf>(A, B, tbl);   /* row_shr:8 */
  il_scan_step<0x142, 0xa>(A, B, tbl);   /* row_bcast:15 into rows 1 and 3 */
  il_scan_step<0x143, 0xc>(A, B, tbl);   /* row_bcast:31 into rows 2 and 3 */
  const float xlast = lse_tbl(B, __fadd_rn(A, carry), tbl);                /* the lane's second cell */
  const float xin = __int_as_float(shr1(__float_as_int(carry), __float_as_int(xlast)));   /* its input */
  xA = lse_tbl(b0, __fadd_rn(a0, xin), tbl);
  xB = lse_tbl(b1, __fadd_rn(a1, xA), tbl);
}

template <int S, bool VSC, int CN, int NR>
__device__ __forceinline__ float il_state_w2(SW_PARAMS, const DState &s, typename SR<S>::T *rows, const int *roff,
                                             int dn, const int *dx, bool hb, const int *bdx, int ev, const bool *vs)
{
  typedef SR<S> R;
  typedef typename R::T T;
  const int init_ok = (s.hdr >> 12) & 1;
  const T bsc = R::bits(s.beginsc);
  const T t0 = R::bits(s.tsc[0]);
  const int bdn = max(rdn, dn);
  int cb[CN];
#pragma unroll
  for (int k = 0; k < CN; k++) cb[k] = s.cbase + (int) (int16_t) (s.coff[k >> 1] >> (16 * (k & 1)));
  T carry[NR];
  int dxm = dx[0];
#pragma unroll
  for (int t = 0; t < NR; t++) {
    carry[t] = ldu(rows, (uint32_t) (roff[t] + s.base + dn - 1));
    dxm = max(dxm, dx[t]);
  }
  float vmax = IMPF;
  for (int d0 = dn; d0 <= dxm; d0 += 128) {
    const int dA = d0 + 2 * lane, dB = dA + 1;
    bool aA[NR], aB[NR];
    int xlA[NR], xlB[NR];
    T cA[NR][CN], cB[NR][CN];
    /* shared offsets, lanes past the widest band reading at its edge, as in
     * il_state */
    const int dcA = min(dA, dxm), dcB = min(dB, dxm);
    const uint32_t xoA = (uint32_t) (dsq_w - dcA), xoB = (uint32_t) (dsq_w - dcB);
#pragma unroll
    for (int t = 0; t < NR; t++) {
      aA[t] = dA <= dx[t]; aB[t] = dB <= dx[t];
      xlA[t] = (int) ldu(dsq_b0 + t, xoA);
      xlB[t] = (int) ldu(dsq_b0 + t, xoB);
      cA[t][0] = R::imp(); cB[t][0] = R::imp();
      const T *rb = rows + roff[t];
#pragma unroll
      for (int k = 1; k < CN; k++) {
        cA[t][k] = ldu(rb, (uint32_t) (cb[k] + dcA - 1));
        cB[t][k] = ldu(rb, (uint32_t) (cb[k] + dcB - 1));
      }
    }
#pragma unroll
    for (int t = 0; t < NR; t++) {
      hold(xlA[t]); hold(xlB[t]);
#pragma unroll
      for (int k = 1; k < CN; k++) { hold(cA[t][k]); hold(cB[t][k]); }
    }
    const T iniA = R::init(init_ok, el, s.endsc, dA - 1), iniB = R::init(init_ok, el, s.endsc, dB - 1);
    T preA[NR], preB[NR], eA[NR], eB[NR], xA[NR], xB[NR];
#pragma unroll
    for (int t = 0; t < NR; t++) {
#pragma unroll
      for (int k = 1; k < CN; k++) { cA[t][k] = R::add(cA[t][k], R::bits(s.tsc[k])); cB[t][k] = R::add(cB[t][k], R::bits(s.tsc[k])); }
      preA[t] = combine_children<S, CN>(cA[t], iniA, false, tbl);
      preB[t] = combine_children<S, CN>(cB[t], iniB, false, tbl);
      eA[t] = pick_esc<S>(ev, xlA[t]);
      eB[t] = pick_esc<S>(ev, xlB[t]);
      il_guess_finside2(preA[t], eA[t], preB[t], eB[t], t0, carry[t], xA[t], xB[t], (const float *) tbl);
    }
    for (int it = 0; it < 66; it++) {
      bool changed = false;
#pragma unroll
      for (int t = 0; t < NR; t++) {
        const T l = R::fromb(shr1(R::tob(carry[t]), R::tob(xB[t])));
        const T nA = R::add(R::op(preA[t], R::add(l, t0), tbl), eA[t]);
        const T nB = R::add(R::op(preB[t], R::add(nA, t0), tbl), eB[t]);
        changed = changed || (aA[t] && !R::same(nA, xA[t])) || (aB[t] && !R::same(nB, xB[t]));
        xA[t] = nA; xB[t] = nB;
      }
      if (!__any(changed)) break;
    }
#pragma unroll
    for (int t = 0; t < NR; t++) {
      if (aA[t]) {
        stu(rows + roff[t], (uint32_t) (s.base + dA), xA[t]);
        begin_update<S>(acc0 + t * acc_n, rdn, hb && dA >= bdn && dA <= bdx[t], dA, xA[t], bsc);
        if (VSC && vs[t]) vmax = fmaxf(vmax, R::score(xA[t]));
      }
      if (aB[t]) {
        stu(rows + roff[t], (uint32_t) (s.base + dB), xB[t]);
        begin_update<S>(acc0 + t * acc_n, rdn, hb && dB >= bdn && dB <= bdx[t], dB, xB[t], bsc);
        if (VSC && vs[t]) vmax = fmaxf(vmax, R::score(xB[t]));
      }
      carry[t] = R::fromb(__builtin_amdgcn_readlane(R::tob(xB[t]), 63));
    }
  }
  return vmax;
}

/* Four k of a bifurcation cell: BEGL deck cell (j-k, d-k) + BEGR cell (j, k),
 * combined into sc in ascending k.  SEL: only k <= lim inside the lane's
 * own kmin..kmax count. */
template <int S, bool SEL>
__device__ __forceinline__ typename SR<S>::T b_step(const typename SR<S>::T *tbl, typename SR<S>::T sc, int yb,
                                                   const typename SR<S>::T *wbase, int jp, int rmask, int rstride,
                                                   int kb, int k, int lim, int dd, int kmin, int kmax)
{
  typedef SR<S> R;
  typedef typename R::T T;
  T bl[4], yv[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int kk = k + u;
    yv[u] = R::fromb(__builtin_amdgcn_readlane(yb, min(kk - kb, 63)));
    const T *rowp = wbase + (((jp - kk) & rmask) * rstride - kk);
    bl[u] = ldu(rowp, (uint32_t) dd);
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int kk = k + u;
    const T nsc = R::op(sc, R::add(bl[u], yv[u]), tbl);
    sc = SEL ? ((kk <= lim && kk >= kmin && kk <= kmax) ? nsc : sc) : nsc;
  }
  return sc;
}

/* B state, one row (src/cm_dpsearch.c:246-284): all lanes walk the union of
 * their k ranges together, so one k reads one BEGL deck row (consecutive
 * columns) and one BEGR cell (v_readlane broadcast); each lane keeps its own
 * kmin(d)..kmax(d) in ascending order. */
template <int S, bool VSC>
__device__ __forceinline__ float b_state(const typename SR<S>::T *tbl, uint32_t el, int lane, typename SR<S>::T *acc, int rdn,
                                         const DState &s, typename SR<S>::T *cur, const typename SR<S>::T *deck,
                                         int jp, int dn, int dx, bool hb, int bdx, bool vs)
{
  typedef SR<S> R;
  typedef typename R::T T;
  const uint32_t hdr = s.hdr;
  const int init_ok = (hdr >> 12) & 1;
  const int banded = (hdr >> 13) & 1;
  const T bsc = R::bits(s.beginsc);
  const int bdn = max(rdn, dn);
  const int kdn_y = s.tsc[0] & 0xffff, kdx_y = s.tsc[0] >> 16;
  const int kdn_w = s.tsc[1] & 0xffff, kdx_w = s.tsc[1] >> 16;
  const int by = (int) s.tsc[3];
  const int rmask = s.aux & 0xffff, rstride = s.aux >> 16;
  const T *wbase = deck + (int) s.tsc[2];
  float vmax = IMPF;
  for (int d0 = dn; d0 <= dx; d0 += 64) {
    const int d = d0 + lane;
    const bool act = d <= dx;
    const int dd = act ? d : dx;
    const int dl = min(d0 + 63, dx);
    int kmin, kmax, K0, K1;
    if (banded) {
      kmin = max(0, max(kdn_y, dd - kdx_w));
      kmax = min(kdx_y, dd - kdn_w);
      K0 = max(0, max(kdn_y, d0 - kdx_w));   /* both bounds grow with d */
      K1 = min(kdx_y, dl - kdn_w);
    } else { kmin = 0; kmax = dd; K0 = 0; K1 = dl; }
    T sc = R::init(init_ok, el, s.endsc, dd);
    /* k in [kin0, kin1] is inside every lane's own range (kmin grows and
     * kmax grows with d): those k need no per-lane test */
    const int kin0 = banded ? max(0, max(kdn_y, dl - kdx_w)) : 0;
    const int kin1 = banded ? min(kdx_y, d0 - kdn_w) : d0;
    for (int kb = K0; kb <= K1; kb += 64) {
      const int ke = min(K1, kb + 63);
      const int yb = R::tob(ldu(cur, (uint32_t) (by + min(kb + lane, ke))));
      /* four k per step, loads first; k past the step's limit is masked
       * (its reads stay inside the slot's guard cells).  The three loops
       * visit kb..ke in ascending order, as each lane's sum requires. */
      const int m0 = max(kb, kin0), m1 = min(ke, kin1);
      int k = kb;
      if (m0 <= m1) {
        for (; k < m0; k += 4) sc = b_step<S, true>(tbl, sc, yb, wbase, jp, rmask, rstride, kb, k, m0 - 1, dd, kmin, kmax);
        k = m0;
        for (; k + 3 <= m1; k += 4) sc = b_step<S, false>(tbl, sc, yb, wbase, jp, rmask, rstride, kb, k, m1, dd, kmin, kmax);
      }
      for (; k <= ke; k += 4) sc = b_step<S, true>(tbl, sc, yb, wbase, jp, rmask, rstride, kb, k, ke, dd, kmin, kmax);
    }
    if (act) {
      stu(cur, (uint32_t) (s.base + d), sc);
      begin_update<S>(acc, rdn, hb && d >= bdn && d <= bdx, d, sc, bsc);
      if (VSC && vs) vmax = fmaxf(vmax, R::score(sc));
    }
  }
  return vmax;
}

/* The local-begin state of reported cells whose root score came from a
 * local begin: the same scan as the accumulation (states in descending v,
 * ties to the later state), on the row that is still in the ring. */
template <int S>
__device__ __forceinline__ int begin_argmax(const DState *stp, int nst, const typename SR<S>::T *cur, int jp, int d)
{
  typedef SR<S> R;
  typename R::T bestv = R::neg_sentinel();
  int br = 0;
  for (int i = 0; i < nst; i++) {
    const DState s = load_state(stp + i);
    if (!((s.hdr >> 11) & 1)) continue;
    const int dn = s.dnx & 0xffff, dx = min(jp, (int) (s.dnx >> 16));
    if (d < dn || d > dx) continue;
    const typename R::T cand = R::add(ldu(cur, (uint32_t) (s.base + d)), R::bits(s.beginsc));
    if (cand >= bestv) { bestv = cand; br = (int) (s.hdr >> 16); }
  }
  return br;
}

template <int S, bool VSC, bool TBL_LDS, int NBC = 1, bool BATCH = false, int NR = 2>
__device__ __forceinline__ void std_scan_body(const ScanArgs *ap_)
{
  /* blocks per step of the generic states (gen_state); with the LDS table
   * Inside keeps 2 blocks of 2 rows, or 1 block of 4 rows, in flight */
  constexpr int NBW = ((S & 7) != SEMI_CYK && TBL_LDS && NR == 2) ? (NBC > 2 ? NBC : 2) : NBC;
  constexpr int RM = 2 * NR - 1;   /* row ring mask */
  typedef SR<S> R;
  typedef typename R::T T;
  extern __shared__ __align__(16) char lds_raw[];
  CArgs *ap = (CArgs *) ap_;

  const int lane = threadIdx.x & 63;
  const int wib  = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   /* wave-uniform: slot pointers live in SGPRs */
  const int slot = blockIdx.x * (blockDim.x >> 6) + wib;

  /* log-sum table (Inside): staged in LDS, or read through L1/L2 */
  const T *tbl = (const T *) fresh(ap)->lsum;
  int lds_off = 0;
  if ((S & 7) != SEMI_CYK && TBL_LDS) {
    T *lt = (T *) lds_raw;
    const T *gt = (const T *) fresh(ap)->lsum;
    for (int i = threadIdx.x; i < INFX_LOGSUM_PAD; i += blockDim.x) lt[i] = gt[i];
    __syncthreads();
    tbl = lt;
    lds_off = INFX_LOGSUM_PAD * 4;
  }
  if (slot >= fresh(ap)->nslots) return;

  const int acc_n = fresh(ap)->acc_n;
  char *sbase = fresh(ap)->scratch + (size_t) slot * fresh(ap)->slot_bytes;
  /* local-begin maxima of rows j .. j+NR-1 (global: int keys) */
  T *acc0 = R::GACC ? (T *) (sbase + fresh(ap)->acc_off) : (T *) (lds_raw + lds_off + (size_t) wib * acc_n * 4 * NR);

  T   *deck = (T *) (sbase + fresh(ap)->deck_off);
  T   *rows = (T *) (sbase + fresh(ap)->row_off);
  float *vscs = (float *) (sbase + fresh(ap)->vsc_off);

  if (VSC) for (int v = lane; v < fresh(ap)->M; v += 64) vscs[v] = IMPF;

  for (int c = slot; c < fresh(ap)->nchunks; c += fresh(ap)->nslots) {
    /* chunk c reports rows ac..bc; its DP starts W-1 rows earlier, but not
     * before the start lo of its window (one window: the whole target) */
    int64_t ac, bc, lo;
    /* the model's parameters: the launch's own, or (batched launches of
     * several models) the chunk's model's; the slot layout, the chunk loop
     * and the per-chunk outputs stay the launch's */
    CArgs *cp = ap;
    int bm = 0;
    if (BATCH) {
      const int64_t *tab = fresh(ap)->chunk_tab;
      ac = tab[4 * c]; bc = tab[4 * c + 1]; lo = tab[4 * c + 2]; bm = (int) tab[4 * c + 3];
      cp = (CArgs *) fresh(ap)->args_tab + bm;
    } else if (const int64_t *tab = fresh(ap)->chunk_tab) {
      ac = tab[3 * c]; bc = tab[3 * c + 1]; lo = tab[3 * c + 2];
    } else {
      ac = fresh(ap)->i0 + (int64_t) c * fresh(ap)->chunk_len;
      bc = min(fresh(ap)->j0, ac + fresh(ap)->chunk_len - 1);
      lo = fresh(ap)->i0;
    }
    const int64_t s0 = max(lo, ac - (int64_t) fresh(cp)->W + 1);
    const int R = fresh(cp)->R;

    {
      const T *rtmpl = (const T *) fresh(cp)->row_tmpl;
      for (int x = lane; x < R; x += 64) {
        const T t = rtmpl[x];
#pragma unroll
        for (int r = 0; r <= RM; r++) rows[r * R + x] = t;
      }
      const T *dtmpl = (const T *) fresh(cp)->deck_tmpl;
      const int RB = fresh(cp)->RB;
      for (int x = lane; x < RB; x += 64) deck[x] = dtmpl[x];
    }
    wave_sync();

    float   best = IMPF;
    int64_t envi = INT64_MAX, envj = INT64_MIN;

    for (int64_t j0 = s0; j0 <= bc; j0 += NR) {
      const int nr = (int) min((int64_t) NR, bc - j0 + 1);   /* rows of this pass inside the chunk */
      const int jp0 = (int) (j0 - s0 + 1);
      const uint8_t *dsq_j = fresh(cp)->dsq + (j0 + 1 - fresh(cp)->dsq_off);
      const int rdn = fresh(cp)->root_dn;
      int roff[NR], xj[NR], rdx[NR];
      bool vs[NR];
#pragma unroll
      for (int t = 0; t < NR; t++) {
        roff[t] = (int) ((j0 + t) & RM) * R;                 /* row j0+t */
        xj[t] = t < nr ? __builtin_amdgcn_readfirstlane((int) dsq_j[t - 1]) : 0;
        rdx[t] = t < nr ? min(jp0 + t, fresh(cp)->root_dx) : rdn - 1;
        vs[t] = t < nr && j0 + t >= ac;
      }
      const int poff = (int) ((j0 + RM) & RM) * R;           /* row j0-1 */
      const int local_begin = fresh(cp)->local_begin;
      const uint32_t el = fresh(cp)->el_selfsc;
      const DState *stp = fresh(cp)->st;
      const T *esc = (const T *) fresh(cp)->esc;
      const int nst = fresh(cp)->nst;
      const int dsq_w = fresh(cp)->W + 64;
      const uint8_t *dsq_b0 = dsq_j - dsq_w;   /* left residue of cell d of row j0+t: dsq_b0[t + dsq_w - d] */

      if (local_begin) {
        const T neg = R::GACC ? R::fromb(R::key(R::neg_sentinel())) : R::neg_sentinel();
#pragma unroll
        for (int t = 0; t < NR; t++)
          for (int d = rdn + lane; d <= rdx[t]; d += 64) acc0[t * acc_n + d - rdn] = neg;
        if (R::GACC) __builtin_amdgcn_s_waitcnt(0);   /* the clears reach L2 before any atomic */
      }

      for (int i = 0; i < nst; i++) {
        const DState s = load_state(stp + i);
        const uint32_t hdr = s.hdr;
        const int kind = hdr & 7;
        const int dn = s.dnx & 0xffff;
        const int sdx = (int) (s.dnx >> 16);
        int dx[NR], bdx[NR];
        int dxm = dn - 1;
#pragma unroll
        for (int t = 0; t < NR; t++) {
          dx[t] = t < nr ? min(jp0 + t, sdx) : dn - 1;
          bdx[t] = min(rdx[t], dx[t]);
          dxm = max(dxm, dx[t]);
        }
        if (dn > dxm) continue;
        const int v = hdr >> 16;
        const bool hb = local_begin && ((hdr >> 11) & 1);
        const int cnum = (hdr >> 8) & 7;
        const int em = (hdr >> 3) & 3;
        /* emission rows: lane y <- score of residue code y (MP: left code y, right code xj) */
        /* Loaded for every state (non-emitting states read entry 0 and never
         * use it): a branch here would join a loaded and a non-loaded value,
         * and the register write on the other path has to wait for any load
         * still in flight -- the previous state's stores included.  One
         * independent load per row (the same row for singlets), so no value
         * is needed before the state's child loads are issued. */
        const int yy = lane < KP ? lane : 0;
        int ev[NR];
#pragma unroll
        for (int t = 0; t < NR; t++) {
          const uint32_t eo = (em == EMITNONE) ? 0u : (uint32_t) (s.aux + (em == EMITPAIR ? yy * KP + xj[t] : yy));
          ev[t] = R::tob(ldu(esc, eo));
        }
        const bool rd = (hdr >> 7) & 1;
        int ioff[NR];   /* children's rows: j0+t, or j0+t-1 for right emitters */
#pragma unroll
        for (int t = 0; t < NR; t++) ioff[t] = rd ? (t == 0 ? poff : roff[t - 1]) : roff[t];
        float vmax;
        if (kind == KIND_GENERIC) {
          /* IR reads its own row j0+t-1 cells for row j0+t: gen_blocks<IR>
           * takes them from the previous row's results in registers */
          const bool seq = (hdr >> 14) & 1;
          if (seq) {
            switch (cnum) {
#define GEN(CN) vmax = gen_state<S, VSC, CN, false, NBW, true, NR>(SW_ARGS, s, rows, ioff, rows, roff, dn, dx, hb, bdx, ev, vs)
            case 2:  GEN(2); break;
            case 3:  GEN(3); break;
            case 4:  GEN(4); break;
            case 5:  GEN(5); break;
            default: GEN(6); break;
#undef GEN
            }
          } else {
            switch (cnum) {
#define GEN(CN) vmax = gen_state<S, VSC, CN, false, NBW, false, NR>(SW_ARGS, s, rows, ioff, rows, roff, dn, dx, hb, bdx, ev, vs)
            case 2:  GEN(2); break;
            case 3:  GEN(3); break;
            case 4:  GEN(4); break;
            case 5:  GEN(5); break;
            default: GEN(6); break;
#undef GEN
            }
          }
        } else if (kind == KIND_IL) {
          switch (cnum) {
#define ILS(CN) if constexpr ((S & 7) == SEMI_FINSIDE) { \
                  vmax = (dxm - dn >= 64) ? il_state_w2<S, VSC, CN, NR>(SW_ARGS, s, rows, roff, dn, dx, hb, bdx, ev[0], vs) \
                                          : il_state<S, VSC, CN, NR>(SW_ARGS, s, rows, roff, dn, dx, hb, bdx, ev[0], vs); \
                } else { \
                  vmax = il_state<S, VSC, CN, NR>(SW_ARGS, s, rows, roff, dn, dx, hb, bdx, ev[0], vs); \
                }
          case 2:  ILS(2); break;
          case 3:  ILS(3); break;
          case 4:  ILS(4); break;
          case 5:  ILS(5); break;
          default: ILS(6); break;
#undef ILS
          }
        } else if (kind == KIND_BEGL) {
          const int rmask = s.aux & 0xffff, rstride = (int) (s.aux >> 16);
          int ooff[NR];
#pragma unroll
          for (int t = 0; t < NR; t++) ooff[t] = ((jp0 + t) & rmask) * rstride;
          switch (cnum) {
#define BGL(CN) vmax = gen_state<S, VSC, CN, true, NBW, false, NR>(SW_ARGS, s, rows, ioff, deck, ooff, dn, dx, false, bdx, ev, vs)
          case 1:  BGL(1); break;
          case 2:  BGL(2); break;
          case 3:  BGL(3); break;
          case 4:  BGL(4); break;
          case 5:  BGL(5); break;
          default: BGL(6); break;
#undef BGL
          }
        } else {
          vmax = IMPF;
#pragma unroll
          for (int t = 0; t < NR; t++) {
            if (t >= nr) break;
            if (t > 0) wave_sync();
            vmax = fmaxf(vmax, b_state<S, VSC>(tbl, el, lane, acc0 + t * acc_n, rdn, s, rows + roff[t], deck, jp0 + t, dn, dx[t], hb, bdx[t], vs[t]));
          }
        }
        if (VSC) {
          for (int o = 32; o >= 1; o >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, o, 64));
          bool any_vs = false;
#pragma unroll
          for (int t = 0; t < NR; t++) any_vs = any_vs || vs[t];
          if (lane == 0 && any_vs) vscs[v] = fmaxf(vscs[v], vmax);
        }
        wave_sync();
      } /* states */

      if (R::GACC && local_begin) __builtin_amdgcn_s_waitcnt(0);   /* every atomic done before the root reads */
      /* ROOT_S (src/cm_dpsearch.c:562-623) and hit candidates, row by row */
#pragma unroll
      for (int t = 0; t < NR; t++) {
        if (t >= nr) break;
        const int64_t j = j0 + t;
        if (j < ac) continue;
        const int jp = jp0 + t;
        const T *cur = rows + roff[t];
        const T *acc = acc0 + t * acc_n;
        const int rdxt = rdx[t];
        CArgs *aq = fresh(cp);
        const int root_cnum = aq->root_cnum;
        const int cand_mode = aq->cand_mode;
        const float cutoff = aq->cutoff, env_cutoff = aq->env_cutoff;
        float carry_max = -INFINITY;   /* prefix max of reportable scores (record filter) */
        for (int d0 = rdn; d0 <= rdxt; d0 += 64) {
          const int d = d0 + lane;
          const bool act = d <= rdxt;
          float bs = IMPF;
          bool lb = false;
          if (act) {
            T al = R::clamp_root(R::add(ldu(cur, (uint32_t) (aq->root_child_base[0] + d)), R::bits(aq->root_tsc[0])));
            for (int k = 1; k < root_cnum; k++)
              al = R::op(al, R::add(ldu(cur, (uint32_t) (aq->root_child_base[k] + d)), R::bits(aq->root_tsc[k])), tbl);
            if (local_begin) {
              const T acv = R::GACC ? R::unkey(__hip_atomic_load((int *) ((char *) acc + (uint32_t) ((d - rdn) * 4)),
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                                    : acc[d - rdn];
              if (acv > al) { al = acv; lb = true; }
            }
            bs = R::score(al);
            best = fmaxf(best, bs);
            if (R::envval(al) >= env_cutoff) { envi = min(envi, j - d + 1); envj = max(envj, j); }
          }
          bool rep;
          const bool q = act && bs >= cutoff && bs > NOTIMP_F;
          if (cand_mode == 1) {
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
          } else if (cand_mode == 2) {
            rep = act && bs > 0.f && bs > NOTIMP_F;
          } else {
            rep = q;
          }
          const unsigned long long m = __ballot(rep);
          if (m) {
            int br = 0;
            if (rep && lb) br = begin_argmax<S>(stp, nst, cur, jp, d);
            const int n = __popcll(m);
            unsigned long long basei = 0;
            if (lane == 0) basei = atomicAdd(aq->ncand, (unsigned long long) n);
            basei = __shfl(basei, 0, 64);
            if (rep) {
              const unsigned long long below = m & ((1ull << lane) - 1ull);
              const unsigned long long idx = basei + __popcll(below);
              if ((int64_t) idx < aq->cand_cap) {
                Cand cd; cd.j = (int32_t) j; cd.d = d; cd.sc = bs; cd.r = BATCH ? (br | (bm << 16)) : br;
                aq->cand[idx] = cd;
              }
            }
          }
        }
      }
      wave_sync();
    } /* j0 */

    for (int o = 32; o >= 1; o >>= 1) {
      best = fmaxf(best, __shfl_xor(best, o, 64));
      envi = min(envi, (int64_t) __shfl_xor((long long) envi, o, 64));
      envj = max(envj, (int64_t) __shfl_xor((long long) envj, o, 64));
    }
    if (lane == 0) { CArgs *aq = fresh(ap); aq->chunk_best[c] = best; aq->chunk_envi[c] = envi; aq->chunk_envj[c] = envj; }
  } /* chunks */
}

/* Kernels.  CYK visits each state for four rows (6 wavefronts per SIMD:
 * 80 VGPRs; measured faster than two rows at 8 wavefronts); Inside
 * with the log-sum table read from L1/L2 is register-allocated for 8
 * wavefronts per SIMD; Inside with the table in LDS is limited to one
 * 16-wavefront workgroup per CU by the table itself. */
#define INFX_NR_CYK 4
template <bool VSC, int G = 0>
__global__ void __launch_bounds__(64 * WPB_MAX) __attribute__((amdgpu_waves_per_eu(6))) std_cyk_kernel(const ScanArgs *a)
{ std_scan_body<SEMI_CYK | G, VSC, false, 1, false, INFX_NR_CYK>(a); }

/* wide-band CYK: four blocks in flight per step, for models whose targets
 * give fewer chunks (of at least 2W residues) than the chip holds wavefronts */
template <bool VSC, int G = 0>
__global__ void __launch_bounds__(64 * WPB_MAX) __attribute__((amdgpu_waves_per_eu(4))) std_cyk_wide_kernel(const ScanArgs *a)
{ std_scan_body<SEMI_CYK | G, VSC, false, 4>(a); }

/* rows per state visit of Inside with the LDS table (one 16-wavefront
 * workgroup per CU: registers for four rows) */
#define INFX_NR_INSIDE_LDS 4
static_assert(INFX_NR_INSIDE_LDS <= INFX_NR_MAX && INFX_NR_CYK <= INFX_NR_MAX, "BEGL deck rings allow for INFX_NR_MAX rows per visit");

template <int S, bool VSC>
__global__ void __launch_bounds__(64 * WPB_MAX) std_inside_lds_kernel(const ScanArgs *a)
{ std_scan_body<S, VSC, true, 1, false, INFX_NR_INSIDE_LDS>(a); }

/* batched launches: chunks of several models' windows in one launch
 * (chunk table rows of four: first row, last row, window start, model) */
__global__ void __launch_bounds__(64 * WPB_MAX) __attribute__((amdgpu_waves_per_eu(8))) std_cyk_batch_kernel(const ScanArgs *a)
{ std_scan_body<SEMI_CYK, false, false, 1, true>(a); }

template <int S>
__global__ void __launch_bounds__(64 * WPB_MAX) std_inside_batch_kernel(const ScanArgs *a)
{ std_scan_body<S, false, true, 1, true>(a); }

template <int S, bool VSC>
__global__ void __launch_bounds__(64 * WPB_MAX) __attribute__((amdgpu_waves_per_eu(8))) std_inside_l2_kernel(const ScanArgs *a)
{ std_scan_body<S, VSC, false>(a); }

/* ------------------------------------------------------------------------ */
/* launch stubs for scan_host.cpp                                            */

static const void *std_kernel_ptr(int semi, int want_vsc, int tbl_lds, int gacc, int wide)
{
  if (wide) {   /* CYK only */
    if (semi != SEMI_CYK) return NULL;
    if (gacc) return want_vsc ? (const void *) std_cyk_wide_kernel<true, SEMI_GACC> : (const void *) std_cyk_wide_kernel<false, SEMI_GACC>;
    return want_vsc ? (const void *) std_cyk_wide_kernel<true> : (const void *) std_cyk_wide_kernel<false>;
  }
  if (gacc) {   /* global accumulators: CYK, and Inside with the table in LDS */
    if (semi == SEMI_CYK) return want_vsc ? (const void *) std_cyk_kernel<true, SEMI_GACC> : (const void *) std_cyk_kernel<false, SEMI_GACC>;
    if (!tbl_lds) return NULL;
    if (semi == SEMI_FINSIDE) return want_vsc ? (const void *) std_inside_lds_kernel<SEMI_FINSIDE | SEMI_GACC, true> : (const void *) std_inside_lds_kernel<SEMI_FINSIDE | SEMI_GACC, false>;
    if (semi == SEMI_IINSIDE) return want_vsc ? (const void *) std_inside_lds_kernel<SEMI_IINSIDE | SEMI_GACC, true> : (const void *) std_inside_lds_kernel<SEMI_IINSIDE | SEMI_GACC, false>;
    return NULL;
  }
  if (semi == SEMI_CYK) return want_vsc ? (const void *) std_cyk_kernel<true> : (const void *) std_cyk_kernel<false>;
  if (semi == SEMI_FINSIDE) {
    if (tbl_lds) return want_vsc ? (const void *) std_inside_lds_kernel<SEMI_FINSIDE, true> : (const void *) std_inside_lds_kernel<SEMI_FINSIDE, false>;
    return want_vsc ? (const void *) std_inside_l2_kernel<SEMI_FINSIDE, true> : (const void *) std_inside_l2_kernel<SEMI_FINSIDE, false>;
  }
  if (semi == SEMI_IINSIDE) {
    if (tbl_lds) return want_vsc ? (const void *) std_inside_lds_kernel<SEMI_IINSIDE, true> : (const void *) std_inside_lds_kernel<SEMI_IINSIDE, false>;
    return want_vsc ? (const void *) std_inside_l2_kernel<SEMI_IINSIDE, true> : (const void *) std_inside_l2_kernel<SEMI_IINSIDE, false>;
  }
  return NULL;
}

/* rows per state visit (NR) of a launch: the slot holds a ring of 2*NR rows
 * and NR rows of local-begin maxima per wavefront (batched launches: 2) */
extern "C" int infx_std_rows(int semi, int tbl_lds, int wide)
{
  if (wide) return 2;
  if (semi == SEMI_CYK) return INFX_NR_CYK;
  return tbl_lds ? INFX_NR_INSIDE_LDS : 2;
}

/* LDS bytes of a workgroup of `wpb` wavefronts (gacc: accumulators in the
 * global slot, LDS holds only the Inside log-sum table) */
extern "C" size_t infx_std_lds_bytes(int semi, int tbl_lds, int wpb, int acc_n, int gacc, int wide)
{
  return (semi != SEMI_CYK && tbl_lds ? (size_t) INFX_LOGSUM_PAD * 4 : 0) +
         (gacc ? 0 : (size_t) wpb * acc_n * 4 * infx_std_rows(semi, tbl_lds, wide));   /* NR rows of 4-byte maxima */
}

/* resident workgroups per CU for this launch shape (0 if it cannot run) */
extern "C" int infx_std_occupancy(int semi, int want_vsc, int tbl_lds, int wpb, int acc_n, int gacc, int wide)
{
  const void *f = std_kernel_ptr(semi, want_vsc, tbl_lds, gacc, wide);
  if (!f || wpb < 1 || wpb > WPB_MAX) return 0;
  const size_t lds = infx_std_lds_bytes(semi, tbl_lds, wpb, acc_n, gacc, wide);
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) != hipSuccess) { (void) hipGetLastError(); return 0; }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * wpb, lds) != hipSuccess) { (void) hipGetLastError(); return 0; }
  return nb;
}

extern "C" hipError_t infx_launch_std(int semi, int want_vsc, int tbl_lds, int gacc, int wide, dim3 grid, dim3 block,
                                      hipStream_t stream, const ScanArgs *a, int acc_n)
{
  const void *f = std_kernel_ptr(semi, want_vsc, tbl_lds, gacc, wide);
  if (!f || block.x % 64 || block.x / 64 > WPB_MAX) return hipErrorInvalidValue;
  const size_t lds = infx_std_lds_bytes(semi, tbl_lds, block.x / 64, acc_n, gacc, wide);
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
  if (e != hipSuccess) return e;
  void *args[] = { (void *) &a };
  e = hipLaunchKernel(f, grid, block, args, lds, stream);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

/* batched launches of several models (CYK; Inside with the table in LDS) */
static const void *std_batch_kernel_ptr(int semi)
{
  if (semi == SEMI_CYK) return (const void *) std_cyk_batch_kernel;
  if (semi == SEMI_FINSIDE) return (const void *) std_inside_batch_kernel<SEMI_FINSIDE>;
  if (semi == SEMI_IINSIDE) return (const void *) std_inside_batch_kernel<SEMI_IINSIDE>;
  return NULL;
}

/* LDS bytes of a batched workgroup (two rows per state visit) */
extern "C" size_t infx_std_batch_lds_bytes(int semi, int wpb, int acc_n)
{
  return (semi != SEMI_CYK ? (size_t) INFX_LOGSUM_PAD * 4 : 0) + (size_t) wpb * acc_n * 4 * 2;
}

extern "C" int infx_std_batch_occupancy(int semi, int wpb, int acc_n)
{
  const void *f = std_batch_kernel_ptr(semi);
  if (!f || wpb < 1 || wpb > WPB_MAX) return 0;
  const size_t lds = infx_std_batch_lds_bytes(semi, wpb, acc_n);
  if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds) != hipSuccess) { (void) hipGetLastError(); return 0; }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 64 * wpb, lds) != hipSuccess) { (void) hipGetLastError(); return 0; }
  return nb;
}

extern "C" hipError_t infx_launch_std_batch(int semi, dim3 grid, dim3 block, hipStream_t stream, const ScanArgs *a, int acc_n)
{
  const void *f = std_batch_kernel_ptr(semi);
  if (!f || block.x % 64 || block.x / 64 > WPB_MAX) return hipErrorInvalidValue;
  const size_t lds = infx_std_batch_lds_bytes(semi, block.x / 64, acc_n);
  hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
  if (e != hipSuccess) return e;
  void *args[] = { (void *) &a };
  e = hipLaunchKernel(f, grid, block, args, lds, stream);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

/* one 64-bit counter set to zero by a kernel (a graph replays kernels
 * reliably; a counter reset is a kernel node, not a memset node) */
__global__ void zero_u64_kernel(unsigned long long *p)
{
  if (threadIdx.x == 0) *p = 0ull;
}
extern "C" hipError_t infx_launch_zero_u64(unsigned long long *p, hipStream_t stream)
{
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, stream, p);
  return hipGetLastError();
}

/* n <= 4096 bytes set to zero by a kernel: the zero pads around an uploaded
 * target (the scanners issue no memset nodes into their stream) */
__global__ void zero_bytes_kernel(unsigned char *p, unsigned n)
{
  for (unsigned i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}
extern "C" hipError_t infx_launch_zero_bytes(void *p, size_t n, hipStream_t stream)
{
  if (n == 0) return hipSuccess;
  if (n > 4096) return hipErrorInvalidValue;
  hipLaunchKernelGGL(zero_bytes_kernel, dim3(1), dim3(64), 0, stream, (unsigned char *) p, (unsigned) n);
  return hipGetLastError();
}

/* ------------------------------------------------------------------------ */
/* Scratch of the no-allocation scanner (infx_FastCYKScanNoAlloc): one fixed
 * device array per GPU, held until the process ends.  In host code the name
 * `scratch` is only a stand-in: its device address comes from
 * hipGetSymbolAddress, looked up once per device with that device current,
 * and only that address is passed to kernels and copies. */
static __device__ __align__(256) unsigned char scratch[INFX_NOALLOC_BYTES];

#include <mutex>
#define INFX_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[INFX_MAX_DEVICES];

int infx_noalloc_scratch(void **ret_ptr, size_t *ret_bytes, char *errbuf)
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= INFX_MAX_DEVICES) {
    (void) hipGetLastError();
    if (errbuf) snprintf(errbuf, 128, "no-allocation scratch: no current HIP device");
    return 100;   /* eslENODEVICE */
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = NULL;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      if (errbuf) snprintf(errbuf, 128, "no-allocation scratch: hipGetSymbolAddress failed (%s)", hipGetErrorString(e));
      return 12;   /* eslESYS */
    }
    g_scratch_ptr[dev] = p;
  }
  *ret_ptr = g_scratch_ptr[dev];
  *ret_bytes = INFX_NOALLOC_BYTES;
  return 0;
}
