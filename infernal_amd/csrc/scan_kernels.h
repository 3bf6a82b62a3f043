/* scan_kernels.h -- device-side data layout shared by the host launchers and
 * the gfx950 kernels of the QDB scanners.
 *
 * Standard scanners.  Device memory of one resident wavefront ("slot"):
 *   guard[G]             G >= W + 128 cells that absorb the out-of-band deck
 *                        reads of masked lanes in the bifurcation k-loop
 *   deck[RB]             BEGL_S decks: state w keeps a ring of ring_w rows
 *                        (a power of two >= the largest split k its B
 *                        parent uses, + 1) of stride_w cells; the row of
 *                        chunk position jp is (jp & (ring_w - 1)) and cell d
 *                        is deck[base_w + row * stride_w + d]
 *   guard[G]
 *   row[2NR][R]          a ring over j of the rows of every other state (NR
 *                        rows per state visit: rows j-1 .. j+NR-1 live);
 *                        state v holds cells d = lo_v..hi_v at row[base_v + d]
 *   vsc[M]               per-state running max (only when requested)
 * lo_v..hi_v covers the band and every cell a parent can read, so reads need
 * no bounds checks: cells outside the band hold their initial value
 * (IMPOSSIBLE, or the d = 0 value), exactly as the reference's full
 * alpha[2][M][W+1] would.
 * LDS: per workgroup the log-sum table (Inside), per wavefront the local-begin
 * maxima of the NR rows of a visit, acc_n cells each for d = root_dn ..
 * root_dn + acc_n - 1 (or in the slot, after vsc, for wide root bands).
 */
#ifndef INFX_SCAN_KERNELS_H
#define INFX_SCAN_KERNELS_H

#include <stdint.h>

#define EMITLEFT  0
#define EMITRIGHT 1
#define EMITPAIR  2
#define EMITNONE  3

#define INFX_LOGSUM_N 23000   /* entries of the FLogsum / ILogsum tables */
#define INFX_LOGSUM_PAD 23008 /* device copies: entries 23000.. are 0 (FLogsum's 'return max' index) */
#define INFX_NOALLOC_BYTES ((size_t) 256 << 20)   /* the no-allocation scanner's fixed scratch per GPU */
#define INFX_NR_MAX 4        /* most rows a standard kernel computes per state visit (BEGL deck rings allow for it) */

enum { KIND_GENERIC = 0, KIND_IL = 1, KIND_B = 2, KIND_BEGL = 3, KIND_SKIP = 4 };
enum { SEMI_CYK = 0, SEMI_FINSIDE = 1, SEMI_IINSIDE = 2 };

/* 64 bytes: one scalar load.  hdr bits: kind 0-2, emitmode 3-4, sd 5-6,
 * rdelta 7 (children read from the j-1 row), cnum 8-10, has_begin 11,
 * init_ok 12 (endsc finite: init(d) = el*d + endsc), banded 13 (B: QDB k
 * bounds, else 0..d), self_prev 14 (the state is its own child on the j-1
 * row: IR), state index v 16-31. */
struct DState {
  uint32_t hdr;
  uint32_t dnx;        /* dn | dx << 16: band at steady state; per-j dx = min(jp, dx) */
  int32_t  base;       /* storage base (row, or deck ring for BEGL) */
  uint32_t aux;        /* emitting: esc_off; BEGL, B: ring_mask | ring_stride << 16 (B: of its BEGL) */
  uint32_t beginsc;    /* float or int bits */
  uint32_t endsc;      /* float or int bits */
  uint32_t tsc[6];     /* float or int bits.  B: [0] kdn_y | kdx_y << 16, [1] kdn_w | kdx_w << 16
                          (min(dmin,W) / min(dmax,W) of BEGR y and BEGL w), [2] BEGL deck base,
                          [3] BEGR row base */
  int32_t  cbase;      /* child k storage base = cbase + coff_k */
  uint32_t coff[3];    /* int16 coff_k: low half (k even) / high half (k odd) of coff[k/2] */
};

struct Cand {
  int32_t j, d;
  float   sc;
  int32_t r;
};

/* launch parameters; the kernel reads them from device memory */
struct ScanArgs {
  const DState  *st;        /* [nst]: every state but E and ROOT_S, v descending */
  int            nst;
  const void    *esc;       /* emission scores, float or int */
  const uint8_t *dsq;       /* dsq[x - dsq_off] for x in [i0, j0] */
  int64_t        dsq_off;
  const void    *row_tmpl;  /* [R] */
  const void    *deck_tmpl; /* [RB] */
  const void    *lsum;      /* log-sum table (Inside) */
  char          *scratch;
  int64_t        slot_bytes;
  int64_t        deck_off, row_off, vsc_off;   /* byte offsets inside a slot */
  int64_t        acc_off;     /* local-begin accumulators in the slot (global-accumulator kernels) */
  int            M, W, R, RB;
  int64_t        i0, j0;
  int64_t        chunk_len;
  const int64_t *chunk_tab; /* null: chunk c covers i0 + c*chunk_len ..; else [3c..3c+2] = first row, last row, window start
                               (batched launches: [4c..4c+3], the fourth the model's index in args_tab) */
  int            nchunks, nslots;
  int            acc_n;       /* local-begin accumulators per row and wavefront */
  /* root state */
  int            root_dn, root_dx;
  int            root_cnum;
  int            root_child_base[6];
  uint32_t       root_tsc[6];
  int            local_begin;
  /* init-score slope (float el_selfsc or int iel_selfsc bits) */
  uint32_t       el_selfsc;
  /* hit reporting */
  int            cand_mode;   /* 0: sc >= thr; 1: record filter (greedy, no NULL3); 2: sc > 0 (gamma) */
  float          cutoff;
  float          env_cutoff;
  Cand          *cand;
  unsigned long long *ncand;
  int64_t        cand_cap;
  float         *chunk_best;  /* [nchunks] */
  int64_t       *chunk_envi;  /* [nchunks] */
  int64_t       *chunk_envj;  /* [nchunks] */
  int            want_vsc;
  const struct ScanArgs *args_tab;   /* batched launches: each model's parameters */
};

/* ---------------------------------------------------------------------- */
/* truncated scanners                                                      */

enum { TSEMI_CYK = 0, TSEMI_IINSIDE = 1 };
#define TRMODE_T_K 0
#define TRMODE_R_K 1
#define TRMODE_L_K 2
#define TRMODE_J_K 3
#define TRMODE_UNKNOWN_K 4

/* 128 bytes */
struct TState {
  int32_t kind;
  int32_t emitmode;
  int32_t cnum;
  int32_t dx;          /* band max (dn = 1) */
  int32_t base;        /* J/L/R rows (or deck rows for BEGL): cell d at base + d */
  int32_t tbase;       /* T rows (B states) */
  int32_t esc_off;     /* oesc/ioesc row */
  int32_t lesc_off;    /* lmesc row (MP) */
  int32_t resc_off;    /* rmesc row (MP) */
  int32_t detached;
  int32_t first_r;     /* first child used in the L/R marginal (1 for IL/IR: no self transition) */
  int32_t init_ok;
  int32_t banded;
  int32_t kdx_y, kdx_w;
  uint32_t endsc;
  uint32_t tsc[6];
  int32_t child_base[6];   /* B: [0] BEGL deck base, [1] BEGR row base, [2] BEGL ring mask, [3] BEGL ring stride */
  int32_t ring_mask;       /* BEGL: deck ring rows - 1 (a power of two minus one) */
  int32_t ring_stride;     /* BEGL: cells per deck ring row */
  int32_t level;           /* same-row dependency level (scheduling) */
  int32_t pad;
};

/* one unit of a row's work: 64-cell blocks b0..b1-1 of state v */
struct TrItem {
  int32_t v, b0, b1;
};

/* a state with a finite truncated-begin penalty, in ascending order.
 * flags: bit 0 a B state, bit 1 its node's first consensus position is 1,
 * bit 2 its last is clen (cm_hit_AllowTruncation's lpos / rpos tests) */
struct TrPen {
  int32_t y, base, tbase, dx, flags;
  uint32_t pen;
};

struct TrWindow {
  int64_t i0, j0;
  int64_t dsq_off;     /* device dsq + dsq_off, indexed by absolute j */
  int32_t id, pad;     /* the caller's window index (windows run longest first) */
};

struct TrCand {
  int32_t win, j, d;
  float   sc;
  int32_t r, mode;
};

struct TrScanArgs {
  const TState  *st;
  const void    *esc;
  const uint8_t *dsq;
  const void    *row_tmpl;   /* [3][R] J, L, R */
  const void    *deck_tmpl;  /* [3][RB] */
  const void    *lsum;
  const TrPen   *pen;
  int            npen;
  const int     *pen_blk;    /* [nblk + 1] offsets, then per 64-cell root block b the
                                indices (ascending) of the begins whose band reaches d = 64b + 1 */
  int            pen_nblk;
  const TrItem  *items;      /* a row's work, level by level; within a level, wave by wave */
  const int     *sched;      /* [nlev * nw + 1]: items of (level l, wave w) are items[sched[l*nw+w] .. sched[l*nw+w+1]) */
  int            nlev, nw;
  const TrWindow *win;       /* longest first */
  int            nwin;
  unsigned      *next_win;   /* work counter: wavefronts take windows in order */
  char          *scratch;
  int64_t        slot_bytes;
  int            nslots;
  int            M, W, R, RB, RT;
  int            root_dx;
  int            fill_L, fill_R, fill_T;
  uint32_t       el_selfsc;
  int            cand_mode;   /* 0: sc >= cutoff; 1: greedy without NULL3: sc >= cutoff, admitted by the
                                 pass (cm_hit_AllowTruncation) and above every admitted d < this one
                                 of its 64-cell block; 2: sc > 0 (gamma) */
  int            pass_any;    /* the pass admits every hit */
  int            local;       /* local begins on */
  float          cutoff, env_cutoff;
  TrCand        *cand;
  unsigned long long *ncand;
  int64_t        cand_cap;
  float         *win_best;
  int           *win_mode;
  int64_t       *win_envi, *win_envj;
  float         *win_vsc;     /* [nwin][M] */
};

#endif
