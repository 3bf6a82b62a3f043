/* scan_kernels.h -- device-side data layout shared by the host launcher and
 * the gfx950 kernels of the standard (non-truncated) QDB scanners.
 *
 * Storage of the DP values (one "slot" per resident wavefront):
 *   row[2][R]            the rolling j / j-1 alpha rows of every non-BEGL
 *                        state; state v holds cells d = lo_v..hi_v at
 *                        row[base_v + d]  (base_v = offset_v - lo_v)
 *   deck[W+1][RB]        the BEGL_S decks, ring over j (row j % (W+1))
 *   acc_sc[W+1], acc_r   per-d best local-begin candidate of the current j
 *   vsc[M]               per-state running max (only when requested)
 * lo_v..hi_v covers the band and every cell a parent can read, so reads need
 * no bounds checks: cells outside the band hold their initial value
 * (IMPOSSIBLE, or the d = 0 value), exactly as the reference's full
 * alpha[2][M][W+1] would.
 */
#ifndef INFX_SCAN_KERNELS_H
#define INFX_SCAN_KERNELS_H

#include <stdint.h>

#define EMITLEFT  0
#define EMITRIGHT 1
#define EMITPAIR  2
#define EMITNONE  3

enum { KIND_GENERIC = 0, KIND_IL = 1, KIND_B = 2, KIND_BEGL = 3, KIND_SKIP = 4 };
enum { SEMI_CYK = 0, SEMI_FINSIDE = 1, SEMI_IINSIDE = 2 };

/* 128 bytes: two scalar-cache lines, read with uniform (SGPR) loads. */
struct DState {
  int32_t kind;
  int32_t emitmode;
  int32_t sd;          /* StateDelta */
  int32_t rdelta;      /* children read from the j-1 row */
  int32_t cnum;        /* children (B: unused) */
  int32_t dn, dx;      /* band at steady state (jp >= W); per-j dx = min(jp, dx) */
  int32_t base;        /* storage base (row, or deck row for BEGL) */
  int32_t esc_off;     /* emission table offset, -1 if non-emitting */
  int32_t has_begin;   /* finite local-begin score */
  int32_t kdn_y, kdx_y, kdn_w, kdx_w;  /* B: min(dmin,W)/min(dmax,W) of BEGR y and BEGL w */
  int32_t banded;      /* B: QDB k bounds (else 0..d) */
  int32_t init_ok;     /* endsc is finite: init(d) = el*d + endsc */
  uint32_t tsc[6];     /* float or int bits */
  uint32_t beginsc;    /* float or int bits */
  uint32_t endsc;      /* float or int bits */
  int32_t child_base[6];  /* storage base of child k (B: [0] = BEGL deck base, [1] = BEGR row base) */
  int32_t pad[2];
};

struct Cand {
  int32_t j, d;
  float   sc;
  int32_t r;
};

struct ScanArgs {
  const DState  *st;        /* [M] */
  const void    *esc;       /* emission scores, float or int */
  const uint8_t *dsq;       /* dsq[x - dsq_off] for x in [i0, j0] */
  int64_t        dsq_off;
  const void    *row_tmpl;  /* [R] */
  const void    *deck_tmpl; /* [RB] */
  const void    *lsum;      /* log-sum table (Inside) */
  char          *scratch;
  int64_t        slot_bytes;
  int            M, W, R, RB;
  int64_t        i0, j0;
  int64_t        chunk_len;
  int            nchunks, nslots;
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
};

#endif
