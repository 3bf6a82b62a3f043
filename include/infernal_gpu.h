/* infernal_gpu.h -- C ABI of the MI355X QDB scanning CYK/Inside library.
 *
 * The five scanner entry points keep the argument lists of Infernal 1.1.5
 * (src/infernal.h:2710-2730) so that src/cm_pipeline.c:pli_dispatch_cm_search
 * (:4187-4210) and src/cmcalibrate.c:process_search_workunit (:2337-2358) can
 * call them unchanged.  The structures below are this library's own mirrors of
 * the fields those scanners read from Infernal's CM_t, CM_SCAN_MX,
 * CM_TR_SCAN_MX and CM_TOPHITS; they carry no Easel types.
 *
 * All scanners run on the GPU (gfx950).  There is no CPU fallback: when no
 * usable HIP device is present a scanner returns eslENODEVICE with a message in
 * errbuf.
 */
#ifndef INFERNAL_GPU_H
#define INFERNAL_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------
 * Status codes.  Values follow Easel's easel.h (not vendored with the
 * reference; numbers taken from Easel's published header, to verify against a
 * real Easel checkout).
 */
#define eslOK           0
#define eslFAIL         1
#define eslEOF          3
#define eslEMEM         5
#define eslENOTFOUND    6
#define eslEFORMAT      7
#define eslEINCOMPAT   10
#define eslEINVAL      11
#define eslESYS        12
#define eslEINCONCEIVABLE 14
#define eslERANGE      16
#define eslENODEVICE  100   /* this library only: no usable GPU */
#define eslERRBUFSIZE 128

typedef uint8_t ESL_DSQ;

/* RNA alphabet, Easel layout "ACGU-RYMKSWHBVDN*~": K=4 canonical residues,
 * code 4 gap, 5..15 degenerate, 16 '*' (no residue), 17 '~' (missing). */
#define INFX_K   4
#define INFX_KP 18

/* Score constants (src/infernal.h:148-177). */
#define IMPOSSIBLE   -1e36
#define INFTY        987654321
#define INTSCALE     1000.0f
#define LOGSUM_TBL   23000
#define NOT_IMPOSSIBLE(x) ((x) > -9.999e35)
#define MAXCONNECT   6

/* State types (src/infernal.h:193-202). */
#define D_st  0
#define MP_st 1
#define ML_st 2
#define MR_st 3
#define IL_st 4
#define IR_st 5
#define S_st  6
#define E_st  7
#define B_st  8
#define EL_st 9

/* Node types (src/infernal.h:208-216). */
#define DUMMY_nd -1
#define BIF_nd    0
#define MATP_nd   1
#define MATL_nd   2
#define MATR_nd   3
#define BEGL_nd   4
#define BEGR_nd   5
#define ROOT_nd   6
#define END_nd    7

/* Unique state ids (src/infernal.h:221-243). */
#define ROOT_S  0
#define ROOT_IL 1
#define ROOT_IR 2
#define BEGL_S  3
#define BEGR_S  4
#define BEGR_IL 5
#define MATP_MP 6
#define MATP_ML 7
#define MATP_MR 8
#define MATP_D  9
#define MATP_IL 10
#define MATP_IR 11
#define MATL_ML 12
#define MATL_D  13
#define MATL_IL 14
#define MATR_MR 15
#define MATR_D  16
#define MATR_IR 17
#define END_E   18
#define BIF_B   19
#define END_EL  20

/* Emit modes (src/infernal.h:180-185). */
#define EMITLEFT  0
#define EMITRIGHT 1
#define EMITPAIR  2
#define EMITNONE  3

/* cm->flags (src/infernal.h:1927-1939, subset). */
#define CMH_BITS         (1<<0)
#define CMH_LOCAL_BEGIN  (1<<10)
#define CMH_LOCAL_END    (1<<11)
#define CM_IS_CONFIGURED (1<<23)

/* cm->config_opts (src/infernal.h:1953-1962, subset). */
#define CM_CONFIG_LOCAL     (1<<0)
#define CM_CONFIG_SCANMX    (1<<6)
#define CM_CONFIG_TRSCANMX  (1<<7)
#define CM_CONFIG_TRUNC     (1<<5)

/* cm->search_opts (src/infernal.h:1990-2001, subset). */
#define CM_SEARCH_INSIDE       (1<<5)
#define CM_SEARCH_CMNOTGREEDY  (1<<8)
#define CM_SEARCH_NULL3        (1<<11)

/* QDB indices (src/infernal.h:1607-1609). */
#define SMX_NOQDB       0
#define SMX_QDB1_TIGHT  1
#define SMX_QDB2_LOOSE  2
#define NSMX_QDB_IDX    3

/* Truncation modes and penalties (src/infernal.h:341-345, :1723-1727). */
#define TRMODE_T        0
#define TRMODE_R        1
#define TRMODE_L        2
#define TRMODE_J        3
#define TRMODE_UNKNOWN  4
#define TRPENALTY_5P_AND_3P 0
#define TRPENALTY_5P_ONLY   1
#define TRPENALTY_3P_ONLY   2
#define NTRPENALTY          3
#define TRPENALTY_NONE     -1

/* Pipeline pass indices (src/infernal.h:2097-2103). */
#define PLI_PASS_CM_SUMMED       0
#define PLI_PASS_STD_ANY         1
#define PLI_PASS_5P_ONLY_FORCE   2
#define PLI_PASS_3P_ONLY_FORCE   3
#define PLI_PASS_5P_AND_3P_FORCE 4
#define PLI_PASS_5P_AND_3P_ANY   5
#define PLI_PASS_HMM_ONLY_ANY    6

/* Hit flags (src/infernal.h:2406-2410, subset). */
#define CM_HIT_FLAGS_DEFAULT         0
#define CM_HIT_IS_INCLUDED           (1<<0)
#define CM_HIT_IS_REPORTED           (1<<1)
#define CM_HIT_IS_REMOVED_DUPLICATE  (1<<2)

/* ---------------------------------------------------------------------------
 * Model.  Field names follow Infernal's CM_t (src/infernal.h:1759-1924).
 * Two-dimensional arrays are row pointers into one contiguous block, as in
 * CreateCMBody (src/cm.c:179).
 */
typedef struct cm_qdbinfo_s {
  int    *dmin1, *dmax1;   /* [0..M-1] tight bands (beta1) */
  int    *dmin2, *dmax2;   /* [0..M-1] loose bands (beta2) */
  double  beta1, beta2;
} CM_QDBINFO;

typedef struct cm_emitmap_s {
  int  clen;
  int *lpos, *rpos, *epos; /* [0..nodes-1] */
} CMEmitMap_t;

typedef struct cm_tr_penalties_s {
  int     M;
  int     ignored_inserts;
  float **g_ptyAA;         /* [NTRPENALTY][0..M-1] global truncated-begin scores */
  float **l_ptyAA;         /* [NTRPENALTY][0..M-1] local truncated-begin scores */
  int   **ig_ptyAA;        /* integer (x1000) versions */
  int   **il_ptyAA;
} CM_TR_PENALTIES;

struct cm_scan_mx_s;
struct cm_tr_scan_mx_s;

typedef struct cm_s {
  char   *name;
  int     M, nodes, clen, W;
  float   pbegin, pend;
  double  beta_W;
  float   null3_omega;
  float  *null;            /* [K] background */
  char   *sttype;          /* [M] */
  char   *stid;            /* [M] */
  int    *ndidx;           /* [M] */
  int    *cfirst, *cnum;   /* [M] */
  int    *plast, *pnum;    /* [M] */
  int    *nodemap;         /* [nodes] */
  char   *ndtype;          /* [nodes] */
  float **t, **e;          /* [M][MAXCONNECT], [M][K*K] probabilities */
  float  *begin, *end;     /* [M] local begin/end probabilities */
  float **tsc, **esc;      /* log-odds bits */
  float **lmesc, **rmesc;  /* [M][Kp] marginal emission scores (truncation) */
  float  *beginsc, *endsc;
  int   **itsc, **iesc, **ilmesc, **irmesc;
  int    *ibeginsc, *iendsc;
  float **oesc;            /* [M] -> Kp (singlets) or Kp*Kp (pairs); NULL for non-emitters */
  int   **ioesc;
  float   el_selfsc;
  int     iel_selfsc;
  float  *root_trans;      /* saved t[0][] before localization */
  int     flags, config_opts, search_opts;
  CM_QDBINFO      *qdbinfo;
  CMEmitMap_t     *emap;
  CM_TR_PENALTIES *trp;
  struct cm_scan_mx_s    *smx;
  struct cm_tr_scan_mx_s *trsmx;
} CM_t;

/* Scan matrix.  Mirrors CM_SCAN_MX (src/infernal.h:1600-1650): the band
 * tables dnAAA/dxAAA are filled exactly as cm_scan_mx_Create
 * (src/cm_mx.c:5725-5739) does; the DP matrices themselves live in device
 * memory owned by `dev`. */
typedef struct cm_scan_mx_s {
  int    M, W;
  int ***dnAAA, ***dxAAA;  /* [qdbidx][1..W][0..M-1] */
  int    floats_valid, ints_valid;
  float *bestsc;           /* [0..W] */
  int   *bestr;            /* [0..W] */
  void  *dev;              /* device resources, owned by this library */
} CM_SCAN_MX;

/* Truncated scan matrix.  Mirrors CM_TR_SCAN_MX; bands as cm_tr_scan_mx_Create
 * (src/cm_mx.c:6428-6441): dn == 1 for every state, only dmax is used. */
typedef struct cm_tr_scan_mx_s {
  int    M, W;
  int ***dnAAA, ***dxAAA;
  int    floats_valid, ints_valid;
  float *bestsc;
  int   *bestr;
  char  *bestmode;
  void  *dev;
} CM_TR_SCAN_MX;

/* Hits.  Mirrors CM_HIT / CM_TOPHITS (src/infernal.h:2424-2480), fields the
 * scanners set. */
typedef struct cm_hit_s {
  int64_t start, stop;
  int64_t srcL;
  int64_t hit_idx;
  int64_t cm_idx, seq_idx;
  int     pass_idx;
  int     in_rc;
  int     root;
  char    mode;
  float   score;
  int     hmmonly;
  int     flags;
} CM_HIT;

typedef struct cm_tophits_s {
  CM_HIT **hit;            /* sorted view */
  CM_HIT  *unsrt;          /* storage, in creation order */
  uint64_t Nalloc;
  uint64_t N;
  int      is_sorted_for_overlap_removal;
} CM_TOPHITS;

/* ---------------------------------------------------------------------------
 * Model input and configuration (host).
 */

/* Read the first model of an INFERNAL1/a save file.  Restates
 * read_asc_1p1_cm (src/cm_file.c:1539-2118) + CMRenormalize (src/cm.c:344) +
 * CreateEmitMap (src/display.c:1110).  The filter HMM that follows the model
 * is skipped. */
int  infx_cm_Read(const char *path, CM_t **ret_cm, char *errbuf);
/* Same, from an in-memory text buffer. */
int  infx_cm_ReadText(const char *text, CM_t **ret_cm, char *errbuf);
void FreeCM(CM_t *cm);

/* cm_Configure (src/cm_modelconfig.c:98, :139-373) for the options the
 * scanners need: CM_CONFIG_LOCAL (cm_localize :467-548), CM_CONFIG_TRUNC
 * (cm_tr_penalties_Create src/cm_trunc.c:51), CM_CONFIG_SCANMX and
 * CM_CONFIG_TRSCANMX (scan matrices), then CMLogoddsify (src/cm.c:784).
 * W_from_cmdline is -1 to keep the file's W.  QDBs are taken from the file. */
int  cm_Configure(CM_t *cm, char *errbuf, int W_from_cmdline);

/* cm_scan_mx_Create (src/cm_mx.c:5686) / cm_tr_scan_mx_Create (:6391). */
int  cm_scan_mx_Create(CM_t *cm, char *errbuf, int do_float, int do_int, CM_SCAN_MX **ret_smx);
void cm_scan_mx_Destroy(CM_t *cm, CM_SCAN_MX *smx);
int  cm_tr_scan_mx_Create(CM_t *cm, char *errbuf, int do_float, int do_int, CM_TR_SCAN_MX **ret_trsmx);
void cm_tr_scan_mx_Destroy(CM_t *cm, CM_TR_SCAN_MX *trsmx);

/* Log-sum tables (src/logsum.c:74-169). */
void  FLogsumInit(void);
float FLogsum(float s1, float s2);
void  init_ilogsum(void);
int   ILogsum(int s1, int s2);

/* Digitize an RNA/DNA text sequence into dsq[0..L+1] with sentinels (Easel
 * esl_abc_Digitize conventions: case-insensitive, T->U, X->N, '.'/'_'->gap).
 * Returns eslEINVAL on an unknown character. */
int  infx_abc_Digitize(const char *seq, int64_t L, ESL_DSQ *dsq);

/* ---------------------------------------------------------------------------
 * Hit lists (src/cm_tophits.c). */
CM_TOPHITS *cm_tophits_Create(void);
void        cm_tophits_Destroy(CM_TOPHITS *th);
int         cm_tophits_CreateNextHit(CM_TOPHITS *th, CM_HIT **ret_hit);
int         cm_tophits_SortForOverlapRemoval(CM_TOPHITS *th);
int         cm_tophits_RemoveOrMarkOverlaps(CM_TOPHITS *th, int do_clans_only, char *errbuf);
int         cm_tophits_CloneHitMostly(CM_TOPHITS *src_th, int h, CM_TOPHITS *dest_th);

/* ---------------------------------------------------------------------------
 * The scanners (GPU).  Contracts, return values and outputs as the reference
 * functions of the same name; hits are appended to <hitlist> after greedy
 * overlap removal (or gamma traceback when cm->search_opts has
 * CM_SEARCH_CMNOTGREEDY).  *ret_vsc, when requested, is malloc'ed and owned
 * by the caller.  Positions must stay below 2^31 - 1 (j0 <= 2147483646):
 * the device's hit candidates carry 32-bit positions, and a larger j0 is
 * refused with eslERANGE (cmsearch scans buffers of one sequence window,
 * whose positions stay below ~10^5).
 */

/* src/cm_dpsearch.c:90 FastCYKScan */
int FastCYKScan(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc);
/* src/cm_dpsearch.c:1731 FastFInsideScan */
int FastFInsideScan(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                    float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                    int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc);
/* src/cm_dpsearch.c:1106 FastIInsideScan */
int FastIInsideScan(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                    float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                    int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc);
/* src/cm_dpsearch_trunc.c:79 RefTrCYKScan */
int RefTrCYKScan(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                 int64_t i0, int64_t j0, float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                 int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, char *ret_mode, float *ret_sc);
/* src/cm_dpsearch_trunc.c:741 RefITrInsideScan */
int RefITrInsideScan(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                     int64_t i0, int64_t j0, float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                     int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, char *ret_mode, float *ret_sc);

/* src/cm_dpsearch_trunc.c:3566 cm_TrFillFromPassIdx, src/cm_trunc.c:464
 * cm_tr_penalties_IdxForPass, src/cm_pipeline.c:2396
 * cm_pli_PassAllowsTruncation, src/cm_tophits.c:3048 cm_hit_AllowTruncation */
int cm_TrFillFromPassIdx(int pass_idx, int *ret_fill_L, int *ret_fill_R, int *ret_fill_T);
int cm_tr_penalties_IdxForPass(int pass_idx);
int cm_pli_PassAllowsTruncation(int pass_idx);
int cm_hit_AllowTruncation(CM_t *cm, int pass_idx, int64_t start, int64_t stop, int64_t i0, int64_t j0, char mode, int b);

/* Replaces cm_CountSearchDPCalcs (src/cm_dpsearch.c:3099-3205): millions of
 * DP calculations per residue of a standard scan of L residues under the
 * bands dmin/dmax (both NULL: no bands), by the reference's definition (per j:
 * 1 + (cnum+1) cells (+cells if emitting) for a non-B state, 1 + k range per
 * B cell, (cnum+1) cells plus the local begins' cells for the root), float
 * sums in the reference's order; per sub-model in *ret_vcalcs (caller frees),
 * whole model in *ret_calcs.  cmcalibrate's forecast calls it
 * (src/cmcalibrate.c:2434-2442). */
int cm_CountSearchDPCalcs(CM_t *cm, char *errbuf, int L, int *dmin, int *dmax, int W, int correct_for_first_W,
                          float **ret_vcalcs, float *ret_calcs);
/* This library's own work measure (no counterpart in the reference; it is
 * not cm_CountSearchDPCalcs): (v,d) cells inside the band per residue at
 * steady state (jp >= W), E states excluded, and algorithmic VALU ops per
 * residue (2*cnum + 1 if emitting per non-B cell, 2*(kmax-kmin+1) per B cell),
 * the numerator of the benchmark's roofline. */
int infx_CountSearchCells(CM_t *cm, int qdbidx, int truncated, double *ret_cells_per_res, double *ret_ops_per_res);
/* As infx_CountSearchCells, plus the max (CYK) / log-sum (Inside) reductions
 * per residue: cnum per non-B cell, kmax-kmin+1 per B cell.  No counterpart
 * in the reference (the Inside roofline counts each reduction as 8 ops). */
int infx_CountSearchWork(CM_t *cm, int qdbidx, int truncated, double *ret_cells, double *ret_ops, double *ret_red);

/* Batched truncated scans (no counterpart in the reference, which makes one
 * RefTrCYKScan / RefITrInsideScan call per window, src/cm_pipeline.c:4187-4195):
 * nwin windows [wi0[w], wj0[w]] of one target dsq, each scanned exactly as an
 * independent call (inside = 0: RefTrCYKScan, 1: RefITrInsideScan), in one
 * launch.  Window w's hits go to hitlists[w] (if hitlists is non-NULL); its
 * best score, mode and envelope to ret_sc[w], ret_mode[w], ret_envi[w],
 * ret_envj[w] (each optional). */
int infx_TrScanWindows(CM_t *cm, char *errbuf, CM_TR_SCAN_MX *trsmx, int inside, int qdbidx, int pass_idx, ESL_DSQ *dsq,
                       int nwin, const int64_t *wi0, const int64_t *wj0, float cutoff, CM_TOPHITS **hitlists,
                       int do_null3, float env_cutoff, float *ret_sc, char *ret_mode, int64_t *ret_envi, int64_t *ret_envj);

/* Exact number of DP cells a scan of i0..j0 computes (E states excluded), from
 * the band tables of <mx> (a CM_SCAN_MX, or a CM_TR_SCAN_MX if truncated). */
int infx_count_scan_cells(const CM_t *cm, const void *mx, int qdbidx, int truncated, int64_t i0, int64_t j0, double *ret_cells);

/* ---------------------------------------------------------------------------
 * Library-level controls and introspection (no counterpart in the reference).
 */
int  infx_gpu_available(char *errbuf);                 /* 1 if a gfx9 HIP device can be used */
/* infx_FastScanBatch through a HIP graph (on != 0): the counter reset (a
 * kernel) and the batched kernel are captured once and the graph is replayed
 * by later calls whose launch shape and device buffers are the same; the
 * parameters are uploaded before each replay.  Results are identical to the
 * direct launch.  No counterpart in the reference. */
void infx_set_batch_graph(int on);
/* FastCYKScan (same arguments, results bit-identical) that allocates no
 * device memory: no hipMalloc, hipFree or allocating library call.  No
 * counterpart in the reference; it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * it).  Every device buffer of the call comes from one fixed 256 MiB array
 * per GPU that the library holds until the process ends; a model and target
 * whose buffers do not fit (profile, residues, chunk outputs, one DP slot and
 * a minimal candidate buffer) are refused with eslERANGE and a message, and
 * a scan whose candidates overflow the rest also returns eslERANGE.  Calls
 * share the array: the library serialises them on a lock, and work issued
 * into the array by other means would race with them. */
int infx_FastCYKScanNoAlloc(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                            float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                            int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc);
/* Keep a device copy of a target's residues dsq[1..L] (checked once here).
 * A later scan of the same buffer (the same dsq pointer, i0..j0 inside
 * 1..L) reads that copy in place instead of uploading its range, so a
 * caller that scans one sequence block several times (CYK filter, Inside,
 * truncated passes) moves it over PCIe once.  The caller must not change
 * dsq[1..L] while it is pinned; pinning another buffer replaces it.
 * eslEINVAL for a non-residue code, eslEMEM, eslENODEVICE.  No counterpart in
 * the reference (its scanners read host memory). */
int infx_target_Pin(const ESL_DSQ *dsq, int64_t L, char *errbuf);
int infx_target_Unpin(void);
/* Timing of the last scanner call on this thread: total wall ms around the
 * call, and the summed device duration of the DP kernel launches in it
 * (HIP events on the library's stream). */
void infx_last_call_timing(double *ret_wall_ms, double *ret_kernel_ms, int64_t *ret_cells, int *ret_launches);
/* The device scan profile the standard scanners would build for a target of
 * length L: up to `cap` 64-byte state descriptors (every state but E and
 * ROOT_S, v descending), the row and deck sizes in cells, the emission table
 * size, and the root band and children (ret_root[0..2] = dn, dx, cnum;
 * [3..8] = child storage bases).  Returns the descriptor count, or -status.
 * No counterpart in the reference: used to check the kernels' address
 * arithmetic on the host. */
int infx_debug_std_profile(CM_t *cm, CM_SCAN_MX *smx, int qdbidx, int semi, int64_t L, void *out_st, int cap,
                           int *ret_R, int *ret_RB, int *ret_nesc, int *ret_root, char *errbuf);
/* The device profile the truncated scanners would build (semi 0 RefTrCYKScan,
 * 1 RefITrInsideScan) for pass pass_idx: up to st_cap 128-byte state
 * descriptors (one per state), up to pen_cap 24-byte truncated-begin entries,
 * the per-64-cell-block begin lists (offsets, then entries), each state's
 * band of stored d cells (lohi[2v], lohi[2v+1]; lo > hi: none) and
 * dims[0..6] = row cells, deck-ring cells, T-row cells, emission entries,
 * root band max, begin entries, begin-list length.  Returns the descriptor count,
 * or -status.  No counterpart in the reference: used to check the truncated
 * kernel's address arithmetic on the host. */
int infx_debug_tr_profile(CM_t *cm, CM_TR_SCAN_MX *trsmx, int qdbidx, int semi, int pass_idx, void *out_st, int st_cap,
                          void *out_pen, int pen_cap, int *out_pblk, int pblk_cap, int *lohi, int *dims, char *errbuf);
/* Attach truncated-scanner device resources to trsmx as a GPU scan would
 * (nprof host-side profiles, no device memory), without a device; FreeCM or
 * cm_tr_scan_mx_Destroy must release them.  eslEINVAL if trsmx already has
 * resources or nprof is outside 0..8.  No counterpart in the reference: lets
 * a CPU test (under AddressSanitizer) check the truncated matrix's teardown. */
int infx_debug_attach_trdev(CM_t *cm, CM_TR_SCAN_MX *trsmx, int nprof);
/* The row schedule of the truncated kernel for nw (1..8) wavefronts per
 * window: each state's same-row dependency level is in the descriptors
 * (infx_debug_tr_profile); out_sched[nlev*nw + 1] holds the item offsets of
 * (level, wavefront) and out_items the items as (state, first block, end
 * block) triples.  Returns the item count, or -status; *ret_nlev = levels.
 * No counterpart in the reference: used to check the schedule on the host. */
int infx_debug_tr_schedule(CM_t *cm, CM_TR_SCAN_MX *trsmx, int qdbidx, int semi, int pass_idx, int nw, int want_vsc,
                           int *out_sched, int sched_cap, int *out_items, int items_cap, int *ret_nlev, char *errbuf);
/* The float truncated-begin penalties of an unconfigured (global) model, as
 * cm_tr_penalties_Create (src/cm_trunc.c:51) builds them with its
 * ignore_inserts flag set or not, into out[2][NTRPENALTY][M] (glocal, then
 * local); the model is not changed.  cm_Configure uses ignore_inserts = 0.
 * Lets a CPU test apply the reference's own check of the penalties,
 * cm_tr_penalties_Validate (src/cm_trunc.c:498), which assumes the flag. */
int infx_debug_tr_penalties(const CM_t *cm, int ignore_inserts, float *out, char *errbuf);
/* Query-dependent bands and W from the model's transition probabilities
 * (local ends off, standard local begins on).  Fills qdbinfo->dmin1/dmax1
 * (beta1) and dmin2/dmax2 (beta2) for every state and *ret_W (beta_W).
 *   infx_CalculateQueryDependentBands  replaces CalculateQueryDependentBands
 *                                      (src/cm_qdband.c:45): Z = 4*clen,
 *                                      doubled until the tail is negligible
 *   infx_BandCalculationEngine         replaces BandCalculationEngine
 *                                      (src/cm_qdband.c:163) for one Z;
 *                                      eslERANGE if Z is too small.  Its
 *                                      ret_gamma, as the reference's, returns
 *                                      the densities gamma_v(0..Z) (here into
 *                                      a caller buffer of M*(Z+1) doubles,
 *                                      or NULL) */
int infx_BandCalculationEngine(CM_t *cm, int Z, CM_QDBINFO *qdbinfo, double beta_W, int *ret_W, double *ret_gamma);
int infx_CalculateQueryDependentBands(CM_t *cm, char *errbuf, CM_QDBINFO *qdbinfo, double beta_W, int *ret_W, int *ret_Z);

/* Many windows wi[w]..wj[w] (sorted, disjoint, 1-based) of one target in one
 * launch, with FastCYKScan (semi 0), FastFInsideScan (1) or FastIInsideScan
 * (2) semantics: the hits appended to hitlist and *ret_sc are those of one
 * standard scan per window.  No counterpart in the reference: for callers
 * that scan many short windows (the pipeline's final stage, the learned
 * gate), where one launch per window would leave the GPU idle. */
int infx_FastScanWindows(int semi, CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq,
                         const int64_t *wi, const int64_t *wj, int nwin, float cutoff, CM_TOPHITS *hitlist,
                         int do_null3, float *ret_sc);

/* Several models in one launch: window w (wi[w]..wj[w] of dsq) is scanned
 * with model wm[w] (cms[wm[w]], smxs[wm[w]]); each model's windows are sorted
 * and disjoint, windows of different models may overlap.  Model m's hits
 * (greedy or gamma as its search_opts say, no NULL3, cutoff cutoffs[m]) are
 * appended to hitlists[m] and equal those of infx_FastScanWindows over its
 * own windows; ret_sc[m] is its best score.  No counterpart in the
 * reference: for callers scanning many models of uneven size (cmscan, a
 * model library against one sequence set) in one launch instead of one per
 * model. */
int infx_FastScanBatch(int semi, int nmodel, CM_t **cms, CM_SCAN_MX **smxs, char *errbuf, int qdbidx, ESL_DSQ *dsq,
                       const int64_t *wi, const int64_t *wj, const int32_t *wm, int nwin, const float *cutoffs,
                       CM_TOPHITS **hitlists, float *ret_sc);

/* Override the work split (0 = automatic): residues per chunk and waves per
 * block.  Used by the autotuner and tests. */
void infx_set_launch_params(int chunk_len, int waves_per_block);

/* ---- learned window gate (no counterpart in the reference) ------------
 * A per-model 1-D convolutional scorer over one-hot ACGU (4 convolutions of
 * width 9 and 128 channels with ReLU, then a per-position linear head), in
 * bf16 on the matrix cores with fp32 accumulation, and its training
 * kernels (width 9, dilation per layer).  It picks the stretches of a target
 * that go to the exact Inside scan; the plain scan remains the fallback and
 * the measure of what the gate misses.
 * Device pointers, default stream; return 0 or a HIP error code.
 * Activations are [position][channel]; a batch is equal-length segments of
 * `seg` positions laid end to end.  Weights: [128][kp] with kp = 64 (first
 * layer, 4 input channels) or 1152. */
/* device memory from the library's own runtime (synchronous) */
int infx_dev_malloc(size_t bytes, void **ptr);
int infx_dev_free(void *ptr);
int infx_dev_memset(void *ptr, int value, size_t bytes);
int infx_dev_upload(void *dst, const void *src, size_t bytes);
int infx_dev_download(void *dst, const void *src, size_t bytes);
int infx_dev_sync(void);
/* training batch: nseg segments of seg residues from starts[] (device) */
int infx_gate_gather(const uint8_t *codes, const float *labels, const int64_t *starts, int nseg, int seg,
                     uint8_t *out_codes, float *out_labels);
int infx_gate_onehot(const uint8_t *codes, int64_t n, void *x_bf16);
/* mode 0: ReLU, bf16 out; 1: bf16 out; 2: fp32 out */
int infx_gate_conv(const void *x_bf16, int64_t n, int seg, int cin, int dil, const void *w_bf16, int kp, const float *bias,
                   void *y, int mode);
/* dw[co][kk] += sum_t g[t][co] * x[t+(k-4)*dil][ci] (fp32 atomics; clear dw first) */
int infx_gate_wgrad(const void *g_bf16, const void *x_bf16, int64_t n, int seg, int cin, int dil, int kp, float *dw);
int infx_gate_relu_bwd(void *g_bf16, const void *y_bf16, int64_t m);
/* out[c] += sum_t g[t][c] (clear out first): bias gradients */
int infx_gate_colsum(const void *g_bf16, int64_t n, float *out);
int infx_gate_head(const void *y_bf16, int64_t n, const float *hw, const float *hb, float *s);
int infx_gate_head_bwd(const float *s, const float *labels, const void *y_bf16, int64_t n, const float *hw, float pos_weight,
                       void *dy_bf16, float *dhw, float *dhb, float *loss);
int infx_gate_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2, float eps, int step);
/* fp32 master [128][kp] -> bf16 forward copy and (cin 128) the k-flipped,
 * transposed copy the data gradient runs with */
int infx_gate_pack(const float *w, int kp, int cin, void *wf_bf16, void *wd_bf16);

#ifdef __cplusplus
}
#endif
#endif /* INFERNAL_GPU_H */
