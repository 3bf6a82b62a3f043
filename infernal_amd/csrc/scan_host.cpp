/* scan_host.cpp -- host side of the GPU scanners: scan matrices (band
 * tables), the device "scan profile" of a configured CM, the launch, and the
 * C-ABI entry points FastCYKScan / FastFInsideScan / FastIInsideScan.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <mutex>
#include <string>
#include <algorithm>
#include <vector>

#include "cm_internal.h"
#include "scan_kernels.h"
#include "scan_host.h"

extern "C" hipError_t infx_launch_std(int semi, int want_vsc, int tbl_lds, int gacc, int wide, dim3 grid, dim3 block,
                                      hipStream_t stream, const ScanArgs *a, int acc_n);
extern "C" size_t infx_std_lds_bytes(int semi, int tbl_lds, int wpb, int acc_n, int gacc, int wide);
extern "C" size_t infx_std_batch_lds_bytes(int semi, int wpb, int acc_n);
extern "C" int infx_std_rows(int semi, int tbl_lds, int wide);
extern "C" int infx_std_occupancy(int semi, int want_vsc, int tbl_lds, int wpb, int acc_n, int gacc, int wide);
extern "C" int infx_std_batch_occupancy(int semi, int wpb, int acc_n);
extern "C" hipError_t infx_launch_zero_u64(unsigned long long *p, hipStream_t stream);
extern "C" hipError_t infx_launch_zero_bytes(void *p, size_t n, hipStream_t stream);
extern "C" hipError_t infx_launch_std_batch(int semi, dim3 grid, dim3 block, hipStream_t stream, const ScanArgs *a, int acc_n);

/* ------------------------------------------------------------------------ */
/* scan matrices (src/cm_mx.c:5686-5836, :6391-6560)                          */

static int ***alloc_bands(int M, int W)
{
  int ***A = (int ***) calloc(NSMX_QDB_IDX, sizeof(int **));
  for (int n = 0; n < NSMX_QDB_IDX; n++) {
    A[n] = (int **) calloc(W + 1, sizeof(int *));
    int *blk = (int *) calloc((size_t) (W + 1) * M, sizeof(int));
    for (int j = 0; j <= W; j++) A[n][j] = blk + (size_t) j * M;
    A[n][0] = blk;  /* row 0 is unused (j == 0 is out of bounds), it owns the block */
  }
  return A;
}
static void free_bands(int ***A)
{
  if (!A) return;
  for (int n = 0; n < NSMX_QDB_IDX; n++) { free(A[n][0]); free(A[n]); }
  free(A);
}

extern "C" int cm_scan_mx_Create(CM_t *cm, char *errbuf, int do_float, int do_int, CM_SCAN_MX **ret_smx)
{
  *ret_smx = NULL;
  if (!do_float && !do_int) return infx_fail(eslEINVAL, errbuf, "cm_scan_mx_Create(), do_float and do_int both FALSE.");
  if (cm->qdbinfo == NULL) return infx_fail(eslEINVAL, errbuf, "cm_scan_mx_Create(), qdbinfo is invalid");
  CM_SCAN_MX *smx = (CM_SCAN_MX *) calloc(1, sizeof(CM_SCAN_MX));
  smx->M = cm->M; smx->W = cm->W;
  smx->dnAAA = alloc_bands(cm->M, cm->W);
  smx->dxAAA = alloc_bands(cm->M, cm->W);
  const CM_QDBINFO *q = cm->qdbinfo;
  for (int j = 1; j <= smx->W; j++)
    for (int v = 0; v < smx->M; v++) {
      int mp = (cm->sttype[v] == MP_st);
      smx->dnAAA[SMX_NOQDB][j][v]      = ESL_MIN(mp ? 2 : 1, smx->W);
      smx->dnAAA[SMX_QDB1_TIGHT][j][v] = ESL_MIN(mp ? ESL_MAX(q->dmin1[v], 2) : ESL_MAX(q->dmin1[v], 1), smx->W);
      smx->dnAAA[SMX_QDB2_LOOSE][j][v] = ESL_MIN(mp ? ESL_MAX(q->dmin2[v], 2) : ESL_MAX(q->dmin2[v], 1), smx->W);
      smx->dxAAA[SMX_NOQDB][j][v]      = j;
      smx->dxAAA[SMX_QDB1_TIGHT][j][v] = ESL_MIN(j, ESL_MIN(q->dmax1[v], smx->W));
      smx->dxAAA[SMX_QDB2_LOOSE][j][v] = ESL_MIN(j, ESL_MIN(q->dmax2[v], smx->W));
    }
  smx->bestsc = (float *) malloc(sizeof(float) * (smx->W + 1));
  smx->bestr  = (int *) malloc(sizeof(int) * (smx->W + 1));
  for (int d = 0; d <= smx->W; d++) { smx->bestsc[d] = IMPOSSIBLE; smx->bestr[d] = 0; }
  smx->floats_valid = do_float ? 1 : 0;
  smx->ints_valid   = do_int ? 1 : 0;
  smx->dev = NULL;
  *ret_smx = smx;
  return eslOK;
}

extern "C" void cm_scan_mx_Destroy(CM_t *cm, CM_SCAN_MX *smx)
{
  (void) cm;
  if (!smx) return;
  infx_dev_release(smx->dev);
  free_bands(smx->dnAAA); free_bands(smx->dxAAA);
  free(smx->bestsc); free(smx->bestr);
  free(smx);
}

extern "C" int cm_tr_scan_mx_Create(CM_t *cm, char *errbuf, int do_float, int do_int, CM_TR_SCAN_MX **ret_trsmx)
{
  *ret_trsmx = NULL;
  if (!do_float && !do_int) return infx_fail(eslEINVAL, errbuf, "cm_tr_scan_mx_Create(), do_float and do_int both FALSE.");
  if (cm->qdbinfo == NULL) return infx_fail(eslEINVAL, errbuf, "cm_tr_scan_mx_Create(), qdbinfo is invalid");
  CM_TR_SCAN_MX *t = (CM_TR_SCAN_MX *) calloc(1, sizeof(CM_TR_SCAN_MX));
  t->M = cm->M; t->W = cm->W;
  t->dnAAA = alloc_bands(cm->M, cm->W);
  t->dxAAA = alloc_bands(cm->M, cm->W);
  const CM_QDBINFO *q = cm->qdbinfo;
  for (int j = 1; j <= t->W; j++)
    for (int v = 0; v < t->M; v++) {
      t->dnAAA[SMX_NOQDB][j][v] = t->dnAAA[SMX_QDB1_TIGHT][j][v] = t->dnAAA[SMX_QDB2_LOOSE][j][v] = 1;
      t->dxAAA[SMX_NOQDB][j][v]      = ESL_MIN(j, t->W);
      t->dxAAA[SMX_QDB1_TIGHT][j][v] = ESL_MIN(j, ESL_MIN(q->dmax1[v], t->W));
      t->dxAAA[SMX_QDB2_LOOSE][j][v] = ESL_MIN(j, ESL_MIN(q->dmax2[v], t->W));
    }
  t->bestsc   = (float *) malloc(sizeof(float) * (t->W + 1));
  t->bestr    = (int *) malloc(sizeof(int) * (t->W + 1));
  t->bestmode = (char *) malloc(t->W + 1);
  t->floats_valid = do_float ? 1 : 0;
  t->ints_valid   = do_int ? 1 : 0;
  *ret_trsmx = t;
  return eslOK;
}

extern "C" void cm_tr_scan_mx_Destroy(CM_t *cm, CM_TR_SCAN_MX *t)
{
  (void) cm;
  if (!t) return;
  infx_trdev_release(t->dev);   /* a truncated matrix owns a TrDev, never an SmxDev */
  free_bands(t->dnAAA); free_bands(t->dxAAA);
  free(t->bestsc); free(t->bestr); free(t->bestmode);
  free(t);
}

/* ------------------------------------------------------------------------ */
/* device context                                                            */

static std::mutex g_mu;
static int  g_state = 0;          /* 0 untried, 1 ok, -1 failed */
static char g_err[eslERRBUFSIZE];
static hipStream_t g_stream;
static hipEvent_t  g_ev0, g_ev1;
static int  g_ncu = 256;
static size_t g_lds_max = 65536;   /* LDS bytes one workgroup may allocate */
static int  g_chunk_override = 0, g_wpb_override = 0;
static int  g_batch_graph = 0;     /* batched launches through a captured graph */

static thread_local double  t_wall_ms = 0, t_kernel_ms = 0;
static thread_local int64_t t_cells = 0;
static thread_local int     t_launches = 0;

extern "C" int infx_dev_init(char *errbuf)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_state == 1) return eslOK;
  if (g_state == -1) return infx_fail(eslENODEVICE, errbuf, "%s", g_err);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    snprintf(g_err, sizeof(g_err), "no HIP device available (%s); this library has no CPU fallback", hipGetErrorString(e));
    g_state = -1;
    return infx_fail(eslENODEVICE, errbuf, "%s", g_err);
  }
  hipDeviceProp_t p;
  if ((e = hipGetDeviceProperties(&p, 0)) != hipSuccess || strncmp(p.gcnArchName, "gfx9", 4) != 0) {
    snprintf(g_err, sizeof(g_err), "device 0 is not a gfx9 (CDNA) GPU: %s", e == hipSuccess ? p.gcnArchName : hipGetErrorString(e));
    g_state = -1;
    return infx_fail(eslENODEVICE, errbuf, "%s", g_err);
  }
  g_ncu = p.multiProcessorCount;
  g_lds_max = p.sharedMemPerBlock;
  if (hipSetDevice(0) != hipSuccess || hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&g_ev0) != hipSuccess || hipEventCreate(&g_ev1) != hipSuccess) {
    snprintf(g_err, sizeof(g_err), "HIP stream/event creation failed");
    g_state = -1;
    return infx_fail(eslENODEVICE, errbuf, "%s", g_err);
  }
  g_state = 1;
  return eslOK;
}

extern "C" int infx_gpu_available(char *errbuf) { return infx_dev_init(errbuf) == eslOK ? 1 : 0; }
hipStream_t infx_stream(void) { return g_stream; }
int infx_num_cu(void) { return g_ncu; }

extern "C" void infx_last_call_timing(double *ret_wall_ms, double *ret_kernel_ms, int64_t *ret_cells, int *ret_launches)
{
  if (ret_wall_ms)   *ret_wall_ms = t_wall_ms;
  if (ret_kernel_ms) *ret_kernel_ms = t_kernel_ms;
  if (ret_cells)     *ret_cells = t_cells;
  if (ret_launches)  *ret_launches = t_launches;
}
void infx_timing_reset(void) { t_wall_ms = t_kernel_ms = 0; t_cells = 0; t_launches = 0; }
void infx_timing_add(double wall, double kern, int64_t cells, int launches)
{
  t_wall_ms += wall; t_kernel_ms += kern; t_cells += cells; t_launches += launches;
}

extern "C" void infx_set_launch_params(int chunk_len, int waves_per_block)
{
  g_chunk_override = chunk_len > 0 ? chunk_len : 0;
  g_wpb_override   = (waves_per_block >= 1 && waves_per_block <= 16) ? waves_per_block : 0;
}
extern "C" void infx_set_batch_graph(int on) { g_batch_graph = on ? 1 : 0; }
int infx_chunk_override(void) { return g_chunk_override; }
int infx_wpb_override(void) { return g_wpb_override; }

/* ------------------------------------------------------------------------ */
/* pinned target: one device copy of a caller's residues dsq[1..L], laid out
 * as 64 zero bytes | residues | 64 zero bytes.  Scans of that same buffer
 * (same dsq pointer, range inside 1..L) read it in place instead of
 * uploading their range; the residues were checked once, when pinned.     */
static const ESL_DSQ *g_pin_host = NULL;
static int64_t        g_pin_L = 0;
static void          *g_pin_dev = NULL;
static size_t         g_pin_cap = 0;

extern "C" int infx_target_Pin(const ESL_DSQ *dsq, int64_t L, char *errbuf)
{
  if (!dsq || L < 1) return infx_fail(eslEINVAL, errbuf, "infx_target_Pin: no residues");
  for (int64_t x = 1; x <= L; x++)
    if (dsq[x] > 15) return infx_fail(eslEINVAL, errbuf, "infx_target_Pin: residue code %d at %lld is not a residue", (int) dsq[x], (long long) x);
  int status;
  if ((status = infx_dev_init(errbuf)) != eslOK) return status;
  std::lock_guard<std::mutex> lk(g_mu);
  const size_t need = (size_t) L + 128;
  if (need > g_pin_cap) {
    if (g_pin_dev) (void) hipFree(g_pin_dev);
    g_pin_dev = NULL; g_pin_cap = 0; g_pin_host = NULL; g_pin_L = 0;
    if (hipMalloc(&g_pin_dev, need) != hipSuccess) { g_pin_dev = NULL; return infx_fail(eslEMEM, errbuf, "infx_target_Pin: %zu bytes of device memory", need); }
    g_pin_cap = need;
  }
  if (hipMemsetAsync(g_pin_dev, 0, need, g_stream) != hipSuccess ||
      hipMemcpyAsync((char *) g_pin_dev + 64, dsq + 1, (size_t) L, hipMemcpyHostToDevice, g_stream) != hipSuccess ||
      hipStreamSynchronize(g_stream) != hipSuccess)
    return infx_fail(eslESYS, errbuf, "infx_target_Pin: upload failed");
  g_pin_host = dsq; g_pin_L = L;
  return eslOK;
}

extern "C" int infx_target_Unpin(void)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_pin_dev) (void) hipFree(g_pin_dev);
  g_pin_dev = NULL; g_pin_cap = 0; g_pin_host = NULL; g_pin_L = 0;
  return eslOK;
}

/* Device address of residue x0 when dsq is the pinned target and x0..x1 lies
 * inside it (reads reach 64 bytes either side: the pads), else NULL. */
const uint8_t *infx_pinned(const ESL_DSQ *dsq, int64_t x0, int64_t x1)
{
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_pin_dev || dsq != g_pin_host || x0 < 1 || x1 > g_pin_L || x0 > x1) return NULL;
  return (const uint8_t *) g_pin_dev + 64 + (x0 - 1);
}

/* growable device buffer */
int DevBuf::reserve(size_t n)
{
  if (n <= cap) return eslOK;
  if (borrowed) return eslERANGE;
  if (ptr) (void) hipFree(ptr);   /* contents are not kept; a failed free leaves nothing to undo */
  ptr = NULL; cap = 0;
  size_t want = n + n / 4 + 256;
  if (hipMalloc(&ptr, want) != hipSuccess) { ptr = NULL; return eslEMEM; }
  cap = want;
  return eslOK;
}
void DevBuf::release() { if (ptr && !borrowed) (void) hipFree(ptr); ptr = NULL; cap = 0; borrowed = false; }

int Arena::give(DevBuf &b, size_t n)
{
  const size_t a = (n + 255) / 256 * 256;
  if (a > left()) return eslERANGE;
  b.ptr = base + used; b.cap = a; b.borrowed = true;
  used += a;
  return eslOK;
}

struct StdProfile {
  int qdbidx, semi, W;
  std::vector<DState> st;
  std::vector<uint32_t> esc, rtmpl, dtmpl;
  int R, RB;
  int root_dn, root_dx, root_cnum;
  int root_child_base[6];
  uint32_t root_tsc[6];
  int local_begin;
  uint32_t el;
  DevBuf d_st, d_esc, d_rtmpl, d_dtmpl;
  bool uploaded = false;
};

struct SmxDev {
  std::vector<StdProfile *> prof;
  DevBuf d_dsq, d_scratch, d_cand, d_chunk, d_ncand, d_lsum_f, d_lsum_i, d_args, d_ctab, d_btab;
  /* the batched launch as a graph (counter reset + kernel), reused while its
   * launch shape and parameter buffers stay the same */
  hipGraph_t graph = NULL;
  hipGraphExec_t gexec = NULL;
  size_t gkey[6] = {0, 0, 0, 0, 0, 0};
  ~SmxDev() {
    if (gexec) (void) hipGraphExecDestroy(gexec);
    if (graph) (void) hipGraphDestroy(graph);
    for (auto *p : prof) { p->d_st.release(); p->d_esc.release(); p->d_rtmpl.release(); p->d_dtmpl.release(); delete p; }
    d_dsq.release(); d_scratch.release(); d_cand.release(); d_chunk.release(); d_ncand.release(); d_args.release(); d_ctab.release();
    d_btab.release(); d_lsum_f.release(); d_lsum_i.release();
  }
};

extern "C" void infx_dev_release(void *dev)
{
  if (!dev) return;
  delete (SmxDev *) dev;
}

static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* d = 0 initial values, as cm_scan_mx_InitializeFloats/Integers
 * (src/cm_mx.c:6039-6138); v = 0..M-1, BEGL_S values for its deck. */
static void d0_init_float(const CM_t *cm, std::vector<float> &a0)
{
  a0.assign(cm->M, (float) IMPOSSIBLE);
  for (int v = cm->M - 1; v >= 0; v--) {
    int st = cm->sttype[v];
    if (cm->stid[v] != BEGL_S) {
      if (st == E_st) a0[v] = 0.;
      else if (st == S_st || st == D_st) {
        int y = cm->cfirst[v];
        float x = cm->endsc[v];
        for (int k = 0; k < cm->cnum[v]; k++) x = ESL_MAX(x, (a0[y + k] + cm->tsc[v][k]));
        x = ESL_MAX(x, IMPOSSIBLE);
        a0[v] = x;
      } else if (st == B_st) {
        a0[v] = a0[cm->cfirst[v]] + a0[cm->cnum[v]];
      }
    } else {
      int y = cm->cfirst[v];
      float x = cm->endsc[v];
      for (int k = 0; k < cm->cnum[v]; k++) x = ESL_MAX(x, (a0[y + k] + cm->tsc[v][k]));
      x = ESL_MAX(x, IMPOSSIBLE);
      a0[v] = x;
    }
  }
}
static void d0_init_int(const CM_t *cm, std::vector<int> &a0)
{
  a0.assign(cm->M, -INFTY);
  for (int v = cm->M - 1; v >= 0; v--) {
    int st = cm->sttype[v];
    if (cm->stid[v] != BEGL_S) {
      if (st == E_st) a0[v] = 0;
      else if (st == S_st || st == D_st) {
        int y = cm->cfirst[v];
        int x = cm->iendsc[v];
        for (int k = 0; k < cm->cnum[v]; k++) x = ESL_MAX(x, (int) ((unsigned) a0[y + k] + (unsigned) cm->itsc[v][k]));
        x = ESL_MAX(x, -INFTY);
        a0[v] = x;
      } else if (st == B_st) {
        a0[v] = (int) ((unsigned) a0[cm->cfirst[v]] + (unsigned) a0[cm->cnum[v]]);
      }
    } else {
      int y = cm->cfirst[v];
      int x = cm->iendsc[v];
      for (int k = 0; k < cm->cnum[v]; k++) x = ESL_MAX(x, (int) ((unsigned) a0[y + k] + (unsigned) cm->itsc[v][k]));
      x = ESL_MAX(x, -INFTY);
      a0[v] = x;
    }
  }
}

/* Build the scan profile of <cm> for bands <qdbidx>, semiring <semi> and
 * effective W (min(smx->W, L)). */
static int build_std_profile(CM_t *cm, CM_SCAN_MX *smx, int qdbidx, int semi, int W, StdProfile &P, char *errbuf)
{
  const int M = cm->M;
  const int *dmin = qdbidx == SMX_QDB1_TIGHT ? cm->qdbinfo->dmin1 : qdbidx == SMX_QDB2_LOOSE ? cm->qdbinfo->dmin2 : NULL;
  const int *dmax = qdbidx == SMX_QDB1_TIGHT ? cm->qdbinfo->dmax1 : qdbidx == SMX_QDB2_LOOSE ? cm->qdbinfo->dmax2 : NULL;
  const int banded = (qdbidx != SMX_NOQDB);
  std::vector<int> dn(M), dx(M), lo(M, INT32_MAX), hi(M, INT32_MIN);
  for (int v = 0; v < M; v++) { dn[v] = smx->dnAAA[qdbidx][W][v]; dx[v] = smx->dxAAA[qdbidx][W][v]; }
  auto extend = [&](int v, int a, int b) { if (a > b) return; lo[v] = ESL_MIN(lo[v], a); hi[v] = ESL_MAX(hi[v], b); };
  auto kbounds = [&](int v, int d, int &kmin, int &kmax) {
    int w = cm->cfirst[v], y = cm->cnum[v];
    if (banded) {
      int dn_y = ESL_MIN(dmin[y], smx->W), dx_y = ESL_MIN(dmax[y], smx->W);
      int dn_w = ESL_MIN(dmin[w], smx->W), dx_w = ESL_MIN(dmax[w], smx->W);
      kmin = ESL_MAX(0, ESL_MAX(dn_y, d - dx_w));
      kmax = ESL_MIN(dx_y, d - dn_w);
    } else { kmin = 0; kmax = d; }
  };
  for (int v = 1; v < M; v++) {
    int st = cm->sttype[v];
    if (st == E_st || dn[v] > dx[v]) continue;
    extend(v, dn[v], dx[v]);
    if (st == B_st) {
      int w = cm->cfirst[v], y = cm->cnum[v];
      for (int d = dn[v]; d <= dx[v]; d++) {
        int kmin, kmax; kbounds(v, d, kmin, kmax);
        if (kmin > kmax) continue;
        extend(y, kmin, kmax);
        extend(w, d - kmax, d - kmin);
      }
    } else {
      int sd = StateDelta(st);
      for (int k = 0; k < cm->cnum[v]; k++) extend(cm->cfirst[v] + k, dn[v] - sd, dx[v] - sd);
    }
  }
  if (dn[0] <= dx[0])
    for (int k = 0; k < cm->cnum[0]; k++) extend(cm->cfirst[0] + k, dn[0], dx[0]);

  std::vector<float> a0f; std::vector<int> a0i;
  if (semi == SEMI_IINSIDE) d0_init_int(cm, a0i); else d0_init_float(cm, a0f);
  const uint32_t impbits = (semi == SEMI_IINSIDE) ? (uint32_t) (-INFTY) : fbits((float) IMPOSSIBLE);

  /* BEGL decks: a ring per BEGL state, rounded up to a power of two.  A
   * state visit writes rows jp .. jp+NR-1 of the deck before its B parent
   * reads rows jp+t-k for k up to the largest split maxk, so rows jp-maxk ..
   * jp+NR-1 must not share a ring row: maxk + NR rows (NR = 4, the most
   * rows any kernel visits at once) */
  std::vector<int> ring(M, 0);
  for (int v = 1; v < M; v++) {
    if (cm->sttype[v] != B_st || dn[v] > dx[v]) continue;
    int w = cm->cfirst[v], maxk = 0;
    for (int d = dn[v]; d <= dx[v]; d++) {
      int kmin, kmax; kbounds(v, d, kmin, kmax);
      if (kmin <= kmax) maxk = ESL_MAX(maxk, kmax);
    }
    int r = 1;
    while (r < maxk + INFX_NR_MAX) r <<= 1;
    ring[w] = r;
  }

  std::vector<int> base(M, 0), stride(M, 0);
  P.R = 0; P.RB = 0;
  P.rtmpl.clear(); P.dtmpl.clear();
  for (int v = 1; v < M; v++) {
    if (lo[v] > hi[v]) continue;
    lo[v] = ESL_MAX(lo[v], 0); hi[v] = ESL_MIN(hi[v], W);
    if (lo[v] > hi[v]) continue;
    const bool begl = (cm->stid[v] == BEGL_S);
    std::vector<uint32_t> &T = begl ? P.dtmpl : P.rtmpl;
    int off = (int) T.size();
    base[v] = off - lo[v];
    stride[v] = hi[v] - lo[v] + 1;
    const int nrows = begl ? ESL_MAX(ring[v], 1) : 1;
    if (begl) ring[v] = nrows;
    for (int r = 0; r < nrows; r++)
      for (int d = lo[v]; d <= hi[v]; d++) {
        uint32_t b = impbits;
        if (d == 0) b = (semi == SEMI_IINSIDE) ? (uint32_t) a0i[v] : fbits(a0f[v]);
        T.push_back(b);
      }
  }
  for (int v = 1; v < M; v++) if (cm->stid[v] == BEGL_S) ring[v] = ESL_MAX(ring[v], 1);
  /* pad to a multiple of 64 words so every slot region stays aligned */
  while (P.rtmpl.size() % 64) P.rtmpl.push_back(impbits);
  while (P.dtmpl.size() % 64) P.dtmpl.push_back(impbits);
  if (P.dtmpl.empty()) for (int i = 0; i < 64; i++) P.dtmpl.push_back(impbits);
  P.R = (int) P.rtmpl.size(); P.RB = (int) P.dtmpl.size();

  /* state descriptors, v = M-1..1 without E states (packed, see scan_kernels.h) */
  P.esc.clear();
  P.st.clear();
  auto fits16 = [](long x) { return x >= 0 && x <= 0xffff; };
  for (int v = M - 1; v >= 1; v--) {
    const int st = cm->sttype[v];
    if (st == E_st) continue;
    if (st == IL_st && cm->cfirst[v] != v)
      return infx_fail(eslEINVAL, errbuf, "IL state %d does not loop on itself", v);
    DState s;
    memset(&s, 0, sizeof(DState));
    const int kind = (st == B_st) ? KIND_B : (cm->stid[v] == BEGL_S) ? KIND_BEGL : (st == IL_st) ? KIND_IL : KIND_GENERIC;
    const int cnum = (st == B_st) ? 0 : cm->cnum[v];
    int has_begin, init_ok;
    if (semi == SEMI_IINSIDE) { has_begin = cm->ibeginsc[v] != -INFTY; init_ok = cm->iendsc[v] != -INFTY; }
    else                      { has_begin = NOT_IMPOSSIBLE(cm->beginsc[v]); init_ok = NOT_IMPOSSIBLE(cm->endsc[v]); }
    if (v > 0xffff || cnum > 7 || !fits16(dn[v]) || !fits16(dx[v]))
      return infx_fail(eslERANGE, errbuf, "state %d does not fit the packed scan descriptor (M, cnum or band too large)", v);
    s.hdr = (uint32_t) kind | ((uint32_t) Emitmode(st) << 3) | ((uint32_t) StateDelta(st) << 5) |
            ((uint32_t) (StateRightDelta(st) > 0) << 7) | ((uint32_t) cnum << 8) | ((uint32_t) has_begin << 11) |
            ((uint32_t) init_ok << 12) | ((uint32_t) banded << 13) |
            ((uint32_t) (StateRightDelta(st) > 0 && cm->cfirst[v] == v) << 14) | ((uint32_t) v << 16);
    s.dnx = (uint32_t) dn[v] | ((uint32_t) dx[v] << 16);
    s.base = base[v];
    if (st == ML_st || st == MR_st || st == IL_st || st == IR_st || st == MP_st) {
      int n = (st == MP_st) ? INFX_KP * INFX_KP : INFX_KP;
      s.aux = (uint32_t) P.esc.size();
      for (int x = 0; x < n; x++) P.esc.push_back(semi == SEMI_IINSIDE ? (uint32_t) cm->ioesc[v][x] : fbits(cm->oesc[v][x]));
    }
    if (semi == SEMI_IINSIDE) { s.beginsc = (uint32_t) cm->ibeginsc[v]; s.endsc = (uint32_t) cm->iendsc[v]; }
    else                      { s.beginsc = fbits(cm->beginsc[v]);       s.endsc = fbits(cm->endsc[v]); }
    if (cm->stid[v] == BEGL_S) {
      if (has_begin) return infx_fail(eslEINVAL, errbuf, "BEGL_S state %d has a local begin score", v);
      if (!fits16(ring[v] - 1) || !fits16(stride[v])) return infx_fail(eslERANGE, errbuf, "BEGL deck of state %d too large", v);
      s.aux = (uint32_t) (ring[v] - 1) | ((uint32_t) stride[v] << 16);
    }
    if (st == B_st) {
      int w = cm->cfirst[v], y = cm->cnum[v];
      int kdn_y = 0, kdx_y = 0, kdn_w = 0, kdx_w = 0;
      if (banded) {
        kdn_y = ESL_MIN(dmin[y], smx->W); kdx_y = ESL_MIN(dmax[y], smx->W);
        kdn_w = ESL_MIN(dmin[w], smx->W); kdx_w = ESL_MIN(dmax[w], smx->W);
      }
      if (!fits16(kdn_y) || !fits16(kdx_y) || !fits16(kdn_w) || !fits16(kdx_w) ||
          !fits16(ESL_MAX(ring[w], 1) - 1) || !fits16(stride[w]))
        return infx_fail(eslERANGE, errbuf, "B state %d does not fit the packed scan descriptor", v);
      s.tsc[0] = (uint32_t) kdn_y | ((uint32_t) kdx_y << 16);
      s.tsc[1] = (uint32_t) kdn_w | ((uint32_t) kdx_w << 16);
      s.tsc[2] = (uint32_t) base[w];
      s.tsc[3] = (uint32_t) base[y];
      s.aux = (uint32_t) (ESL_MAX(ring[w], 1) - 1) | ((uint32_t) stride[w] << 16);
    } else {
      for (int k = 0; k < cnum && k < 6; k++)
        s.tsc[k] = (semi == SEMI_IINSIDE) ? (uint32_t) cm->itsc[v][k] : fbits(cm->tsc[v][k]);
      s.cbase = base[cm->cfirst[v]];
      for (int k = 0; k < cnum && k < 6; k++) {
        long off = (long) base[cm->cfirst[v] + k] - s.cbase;
        if (off < -32768 || off > 32767) return infx_fail(eslERANGE, errbuf, "children of state %d are too far apart for the packed scan descriptor", v);
        s.coff[k >> 1] |= ((uint32_t) (uint16_t) (int16_t) off) << (16 * (k & 1));
      }
    }
    P.st.push_back(s);
  }
  P.root_dn = dn[0]; P.root_dx = dx[0];
  P.root_cnum = cm->cnum[0];
  if (P.root_cnum < 1 || P.root_cnum > 6) return infx_fail(eslEINVAL, errbuf, "ROOT_S has %d children", P.root_cnum);
  for (int k = 0; k < 6; k++) {
    P.root_child_base[k] = (k < P.root_cnum) ? base[cm->cfirst[0] + k] : 0;
    P.root_tsc[k] = (k < P.root_cnum) ? (semi == SEMI_IINSIDE ? (uint32_t) cm->itsc[0][k] : fbits(cm->tsc[0][k])) : 0;
  }
  P.local_begin = (cm->flags & CMH_LOCAL_BEGIN) ? 1 : 0;
  P.el = (semi == SEMI_IINSIDE) ? (uint32_t) cm->iel_selfsc : fbits(cm->el_selfsc);
  if (P.esc.empty()) P.esc.push_back(0);
  P.qdbidx = qdbidx; P.semi = semi; P.W = W;
  return eslOK;
}

static hipError_t g_last_hip = hipSuccess;

/* copy n bytes to a device buffer of at least n+pad bytes */
static int upload(DevBuf &b, const void *src, size_t n, size_t pad = 0)
{
  int st;
  if ((st = b.reserve(n + pad)) != eslOK) return st;
  if (n == 0) return eslOK;
  g_last_hip = hipMemcpyAsync(b.ptr, src, n, hipMemcpyHostToDevice, g_stream);
  return g_last_hip == hipSuccess ? eslOK : eslESYS;
}

static bool same_profile(const StdProfile &a, const StdProfile &b)
{
  return a.st.size() == b.st.size() && memcmp(a.st.data(), b.st.data(), a.st.size() * sizeof(DState)) == 0 &&
         a.esc == b.esc && a.rtmpl == b.rtmpl && a.dtmpl == b.dtmpl &&
         a.root_dn == b.root_dn && a.root_dx == b.root_dx && a.root_cnum == b.root_cnum &&
         memcmp(a.root_child_base, b.root_child_base, sizeof(a.root_child_base)) == 0 &&
         memcmp(a.root_tsc, b.root_tsc, sizeof(a.root_tsc)) == 0 && a.local_begin == b.local_begin && a.el == b.el;
}

/* ------------------------------------------------------------------------ */
/* the standard scanners                                                      */

/* One scan over the windows wi[w]..wj[w] (sorted, disjoint) of one target:
 * the cells, candidates and hits are those of scanning each window on its
 * own (a chunk's DP never starts before its window), in one launch. */
static int std_scan(int semi, const char *fname, CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq,
                    const int64_t *wi, const int64_t *wj, int nwin, float cutoff, CM_TOPHITS *hitlist, int do_null3,
                    float env_cutoff, int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc,
                    Arena *ar = nullptr)
{
  int status;
  auto t_start = std::chrono::steady_clock::now();
  infx_timing_reset();
  /* contract checks (src/cm_dpsearch.c:129-135, :1146-1151, :1771-1776) */
  if (nwin < 1)                       return infx_fail(eslEINCOMPAT, errbuf, "%s, no window", fname);
  const int64_t i0 = wi[0], j0 = wj[nwin - 1];   /* the span the windows lie in */
  if (!(cm->flags & CMH_BITS))        return infx_fail(eslEINCOMPAT, errbuf, "%s, CMH_BITS flag is not raised.\n", fname);
  for (int w = 0; w < nwin; w++) {
    if (wj[w] < wi[w])                return infx_fail(eslEINCOMPAT, errbuf, "%s, i0: %lld j0: %lldd\n", fname, (long long) wi[w], (long long) wj[w]);
    if (w > 0 && wi[w] <= wj[w - 1])  return infx_fail(eslEINCOMPAT, errbuf, "%s, windows must be sorted and disjoint", fname);
  }
  if (dsq == NULL)                    return infx_fail(eslEINCOMPAT, errbuf, "%s, dsq is NULL\n", fname);
  if (semi == SEMI_CYK && (cm->search_opts & CM_SEARCH_INSIDE))     return infx_fail(eslEINCOMPAT, errbuf, "%s, CM_SEARCH_INSIDE flag raised", fname);
  if (semi != SEMI_CYK && !(cm->search_opts & CM_SEARCH_INSIDE))    return infx_fail(eslEINCOMPAT, errbuf, "%s, CM_SEARCH_INSIDE flag not raised", fname);
  if (smx == NULL)                    return infx_fail(eslEINCOMPAT, errbuf, "%s, smx == NULL\n", fname);
  if (semi == SEMI_CYK && !smx->floats_valid) return infx_fail(eslEINCOMPAT, errbuf, "%s, smx->floats_valid if FALSE", fname);
  if (semi != SEMI_CYK && !smx->ints_valid)   return infx_fail(eslEINCOMPAT, errbuf, "%s, smx->ints_valid if FALSE", fname);
  if (cm->qdbinfo == NULL)            return infx_fail(eslEINCOMPAT, errbuf, "%s, cm->qdbinfo == NULL\n", fname);
  if (qdbidx < 0 || qdbidx > 2)       return infx_fail(eslEINCOMPAT, errbuf, "%s, qdbidx is invalid", fname);
  if (i0 < 1)                         return infx_fail(eslEINCOMPAT, errbuf, "%s, i0 must be >= 1", fname);
  if (j0 - i0 + 1 > (int64_t) INT32_MAX - 1) return infx_fail(eslERANGE, errbuf, "%s, target range too long for one call", fname);
  /* hit candidates carry 32-bit positions (cmsearch scans buffers of one
   * sequence window, positions below ~10^5) */
  if (j0 > (int64_t) INT32_MAX - 1) return infx_fail(eslERANGE, errbuf, "%s, j0 = %lld: positions past %d do not fit the hit candidates", fname, (long long) j0, INT32_MAX - 1);
  int64_t L = 0, Lmax = 0;      /* residues scanned; the longest window */
  const uint8_t *pinned = infx_pinned(dsq, i0, j0);   /* residues already on the device (checked when pinned) */
  for (int w = 0; w < nwin; w++) {
    for (int64_t x = wi[w]; !pinned && x <= wj[w]; x++)
      if (dsq[x] > 15) return infx_fail(eslEINVAL, errbuf, "%s, residue code %d at %lld is not a residue", fname, (int) dsq[x], (long long) x);
    L += wj[w] - wi[w] + 1;
    Lmax = ESL_MAX(Lmax, wj[w] - wi[w] + 1);
  }
  const int64_t Lspan = j0 - i0 + 1;
  const int W = (smx->W > Lmax) ? (int) Lmax : smx->W;

  /* profile (host only: its errors are reported with or without a device) */
  StdProfile fresh;
  if ((status = build_std_profile(cm, smx, qdbidx, semi, W, fresh, errbuf)) != eslOK) return status;

  /* DP slot of one wavefront: guard | deck | guard | rows[2 NR] (ring over j) |
   * guard | vsc | global accumulators (rows j .. j+NR-1; 4-byte cells), with
   * NR the rows per state visit of the kernel */
  auto slot_layout = [&](int R, int RB, int gacc_, int acc_n_, int nr_, size_t *deck_off_, size_t *row_off_, size_t *vsc_off_, size_t *acc_off_) {
    const size_t G = (size_t) ((W + 128 + 63) / 64) * 64;
    *deck_off_ = G * 4;
    *row_off_  = *deck_off_ + ((size_t) RB + G) * 4;
    *vsc_off_  = *row_off_ + ((size_t) 2 * nr_ * R + G) * 4;
    *acc_off_  = ((*vsc_off_ + (size_t) cm->M * 4) + 255) / 256 * 256;
    return ((*acc_off_ + (gacc_ ? (size_t) acc_n_ * 4 * nr_ : 0)) + 255) / 256 * 256;
  };
  auto r256 = [](size_t n) { return (n + 255) / 256 * 256; };
  /* the no-allocation path: everything must fit the caller's fixed region
   * (profile, residues, parameters, chunk outputs, one DP slot, a minimal
   * candidate buffer), checked before any device work */
  const int64_t ar_cand_min = ESL_MAX((int64_t) 4096, ESL_MIN((int64_t) 4 * L, (int64_t) 1 << 20));
  if (ar) {
    size_t d_, r_, v_, a_;
    const int64_t nch_max = L / ESL_MAX((int64_t) 2 * W, (int64_t) 64) + nwin + 1;
    const size_t need = r256(fresh.st.size() * sizeof(DState)) + r256(fresh.esc.size() * 4) + r256(fresh.rtmpl.size() * 4 + 4) +
                        r256(fresh.dtmpl.size() * 4 + 4) + (semi != SEMI_CYK ? r256(INFX_LOGSUM_PAD * 4) : 0) +
                        (pinned ? 256 : r256((size_t) Lspan + 128)) + r256(sizeof(ScanArgs)) + 256 +
                        r256(ESL_MAX((size_t) nch_max * 24 + 64, r256((size_t) nch_max * 4) + (size_t) nch_max * 16 + 64)) +
                        (nwin > 1 ? r256((size_t) nch_max * 24) : 0) +
                        slot_layout(fresh.R, fresh.RB, 0, 0, infx_std_rows(semi, 1, 0), &d_, &r_, &v_, &a_) +
                        r256((size_t) ar_cand_min * sizeof(Cand));
    if (need > ar->left())
      return infx_fail(eslERANGE, errbuf, "%s: this model and target need %zu bytes of device scratch; the no-allocation scratch holds %zu",
                       fname, need, ar->left());
  }

  if ((status = infx_dev_init(errbuf)) != eslOK) return status;
  SmxDev Dloc;           /* the no-allocation path's buffers (borrowed from the arena) */
  StdProfile Ploc;
  SmxDev *D = NULL;
  StdProfile *P = NULL;
  if (ar) { D = &Dloc; P = &Ploc; }
  else {
    if (!smx->dev) smx->dev = new SmxDev();
    D = (SmxDev *) smx->dev;
    for (auto *p : D->prof) if (p->qdbidx == qdbidx && p->semi == semi && p->W == W) P = p;
    if (!P) { P = new StdProfile(); D->prof.push_back(P); }
  }
  if (ar && (ar->give(P->d_st, fresh.st.size() * sizeof(DState)) != eslOK || ar->give(P->d_esc, fresh.esc.size() * 4) != eslOK ||
             ar->give(P->d_rtmpl, fresh.rtmpl.size() * 4 + 4) != eslOK || ar->give(P->d_dtmpl, fresh.dtmpl.size() * 4 + 4) != eslOK))
    return infx_fail(eslERANGE, errbuf, "%s: profile does not fit the no-allocation scratch", fname);
  if (!P->uploaded || !same_profile(*P, fresh)) {
    P->st = fresh.st; P->esc = fresh.esc; P->rtmpl = fresh.rtmpl; P->dtmpl = fresh.dtmpl;
    P->R = fresh.R; P->RB = fresh.RB; P->root_dn = fresh.root_dn; P->root_dx = fresh.root_dx; P->root_cnum = fresh.root_cnum;
    memcpy(P->root_child_base, fresh.root_child_base, sizeof(P->root_child_base));
    memcpy(P->root_tsc, fresh.root_tsc, sizeof(P->root_tsc));
    P->local_begin = fresh.local_begin; P->el = fresh.el;
    P->qdbidx = qdbidx; P->semi = semi; P->W = W;
    if ((status = upload(P->d_st, P->st.data(), P->st.size() * sizeof(DState))) != eslOK ||
        (status = upload(P->d_esc, P->esc.data(), P->esc.size() * 4)) != eslOK ||
        (status = upload(P->d_rtmpl, P->rtmpl.data(), P->rtmpl.size() * 4, 4)) != eslOK ||
        (status = upload(P->d_dtmpl, P->dtmpl.data(), P->dtmpl.size() * 4, 4)) != eslOK)
      return infx_fail(status, errbuf, "%s: device upload failed: %s", fname, hipGetErrorString(g_last_hip));
    P->uploaded = true;
  }
  /* device tables carry INFX_LOGSUM_PAD entries, the ones past the
   * reference's 23000 zero */
  if (ar && semi == SEMI_FINSIDE && ar->give(D->d_lsum_f, INFX_LOGSUM_PAD * 4) != eslOK) return infx_fail(eslERANGE, errbuf, "%s: scratch full", fname);
  if (ar && semi == SEMI_IINSIDE && ar->give(D->d_lsum_i, INFX_LOGSUM_PAD * 4) != eslOK) return infx_fail(eslERANGE, errbuf, "%s: scratch full", fname);
  if (semi == SEMI_FINSIDE && (ar || D->d_lsum_f.cap == 0)) {
    std::vector<float> t(INFX_LOGSUM_PAD, 0.f);
    memcpy(t.data(), infx_flogsum_table(), LOGSUM_TBL * 4);
    if ((status = upload(D->d_lsum_f, t.data(), t.size() * 4)) != eslOK) return infx_fail(status, errbuf, "upload failed");
  }
  if (semi == SEMI_IINSIDE && (ar || D->d_lsum_i.cap == 0)) {
    std::vector<int> t(INFX_LOGSUM_PAD, 0);
    memcpy(t.data(), infx_ilogsum_table(), LOGSUM_TBL * 4);
    if ((status = upload(D->d_lsum_i, t.data(), t.size() * 4)) != eslOK) return infx_fail(status, errbuf, "upload failed");
  }

  /* launch shape: wavefronts per workgroup (LDS: log-sum table + per-wave
   * local-begin accumulators), workgroups resident per CU */
  const int acc_n = (P->root_dn <= P->root_dx) ? P->root_dx - P->root_dn + 1 : 0;
  int tbl_lds = 1;
  if (const char *e = getenv("INFX_LOGSUM_TABLE")) tbl_lds = strcmp(e, "l2") != 0;
  /* accumulators in LDS, or in the global slot when that keeps more
   * wavefronts resident (wide root bands) */
  int gacc_mode = -1;   /* -1: choose, 0: LDS, 1: global */
  if (const char *e = getenv("INFX_BEGIN_ACC")) gacc_mode = (strcmp(e, "global") == 0) ? 1 : (strcmp(e, "lds") == 0) ? 0 : -1;
  /* chunks of the shortest length (2W) the target gives: more resident
   * wavefronts than that find no work */
  int64_t units = 0;
  for (int w = 0; w < nwin; w++) units += (wj[w] - wi[w] + 1 + 2 * (int64_t) W - 1) / (2 * (int64_t) W);
  auto pick = [&](int wide, int *wpb_, int *occ_, int *gacc_) {
    int wpb = 0, occ = 0, gacc = 0;
    for (int g = 0; g <= 1; g++) {
      if ((gacc_mode >= 0 && g != gacc_mode) || (g && acc_n == 0)) continue;
      int bw = 0, bo = 0;   /* the workgroup size that keeps the most wavefronts resident (ties: larger) */
      for (int w = 16; w >= 1; w >>= 1) {
        if (infx_wpb_override() && w != infx_wpb_override()) continue;
        if (infx_std_lds_bytes(semi, tbl_lds, w, acc_n, g, wide) > (size_t) g_lds_max) continue;
        const int o = infx_std_occupancy(semi, ret_vsc != NULL, tbl_lds, w, acc_n, g, wide);
        if (o * w > bo * bw) { bw = w; bo = o; }
      }
      if (bw == 0 || bo == 0) continue;
      /* global accumulators cost an L2 atomic per begin cell: taken only for
       * at least 5/4 of the busy wavefronts LDS accumulators allow */
      const int64_t busy_new = ESL_MIN(units, (int64_t) bo * bw * g_ncu), busy_old = ESL_MIN(units, (int64_t) occ * wpb * g_ncu);
      if (wpb == 0 || busy_new * 4 >= busy_old * 5) { wpb = bw; occ = bo; gacc = g; }
    }
    *wpb_ = wpb; *occ_ = occ; *gacc_ = gacc;
  };
  int wpb = 0, occ = 0, gacc = 0, wide = 0;
  pick(0, &wpb, &occ, &gacc);
  /* wide-band CYK (four blocks of loads in flight per wavefront) when the
   * target gives fewer chunks of 2W residues than half the wavefronts the
   * narrow kernel keeps resident: fewer, busier wavefronts */
  {
    int wide_mode = -1;
    if (const char *e = getenv("INFX_WIDE")) wide_mode = atoi(e) ? 1 : 0;
    const int want = (semi == SEMI_CYK) && (wide_mode == 1 || (wide_mode < 0 && units * 2 < (int64_t) occ * wpb * g_ncu));
    if (want) {
      int w2 = 0, o2 = 0, g2 = 0;
      pick(1, &w2, &o2, &g2);
      if (w2 > 0 && o2 > 0) { wpb = w2; occ = o2; gacc = g2; wide = 1; }
    }
  }
  if (wpb < 1 || occ < 1)
    return infx_fail(eslEINVAL, errbuf, "%s: no workgroup shape fits (W = %d, LDS limit %zu bytes)", fname, W, (size_t) g_lds_max);
  const int64_t resident = (int64_t) occ * wpb * g_ncu;

  /* work split.  A chunk's wavefront sweeps chunk + W - 1 rows (its W - 1
   * warm-up rows rebuild the cells its first rows read), so a call takes
   * about that many row times once every chunk has a wavefront.  Targets that
   * fill the chip with chunks of 2W or more: one chunk per resident
   * wavefront (the warm-up at most a third of the work).  Shorter targets:
   * shorter chunks (at least 64 rows, at most 2W), up to two wavefronts per
   * SIMD, which trades extra warm-up work on an idle chip for latency: a
   * 2,300-nt target and W = 367 sweeps at most 430 rows per wavefront
   * instead of 1,100. */
  int64_t chunk = infx_chunk_override();
  if (chunk <= 0) {
    const int64_t fill = (L + resident - 1) / resident;
    if (fill >= (int64_t) 2 * W) {
      chunk = fill;
    } else {
      const int64_t conc = ESL_MIN(resident, (int64_t) 8 * g_ncu);
      chunk = ESL_MIN((int64_t) 2 * W, ESL_MAX((L + conc - 1) / conc, (int64_t) 64));
    }
    chunk = ESL_MAX(chunk, (int64_t) 64);
  }
  int64_t nchunks64 = 0;
  for (int w = 0; w < nwin; w++) nchunks64 += (wj[w] - wi[w] + chunk) / chunk;
  if (nchunks64 > INT32_MAX / 4) return infx_fail(eslERANGE, errbuf, "%s: too many chunks", fname);
  const int nchunks = (int) nchunks64;
  /* several windows: the chunk table (first row, last row, window start) */
  std::vector<int64_t> ctab;
  if (nwin > 1) {
    ctab.reserve((size_t) nchunks * 3);
    for (int w = 0; w < nwin; w++)
      for (int64_t a = wi[w]; a <= wj[w]; a += chunk) { ctab.push_back(a); ctab.push_back(ESL_MIN(wj[w], a + chunk - 1)); ctab.push_back(wi[w]); }
  }
  size_t deck_off, row_off, vsc_off, acc_off;
  const int nrows = infx_std_rows(semi, tbl_lds, wide);
  const size_t slot_bytes = slot_layout(P->R, P->RB, gacc, acc_n, nrows, &deck_off, &row_off, &vsc_off, &acc_off);
  const size_t chunk_out_bytes = ESL_MAX((size_t) nchunks * 24 + 64, ((((size_t) nchunks * 4 + 255) / 256) * 256) + (size_t) nchunks * 16 + 64);
  size_t scratch_budget;
  if (ar) {
    /* the arena's rest, less the fixed buffers and a minimal candidate buffer */
    const size_t fixed = (pinned ? 256 : r256((size_t) Lspan + 128)) + r256(chunk_out_bytes) + 256 + r256(sizeof(ScanArgs)) +
                         (nwin > 1 ? r256(ctab.size() * sizeof(int64_t)) : 0) + r256((size_t) ar_cand_min * sizeof(Cand));
    scratch_budget = ar->left() > fixed ? ar->left() - fixed : 0;
  } else {
    size_t mem_free = 0, mem_total = 0;
    if (hipMemGetInfo(&mem_free, &mem_total) != hipSuccess) { (void) hipGetLastError(); mem_free = (size_t) 16 << 30; }
    scratch_budget = ESL_MIN((size_t) 96 << 30, mem_free > ((size_t) 8 << 30) ? mem_free - ((size_t) 8 << 30) : mem_free / 2)
                     + D->d_scratch.cap;
  }
  int nslots = (int) ESL_MIN((int64_t) nchunks, (int64_t) (scratch_budget / slot_bytes));
  nslots = (int) ESL_MIN((int64_t) nslots, resident);
  if (nslots < 1) return infx_fail(ar ? eslERANGE : eslEMEM, errbuf, "%s: one DP slot needs %zu bytes", fname, slot_bytes);
  /* Few slots (a short target, a few windows): workgroups small enough to
   * reach every CU, since a workgroup runs on one CU and 16-wavefront
   * workgroups would leave most of the chip idle.  Cells and hits do not
   * depend on the shape. */
  if (!infx_wpb_override())
    while (wpb > 1 && (int64_t) ((nslots + wpb - 1) / wpb) < (int64_t) g_ncu) wpb >>= 1;
  if (ar && (ar->give(D->d_dsq, pinned ? 64 : (size_t) Lspan + 128) != eslOK || ar->give(D->d_scratch, slot_bytes * nslots) != eslOK ||
             ar->give(D->d_chunk, chunk_out_bytes) != eslOK || ar->give(D->d_ncand, 64) != eslOK ||
             ar->give(D->d_args, sizeof(ScanArgs)) != eslOK ||
             (nwin > 1 && ar->give(D->d_ctab, ctab.size() * sizeof(int64_t)) != eslOK) ||
             ar->give(D->d_cand, ar->left() / 256 * 256) != eslOK))
    return infx_fail(eslERANGE, errbuf, "%s: buffers do not fit the no-allocation scratch", fname);

  if ((status = D->d_dsq.reserve(pinned ? 64 : (size_t) Lspan + 128)) != eslOK ||
      (status = D->d_scratch.reserve(slot_bytes * nslots)) != eslOK ||
      (status = D->d_chunk.reserve((size_t) nchunks * 24 + 64)) != eslOK ||
      (status = D->d_ncand.reserve(64)) != eslOK)
    return infx_fail(eslEMEM, errbuf, "%s: device allocation failed", fname);
  /* candidate buffer: four per residue (the greedy record filter keeps a
   * handful per j), at most 8 GB; an overflow re-runs the scan once with the
   * exact size */
  int64_t cand_cap = ESL_MAX((int64_t) 1 << 16, ESL_MIN((int64_t) 4 * L, ((int64_t) 8 << 30) / (int64_t) sizeof(Cand)));
  cand_cap = ESL_MAX(cand_cap, (int64_t) D->d_cand.cap / (int64_t) sizeof(Cand));
  if (ar) cand_cap = (int64_t) (D->d_cand.cap / sizeof(Cand));
  if ((status = D->d_cand.reserve((size_t) cand_cap * sizeof(Cand))) != eslOK) return infx_fail(eslEMEM, errbuf, "%s: device allocation failed", fname);
  cand_cap = (int64_t) (D->d_cand.cap / sizeof(Cand));
  /* target residues i0..j0 behind and before 64 zero bytes (residue codes
   * past j0 are read, and ignored, by idle lanes of the last row pair); the
   * pads are cleared by a kernel, so a call leaves no memset in the stream */
  if (!pinned &&
      (infx_launch_zero_bytes(D->d_dsq.ptr, 64, g_stream) != hipSuccess ||
       infx_launch_zero_bytes((char *) D->d_dsq.ptr + 64 + Lspan, 64, g_stream) != hipSuccess ||
       hipMemcpyAsync((char *) D->d_dsq.ptr + 64, dsq + i0, (size_t) Lspan, hipMemcpyHostToDevice, g_stream) != hipSuccess))
    return infx_fail(eslESYS, errbuf, "%s: dsq upload failed", fname);

  ScanArgs A;
  memset(&A, 0, sizeof(A));
  A.st = (const DState *) P->d_st.ptr;
  A.nst = (int) P->st.size();
  A.esc = P->d_esc.ptr;
  A.dsq = pinned ? pinned : (const uint8_t *) D->d_dsq.ptr + 64;
  A.dsq_off = i0;
  A.row_tmpl = P->d_rtmpl.ptr;
  A.deck_tmpl = P->d_dtmpl.ptr;
  A.lsum = (semi == SEMI_FINSIDE) ? D->d_lsum_f.ptr : (semi == SEMI_IINSIDE) ? D->d_lsum_i.ptr : NULL;
  A.scratch = (char *) D->d_scratch.ptr;
  A.slot_bytes = (int64_t) slot_bytes;
  A.deck_off = (int64_t) deck_off; A.row_off = (int64_t) row_off; A.vsc_off = (int64_t) vsc_off;
  A.acc_off = (int64_t) acc_off;
  A.acc_n = acc_n;
  A.M = cm->M; A.W = W; A.R = P->R; A.RB = P->RB;
  A.i0 = i0; A.j0 = j0;
  A.chunk_len = chunk; A.nchunks = nchunks; A.nslots = nslots;
  A.chunk_tab = NULL;
  if (nwin > 1) {
    if ((status = upload(D->d_ctab, ctab.data(), ctab.size() * sizeof(int64_t))) != eslOK)
      return infx_fail(eslEMEM, errbuf, "%s: chunk table upload failed", fname);
    A.chunk_tab = (const int64_t *) D->d_ctab.ptr;
  }
  A.root_dn = P->root_dn; A.root_dx = P->root_dx; A.root_cnum = P->root_cnum;
  memcpy(A.root_child_base, P->root_child_base, sizeof(A.root_child_base));
  memcpy(A.root_tsc, P->root_tsc, sizeof(A.root_tsc));
  A.local_begin = P->local_begin;
  A.el_selfsc = P->el;
  const int greedy = !(cm->search_opts & CM_SEARCH_CMNOTGREEDY);
  A.cand_mode = (hitlist == NULL) ? 0 : (!greedy ? 2 : (do_null3 ? 0 : 1));
  A.cutoff = (hitlist == NULL) ? INFINITY : cutoff;
  A.env_cutoff = env_cutoff;
  A.cand = (Cand *) D->d_cand.ptr;
  A.ncand = (unsigned long long *) D->d_ncand.ptr;
  A.cand_cap = cand_cap;
  A.chunk_best = (float *) D->d_chunk.ptr;
  A.chunk_envi = (int64_t *) ((char *) D->d_chunk.ptr + (((size_t) nchunks * 4 + 255) / 256) * 256);
  A.chunk_envj = A.chunk_envi + nchunks;
  A.want_vsc = (ret_vsc != NULL);
  if ((status = D->d_chunk.reserve(((((size_t) nchunks * 4 + 255) / 256) * 256) + (size_t) nchunks * 16 + 64)) != eslOK)
    return infx_fail(eslEMEM, errbuf, "%s: device allocation failed", fname);
  A.chunk_best = (float *) D->d_chunk.ptr;
  A.chunk_envi = (int64_t *) ((char *) D->d_chunk.ptr + (((size_t) nchunks * 4 + 255) / 256) * 256);
  A.chunk_envj = A.chunk_envi + nchunks;

  dim3 block(64 * wpb), grid((nslots + wpb - 1) / wpb);
  if (getenv("INFX_VERBOSE"))
    fprintf(stderr, "[infx] %s: L %lld W %d chunk %lld nchunks %d nslots %d wpb %d blocks/CU %d slot %zu B table %s begin-acc %s%s\n", fname,
            (long long) L, W, (long long) chunk, nchunks, nslots, wpb, occ, slot_bytes, tbl_lds ? "lds" : "l2", gacc ? "global" : "lds",
            wide ? " wide" : "");
  unsigned long long ncand = 0;
  float kernel_ms_total = 0;
  int launches = 0;
  for (int attempt = 0; attempt < 3; attempt++) {
    if (infx_launch_zero_u64(A.ncand, g_stream) != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: counter reset failed", fname);
    (void) hipEventRecord(g_ev0, g_stream);   /* timing only */
    if (upload(D->d_args, &A, sizeof(A)) != eslOK) return infx_fail(eslESYS, errbuf, "%s: parameter upload failed", fname);
    hipError_t e = infx_launch_std(semi, A.want_vsc, tbl_lds, gacc, wide, grid, block, g_stream, (const ScanArgs *) D->d_args.ptr, acc_n);
    (void) hipEventRecord(g_ev1, g_stream);
    if (e != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: kernel launch failed: %s", fname, hipGetErrorString(e));
    if (hipMemcpyAsync(&ncand, A.ncand, sizeof(ncand), hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
        hipStreamSynchronize(g_stream) != hipSuccess)
      return infx_fail(eslESYS, errbuf, "%s: kernel execution failed: %s", fname, hipGetErrorString(hipGetLastError()));
    float ms = 0;
    if (hipEventElapsedTime(&ms, g_ev0, g_ev1) != hipSuccess) { (void) hipGetLastError(); ms = 0; }
    kernel_ms_total += ms; launches++;
    if ((int64_t) ncand <= cand_cap) break;
    if ((status = D->d_cand.reserve((size_t) ncand * sizeof(Cand) + 4096)) != eslOK)
      return infx_fail(ar ? eslERANGE : eslEMEM, errbuf, ar ? "%s: %llu candidate cells do not fit the no-allocation scratch (room for %lld)"
                                                            : "%s: candidate buffer allocation failed (%llu candidates, %lld allocated)",
                       fname, (unsigned long long) ncand, (long long) cand_cap);
    cand_cap = (int64_t) (D->d_cand.cap / sizeof(Cand));
    A.cand = (Cand *) D->d_cand.ptr; A.cand_cap = cand_cap;
  }
  if ((int64_t) ncand > cand_cap) return infx_fail(eslEMEM, errbuf, "%s: candidate buffer overflow", fname);

  std::vector<Cand> hc(ncand);
  std::vector<float> cbest(nchunks);
  std::vector<int64_t> cenvi(nchunks), cenvj(nchunks);
  if ((ncand && hipMemcpyAsync(hc.data(), A.cand, ncand * sizeof(Cand), hipMemcpyDeviceToHost, g_stream) != hipSuccess) ||
      hipMemcpyAsync(cbest.data(), A.chunk_best, nchunks * 4, hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
      hipMemcpyAsync(cenvi.data(), A.chunk_envi, nchunks * 8, hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
      hipMemcpyAsync(cenvj.data(), A.chunk_envj, nchunks * 8, hipMemcpyDeviceToHost, g_stream) != hipSuccess)
    return infx_fail(eslESYS, errbuf, "%s: result download failed", fname);
  std::vector<float> vsc_slots;
  if (ret_vsc) {
    vsc_slots.resize((size_t) nslots * cm->M);
    if (hipMemcpy2DAsync(vsc_slots.data(), (size_t) cm->M * 4, A.scratch + vsc_off, slot_bytes, (size_t) cm->M * 4, nslots,
                         hipMemcpyDeviceToHost, g_stream) != hipSuccess)
      return infx_fail(eslESYS, errbuf, "%s: vsc download failed", fname);
  }
  if (hipStreamSynchronize(g_stream) != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: download failed", fname);

  float vsc_root = IMPOSSIBLE;
  int64_t envi = INT64_MAX, envj = INT64_MIN;
  for (int c = 0; c < nchunks; c++) {
    vsc_root = ESL_MAX(vsc_root, cbest[c]);
    envi = ESL_MIN(envi, cenvi[c]);
    envj = ESL_MAX(envj, cenvj[c]);
  }
  if (ret_envi) *ret_envi = (envi == INT64_MAX) ? -1 : envi;
  if (ret_envj) *ret_envj = (envj == INT64_MIN) ? -1 : envj;
  if (ret_sc) *ret_sc = vsc_root;
  if (ret_vsc) {
    float *vsc = (float *) malloc(sizeof(float) * cm->M);
    for (int v = 0; v < cm->M; v++) {
      float m = IMPOSSIBLE;
      for (int s = 0; s < nslots; s++) m = ESL_MAX(m, vsc_slots[(size_t) s * cm->M + v]);
      vsc[v] = m;
    }
    vsc[0] = vsc_root;
    *ret_vsc = vsc;
  }

  auto t_resolve = std::chrono::steady_clock::now();
  if (hitlist) {
    std::vector<INFX_CAND> cands(ncand);
    for (size_t k = 0; k < ncand; k++) { cands[k].j = hc[k].j; cands[k].d = hc[k].d; cands[k].sc = hc[k].sc; cands[k].r = hc[k].r; cands[k].mode = TRMODE_J; }
    if ((status = infx_resolve_hits(cm, errbuf, PLI_PASS_STD_ANY, 0, cands.data(), (int64_t) ncand, dsq, i0, j0, W, cutoff, do_null3,
                                    greedy, hitlist)) != eslOK)
      return status;
  }

  double cells = 0;
  {
    CellRows cr;
    infx_cell_rows(cm, smx, qdbidx, 0, &cr);
    for (int w = 0; w < nwin; w++) cells += cr.cells(wj[w] - wi[w] + 1);
  }
  double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (getenv("INFX_VERBOSE"))
    fprintf(stderr, "[infx] %s: wall %.1f ms = kernel %.1f ms (%d launch(es)) + hit replay %.1f ms (%llu candidates) + rest\n", fname,
            wall, kernel_ms_total, launches,
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_resolve).count(), (unsigned long long) ncand);
  infx_timing_add(wall, kernel_ms_total, (int64_t) cells, launches);
  return eslOK;
}

/* The device scan profile FastCYKScan/FastFInsideScan/FastIInsideScan would
 * upload for a target of length L (for layout checks on the host):
 * descriptors (up to cap), row/deck sizes, emission-table size, root band
 * and children.  Returns the number of descriptors, or -status on error. */
extern "C" int infx_debug_std_profile(CM_t *cm, CM_SCAN_MX *smx, int qdbidx, int semi, int64_t L, void *out_st, int cap,
                                      int *ret_R, int *ret_RB, int *ret_nesc, int *ret_root /* [3 + 6]: dn, dx, cnum, child bases */,
                                      char *errbuf)
{
  if (!smx || L < 1) return -eslEINVAL;
  const int W = (smx->W > L) ? (int) L : smx->W;
  StdProfile P;
  int status = build_std_profile(cm, smx, qdbidx, semi, W, P, errbuf);
  if (status != eslOK) return -status;
  for (int i = 0; i < (int) P.st.size() && i < cap; i++) memcpy((char *) out_st + (size_t) i * sizeof(DState), &P.st[i], sizeof(DState));
  *ret_R = P.R; *ret_RB = P.RB; *ret_nesc = (int) P.esc.size();
  ret_root[0] = P.root_dn; ret_root[1] = P.root_dx; ret_root[2] = P.root_cnum;
  for (int k = 0; k < 6; k++) ret_root[3 + k] = P.root_child_base[k];
  return (int) P.st.size();
}

/* Exact number of (v, j, d) cells the reference computes over j = i0..j0
 * (E states excluded, v = 0 included): bands clipped by jp for jp < W. */
void infx_cell_rows(const CM_t *cm, const void *mx, int qdbidx, int truncated, CellRows *out)
{
  int M, W; int ***dnA, ***dxA;
  if (truncated) { const CM_TR_SCAN_MX *t = (const CM_TR_SCAN_MX *) mx; M = t->M; W = t->W; dnA = t->dnAAA; dxA = t->dxAAA; }
  else           { const CM_SCAN_MX *s = (const CM_SCAN_MX *) mx;    M = s->M; W = s->W; dnA = s->dnAAA; dxA = s->dxAAA; }
  out->W = W;
  out->row.assign((size_t) W + 1, 0.0);
  out->pre.assign((size_t) W + 1, 0.0);
  for (int jp = 1; jp <= W; jp++) {
    double row = 0;
    for (int v = 0; v < M; v++) {
      if (cm->sttype[v] == E_st) continue;
      const int dn = dnA[qdbidx][jp][v], dx = dxA[qdbidx][jp][v];
      if (dx >= dn) row += dx - dn + 1;
    }
    out->row[jp] = row;
    out->pre[jp] = out->pre[jp - 1] + row;
  }
}

extern "C" int infx_count_scan_cells(const CM_t *cm, const void *mx, int qdbidx, int truncated, int64_t i0, int64_t j0, double *ret_cells)
{
  CellRows cr;
  infx_cell_rows(cm, mx, qdbidx, truncated, &cr);
  *ret_cells = cr.cells(j0 - i0 + 1);
  return eslOK;
}

extern "C" int FastCYKScan(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                           float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                           int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc)
{
  return std_scan(SEMI_CYK, "FastCYKScan", cm, errbuf, smx, qdbidx, dsq, &i0, &j0, 1, cutoff, hitlist, do_null3, env_cutoff,
                  ret_envi, ret_envj, ret_vsc, ret_sc);
}
extern "C" int FastFInsideScan(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                               float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                               int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc)
{
  return std_scan(SEMI_FINSIDE, "FastFInsideScan", cm, errbuf, smx, qdbidx, dsq, &i0, &j0, 1, cutoff, hitlist, do_null3, env_cutoff,
                  ret_envi, ret_envj, ret_vsc, ret_sc);
}
/* FastCYKScan without device allocation: every device buffer of the call
 * (profile, residues, parameters, DP slots, candidates, chunk outputs) is
 * carved from the library's fixed scratch array on the current device, so
 * the call makes no hipMalloc / hipFree.  Calls are serialised on a mutex:
 * they share the array, which keeps the last call's contents (the counters a
 * launch needs zeroed are cleared by the call).  A model and target whose
 * buffers do not fit even one DP slot are refused with eslERANGE before any
 * device work; the DP slots in use (and so the parallelism) are what the
 * array holds after the fixed buffers. */
static std::mutex g_noalloc_mu;
extern "C" int infx_FastCYKScanNoAlloc(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                                       float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                                       int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc)
{
  int status;
  Arena ar;
  void *base = NULL;
  size_t bytes = 0;
  if ((status = infx_dev_init(errbuf)) == eslOK && (status = infx_noalloc_scratch(&base, &bytes, errbuf)) != eslOK) return status;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (status == eslOK) { ar.base = (char *) base; ar.cap = bytes; }
  else {
    /* no device: the size check still runs (against the array's size), then
     * the scan reports the missing device */
    ar.base = NULL; ar.cap = INFX_NOALLOC_BYTES;
  }
  return std_scan(SEMI_CYK, "FastCYKScan", cm, errbuf, smx, qdbidx, dsq, &i0, &j0, 1, cutoff, hitlist, do_null3, env_cutoff,
                  ret_envi, ret_envj, ret_vsc, ret_sc, &ar);
}

/* Many windows of one target in one launch (no counterpart in the
 * reference: the windows' hits are those of one standard scan per window). */
extern "C" int infx_FastScanWindows(int semi, CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq,
                                    const int64_t *wi, const int64_t *wj, int nwin, float cutoff, CM_TOPHITS *hitlist,
                                    int do_null3, float *ret_sc)
{
  static const char *names[3] = { "FastCYKScan", "FastFInsideScan", "FastIInsideScan" };
  if (semi < SEMI_CYK || semi > SEMI_IINSIDE) return infx_fail(eslEINVAL, errbuf, "infx_FastScanWindows: bad algorithm %d", semi);
  if (nwin < 1 || wi == NULL || wj == NULL) return infx_fail(eslEINCOMPAT, errbuf, "infx_FastScanWindows: no windows");
  return std_scan(semi, names[semi], cm, errbuf, smx, qdbidx, dsq, wi, wj, nwin, cutoff, hitlist, do_null3, 0.f,
                  NULL, NULL, NULL, ret_sc);
}

/* Several models' windows of one target in one launch.  Window w belongs to
 * model wm[w]; each model's windows are sorted and disjoint (windows of
 * different models may overlap).  Every model's hits are those of
 * infx_FastScanWindows over its own windows (greedy reporting, no NULL3),
 * appended to hitlists[m]; ret_sc[m] is its best score.  Chunks of every
 * model share one launch: the slot is sized for the largest model and each
 * chunk reads its own model's parameters. */
extern "C" int infx_FastScanBatch(int semi, int nmodel, CM_t **cms, CM_SCAN_MX **smxs, char *errbuf, int qdbidx, ESL_DSQ *dsq,
                                  const int64_t *wi, const int64_t *wj, const int32_t *wm, int nwin, const float *cutoffs,
                                  CM_TOPHITS **hitlists, float *ret_sc)
{
  static const char *names[3] = { "FastCYKScan", "FastFInsideScan", "FastIInsideScan" };
  int status;
  auto t_start = std::chrono::steady_clock::now();
  infx_timing_reset();
  if (semi < SEMI_CYK || semi > SEMI_IINSIDE) return infx_fail(eslEINVAL, errbuf, "infx_FastScanBatch: bad algorithm %d", semi);
  const char *fname = names[semi];
  if (nmodel < 1 || nmodel > 32767 || !cms || !smxs) return infx_fail(eslEINCOMPAT, errbuf, "infx_FastScanBatch: bad model count %d", nmodel);
  if (nwin < 1 || !wi || !wj || !wm || !dsq) return infx_fail(eslEINCOMPAT, errbuf, "infx_FastScanBatch: no windows");
  /* per-model contract checks, as the single-model scanners make them */
  std::vector<int64_t> lo_m(nmodel, INT64_MAX), hi_m(nmodel, 0), lmax_m(nmodel, 0), last_j(nmodel, 0);
  int64_t span_lo = INT64_MAX, span_hi = 0, Ltot = 0;
  for (int w = 0; w < nwin; w++) {
    const int m = wm[w];
    if (m < 0 || m >= nmodel) return infx_fail(eslEINVAL, errbuf, "%s, window %d: model index %d out of range", fname, w, m);
    if (wi[w] < 1 || wj[w] < wi[w]) return infx_fail(eslEINCOMPAT, errbuf, "%s, i0: %lld j0: %lldd\n", fname, (long long) wi[w], (long long) wj[w]);
    if (last_j[m] && wi[w] <= last_j[m]) return infx_fail(eslEINCOMPAT, errbuf, "%s, windows of one model must be sorted and disjoint", fname);
    last_j[m] = wj[w];
    lo_m[m] = ESL_MIN(lo_m[m], wi[w]); hi_m[m] = ESL_MAX(hi_m[m], wj[w]);
    lmax_m[m] = ESL_MAX(lmax_m[m], wj[w] - wi[w] + 1);
    span_lo = ESL_MIN(span_lo, wi[w]); span_hi = ESL_MAX(span_hi, wj[w]);
    Ltot += wj[w] - wi[w] + 1;
  }
  if (span_hi - span_lo + 1 > (int64_t) INT32_MAX - 1) return infx_fail(eslERANGE, errbuf, "%s, target range too long for one call", fname);
  if (span_hi > (int64_t) INT32_MAX - 1) return infx_fail(eslERANGE, errbuf, "%s, j0 = %lld: positions past %d do not fit the hit candidates", fname, (long long) span_hi, INT32_MAX - 1);
  const uint8_t *pinned = infx_pinned(dsq, span_lo, span_hi);   /* residues already on the device (checked when pinned) */
  for (int64_t x = span_lo; !pinned && x <= span_hi; x++)
    if (dsq[x] > 15) return infx_fail(eslEINVAL, errbuf, "%s, residue code %d at %lld is not a residue", fname, (int) dsq[x], (long long) x);
  std::vector<int> Wm(nmodel, 0);
  std::vector<StdProfile> fresh(nmodel);
  for (int m = 0; m < nmodel; m++) {
    CM_t *cm = cms[m]; CM_SCAN_MX *smx = smxs[m];
    if (!cm || !smx)                    return infx_fail(eslEINCOMPAT, errbuf, "%s, model %d: cm or smx is NULL", fname, m);
    if (!(cm->flags & CMH_BITS))        return infx_fail(eslEINCOMPAT, errbuf, "%s, CMH_BITS flag is not raised.\n", fname);
    if (semi == SEMI_CYK && (cm->search_opts & CM_SEARCH_INSIDE))  return infx_fail(eslEINCOMPAT, errbuf, "%s, CM_SEARCH_INSIDE flag raised", fname);
    if (semi != SEMI_CYK && !(cm->search_opts & CM_SEARCH_INSIDE)) return infx_fail(eslEINCOMPAT, errbuf, "%s, CM_SEARCH_INSIDE flag not raised", fname);
    if (semi == SEMI_CYK && !smx->floats_valid) return infx_fail(eslEINCOMPAT, errbuf, "%s, smx->floats_valid if FALSE", fname);
    if (semi != SEMI_CYK && !smx->ints_valid)   return infx_fail(eslEINCOMPAT, errbuf, "%s, smx->ints_valid if FALSE", fname);
    if (cm->qdbinfo == NULL)            return infx_fail(eslEINCOMPAT, errbuf, "%s, cm->qdbinfo == NULL\n", fname);
    if (qdbidx < 0 || qdbidx > 2)       return infx_fail(eslEINCOMPAT, errbuf, "%s, qdbidx is invalid", fname);
    if (cm->M > 65535)                  return infx_fail(eslERANGE, errbuf, "%s, model %d: %d states (a candidate keeps its state in 16 bits)", fname, m, cm->M);
    if (lmax_m[m] == 0) continue;       /* a model without windows */
    Wm[m] = (smx->W > lmax_m[m]) ? (int) lmax_m[m] : smx->W;
    if ((status = build_std_profile(cm, smx, qdbidx, semi, Wm[m], fresh[m], errbuf)) != eslOK) return status;
  }
  if ((status = infx_dev_init(errbuf)) != eslOK) return status;

  /* each model's profile on the device (cached with its scan matrix) */
  std::vector<StdProfile *> prof(nmodel, (StdProfile *) NULL);
  for (int m = 0; m < nmodel; m++) {
    if (lmax_m[m] == 0) continue;
    CM_SCAN_MX *smx = smxs[m];
    if (!smx->dev) smx->dev = new SmxDev();
    SmxDev *Dm = (SmxDev *) smx->dev;
    StdProfile *P = NULL;
    for (auto *p : Dm->prof) if (p->qdbidx == qdbidx && p->semi == semi && p->W == Wm[m]) P = p;
    if (!P) { P = new StdProfile(); Dm->prof.push_back(P); }
    if (!P->uploaded || !same_profile(*P, fresh[m])) {
      const StdProfile &F = fresh[m];
      P->st = F.st; P->esc = F.esc; P->rtmpl = F.rtmpl; P->dtmpl = F.dtmpl;
      P->R = F.R; P->RB = F.RB; P->root_dn = F.root_dn; P->root_dx = F.root_dx; P->root_cnum = F.root_cnum;
      memcpy(P->root_child_base, F.root_child_base, sizeof(P->root_child_base));
      memcpy(P->root_tsc, F.root_tsc, sizeof(P->root_tsc));
      P->local_begin = F.local_begin; P->el = F.el;
      P->qdbidx = qdbidx; P->semi = semi; P->W = Wm[m];
      if ((status = upload(P->d_st, P->st.data(), P->st.size() * sizeof(DState))) != eslOK ||
          (status = upload(P->d_esc, P->esc.data(), P->esc.size() * 4)) != eslOK ||
          (status = upload(P->d_rtmpl, P->rtmpl.data(), P->rtmpl.size() * 4, 4)) != eslOK ||
          (status = upload(P->d_dtmpl, P->dtmpl.data(), P->dtmpl.size() * 4, 4)) != eslOK)
        return infx_fail(status, errbuf, "%s: device upload failed: %s", fname, hipGetErrorString(g_last_hip));
      P->uploaded = true;
    }
    prof[m] = P;
  }
  /* launch-wide buffers live with the first model's scan matrix */
  int m0 = 0;
  while (m0 < nmodel && !prof[m0]) m0++;
  SmxDev *D = (SmxDev *) smxs[m0]->dev;
  if (semi == SEMI_FINSIDE && D->d_lsum_f.cap == 0) {
    std::vector<float> t(INFX_LOGSUM_PAD, 0.f);
    memcpy(t.data(), infx_flogsum_table(), LOGSUM_TBL * 4);
    if ((status = upload(D->d_lsum_f, t.data(), t.size() * 4)) != eslOK) return infx_fail(status, errbuf, "upload failed");
  }
  if (semi == SEMI_IINSIDE && D->d_lsum_i.cap == 0) {
    std::vector<int> t(INFX_LOGSUM_PAD, 0);
    memcpy(t.data(), infx_ilogsum_table(), LOGSUM_TBL * 4);
    if ((status = upload(D->d_lsum_i, t.data(), t.size() * 4)) != eslOK) return infx_fail(status, errbuf, "upload failed");
  }

  /* launch shape: LDS accumulators for the widest root band */
  int acc_max = 0, R_max = 0, RB_max = 0, M_max = 0, W_max = 0;
  for (int m = 0; m < nmodel; m++) {
    if (!prof[m]) continue;
    const StdProfile *P = prof[m];
    acc_max = ESL_MAX(acc_max, (P->root_dn <= P->root_dx) ? P->root_dx - P->root_dn + 1 : 0);
    R_max = ESL_MAX(R_max, P->R); RB_max = ESL_MAX(RB_max, P->RB);
    M_max = ESL_MAX(M_max, cms[m]->M); W_max = ESL_MAX(W_max, Wm[m]);
  }
  int wpb = 0, occ = 0;
  for (int w = 16; w >= 1; w >>= 1) {
    if (infx_wpb_override() && w != infx_wpb_override()) continue;
    if (infx_std_batch_lds_bytes(semi, w, acc_max) > (size_t) g_lds_max) continue;
    const int o = infx_std_batch_occupancy(semi, w, acc_max);
    if (o * w > occ * wpb) { wpb = w; occ = o; }
  }
  if (wpb < 1 || occ < 1)
    return infx_fail(eslEINVAL, errbuf, "%s: no workgroup shape fits (W = %d, LDS limit %zu bytes)", fname, W_max, (size_t) g_lds_max);
  const int64_t resident = (int64_t) occ * wpb * g_ncu;

  /* chunks: each at least 2W of its model long; the costliest (rows x
   * cells per row) first, so the long ones do not trail the launch */
  struct Ch { int64_t a, b, lo; int m; double cost; };
  std::vector<Ch> ch;
  std::vector<double> cpr(nmodel, 0.0);   /* cells per residue at full band */
  std::vector<CellRows> crows(nmodel);
  for (int m = 0; m < nmodel; m++) if (prof[m]) { infx_cell_rows(cms[m], smxs[m], qdbidx, 0, &crows[m]); cpr[m] = crows[m].cells(Wm[m]) / Wm[m]; }
  int64_t base_chunk = infx_chunk_override() > 0 ? infx_chunk_override() : (Ltot + resident - 1) / resident;
  for (int w = 0; w < nwin; w++) {
    const int m = wm[w];
    const int64_t chunk = ESL_MAX(ESL_MAX(base_chunk, (int64_t) 2 * Wm[m]), (int64_t) 64);
    for (int64_t a = wi[w]; a <= wj[w]; a += chunk) {
      const int64_t b = ESL_MIN(wj[w], a + chunk - 1);
      const int64_t s0 = ESL_MAX(wi[w], a - Wm[m] + 1);
      ch.push_back(Ch{a, b, wi[w], m, (double) (b - s0 + 1) * cpr[m]});
    }
  }
  std::stable_sort(ch.begin(), ch.end(), [](const Ch &x, const Ch &y) { return x.cost > y.cost; });
  if (ch.size() > (size_t) INT32_MAX / 4) return infx_fail(eslERANGE, errbuf, "%s: too many chunks", fname);
  const int nchunks = (int) ch.size();
  std::vector<int64_t> ctab((size_t) nchunks * 4);
  for (int c = 0; c < nchunks; c++) { ctab[4 * c] = ch[c].a; ctab[4 * c + 1] = ch[c].b; ctab[4 * c + 2] = ch[c].lo; ctab[4 * c + 3] = ch[c].m; }

  /* one slot layout for every model: the largest deck, ring and vsc */
  const size_t G = (size_t) ((W_max + 128 + 63) / 64) * 64;
  const size_t deck_off = G * 4;
  const size_t row_off  = deck_off + ((size_t) RB_max + G) * 4;
  const size_t vsc_off  = row_off + ((size_t) 4 * R_max + G) * 4;
  const size_t slot_bytes = ((vsc_off + (size_t) M_max * 4) + 255) / 256 * 256;
  size_t mem_free = 0, mem_total = 0;
  if (hipMemGetInfo(&mem_free, &mem_total) != hipSuccess) { (void) hipGetLastError(); mem_free = (size_t) 16 << 30; }
  const size_t scratch_budget = ESL_MIN((size_t) 96 << 30, mem_free > ((size_t) 8 << 30) ? mem_free - ((size_t) 8 << 30) : mem_free / 2)
                                + D->d_scratch.cap;
  int nslots = (int) ESL_MIN((int64_t) nchunks, (int64_t) (scratch_budget / slot_bytes));
  nslots = (int) ESL_MIN((int64_t) nslots, resident);
  if (nslots < 1) return infx_fail(eslEMEM, errbuf, "%s: one DP slot needs %zu bytes", fname, slot_bytes);
  /* few slots: workgroups small enough to reach every CU (as in std_scan) */
  if (!infx_wpb_override())
    while (wpb > 1 && (int64_t) ((nslots + wpb - 1) / wpb) < (int64_t) g_ncu) wpb >>= 1;
  const int64_t Lspan = span_hi - span_lo + 1;
  int64_t cand_cap = ESL_MAX((int64_t) 1 << 16, ESL_MIN((int64_t) 4 * Ltot, ((int64_t) 8 << 30) / (int64_t) sizeof(Cand)));
  cand_cap = ESL_MAX(cand_cap, (int64_t) D->d_cand.cap / (int64_t) sizeof(Cand));
  if ((status = D->d_dsq.reserve(pinned ? 64 : (size_t) Lspan + 128)) != eslOK ||
      (status = D->d_scratch.reserve(slot_bytes * nslots)) != eslOK ||
      (status = D->d_ncand.reserve(64)) != eslOK ||
      (status = D->d_cand.reserve((size_t) cand_cap * sizeof(Cand))) != eslOK ||
      (status = D->d_chunk.reserve(((((size_t) nchunks * 4 + 255) / 256) * 256) + (size_t) nchunks * 16 + 64)) != eslOK)
    return infx_fail(eslEMEM, errbuf, "%s: device allocation failed", fname);
  cand_cap = (int64_t) (D->d_cand.cap / sizeof(Cand));
  if (!pinned &&
      (hipMemsetAsync(D->d_dsq.ptr, 0, (size_t) Lspan + 128, g_stream) != hipSuccess ||
       hipMemcpyAsync((char *) D->d_dsq.ptr + 64, dsq + span_lo, (size_t) Lspan, hipMemcpyHostToDevice, g_stream) != hipSuccess))
    return infx_fail(eslESYS, errbuf, "%s: dsq upload failed", fname);

  /* each model's parameters, then the launch's */
  std::vector<ScanArgs> args(nmodel);
  for (int m = 0; m < nmodel; m++) {
    ScanArgs &A = args[m];
    memset(&A, 0, sizeof(A));
    const StdProfile *P = prof[m] ? prof[m] : prof[m0];
    const CM_t *cm = prof[m] ? cms[m] : cms[m0];
    A.st = (const DState *) P->d_st.ptr; A.nst = (int) P->st.size(); A.esc = P->d_esc.ptr;
    A.dsq = pinned ? pinned : (const uint8_t *) D->d_dsq.ptr + 64; A.dsq_off = span_lo;
    A.row_tmpl = P->d_rtmpl.ptr; A.deck_tmpl = P->d_dtmpl.ptr;
    A.lsum = (semi == SEMI_FINSIDE) ? D->d_lsum_f.ptr : (semi == SEMI_IINSIDE) ? D->d_lsum_i.ptr : NULL;
    A.scratch = (char *) D->d_scratch.ptr; A.slot_bytes = (int64_t) slot_bytes;
    A.deck_off = (int64_t) deck_off; A.row_off = (int64_t) row_off; A.vsc_off = (int64_t) vsc_off;
    A.acc_n = acc_max;
    A.M = cm->M; A.W = prof[m] ? Wm[m] : Wm[m0]; A.R = P->R; A.RB = P->RB;
    A.root_dn = P->root_dn; A.root_dx = P->root_dx; A.root_cnum = P->root_cnum;
    memcpy(A.root_child_base, P->root_child_base, sizeof(A.root_child_base));
    memcpy(A.root_tsc, P->root_tsc, sizeof(A.root_tsc));
    A.local_begin = P->local_begin; A.el_selfsc = P->el;
    const int greedy = !(cm->search_opts & CM_SEARCH_CMNOTGREEDY);
    A.cand_mode = (hitlists == NULL || hitlists[m] == NULL) ? 0 : (greedy ? 1 : 2);
    A.cutoff = (hitlists == NULL || hitlists[m] == NULL) ? INFINITY : cutoffs[m];
    A.env_cutoff = INFINITY;
    A.cand = (Cand *) D->d_cand.ptr; A.ncand = (unsigned long long *) D->d_ncand.ptr; A.cand_cap = cand_cap;
    A.chunk_best = (float *) D->d_chunk.ptr;
    A.chunk_envi = (int64_t *) ((char *) D->d_chunk.ptr + (((size_t) nchunks * 4 + 255) / 256) * 256);
    A.chunk_envj = A.chunk_envi + nchunks;
  }
  ScanArgs Am = args[m0];
  Am.nchunks = nchunks; Am.nslots = nslots; Am.acc_n = acc_max;
  if ((status = upload(D->d_ctab, ctab.data(), ctab.size() * sizeof(int64_t))) != eslOK)
    return infx_fail(eslEMEM, errbuf, "%s: chunk table upload failed", fname);
  Am.chunk_tab = (const int64_t *) D->d_ctab.ptr;

  dim3 block(64 * wpb), grid((nslots + wpb - 1) / wpb);
  unsigned long long ncand = 0;
  float kernel_ms_total = 0;
  int launches = 0;
  for (int attempt = 0; attempt < 3; attempt++) {
    if ((status = upload(D->d_btab, args.data(), args.size() * sizeof(ScanArgs))) != eslOK)
      return infx_fail(eslEMEM, errbuf, "%s: parameter upload failed", fname);
    Am.args_tab = (const ScanArgs *) D->d_btab.ptr;
    if (upload(D->d_args, &Am, sizeof(Am)) != eslOK) return infx_fail(eslESYS, errbuf, "%s: parameter upload failed", fname);
    hipError_t e = hipSuccess;
    if (g_batch_graph) {
      /* counter reset and kernel as one graph; the kernel reads its
       * parameters from device memory (uploaded above), so the graph is
       * reused while the launch shape and the buffers it names stay put */
      const size_t key[6] = {(size_t) semi, (size_t) grid.x, (size_t) block.x, (size_t) acc_max, (size_t) D->d_args.ptr, (size_t) Am.ncand};
      if (!D->gexec || memcmp(key, D->gkey, sizeof(key)) != 0) {
        if (D->gexec) { (void) hipGraphExecDestroy(D->gexec); D->gexec = NULL; }
        if (D->graph) { (void) hipGraphDestroy(D->graph); D->graph = NULL; }
        hipGraph_t g = NULL;
        if ((e = hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal)) == hipSuccess) {
          hipError_t e1 = infx_launch_zero_u64(Am.ncand, g_stream);
          hipError_t e2 = (e1 == hipSuccess) ? infx_launch_std_batch(semi, grid, block, g_stream, (const ScanArgs *) D->d_args.ptr, acc_max) : e1;
          e = hipStreamEndCapture(g_stream, &g);
          if (e == hipSuccess && e2 != hipSuccess) e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&D->gexec, g, NULL, NULL, 0);
        if (e != hipSuccess) {
          if (g) (void) hipGraphDestroy(g);
          D->gexec = NULL;
          return infx_fail(eslESYS, errbuf, "%s: graph capture failed: %s", fname, hipGetErrorString(e));
        }
        D->graph = g;
        memcpy(D->gkey, key, sizeof(key));
      }
      (void) hipEventRecord(g_ev0, g_stream);   /* timing only */
      e = hipGraphLaunch(D->gexec, g_stream);
      (void) hipEventRecord(g_ev1, g_stream);
    } else {
      if (hipMemsetAsync(Am.ncand, 0, sizeof(unsigned long long), g_stream) != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: memset failed", fname);
      (void) hipEventRecord(g_ev0, g_stream);   /* timing only */
      e = infx_launch_std_batch(semi, grid, block, g_stream, (const ScanArgs *) D->d_args.ptr, acc_max);
      (void) hipEventRecord(g_ev1, g_stream);
    }
    if (e != hipSuccess) return infx_fail(eslESYS, errbuf, "%s: kernel launch failed: %s", fname, hipGetErrorString(e));
    if (hipMemcpyAsync(&ncand, Am.ncand, sizeof(ncand), hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
        hipStreamSynchronize(g_stream) != hipSuccess)
      return infx_fail(eslESYS, errbuf, "%s: kernel execution failed: %s", fname, hipGetErrorString(hipGetLastError()));
    float ms = 0;
    if (hipEventElapsedTime(&ms, g_ev0, g_ev1) != hipSuccess) { (void) hipGetLastError(); ms = 0; }
    kernel_ms_total += ms; launches++;
    if ((int64_t) ncand <= cand_cap) break;
    if ((status = D->d_cand.reserve((size_t) ncand * sizeof(Cand) + 4096)) != eslOK) return infx_fail(eslEMEM, errbuf, "%s: candidate buffer allocation failed", fname);
    cand_cap = (int64_t) (D->d_cand.cap / sizeof(Cand));
    for (auto &A : args) { A.cand = (Cand *) D->d_cand.ptr; A.cand_cap = cand_cap; }
    Am.cand = (Cand *) D->d_cand.ptr; Am.cand_cap = cand_cap;
  }
  if ((int64_t) ncand > cand_cap) return infx_fail(eslEMEM, errbuf, "%s: candidate buffer overflow", fname);

  std::vector<Cand> hc(ncand);
  std::vector<float> cbest(nchunks);
  if ((ncand && hipMemcpyAsync(hc.data(), Am.cand, ncand * sizeof(Cand), hipMemcpyDeviceToHost, g_stream) != hipSuccess) ||
      hipMemcpyAsync(cbest.data(), Am.chunk_best, nchunks * 4, hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
      hipStreamSynchronize(g_stream) != hipSuccess)
    return infx_fail(eslESYS, errbuf, "%s: result download failed", fname);
  if (ret_sc) {
    for (int m = 0; m < nmodel; m++) ret_sc[m] = IMPOSSIBLE;
    for (int c = 0; c < nchunks; c++) ret_sc[ch[c].m] = ESL_MAX(ret_sc[ch[c].m], cbest[c]);
  }
  if (hitlists) {
    std::vector<std::vector<INFX_CAND>> per(nmodel);
    for (size_t k = 0; k < ncand; k++) {
      const int m = (int) ((uint32_t) hc[k].r >> 16);
      if (m >= nmodel) return infx_fail(eslEINCONCEIVABLE, errbuf, "%s: candidate of model %d", fname, m);
      INFX_CAND x; x.j = hc[k].j; x.d = hc[k].d; x.sc = hc[k].sc; x.r = hc[k].r & 0xffff; x.mode = TRMODE_J;
      per[m].push_back(x);
    }
    for (int m = 0; m < nmodel; m++) {
      if (!hitlists[m] || !prof[m]) continue;
      const int greedy = !(cms[m]->search_opts & CM_SEARCH_CMNOTGREEDY);
      if ((status = infx_resolve_hits(cms[m], errbuf, PLI_PASS_STD_ANY, 0, per[m].data(), (int64_t) per[m].size(), dsq, lo_m[m], hi_m[m],
                                      Wm[m], cutoffs[m], 0, greedy, hitlists[m])) != eslOK)
        return status;
    }
  }
  double cells = 0;
  for (int w = 0; w < nwin; w++) cells += crows[wm[w]].cells(wj[w] - wi[w] + 1);
  double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (getenv("INFX_VERBOSE"))
    fprintf(stderr, "[infx] %s batch: %d models, %d windows, %d chunks, %d slots x %zu B, wpb %d blocks/CU %d: wall %.1f ms, kernel %.1f ms, %llu candidates\n",
            fname, nmodel, nwin, nchunks, nslots, slot_bytes, wpb, occ, wall, kernel_ms_total, (unsigned long long) ncand);
  infx_timing_add(wall, kernel_ms_total, (int64_t) cells, launches);
  return eslOK;
}

extern "C" int FastIInsideScan(CM_t *cm, char *errbuf, CM_SCAN_MX *smx, int qdbidx, ESL_DSQ *dsq, int64_t i0, int64_t j0,
                               float cutoff, CM_TOPHITS *hitlist, int do_null3, float env_cutoff,
                               int64_t *ret_envi, int64_t *ret_envj, float **ret_vsc, float *ret_sc)
{
  return std_scan(SEMI_IINSIDE, "FastIInsideScan", cm, errbuf, smx, qdbidx, dsq, &i0, &j0, 1, cutoff, hitlist, do_null3, env_cutoff,
                  ret_envi, ret_envj, ret_vsc, ret_sc);
}
