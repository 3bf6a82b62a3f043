/* scan_host.h -- host-side launch helpers shared by the standard and the
 * truncated GPU scanners. */
#ifndef INFX_SCAN_HOST_H
#define INFX_SCAN_HOST_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <vector>

struct DevBuf {
  void  *ptr = nullptr;
  size_t cap = 0;
  bool   borrowed = false;  /* memory lent by an arena: never grown or freed here */
  int  reserve(size_t n);   /* grow (contents not kept); a borrowed buffer cannot grow: eslERANGE */
  void release();
};

/* A fixed device region that one call's buffers are carved from (the
 * no-allocation scanner's static scratch array): bump allocation, 256-byte
 * aligned, nothing freed. */
struct Arena {
  char  *base = nullptr;
  size_t cap = 0, used = 0;
  size_t left() const { return cap > used ? cap - used : 0; }
  int give(DevBuf &b, size_t n);
};

/* The no-allocation scanner's scratch array on the current device: its
 * device address (looked up once per device) and size. */
int infx_noalloc_scratch(void **ret_ptr, size_t *ret_bytes, char *errbuf);

extern "C" int  infx_dev_init(char *errbuf);
extern "C" void infx_dev_release(void *dev);
extern "C" void infx_trdev_release(void *trdev);
extern "C" int  infx_count_scan_cells(const CM_t *cm, const void *mx, int qdbidx, int truncated, int64_t i0, int64_t j0, double *ret_cells);
/* banded cells of a scan matrix's rows at chunk positions jp = 1..W (E states
 * skipped): row[jp], and pre[k] = row[1] + ... + row[k].  A window of L
 * residues scans rows 1..W'-1 once and row W' = min(W, L) for its other
 * L - W' + 1 positions: cells(L) = pre[W'-1] + (L - W' + 1) row[W'].  Built
 * once per model and call, O(1) per window. */
struct CellRows {
  std::vector<double> row, pre;
  int W = 0;
  double cells(int64_t L) const {
    if (L <= 0 || W <= 0) return 0.0;
    const int w = (int) (L < W ? L : W);
    return pre[w - 1] + (double) (L - w + 1) * row[w];
  }
};
void infx_cell_rows(const CM_t *cm, const void *mx, int qdbidx, int truncated, CellRows *out);

hipStream_t infx_stream(void);
const uint8_t *infx_pinned(const uint8_t *dsq, int64_t x0, int64_t x1);
int  infx_num_cu(void);
int  infx_chunk_override(void);
int  infx_wpb_override(void);
void infx_timing_reset(void);
void infx_timing_add(double wall, double kern, int64_t cells, int launches);

#endif
