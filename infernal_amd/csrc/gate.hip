/* gate.hip -- the learned window gate: a per-model 1-D convolutional scorer
 * over one-hot ACGU, in bf16 on the matrix cores with fp32 accumulation, with
 * its own training kernels.  No counterpart in the reference: it sits in
 * front of the exact Inside scan (which stays the judge of its misses).
 *
 * Network: 4 convolutions (kernel 9, dilation dil, padding 4*dil, 128
 * channels, ReLU), then a per-position linear head (128 -> 1 logit).  Activations are channels-last
 * [position][channel] bf16; a batch is a run of equal-length segments laid
 * end to end (no convolution reads across a segment edge).
 *
 * A convolution is an implicit GEMM: out[t][co] = sum over (k, ci) of
 * in[t+(k-4)*dil][ci] * w[co][k*Cin+ci].  The reduction axis kk = k*Cin+ci is
 * padded to Kp (a multiple of 16) with zero weights.  One wavefront computes
 * a 16-position x 128-channel tile with v_mfma_f32_16x16x16_bf16
 * (A: lane l holds rows l%16, k 4*(l/16)..+3; B: columns l%16, same k;
 * D: lane l holds rows 4*(l/16)..+3 of column l%16).
 *
 * Training: the data gradient of a convolution is the same kernel run on
 * the output gradient with the weights flipped in k and transposed
 * (wd[ci][k'*Cout+co] = w[co][(8-k')*Cin+ci]); the weight gradient is a
 * GEMM over positions (rows co, columns kk) split over workgroups and summed
 * with fp32 atomics; the head, the loss (logistic, per position) and Adam
 * are small fp32 kernels.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint16_t bf16_t;
typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

#define GATE_C   128   /* channels of the hidden layers */
#define GATE_KW  9     /* kernel width */

/* fp32 -> bf16, round to nearest even (no NaN here) */
__device__ __forceinline__ bf16_t f2bf(float x)
{
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t) (u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(((uint32_t) h) << 16); }

/* codes (0..3 = ACGU, anything else: all-zero vector) -> one-hot [N][4] bf16 */
__global__ void gate_onehot_kernel(const uint8_t *codes, int64_t n, bf16_t *x)
{
  const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int c = codes[t];
  uint64_t v = 0;
  if (c < 4) v = (uint64_t) 0x3f80u << (16 * c);   /* 1.0 in bf16 */
  *(uint64_t *) (x + 4 * t) = v;
}

/* Forward convolution.  grid.x = ceil(n / 64), block = 256 (4 waves).
 * x [n][cin] bf16, w [128][kp] bf16, bias [128] fp32 (may be null),
 * y [n][128] bf16 or yf [n][128] fp32.  cin is 4 or 128. */
template <bool RELU, bool OUTF32>
__global__ void __launch_bounds__(256) gate_conv_kernel(const bf16_t *x, int64_t n, int seg, int cin, int dil,
                                                        const bf16_t *w, int kp, const float *bias,
                                                        bf16_t *y, float *yf)
{
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t t0 = (int64_t) blockIdx.x * 64 + wv * 16;
  if (t0 >= n) return;
  const int r = lane & 15, q = lane >> 4;
  const int64_t t = t0 + r;                        /* this lane's A row */
  const int64_t sbeg = (t / seg) * seg;            /* its segment */
  const int64_t send = min(sbeg + seg, n);
  f4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; j++) acc[j] = (f4){0.f, 0.f, 0.f, 0.f};
  const int kmax = GATE_KW * cin;
  for (int kk0 = 0; kk0 < kp; kk0 += 16) {
    const int kk = kk0 + 4 * q;
    s4 a = (s4){0, 0, 0, 0};
    if (kk < kmax && t < n) {
      const int k = kk / cin, ci = kk - k * cin;
      const int64_t src = t + (int64_t) (k - 4) * dil;
      if (src >= sbeg && src < send) a = *(const s4 *) (x + src * cin + ci);
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const s4 b = *(const s4 *) (w + (size_t) (16 * j + r) * kp + kk);
      acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int co = 16 * j + r;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t to = t0 + 4 * q + i;
      if (to >= n) continue;
      float v = acc[j][i] + bb;
      if (RELU) v = fmaxf(v, 0.f);
      if (OUTF32) yf[to * GATE_C + co] = v;
      else        y[to * GATE_C + co] = f2bf(v);
    }
  }
}

/* Weight gradient: dw[co][kk] += sum over t of g[t][co] * x[t+k-4][ci].
 * grid = (128/16, kp/64, nsplit), block 64; each workgroup sums positions
 * [z*tchunk, (z+1)*tchunk). */
__global__ void __launch_bounds__(64) gate_wgrad_kernel(const bf16_t *g, const bf16_t *x, int64_t n, int seg, int cin,
                                                        int dil, int kp, int64_t tchunk, float *dw)
{
  const int lane = threadIdx.x, r = lane & 15, q = lane >> 4;
  const int co0 = blockIdx.x * 16, kb = blockIdx.y * 64;
  const int64_t tz0 = (int64_t) blockIdx.z * tchunk;
  const int64_t tz1 = min(n, tz0 + tchunk);
  const int kmax = GATE_KW * cin;
  /* the B columns of this lane: kk = kb + 16*c + r */
  int kcol[4], cicol[4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    const int kk = kb + 16 * c + r;
    kcol[c] = kk < kmax ? kk / cin : -100;
    cicol[c] = kk < kmax ? kk - (kk / cin) * cin : 0;
  }
  f4 acc[4];
#pragma unroll
  for (int c = 0; c < 4; c++) acc[c] = (f4){0.f, 0.f, 0.f, 0.f};
  for (int64_t tb = tz0; tb < tz1; tb += 16) {
    s4 a;
    bf16_t av[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int64_t t = tb + 4 * q + i;
      av[i] = (t < tz1) ? g[t * GATE_C + co0 + r] : (bf16_t) 0;
    }
    a = (s4){(short) av[0], (short) av[1], (short) av[2], (short) av[3]};
#pragma unroll
    for (int c = 0; c < 4; c++) {
      bf16_t bv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int64_t t = tb + 4 * q + i;
        bv[i] = 0;
        if (t < tz1 && kcol[c] >= 0) {
          const int64_t sbeg = (t / seg) * seg, send = min(sbeg + seg, n);
          const int64_t src = t + (int64_t) (kcol[c] - 4) * dil;
          if (src >= sbeg && src < send) bv[i] = x[src * cin + cicol[c]];
        }
      }
      const s4 b = (s4){(short) bv[0], (short) bv[1], (short) bv[2], (short) bv[3]};
      acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[c], 0, 0, 0);
    }
  }
  /* D: lane holds rows (co) 4q..4q+3 of column (kk) r */
#pragma unroll
  for (int c = 0; c < 4; c++)
#pragma unroll
    for (int i = 0; i < 4; i++)
      atomicAdd(dw + (size_t) (co0 + 4 * q + i) * kp + kb + 16 * c + r, acc[c][i]);
}

/* grad *= (y > 0), both [n][128] bf16 */
__global__ void gate_relu_bwd_kernel(bf16_t *g, const bf16_t *y, int64_t m)
{
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m && !(bf2f(y[i]) > 0.f)) g[i] = 0;
}

/* head: s[t] = b + sum_c y[t][c] * hw[c]; one thread per position */
__global__ void gate_head_kernel(const bf16_t *y, int64_t n, const float *hw, const float *hb, float *s)
{
  const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const bf16_t *row = y + t * GATE_C;
  float acc = 0.f;
  for (int c = 0; c < GATE_C; c += 4) {
    const uint64_t u = *(const uint64_t *) (row + c);
#pragma unroll
    for (int i = 0; i < 4; i++) acc += bf2f((bf16_t) (u >> (16 * i))) * hw[c + i];
  }
  s[t] = acc + hb[0];
}

/* loss and head gradients.  Logistic loss per position with weight pw on
 * positives: l = pw*y*softplus(-s) + (1-y)*softplus(s); ds = (sigma(s) - y)
 * scaled as l, divided by n.  Writes ds[t], dy[t][c] = ds[t]*hw[c] (bf16),
 * and accumulates loss, dhw[c] and dhb. */
__global__ void gate_head_bwd_kernel(const float *s, const float *lab, const bf16_t *y, int64_t n, const float *hw,
                                     float pw, bf16_t *dy, float *dhw, float *dhb, float *loss)
{
  const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ float red[GATE_C + 2];
  if (threadIdx.x < GATE_C + 2) red[threadIdx.x] = 0.f;
  __syncthreads();
  if (t < n) {
    const float z = s[t], yl = lab[t];
    const float sig = 1.f / (1.f + __expf(-z));
    const float wgt = yl > 0.5f ? pw : 1.f;
    const float sp_pos = fmaxf(z, 0.f) + log1pf(__expf(-fabsf(z)));   /* softplus(z) */
    const float l = yl > 0.5f ? wgt * (sp_pos - z) : sp_pos;
    const float ds = wgt * (sig - yl) / (float) n;
    atomicAdd(&red[GATE_C], l / (float) n);
    atomicAdd(&red[GATE_C + 1], ds);
    const bf16_t *row = y + t * GATE_C;
    for (int c = 0; c < GATE_C; c++) {
      atomicAdd(&red[c], ds * bf2f(row[c]));
      dy[t * GATE_C + c] = f2bf(ds * hw[c]);
    }
  }
  __syncthreads();
  if (threadIdx.x < GATE_C) atomicAdd(&dhw[threadIdx.x], red[threadIdx.x]);
  if (threadIdx.x == GATE_C) atomicAdd(dhb, red[GATE_C + 1]);
  if (threadIdx.x == GATE_C + 1) atomicAdd(loss, red[GATE_C]);
}

/* column sums of a [n][128] bf16 array into out[128] (fp32 atomics; clear
 * out first): the bias gradient */
__global__ void gate_colsum_kernel(const bf16_t *g, int64_t n, float *out)
{
  const int c = threadIdx.x & (GATE_C - 1);
  const int64_t r0 = (int64_t) blockIdx.x * 256 + (threadIdx.x >> 7);
  float acc = 0.f;
  for (int64_t r = r0; r < min(n, (int64_t) (blockIdx.x + 1) * 256); r += blockDim.x >> 7) acc += bf2f(g[r * GATE_C + c]);
  atomicAdd(out + c, acc);
}

/* Adam on fp32 master weights */
__global__ void gate_adam_kernel(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                                 float eps, float c1, float c2)
{
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi; v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

/* master weights [128][kp] fp32 -> forward copy [128][kp] bf16 and (cin ==
 * 128) the data-gradient copy wd[ci][k'*128+co] = w[co][(8-k')*128+ci] */
__global__ void gate_pack_kernel(const float *w, int kp, int cin, bf16_t *wf, bf16_t *wd)
{
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t) GATE_C * kp) return;
  const int co = (int) (i / kp), kk = (int) (i % kp);
  const bf16_t h = f2bf(w[i]);
  wf[i] = h;
  if (wd && kk < GATE_KW * cin) {
    const int k = kk / cin, ci = kk % cin;
    wd[(size_t) ci * kp + (8 - k) * GATE_C + co] = h;
  }
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned) ((n + b - 1) / b); }

/* training batch: nseg segments of seg residues starting at starts[s] of a
 * device-resident target and its labels */
__global__ void gate_gather_kernel(const uint8_t *codes, const float *lab, const int64_t *starts, int seg, int64_t m,
                                   uint8_t *oc, float *ol)
{
  const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t src = starts[i / seg] + i % seg;
  oc[i] = codes[src];
  ol[i] = lab[src];
}

/* ------------------------------------------------------------------------ */
/* C ABI (stream 0; all pointers are device pointers)                        */

/* Device memory for the gate's callers, from this library's runtime (the
 * kernels below and these buffers share one runtime and its null stream, so
 * every call is ordered after the previous one). */
extern "C" int infx_dev_malloc(size_t bytes, void **ptr) { return (int) hipMalloc(ptr, bytes > 0 ? bytes : 1); }
extern "C" int infx_dev_free(void *ptr) { return (int) hipFree(ptr); }
extern "C" int infx_dev_memset(void *ptr, int value, size_t bytes) { return (int) hipMemset(ptr, value, bytes); }
extern "C" int infx_dev_upload(void *dst, const void *src, size_t bytes) { return (int) hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); }
extern "C" int infx_dev_download(void *dst, const void *src, size_t bytes) { return (int) hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost); }
extern "C" int infx_dev_sync(void) { return (int) hipDeviceSynchronize(); }

extern "C" int infx_gate_gather(const uint8_t *codes, const float *lab, const int64_t *starts, int nseg, int seg,
                                uint8_t *oc, float *ol)
{
  const int64_t m = (int64_t) nseg * seg;
  if (m <= 0) return 0;
  hipLaunchKernelGGL(gate_gather_kernel, dim3(nblk(m, 256)), dim3(256), 0, 0, codes, lab, starts, seg, m, oc, ol);
  return (int) hipGetLastError();
}


extern "C" int infx_gate_onehot(const uint8_t *codes, int64_t n, void *x)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gate_onehot_kernel, dim3(nblk(n, 256)), dim3(256), 0, 0, codes, n, (bf16_t *) x);
  return (int) hipGetLastError();
}

/* mode 0: bf16 out + ReLU; 1: bf16 out, no ReLU; 2: fp32 out, no ReLU */
extern "C" int infx_gate_conv(const void *x, int64_t n, int seg, int cin, int dil, const void *w, int kp, const float *bias,
                              void *y, int mode)
{
  if (n <= 0) return 0;
  if ((cin != 4 && cin != GATE_C) || kp % 16 || kp < GATE_KW * cin || seg <= 0 || dil < 1) return (int) hipErrorInvalidValue;
  const dim3 g(nblk(n, 64)), b(256);
  if (mode == 0) hipLaunchKernelGGL((gate_conv_kernel<true, false>), g, b, 0, 0, (const bf16_t *) x, n, seg, cin, dil, (const bf16_t *) w, kp, bias, (bf16_t *) y, (float *) nullptr);
  else if (mode == 1) hipLaunchKernelGGL((gate_conv_kernel<false, false>), g, b, 0, 0, (const bf16_t *) x, n, seg, cin, dil, (const bf16_t *) w, kp, bias, (bf16_t *) y, (float *) nullptr);
  else hipLaunchKernelGGL((gate_conv_kernel<false, true>), g, b, 0, 0, (const bf16_t *) x, n, seg, cin, dil, (const bf16_t *) w, kp, bias, (bf16_t *) nullptr, (float *) y);
  return (int) hipGetLastError();
}

extern "C" int infx_gate_wgrad(const void *g, const void *x, int64_t n, int seg, int cin, int dil, int kp, float *dw)
{
  if (n <= 0) return 0;
  if ((cin != 4 && cin != GATE_C) || kp % 64 || kp < GATE_KW * cin || seg <= 0 || dil < 1) return (int) hipErrorInvalidValue;
  const int64_t tchunk = 1024;
  const dim3 grid(GATE_C / 16, kp / 64, nblk(n, (int) tchunk));
  hipLaunchKernelGGL(gate_wgrad_kernel, grid, dim3(64), 0, 0, (const bf16_t *) g, (const bf16_t *) x, n, seg, cin, dil, kp, tchunk, dw);
  return (int) hipGetLastError();
}

extern "C" int infx_gate_relu_bwd(void *g, const void *y, int64_t m)
{
  if (m <= 0) return 0;
  hipLaunchKernelGGL(gate_relu_bwd_kernel, dim3(nblk(m, 256)), dim3(256), 0, 0, (bf16_t *) g, (const bf16_t *) y, m);
  return (int) hipGetLastError();
}

extern "C" int infx_gate_head(const void *y, int64_t n, const float *hw, const float *hb, float *s)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gate_head_kernel, dim3(nblk(n, 256)), dim3(256), 0, 0, (const bf16_t *) y, n, hw, hb, s);
  return (int) hipGetLastError();
}

extern "C" int infx_gate_head_bwd(const float *s, const float *lab, const void *y, int64_t n, const float *hw, float pw,
                                  void *dy, float *dhw, float *dhb, float *loss)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gate_head_bwd_kernel, dim3(nblk(n, 256)), dim3(256), 0, 0, s, lab, (const bf16_t *) y, n, hw, pw,
                     (bf16_t *) dy, dhw, dhb, loss);
  return (int) hipGetLastError();
}

extern "C" int infx_gate_colsum(const void *g, int64_t n, float *out)
{
  if (n <= 0) return 0;
  hipLaunchKernelGGL(gate_colsum_kernel, dim3(nblk(n, 256)), dim3(256), 0, 0, (const bf16_t *) g, n, out);
  return (int) hipGetLastError();
}

extern "C" int infx_gate_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                              float eps, int step)
{
  if (n <= 0) return 0;
  const float c1 = 1.f - powf(b1, (float) step), c2 = 1.f - powf(b2, (float) step);
  hipLaunchKernelGGL(gate_adam_kernel, dim3(nblk(n, 256)), dim3(256), 0, 0, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
  return (int) hipGetLastError();
}

extern "C" int infx_gate_pack(const float *w, int kp, int cin, void *wf, void *wd)
{
  hipLaunchKernelGGL(gate_pack_kernel, dim3(nblk((int64_t) GATE_C * kp, 256)), dim3(256), 0, 0, w, kp, cin,
                     (bf16_t *) wf, (bf16_t *) wd);
  return (int) hipGetLastError();
}
