// wg_mlp.hip -- the calibration surrogate: a 16-128-128-128-1 GELU perceptron
// trained and evaluated on the GPU with hand-written BF16 MFMA kernels (FP32
// accumulation).  No counterpart in the reference: it stands in for the γ search
// of calibration.cpp:266-636 by proposing γ from basin features.
//
// Precision: FP32 master parameters; BF16 copies of the weights and of every
// GEMM operand (inputs, hidden activations, back-propagated deltas); FP32
// accumulation, biases, pre-activations, GELU, loss and Adam.
//
// All GEMMs are C[M][N] = sum_k A[m][k] * Bt[n][k] with both operands K-contiguous
// BF16, on v_mfma_f32_16x16x32_bf16 (lane l: A row l%16 and B column l%16 with
// k = (l/16)*8 .. +7; C column l%16, rows (l/16)*4 .. +3).  The backward pass
// uses transposed copies written by the producing epilogues, so every operand
// load is one 16-byte vector load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/watergap2_amd.h"

namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;
using bf16 = __bf16;

constexpr int kHid = 128;   // hidden units
constexpr int kIn = 16;     // features
constexpr int kInPad = 32;  // features padded to one MFMA K step

// parameter offsets in the flat FP32 vector
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + int64_t(kHid) * kIn;
constexpr int64_t kOffW2 = kOffB1 + kHid;
constexpr int64_t kOffB2 = kOffW2 + int64_t(kHid) * kHid;
constexpr int64_t kOffW3 = kOffB2 + kHid;
constexpr int64_t kOffB3 = kOffW3 + int64_t(kHid) * kHid;
constexpr int64_t kOffW4 = kOffB3 + kHid;
constexpr int64_t kOffB4 = kOffW4 + kHid;
constexpr int64_t kNParams = kOffB4 + 1;

__device__ inline float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

struct Epilogue {
  int mode;           // 0: C; 1: hidden forward; 2: delta of the previous layer
  float *out_f32;     // mode 0: C; 1: z (may be null: forward only); 2: dz
  const float *bias;  // mode 1
  const float *zprev; // mode 2: pre-activation whose GELU derivative scales C
  bf16 *out_b;        // modes 1, 2: bf16 [M][N]
  bf16 *out_bt;       // modes 1, 2: bf16 [N][M] (mode 1: may be null)
};

// One wavefront per 32 x 32 tile of C (2 x 2 MFMA blocks); M, N, K multiples of 32.
__global__ __launch_bounds__(64) void mlp_gemm(const bf16 *__restrict__ A, const bf16 *__restrict__ Bt, int M,
                                               int N, int K, Epilogue ep) {
  const int l = threadIdx.x;
  const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
  if (m0 >= M || n0 >= N) return;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int r = l & 15, kq = (l >> 4) * 8;
  for (int k0 = 0; k0 < K; k0 += 32) {
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      a[i] = *reinterpret_cast<const bf16x8 *>(A + int64_t(m0 + 16 * i + r) * K + k0 + kq);
      b[i] = *reinterpret_cast<const bf16x8 *>(Bt + int64_t(n0 + 16 * i + r) * K + k0 + kq);
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int row = m0 + 16 * i + (l >> 4) * 4 + e, col = n0 + 16 * j + r;
        const int64_t o = int64_t(row) * N + col;
        float c = acc[i][j][e];
        if (ep.mode == 0) {
          ep.out_f32[o] = c;
        } else if (ep.mode == 1) {
          const float z = c + ep.bias[col];
          if (ep.out_f32) ep.out_f32[o] = z;
          const bf16 h = bf16(gelu(z));
          ep.out_b[o] = h;
          if (ep.out_bt) ep.out_bt[int64_t(col) * M + row] = h;
        } else {
          const float dz = c * gelu_grad(ep.zprev[o]);
          ep.out_f32[o] = dz;
          const bf16 d = bf16(dz);
          ep.out_b[o] = d;
          ep.out_bt[int64_t(col) * M + row] = d;
        }
      }
}

// Batch rows from the resident sample set: X [R][32] bf16 (features padded with
// zeros), its transpose [32][R], targets.  Rows beyond n_valid are zero.  A null
// idx takes rows 0 .. n_valid - 1 in order; a null xt or tb is not written.
__global__ void mlp_gather(const float *__restrict__ xs, const float *__restrict__ ts, const int32_t *__restrict__ idx,
                           int n_valid, int R, bf16 *__restrict__ xb, bf16 *__restrict__ xt, float *__restrict__ tb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const bool ok = r < n_valid;
  const int64_t s = ok ? (idx ? idx[r] : r) : 0;
  for (int k = 0; k < kInPad; k++) {
    const float v = (ok && k < kIn) ? xs[s * kIn + k] : 0.f;
    const bf16 b = bf16(v);
    xb[int64_t(r) * kInPad + k] = b;
    if (xt) xt[int64_t(k) * R + r] = b;
  }
  if (tb) tb[r] = ok ? ts[s] : 0.f;
}

// Output layer forward: y = h3 . w4 + b4 (FP32, sequential over units), rows
// 0 .. n_out - 1 of the batch.
__global__ void mlp_out_fwd(const bf16 *__restrict__ h3, const bf16 *__restrict__ w4, const float *__restrict__ b4,
                            int n_out, float *__restrict__ y) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_out) return;
  float s = 0.f;
  for (int k = 0; k < kHid; k++) s += float(h3[int64_t(r) * kHid + k]) * float(w4[k]);
  y[r] = s + b4[0];
}

// Loss gradient and the output layer's backward: dy = 2 (y - t) / n_valid,
// dz3 = dy w4 * GELU'(z3) (FP32 and BF16 copies).
__global__ void mlp_out_bwd(const float *__restrict__ y, const float *__restrict__ t, const bf16 *__restrict__ w4,
                            const float *__restrict__ z3, int n_valid, int R, float *__restrict__ dy,
                            float *__restrict__ dz, bf16 *__restrict__ dzb, bf16 *__restrict__ dzt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  const float d = r < n_valid ? 2.f * (y[r] - t[r]) / float(n_valid) : 0.f;
  dy[r] = d;
  for (int k = 0; k < kHid; k++) {
    const int64_t o = int64_t(r) * kHid + k;
    const float v = d * float(w4[k]) * gelu_grad(z3[o]);
    dz[o] = v;
    const bf16 b = bf16(v);
    dzb[o] = b;
    dzt[int64_t(k) * R + r] = b;
  }
}

// Column sums over the batch (biases, output weights), in row order.
__global__ void mlp_colsum(const float *__restrict__ dz, int R, int N, float *__restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int r = 0; r < R; r++) s += dz[int64_t(r) * N + n];
  out[n] = s;
}

__global__ void mlp_dw4(const float *__restrict__ dy, const bf16 *__restrict__ h3, int R, float *__restrict__ g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > kHid) return;
  float s = 0.f;
  if (k < kHid) {
    for (int r = 0; r < R; r++) s += dy[r] * float(h3[int64_t(r) * kHid + k]);
    g[kOffW4 + k] = s;
  } else {
    for (int r = 0; r < R; r++) s += dy[r];
    g[kOffB4] = s;
  }
}

// First-layer weight gradient: the padded [128][32] GEMM result without the padding columns.
__global__ void mlp_unpad_w1(const float *__restrict__ gpad, float *__restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kHid * kIn) return;
  g[kOffW1 + i] = gpad[(i / kIn) * kInPad + (i % kIn)];
}

// Mean squared error of the valid rows (one thread, row order).
__global__ void mlp_loss(const float *__restrict__ y, const float *__restrict__ t, int n_valid, float *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.f;
  for (int r = 0; r < n_valid; r++) s += (y[r] - t[r]) * (y[r] - t[r]);
  out[0] += s / float(n_valid);
}

// Adam step on the FP32 master parameters.
__global__ void mlp_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                         float *__restrict__ m2, int64_t n, float lr, float b1, float b2, float eps, float c1,
                         float c2) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 weight copies for the GEMMs: W1 padded [128][32] and its transpose [32][128],
// W2, W3 and their transposes, w4.
__global__ void mlp_pack(const float *__restrict__ p, bf16 *__restrict__ w1, bf16 *__restrict__ w1t,
                         bf16 *__restrict__ w2, bf16 *__restrict__ w2t, bf16 *__restrict__ w3, bf16 *__restrict__ w3t,
                         bf16 *__restrict__ w4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kHid * kInPad) {
    const int n = i / kInPad, k = i % kInPad;
    const bf16 v = bf16(k < kIn ? p[kOffW1 + n * kIn + k] : 0.f);
    w1[i] = v;
    w1t[k * kHid + n] = v;
  }
  if (i < kHid * kHid) {
    const int n = i / kHid, k = i % kHid;
    const bf16 a = bf16(p[kOffW2 + i]), b = bf16(p[kOffW3 + i]);
    w2[i] = a;
    w2t[k * kHid + n] = a;
    w3[i] = b;
    w3t[k * kHid + n] = b;
  }
  if (i < kHid) w4[i] = bf16(p[kOffW4 + i]);
}

// Activations of wg_mlp_forward_noalloc: X [R][32] bf16 and the three hidden
// layers [R][128] bf16 for up to WG_MLP_NOALLOC_MAX_ROWS rows (6.5 MiB), one copy
// per GPU for the life of the process.  Host code uses the per-device address
// that hipGetSymbolAddress returns, never this name.
constexpr size_t kScratchXBytes = size_t(WG_MLP_NOALLOC_MAX_ROWS) * kInPad * sizeof(bf16);
constexpr size_t kScratchHBytes = size_t(WG_MLP_NOALLOC_MAX_ROWS) * kHid * sizeof(bf16);
constexpr size_t kScratchBytes = kScratchXBytes + 3 * kScratchHBytes;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace

struct wg_mlp {
  std::string err;
  hipStream_t stream = nullptr;
  int device = 0;
  unsigned char *scratch = nullptr;  // this device's address of the scratch array
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr, *gpad = nullptr;
  bf16 *w1 = nullptr, *w1t = nullptr, *w2 = nullptr, *w2t = nullptr, *w3 = nullptr, *w3t = nullptr, *w4 = nullptr;
  int64_t adam_t = 0;
  // batch buffers (rows R)
  int R = 0;
  bf16 *xb = nullptr, *xt = nullptr, *h[3] = {}, *ht[3] = {}, *dzb = nullptr, *dzt = nullptr, *dzb2 = nullptr,
       *dzt2 = nullptr;
  float *z[3] = {}, *dz = nullptr, *dz2 = nullptr, *y = nullptr, *t = nullptr, *dy = nullptr, *loss = nullptr;
  int32_t *idx = nullptr;
  std::vector<void *> batch_bufs;
};

namespace {

int mlp_fail(wg_mlp *h, int code, const std::string &msg) {
  if (h) h->err = msg;
  return code;
}

#define MLP_CHECK(h, x)                                                              \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) return mlp_fail(h, WG_E_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int mlp_alloc(wg_mlp *h, T **p, size_t n, std::vector<void *> *track = nullptr) {
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return mlp_fail(h, WG_E_NOMEM, "hipMalloc failed");
  e = hipMemsetAsync(*p, 0, std::max<size_t>(n, 1) * sizeof(T), h->stream);
  if (e != hipSuccess) return mlp_fail(h, WG_E_HIP, "hipMemsetAsync failed");
  if (track) track->push_back(*p);
  return WG_OK;
}

int ensure_batch(wg_mlp *h, int R) {
  if (R <= h->R) return WG_OK;
  MLP_CHECK(h, hipStreamSynchronize(h->stream));
  for (void *q : h->batch_bufs) (void)hipFree(q);
  h->batch_bufs.clear();
  auto *t = &h->batch_bufs;
  int rc = mlp_alloc(h, &h->xb, size_t(R) * kInPad, t);
  if (!rc) rc = mlp_alloc(h, &h->xt, size_t(R) * kInPad, t);
  for (int l = 0; l < 3 && !rc; l++) {
    rc = mlp_alloc(h, &h->h[l], size_t(R) * kHid, t);
    if (!rc) rc = mlp_alloc(h, &h->ht[l], size_t(R) * kHid, t);
    if (!rc) rc = mlp_alloc(h, &h->z[l], size_t(R) * kHid, t);
  }
  if (!rc) rc = mlp_alloc(h, &h->dzb, size_t(R) * kHid, t);
  if (!rc) rc = mlp_alloc(h, &h->dzt, size_t(R) * kHid, t);
  if (!rc) rc = mlp_alloc(h, &h->dzb2, size_t(R) * kHid, t);
  if (!rc) rc = mlp_alloc(h, &h->dzt2, size_t(R) * kHid, t);
  if (!rc) rc = mlp_alloc(h, &h->dz, size_t(R) * kHid, t);
  if (!rc) rc = mlp_alloc(h, &h->dz2, size_t(R) * kHid, t);
  if (!rc) rc = mlp_alloc(h, &h->y, size_t(R), t);
  if (!rc) rc = mlp_alloc(h, &h->t, size_t(R), t);
  if (!rc) rc = mlp_alloc(h, &h->dy, size_t(R), t);
  if (!rc) rc = mlp_alloc(h, &h->idx, size_t(R), t);
  if (rc) return rc;
  h->R = R;
  return WG_OK;
}

int launch_gemm_on(wg_mlp *h, hipStream_t stream, const bf16 *A, const bf16 *Bt, int M, int N, int K,
                   const Epilogue &ep) {
  if (M % 32 || N % 32 || K % 32) return mlp_fail(h, WG_E_ARG, "GEMM shape is not a multiple of 32");
  hipLaunchKernelGGL(mlp_gemm, dim3(N / 32, M / 32), dim3(64), 0, stream, A, Bt, M, N, K, ep);
  MLP_CHECK(h, hipGetLastError());
  return WG_OK;
}

int launch_gemm(wg_mlp *h, const bf16 *A, const bf16 *Bt, int M, int N, int K, const Epilogue &ep) {
  return launch_gemm_on(h, h->stream, A, Bt, M, N, K, ep);
}

// Per-device address of the scratch array, resolved once per device with that
// device current (each GPU holds its own copy at its own address).
std::mutex g_scratch_mu;
unsigned char *g_scratch_addr[64] = {};

int scratch_address(wg_mlp *h) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (h->device < 0 || h->device >= 64) return mlp_fail(h, WG_E_ARG, "device ordinal out of range");
  if (!g_scratch_addr[h->device]) {
    void *ptr = nullptr;
    MLP_CHECK(h, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_addr[h->device] = static_cast<unsigned char *>(ptr);
  }
  h->scratch = g_scratch_addr[h->device];
  return WG_OK;
}

int pack(wg_mlp *h) {
  hipLaunchKernelGGL(mlp_pack, dim3((kHid * kHid + 255) / 256), dim3(256), 0, h->stream, h->p, h->w1, h->w1t, h->w2,
                     h->w2t, h->w3, h->w3t, h->w4);
  MLP_CHECK(h, hipGetLastError());
  return WG_OK;
}

// Forward pass over R rows already in xb/xt.
int forward(wg_mlp *h, int R) {
  const bf16 *in[3] = {h->xb, nullptr, nullptr};
  const bf16 *w[3] = {h->w1, h->w2, h->w3};
  const float *b[3] = {h->p + kOffB1, h->p + kOffB2, h->p + kOffB3};
  const int K[3] = {kInPad, kHid, kHid};
  for (int l = 0; l < 3; l++) {
    Epilogue ep{1, h->z[l], b[l], nullptr, h->h[l], h->ht[l]};
    int rc = launch_gemm(h, l == 0 ? in[0] : h->h[l - 1], w[l], R, kHid, K[l], ep);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(mlp_out_fwd, dim3((R + 255) / 256), dim3(256), 0, h->stream, h->h[2], h->w4, h->p + kOffB4, R,
                     h->y);
  MLP_CHECK(h, hipGetLastError());
  return WG_OK;
}

// Backward pass for the rows of the last forward (targets in h->t); gradients into h->g.
int backward(wg_mlp *h, int R, int n_valid) {
  hipLaunchKernelGGL(mlp_out_bwd, dim3((R + 255) / 256), dim3(256), 0, h->stream, h->y, h->t, h->w4, h->z[2],
                     n_valid, R, h->dy, h->dz, h->dzb, h->dzt);
  MLP_CHECK(h, hipGetLastError());
  hipLaunchKernelGGL(mlp_dw4, dim3(1), dim3(kHid + 64), 0, h->stream, h->dy, h->h[2], R, h->g);
  MLP_CHECK(h, hipGetLastError());
  // layer 3: dW3 = dz3^T h2, db3, dz2 = (dz3 W3) * GELU'(z2)
  int rc = launch_gemm(h, h->dzt, h->ht[1], kHid, kHid, R, Epilogue{0, h->g + kOffW3, nullptr, nullptr, nullptr, nullptr});
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_colsum, dim3(1), dim3(kHid), 0, h->stream, h->dz, R, kHid, h->g + kOffB3);
  rc = launch_gemm(h, h->dzb, h->w3t, R, kHid, kHid, Epilogue{2, h->dz2, nullptr, h->z[1], h->dzb2, h->dzt2});
  if (rc) return rc;
  // layer 2
  rc = launch_gemm(h, h->dzt2, h->ht[0], kHid, kHid, R, Epilogue{0, h->g + kOffW2, nullptr, nullptr, nullptr, nullptr});
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_colsum, dim3(1), dim3(kHid), 0, h->stream, h->dz2, R, kHid, h->g + kOffB2);
  rc = launch_gemm(h, h->dzb2, h->w2t, R, kHid, kHid, Epilogue{2, h->dz, nullptr, h->z[0], h->dzb, h->dzt});
  if (rc) return rc;
  // layer 1
  rc = launch_gemm(h, h->dzt, h->xt, kHid, kInPad, R, Epilogue{0, h->gpad, nullptr, nullptr, nullptr, nullptr});
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_unpad_w1, dim3((kHid * kIn + 255) / 256), dim3(256), 0, h->stream, h->gpad, h->g);
  hipLaunchKernelGGL(mlp_colsum, dim3(1), dim3(kHid), 0, h->stream, h->dz, R, kHid, h->g + kOffB1);
  MLP_CHECK(h, hipGetLastError());
  return WG_OK;
}

int round32(int r) { return (r + 31) / 32 * 32; }

// Uploads rows [r0, r0 + n) of host features/targets as a batch (identity index).
int upload_rows(wg_mlp *h, const float *x, const float *t, int n, float **dx, float **dt) {
  int rc = mlp_alloc(h, dx, size_t(n) * kIn);
  if (!rc && t) rc = mlp_alloc(h, dt, size_t(n));
  if (rc) return rc;
  MLP_CHECK(h, hipMemcpyAsync(*dx, x, size_t(n) * kIn * sizeof(float), hipMemcpyHostToDevice, h->stream));
  if (t) MLP_CHECK(h, hipMemcpyAsync(*dt, t, size_t(n) * sizeof(float), hipMemcpyHostToDevice, h->stream));
  return WG_OK;
}

}  // namespace

extern "C" int wg_mlp_create(uint64_t seed, wg_mlp **out) {
  if (!out) return WG_E_ARG;
  *out = nullptr;
  wg_mlp *h = new (std::nothrow) wg_mlp();
  if (!h) return WG_E_NOMEM;
  *out = h;
  MLP_CHECK(h, hipGetDevice(&h->device));
  MLP_CHECK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  int rc = scratch_address(h);
  if (!rc) rc = mlp_alloc(h, &h->p, kNParams);
  if (!rc) rc = mlp_alloc(h, &h->g, kNParams);
  if (!rc) rc = mlp_alloc(h, &h->m1, kNParams);
  if (!rc) rc = mlp_alloc(h, &h->m2, kNParams);
  if (!rc) rc = mlp_alloc(h, &h->gpad, size_t(kHid) * kInPad);
  if (!rc) rc = mlp_alloc(h, &h->w1, size_t(kHid) * kInPad);
  if (!rc) rc = mlp_alloc(h, &h->w1t, size_t(kHid) * kInPad);
  if (!rc) rc = mlp_alloc(h, &h->w2, size_t(kHid) * kHid);
  if (!rc) rc = mlp_alloc(h, &h->w2t, size_t(kHid) * kHid);
  if (!rc) rc = mlp_alloc(h, &h->w3, size_t(kHid) * kHid);
  if (!rc) rc = mlp_alloc(h, &h->w3t, size_t(kHid) * kHid);
  if (!rc) rc = mlp_alloc(h, &h->w4, kHid);
  if (!rc) rc = mlp_alloc(h, &h->loss, 1);
  if (rc) return rc;
  // He-uniform initial weights from a 64-bit LCG (host), zero biases
  std::vector<float> p(kNParams, 0.f);
  uint64_t s = seed * 6364136223846793005ULL + 1442695040888963407ULL;
  auto uni = [&]() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return float((s >> 40) * (1.0 / 16777216.0)) * 2.f - 1.f;
  };
  auto fill = [&](int64_t off, int n_out, int n_in) {
    const float a = std::sqrt(6.f / float(n_in));
    for (int64_t i = 0; i < int64_t(n_out) * n_in; i++) p[off + i] = a * uni();
  };
  fill(kOffW1, kHid, kIn);
  fill(kOffW2, kHid, kHid);
  fill(kOffW3, kHid, kHid);
  fill(kOffW4, 1, kHid);
  MLP_CHECK(h, hipMemcpyAsync(h->p, p.data(), kNParams * sizeof(float), hipMemcpyHostToDevice, h->stream));
  rc = pack(h);
  if (rc) return rc;
  MLP_CHECK(h, hipStreamSynchronize(h->stream));
  return WG_OK;
}

extern "C" int wg_mlp_destroy(wg_mlp *h) {
  if (!h) return WG_E_ARG;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void *q : h->batch_bufs) (void)hipFree(q);
  void *bufs[] = {h->p, h->g, h->m1, h->m2, h->gpad, h->w1, h->w1t, h->w2, h->w2t, h->w3, h->w3t, h->w4, h->loss};
  for (void *q : bufs)
    if (q) (void)hipFree(q);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return WG_OK;
}

extern "C" const char *wg_mlp_last_error(const wg_mlp *h) { return h ? h->err.c_str() : "null handle"; }

extern "C" int64_t wg_mlp_n_params(void) { return kNParams; }

extern "C" int wg_mlp_get_params(wg_mlp *h, float *out) {
  if (!h || !out) return WG_E_ARG;
  MLP_CHECK(h, hipMemcpyAsync(out, h->p, kNParams * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  MLP_CHECK(h, hipStreamSynchronize(h->stream));
  return WG_OK;
}

extern "C" int wg_mlp_set_params(wg_mlp *h, const float *in) {
  if (!h || !in) return WG_E_ARG;
  MLP_CHECK(h, hipMemcpyAsync(h->p, in, kNParams * sizeof(float), hipMemcpyHostToDevice, h->stream));
  int rc = pack(h);
  if (rc) return rc;
  MLP_CHECK(h, hipStreamSynchronize(h->stream));
  return WG_OK;
}

extern "C" int wg_mlp_forward(wg_mlp *h, int32_t rows, const float *x, float *y) {
  if (!h || rows <= 0 || !x || !y) return WG_E_ARG;
  const int chunk = 8192;
  float *dx = nullptr;
  int rc = upload_rows(h, x, nullptr, rows, &dx, nullptr);
  if (rc) return rc;
  rc = ensure_batch(h, std::min(round32(rows), chunk));
  std::vector<int32_t> iota;
  for (int r0 = 0; !rc && r0 < rows; r0 += chunk) {
    const int n = std::min(chunk, rows - r0), R = round32(n);
    iota.resize(n);
    for (int i = 0; i < n; i++) iota[i] = r0 + i;
    hipError_t e = hipMemcpyAsync(h->idx, iota.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) { rc = mlp_fail(h, WG_E_HIP, "hipMemcpyAsync"); break; }
    hipLaunchKernelGGL(mlp_gather, dim3((R + 255) / 256), dim3(256), 0, h->stream, dx, nullptr, h->idx, n, R, h->xb,
                       h->xt, nullptr);
    rc = forward(h, R);
    if (rc) break;
    e = hipMemcpyAsync(y + r0, h->y, n * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = mlp_fail(h, WG_E_HIP, "forward copy");
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(dx);
  return rc;
}

extern "C" int wg_mlp_forward_noalloc(wg_mlp *h, int32_t rows, const float *x, float *y, void *stream) {
  if (!h || !x || !y) return WG_E_ARG;
  if (rows <= 0 || rows > WG_MLP_NOALLOC_MAX_ROWS)
    return mlp_fail(h, WG_E_ARG, "wg_mlp_forward_noalloc: rows must be 1.." + std::to_string(WG_MLP_NOALLOC_MAX_ROWS) +
                                     " (the size of its scratch array), got " + std::to_string(rows));
  if (!h->scratch) return mlp_fail(h, WG_E_STATE, "wg_mlp_forward_noalloc: scratch address not resolved");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : h->stream;
  const int R = round32(rows);
  bf16 *xb = reinterpret_cast<bf16 *>(h->scratch);
  bf16 *hb[3];
  for (int l = 0; l < 3; l++) hb[l] = reinterpret_cast<bf16 *>(h->scratch + kScratchXBytes + l * kScratchHBytes);
  // every row of the batch, padding included, is written before it is read:
  // nothing in the scratch array needs clearing between calls
  hipLaunchKernelGGL(mlp_gather, dim3((R + 255) / 256), dim3(256), 0, s, x, nullptr, nullptr, rows, R, xb, nullptr,
                     nullptr);
  MLP_CHECK(h, hipGetLastError());
  const bf16 *w[3] = {h->w1, h->w2, h->w3};
  const float *b[3] = {h->p + kOffB1, h->p + kOffB2, h->p + kOffB3};
  const int K[3] = {kInPad, kHid, kHid};
  for (int l = 0; l < 3; l++) {
    Epilogue ep{1, nullptr, b[l], nullptr, hb[l], nullptr};
    int rc = launch_gemm_on(h, s, l == 0 ? xb : hb[l - 1], w[l], R, kHid, K[l], ep);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(mlp_out_fwd, dim3((rows + 255) / 256), dim3(256), 0, s, hb[2], h->w4, h->p + kOffB4, rows, y);
  MLP_CHECK(h, hipGetLastError());
  return WG_OK;
}

extern "C" int wg_mlp_gradient(wg_mlp *h, int32_t rows, const float *x, const float *t, float *loss, float *grad) {
  if (!h || rows <= 0 || !x || !t || !grad) return WG_E_ARG;
  const int R = round32(rows);
  float *dx = nullptr, *dt = nullptr;
  int rc = upload_rows(h, x, t, rows, &dx, &dt);
  if (!rc) rc = ensure_batch(h, R);
  if (!rc) {
    std::vector<int32_t> iota(rows);
    for (int i = 0; i < rows; i++) iota[i] = i;
    hipError_t e = hipMemcpyAsync(h->idx, iota.data(), rows * sizeof(int32_t), hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) rc = mlp_fail(h, WG_E_HIP, "hipMemcpyAsync");
  }
  if (!rc) {
    hipLaunchKernelGGL(mlp_gather, dim3((R + 255) / 256), dim3(256), 0, h->stream, dx, dt, h->idx, rows, R, h->xb,
                       h->xt, h->t);
    rc = forward(h, R);
  }
  if (!rc) rc = backward(h, R, rows);
  if (!rc) {
    (void)hipMemsetAsync(h->loss, 0, sizeof(float), h->stream);
    hipLaunchKernelGGL(mlp_loss, dim3(1), dim3(64), 0, h->stream, h->y, h->t, rows, h->loss);
    hipError_t e = hipMemcpyAsync(grad, h->g, kNParams * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) rc = mlp_fail(h, WG_E_HIP, "gradient copy");
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(dx);
  (void)hipFree(dt);
  return rc;
}

extern "C" int wg_mlp_train(wg_mlp *h, int32_t n_samples, const float *x, const float *t, int32_t epochs,
                            int32_t batch, float lr, uint64_t seed, float *epoch_loss) {
  if (!h || n_samples <= 0 || !x || !t || epochs <= 0 || batch < 32 || batch % 32) return WG_E_ARG;
  float *dx = nullptr, *dt = nullptr;
  int rc = upload_rows(h, x, t, n_samples, &dx, &dt);
  if (!rc) rc = ensure_batch(h, batch);
  std::vector<int32_t> perm(n_samples);
  for (int i = 0; i < n_samples; i++) perm[i] = i;
  uint64_t s = seed ^ 0x9E3779B97F4A7C15ULL;
  int32_t *dperm = nullptr;
  if (!rc) rc = mlp_alloc(h, &dperm, size_t(n_samples));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; !rc && ep < epochs; ep++) {
    // Fisher-Yates shuffle of the sample order (host)
    for (int i = n_samples - 1; i > 0; i--) {
      s = s * 6364136223846793005ULL + 1442695040888963407ULL;
      const int j = int((s >> 33) % uint64_t(i + 1));
      std::swap(perm[i], perm[j]);
    }
    hipError_t e = hipMemcpyAsync(dperm, perm.data(), n_samples * sizeof(int32_t), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->loss, 0, sizeof(float), h->stream);
    if (e != hipSuccess) { rc = mlp_fail(h, WG_E_HIP, "epoch setup"); break; }
    int n_batches = 0;
    for (int r0 = 0; r0 < n_samples; r0 += batch) {
      const int n = std::min(batch, n_samples - r0);
      const int R = round32(n);
      hipLaunchKernelGGL(mlp_gather, dim3((R + 255) / 256), dim3(256), 0, h->stream, dx, dt, dperm + r0, n, R,
                         h->xb, h->xt, h->t);
      rc = forward(h, R);
      if (!rc) rc = backward(h, R, n);
      if (rc) break;
      hipLaunchKernelGGL(mlp_loss, dim3(1), dim3(64), 0, h->stream, h->y, h->t, n, h->loss);
      h->adam_t++;
      const float c1 = 1.f - std::pow(b1, float(h->adam_t)), c2 = 1.f - std::pow(b2, float(h->adam_t));
      hipLaunchKernelGGL(mlp_adam, dim3(unsigned((kNParams + 255) / 256)), dim3(256), 0, h->stream, h->p, h->g,
                         h->m1, h->m2, kNParams, lr, b1, b2, eps, c1, c2);
      rc = pack(h);
      if (rc) break;
      n_batches++;
    }
    if (rc) break;
    float l = 0.f;
    e = hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) { rc = mlp_fail(h, WG_E_HIP, "loss copy"); break; }
    if (epoch_loss) epoch_loss[ep] = l / float(std::max(n_batches, 1));
  }
  (void)hipStreamSynchronize(h->stream);
  (void)hipFree(dx);
  (void)hipFree(dt);
  (void)hipFree(dperm);
  return rc;
}
