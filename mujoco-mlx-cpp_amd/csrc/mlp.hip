// Learned warm start: a small MLP (in -> H -> H -> out, SiLU) trained and run
// on the GPU with BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate).
//
// The reference has no learned warm start (its solver starts from qacc = 0,
// reference src/compat/mkx_kernels/solver.hpp:592); these kernels are new.
//
// GEMM tile: 64 x 64 output per 256-thread workgroup, K step 32, A and B
// staged in LDS as [row][k] so each lane's 8 consecutive K values are one
// 16-byte LDS read. Wave w owns output rows 16w..16w+15 and four 16 x 16 MFMA
// tiles across N.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mjmlx/mjmlx_hip.h"
#include "mlp.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace {

#define HIPOK(expr)                                                                                   \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " (" #expr ")"); \
  } while (0)

__device__ __forceinline__ float bf2f(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}
// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float silu(float z) { return z / (1.f + expf(-z)); }

enum { BM = 64, BN = 64, BK = 32, KPAD = BK + 8 };

// C = op(A) op(B) (+ bias, activation). op(A) is M x K, op(B) is K x N.
// TA: A stored K x M (lda = row stride); TB: B stored N x K.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, const float* __restrict__ bias,
                                              int act, float* __restrict__ Z, uint16_t* __restrict__ Y,
                                              float* __restrict__ Yf, int ldc) {
  __shared__ __align__(16) uint16_t As[BM][KPAD];
  __shared__ __align__(16) uint16_t Bs[BN][KPAD];
  int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  v4f acc[4];
  for (int t = 0; t < 4; t++) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += BK) {
    for (int e = tid; e < BM * BK; e += 256) {
      int mm, kk;
      if (TA) { kk = e / BM; mm = e % BM; } else { mm = e / BK; kk = e % BK; }
      int gm = m0 + mm, gk = k0 + kk;
      uint16_t v = 0;
      if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[mm][kk] = v;
    }
    for (int e = tid; e < BN * BK; e += 256) {
      int nn, kk;
      if (TB) { nn = e / BK; kk = e % BK; } else { kk = e / BN; nn = e % BN; }
      int gn = n0 + nn, gk = k0 + kk;
      uint16_t v = 0;
      if (gn < N && gk < K) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      Bs[nn][kk] = v;
    }
    __syncthreads();
    v8bf a = *reinterpret_cast<const v8bf*>(&As[16 * wave + (lane & 15)][8 * (lane >> 4)]);
    for (int t = 0; t < 4; t++) {
      v8bf b = *reinterpret_cast<const v8bf*>(&Bs[16 * t + (lane & 15)][8 * (lane >> 4)]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
    __syncthreads();
  }
  for (int t = 0; t < 4; t++) {
    int c = n0 + 16 * t + (lane & 15);
    if (c >= N) continue;
    float bb = bias ? bias[c] : 0.f;
    for (int i = 0; i < 4; i++) {
      int r = m0 + 16 * wave + 4 * (lane >> 4) + i;
      if (r >= M) continue;
      float z = acc[t][i] + bb;
      size_t o = (size_t)r * ldc + c;
      if (Z) Z[o] = z;
      float y = act ? silu(z) : z;
      if (Y) Y[o] = f2bf(y);
      if (Yf) Yf[o] = y;
    }
  }
}

// dZ = dY * silu'(Z); writes bf16 (for the next GEMMs) and fp32 (for the bias gradient)
__global__ void k_silu_bwd(int n, const float* __restrict__ dY, const float* __restrict__ Z, uint16_t* __restrict__ dZb,
                           float* __restrict__ dZf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float z = Z[i], s = 1.f / (1.f + expf(-z));
  float g = dY[i] * s * (1.f + z * (1.f - s));
  dZf[i] = g;
  dZb[i] = f2bf(g);
}

// column sums of an R x C row-major fp32 matrix (bias gradients), one thread per column
__global__ void k_colsum(int R, int C, const float* __restrict__ X, float* __restrict__ out) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  float s = 0.f;
  for (int r = 0; r < R; r++) s += X[(size_t)r * C + c];
  out[c] = s;
}

// mean squared error over an R x C batch: gradient (fp32 + bf16) and per-row loss
__global__ void k_mse(int R, int C, const float* __restrict__ pred, const float* __restrict__ tgt, float scale,
                      float* __restrict__ dP, uint16_t* __restrict__ dPb, float* __restrict__ rowloss) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  float l = 0.f;
  for (int c = 0; c < C; c++) {
    size_t o = (size_t)r * C + c;
    float d = pred[o] - tgt[o];
    l += d * d;
    float g = 2.f * scale * d;
    dP[o] = g;
    dPb[o] = f2bf(g);
  }
  rowloss[r] = l;
}

// AdamW on fp32 master weights; refreshes the bf16 copy used by the GEMMs
__global__ void k_adamw(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                        float* __restrict__ v, uint16_t* __restrict__ pb, float lr, float b1, float b2, float eps,
                        float wd, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float pi = p[i];
  pi -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * pi);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

}  // namespace

// ------------------------------------------------------------------ host object
struct MjmlxMLP {
  int nin = 0, nh = 0, nout = 0;
  hipStream_t stream = nullptr;
  // parameters: W1 [nin x nh], b1, W2 [nh x nh], b2, W3 [nh x nout], b3 (row-major, x W convention)
  size_t off[6] = {}, len[6] = {}, nparam = 0;
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t* pb = nullptr;
  // normalisation (host-visible copies kept on device for the feature kernel)
  float *in_mean = nullptr, *in_istd = nullptr, *out_mean = nullptr, *out_std = nullptr;
  std::vector<float> h_in_mean, h_in_istd, h_out_mean, h_out_std;
  // activation buffers sized for cap rows
  int cap = 0;
  uint16_t *xb = nullptr, *h1b = nullptr, *h2b = nullptr, *d3b = nullptr, *d2b = nullptr, *d1b = nullptr;
  float *z1 = nullptr, *z2 = nullptr, *out = nullptr, *tgt = nullptr, *dout = nullptr, *dh2 = nullptr, *dh1 = nullptr,
        *d2f = nullptr, *d1f = nullptr, *rowloss = nullptr;
  long step = 0;

  const uint16_t* Wb(int i) const { return pb + off[2 * i]; }
  const float* bias(int i) const { return p + off[2 * i + 1]; }
  float* gW(int i) { return g + off[2 * i]; }
  float* gb(int i) { return g + off[2 * i + 1]; }

  void ensure(int rows) {
    if (rows <= cap) return;
    void* old[] = {xb, h1b, h2b, d3b, d2b, d1b, z1, z2, out, tgt, dout, dh2, dh1, d2f, d1f, rowloss};
    for (void* q : old) if (q) (void)hipFree(q);
    cap = rows;
    size_t R = rows;
    HIPOK(hipMalloc(&xb, R * nin * 2));
    HIPOK(hipMalloc(&h1b, R * nh * 2));
    HIPOK(hipMalloc(&h2b, R * nh * 2));
    HIPOK(hipMalloc(&d3b, R * nout * 2));
    HIPOK(hipMalloc(&d2b, R * nh * 2));
    HIPOK(hipMalloc(&d1b, R * nh * 2));
    HIPOK(hipMalloc(&z1, R * nh * 4));
    HIPOK(hipMalloc(&z2, R * nh * 4));
    HIPOK(hipMalloc(&out, R * nout * 4));
    HIPOK(hipMalloc(&tgt, R * nout * 4));
    HIPOK(hipMalloc(&dout, R * nout * 4));
    HIPOK(hipMalloc(&dh2, R * nh * 4));
    HIPOK(hipMalloc(&dh1, R * nh * 4));
    HIPOK(hipMalloc(&d2f, R * nh * 4));
    HIPOK(hipMalloc(&d1f, R * nh * 4));
    HIPOK(hipMalloc(&rowloss, R * 4));
  }
};

namespace mjmlp {

void gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
          const float* bias, int act, float* Z, uint16_t* Y, float* Yf, int ldc) {
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, act, Z, Y, Yf, ldc);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, act, Z, Y, Yf, ldc);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, act, Z, Y, Yf, ldc);
  else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, bias, act, Z, Y, Yf, ldc);
  HIPOK(hipGetLastError());
}

// forward of `rows` bf16 inputs already in mlp->xb (or x); output fp32 (normalised units) into out
void forward(MjmlxMLP* n, const uint16_t* x, int rows, float* out, hipStream_t s, bool keep_z) {
  gemm(s, false, false, rows, n->nh, n->nin, x, n->nin, n->Wb(0), n->nh, n->bias(0), 1, keep_z ? n->z1 : nullptr, n->h1b, nullptr, n->nh);
  gemm(s, false, false, rows, n->nh, n->nh, n->h1b, n->nh, n->Wb(1), n->nh, n->bias(1), 1, keep_z ? n->z2 : nullptr, n->h2b, nullptr, n->nh);
  gemm(s, false, false, rows, n->nout, n->nh, n->h2b, n->nh, n->Wb(2), n->nout, n->bias(2), 0, nullptr, nullptr, out, n->nout);
}

// gradient of mean((out - tgt)^2) w.r.t. all parameters into n->g; returns nothing (loss in rowloss)
void backward(MjmlxMLP* n, const uint16_t* x, int rows, hipStream_t s) {
  int B = rows, H = n->nh, O = n->nout, I = n->nin;
  float scale = 1.f / (float)(B * O);
  hipLaunchKernelGGL(k_mse, dim3((B + 255) / 256), dim3(256), 0, s, B, O, n->out, n->tgt, scale, n->dout, n->d3b, n->rowloss);
  // layer 3: dW3 = h2^T dout, db3 = colsum(dout), dh2 = dout W3^T
  gemm(s, true, false, H, O, B, n->h2b, H, n->d3b, O, nullptr, 0, nullptr, nullptr, n->gW(2), O);
  hipLaunchKernelGGL(k_colsum, dim3((O + 63) / 64), dim3(64), 0, s, B, O, n->dout, n->gb(2));
  gemm(s, false, true, B, H, O, n->d3b, O, n->Wb(2), O, nullptr, 0, nullptr, nullptr, n->dh2, H);
  hipLaunchKernelGGL(k_silu_bwd, dim3((B * H + 255) / 256), dim3(256), 0, s, B * H, n->dh2, n->z2, n->d2b, n->d2f);
  // layer 2
  gemm(s, true, false, H, H, B, n->h1b, H, n->d2b, H, nullptr, 0, nullptr, nullptr, n->gW(1), H);
  hipLaunchKernelGGL(k_colsum, dim3((H + 63) / 64), dim3(64), 0, s, B, H, n->d2f, n->gb(1));
  gemm(s, false, true, B, H, H, n->d2b, H, n->Wb(1), H, nullptr, 0, nullptr, nullptr, n->dh1, H);
  hipLaunchKernelGGL(k_silu_bwd, dim3((B * H + 255) / 256), dim3(256), 0, s, B * H, n->dh1, n->z1, n->d1b, n->d1f);
  // layer 1
  gemm(s, true, false, I, H, B, x, I, n->d1b, H, nullptr, 0, nullptr, nullptr, n->gW(0), H);
  hipLaunchKernelGGL(k_colsum, dim3((H + 63) / 64), dim3(64), 0, s, B, H, n->d1f, n->gb(0));
  HIPOK(hipGetLastError());
}

}  // namespace mjmlp

// ------------------------------------------------------------------ C API
extern "C" {

MJMLX_API MjmlxMLP* mjmlx_mlp_create(int nin, int nhidden, int nout, int seed) {
  if (nin <= 0 || nhidden <= 0 || nout <= 0) return nullptr;
  MjmlxMLP* n = nullptr;
  try {
    n = new MjmlxMLP();
    n->nin = nin; n->nh = nhidden; n->nout = nout;
    size_t shapes[6] = {(size_t)nin * nhidden, (size_t)nhidden, (size_t)nhidden * nhidden, (size_t)nhidden,
                        (size_t)nhidden * nout, (size_t)nout};
    size_t o = 0;
    for (int i = 0; i < 6; i++) { n->off[i] = o; n->len[i] = shapes[i]; o += (shapes[i] + 7) & ~size_t(7); }
    n->nparam = o;
    // He-uniform init for SiLU layers, small last layer; zero biases
    std::vector<float> h(o, 0.f);
    std::mt19937 rng(seed);
    int fan[3] = {nin, nhidden, nhidden};
    for (int l = 0; l < 3; l++) {
      float lim = std::sqrt(6.f / (float)fan[l]) * (l == 2 ? 0.1f : 1.f);
      std::uniform_real_distribution<float> U(-lim, lim);
      for (size_t i = 0; i < shapes[2 * l]; i++) h[n->off[2 * l] + i] = U(rng);
    }
    HIPOK(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
    HIPOK(hipMalloc(&n->p, o * 4));
    HIPOK(hipMalloc(&n->g, o * 4));
    HIPOK(hipMalloc(&n->m, o * 4));
    HIPOK(hipMalloc(&n->v, o * 4));
    HIPOK(hipMalloc(&n->pb, o * 2));
    HIPOK(hipMemcpy(n->p, h.data(), o * 4, hipMemcpyHostToDevice));
    HIPOK(hipMemset(n->g, 0, o * 4));
    HIPOK(hipMemset(n->m, 0, o * 4));
    HIPOK(hipMemset(n->v, 0, o * 4));
    hipLaunchKernelGGL(k_to_bf16, dim3((o + 255) / 256), dim3(256), 0, n->stream, o, n->p, n->pb);
    HIPOK(hipGetLastError());
    n->h_in_mean.assign(nin, 0.f); n->h_in_istd.assign(nin, 1.f);
    n->h_out_mean.assign(nout, 0.f); n->h_out_std.assign(nout, 1.f);
    HIPOK(hipMalloc(&n->in_mean, nin * 4));
    HIPOK(hipMalloc(&n->in_istd, nin * 4));
    HIPOK(hipMalloc(&n->out_mean, nout * 4));
    HIPOK(hipMalloc(&n->out_std, nout * 4));
    mjmlx_mlp_set_normalization(n, nullptr, nullptr, nullptr, nullptr);
    HIPOK(hipStreamSynchronize(n->stream));
    return n;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_create error: %s\n", e.what());
    mjmlx_mlp_free(n);
    return nullptr;
  }
}

MJMLX_API void mjmlx_mlp_free(MjmlxMLP* n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  void* ptrs[] = {n->p, n->g, n->m, n->v, n->pb, n->in_mean, n->in_istd, n->out_mean, n->out_std,
                  n->xb, n->h1b, n->h2b, n->d3b, n->d2b, n->d1b, n->z1, n->z2, n->out, n->tgt, n->dout, n->dh2,
                  n->dh1, n->d2f, n->d1f, n->rowloss};
  for (void* q : ptrs) if (q) (void)hipFree(q);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

MJMLX_API int mjmlx_mlp_dims(const MjmlxMLP* n, int* nin, int* nh, int* nout, int* nparam) {
  if (!n) return -1;
  if (nin) *nin = n->nin;
  if (nh) *nh = n->nh;
  if (nout) *nout = n->nout;
  if (nparam) *nparam = (int)n->nparam;
  return 0;
}

MJMLX_API int mjmlx_mlp_set_normalization(MjmlxMLP* n, const float* in_mean, const float* in_std, const float* out_mean,
                                          const float* out_std) {
  if (!n) return -1;
  try {
    for (int i = 0; i < n->nin; i++) {
      if (in_mean) n->h_in_mean[i] = in_mean[i];
      if (in_std) n->h_in_istd[i] = in_std[i] > 1e-12f ? 1.f / in_std[i] : 1.f;
    }
    for (int i = 0; i < n->nout; i++) {
      if (out_mean) n->h_out_mean[i] = out_mean[i];
      if (out_std) n->h_out_std[i] = out_std[i] > 1e-12f ? out_std[i] : 1.f;
    }
    HIPOK(hipMemcpy(n->in_mean, n->h_in_mean.data(), n->nin * 4, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(n->in_istd, n->h_in_istd.data(), n->nin * 4, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(n->out_mean, n->h_out_mean.data(), n->nout * 4, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(n->out_std, n->h_out_std.data(), n->nout * 4, hipMemcpyHostToDevice));
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_set_normalization error: %s\n", e.what());
    return -1;
  }
}

// fp32 master parameters in the order W1 b1 W2 b2 W3 b3 (packed, no padding)
MJMLX_API int mjmlx_mlp_get_params(const MjmlxMLP* n, float* out) {
  if (!n || !out) return -1;
  size_t o = 0;
  for (int i = 0; i < 6; i++) {
    if (hipMemcpy(out + o, n->p + n->off[i], n->len[i] * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
    o += n->len[i];
  }
  return (int)o;
}

MJMLX_API int mjmlx_mlp_set_params(MjmlxMLP* n, const float* in) {
  if (!n || !in) return -1;
  size_t o = 0;
  for (int i = 0; i < 6; i++) {
    if (hipMemcpy(n->p + n->off[i], in + o, n->len[i] * 4, hipMemcpyHostToDevice) != hipSuccess) return -1;
    o += n->len[i];
  }
  hipLaunchKernelGGL(k_to_bf16, dim3((n->nparam + 255) / 256), dim3(256), 0, n->stream, n->nparam, n->p, n->pb);
  if (hipStreamSynchronize(n->stream) != hipSuccess) return -1;
  return (int)o;
}

// y = MLP(x) for `rows` rows of already-normalised inputs (host buffers; output in normalised units)
MJMLX_API int mjmlx_mlp_forward(MjmlxMLP* n, const float* x, int rows, float* y) {
  if (!n || !x || !y || rows <= 0) return -1;
  try {
    n->ensure(rows);
    float* dx;
    HIPOK(hipMalloc(&dx, (size_t)rows * n->nin * 4));
    HIPOK(hipMemcpyAsync(dx, x, (size_t)rows * n->nin * 4, hipMemcpyHostToDevice, n->stream));
    size_t tot = (size_t)rows * n->nin;
    hipLaunchKernelGGL(k_to_bf16, dim3((tot + 255) / 256), dim3(256), 0, n->stream, tot, dx, n->xb);
    mjmlp::forward(n, n->xb, rows, n->out, n->stream, false);
    HIPOK(hipMemcpyAsync(y, n->out, (size_t)rows * n->nout * 4, hipMemcpyDeviceToHost, n->stream));
    HIPOK(hipStreamSynchronize(n->stream));
    (void)hipFree(dx);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_forward error: %s\n", e.what());
    return -1;
  }
}

// gradient of mean((MLP(x) - y)^2) w.r.t. the parameters (packed like get_params); returns the loss
MJMLX_API float mjmlx_mlp_grad(MjmlxMLP* n, const float* x, const float* y, int rows, float* grad_out) {
  if (!n || !x || !y || rows <= 0) return -1.f;
  try {
    n->ensure(rows);
    float* dx;
    HIPOK(hipMalloc(&dx, (size_t)rows * n->nin * 4));
    HIPOK(hipMemcpyAsync(dx, x, (size_t)rows * n->nin * 4, hipMemcpyHostToDevice, n->stream));
    HIPOK(hipMemcpyAsync(n->tgt, y, (size_t)rows * n->nout * 4, hipMemcpyHostToDevice, n->stream));
    size_t tot = (size_t)rows * n->nin;
    hipLaunchKernelGGL(k_to_bf16, dim3((tot + 255) / 256), dim3(256), 0, n->stream, tot, dx, n->xb);
    mjmlp::forward(n, n->xb, rows, n->out, n->stream, true);
    mjmlp::backward(n, n->xb, rows, n->stream);
    std::vector<float> rl(rows);
    HIPOK(hipMemcpyAsync(rl.data(), n->rowloss, rows * 4, hipMemcpyDeviceToHost, n->stream));
    if (grad_out) {
      size_t o = 0;
      for (int i = 0; i < 6; i++) {
        HIPOK(hipMemcpyAsync(grad_out + o, n->g + n->off[i], n->len[i] * 4, hipMemcpyDeviceToHost, n->stream));
        o += n->len[i];
      }
    }
    HIPOK(hipStreamSynchronize(n->stream));
    (void)hipFree(dx);
    double s = 0;
    for (float v : rl) s += v;
    return (float)(s / ((double)rows * n->nout));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_grad error: %s\n", e.what());
    return -1.f;
  }
}

// AdamW training on (x, y) pairs held on the host (normalised units); shuffles
// with `seed` each epoch; returns the mean loss of the last epoch
MJMLX_API float mjmlx_mlp_train(MjmlxMLP* n, const float* x, const float* y, long rows, int batch, int epochs, float lr,
                                float weight_decay, int seed) {
  if (!n || !x || !y || rows <= 0 || batch <= 0) return -1.f;
  try {
    n->ensure(batch);
    size_t I = n->nin, O = n->nout;
    uint16_t* dX;
    float *dY, *stage;
    HIPOK(hipMalloc(&dX, (size_t)rows * I * 2));
    HIPOK(hipMalloc(&dY, (size_t)rows * O * 4));
    HIPOK(hipMalloc(&stage, (size_t)std::min<long>(rows, 1 << 18) * I * 4));
    for (long r0 = 0; r0 < rows; r0 += (1 << 18)) {
      long nr = std::min<long>(rows - r0, 1 << 18);
      HIPOK(hipMemcpyAsync(stage, x + r0 * I, nr * I * 4, hipMemcpyHostToDevice, n->stream));
      size_t tot = nr * I;
      hipLaunchKernelGGL(k_to_bf16, dim3((tot + 255) / 256), dim3(256), 0, n->stream, tot, stage, dX + r0 * I);
    }
    HIPOK(hipMemcpyAsync(dY, y, rows * O * 4, hipMemcpyHostToDevice, n->stream));
    std::vector<long> perm(rows);
    for (long i = 0; i < rows; i++) perm[i] = i;
    std::mt19937_64 rng(seed);
    // gathered batches: a permutation of whole batches of contiguous rows keeps the gather cheap
    long nb = rows / batch;
    std::vector<long> order(nb);
    for (long i = 0; i < nb; i++) order[i] = i;
    double last = 0;
    std::vector<float> rl(batch);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ep++) {
      std::shuffle(order.begin(), order.end(), rng);
      double tot = 0;
      for (long bi = 0; bi < nb; bi++) {
        long r0 = order[bi] * batch;
        const uint16_t* xb = dX + r0 * I;
        HIPOK(hipMemcpyAsync(n->tgt, dY + r0 * O, batch * O * 4, hipMemcpyDeviceToDevice, n->stream));
        mjmlp::forward(n, xb, batch, n->out, n->stream, true);
        mjmlp::backward(n, xb, batch, n->stream);
        n->step++;
        float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
        hipLaunchKernelGGL(k_adamw, dim3((n->nparam + 255) / 256), dim3(256), 0, n->stream, (int)n->nparam, n->p, n->g,
                           n->m, n->v, n->pb, lr, b1, b2, eps, weight_decay, c1, c2);
        if (bi == nb - 1 || (bi % 64) == 0) {
          HIPOK(hipMemcpyAsync(rl.data(), n->rowloss, batch * 4, hipMemcpyDeviceToHost, n->stream));
          HIPOK(hipStreamSynchronize(n->stream));
          double s = 0;
          for (float v : rl) s += v;
          tot = s / ((double)batch * O);
        }
      }
      last = tot;
    }
    HIPOK(hipStreamSynchronize(n->stream));
    (void)hipFree(dX);
    (void)hipFree(dY);
    (void)hipFree(stage);
    return (float)last;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_train error: %s\n", e.what());
    return -1.f;
  }
}

}  // extern "C"
