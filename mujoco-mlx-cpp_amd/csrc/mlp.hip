// Learned warm start: a small MLP (nin -> nh -> nh -> nout, SiLU) trained and
// run on the GPU with BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate).
//
// The reference has no learned warm start (its solver starts from qacc = 0,
// reference src/compat/mkx_kernels/solver.hpp:592); these kernels are new.
//
// GEMM: 64 x 64 output tile per 256-thread workgroup, K step 32. Both operands
// are staged in LDS as [row][k] (row = M index for A, N index for B) so each
// lane's 8 consecutive K values are one 16-byte LDS read; global loads are
// 16 bytes along whichever dimension is contiguous in memory. Wave w owns the
// 32 x 32 quadrant (w / 2, w % 2) as 2 x 2 MFMA tiles.
//
// MFMA operand layout (16 x 16 x 32): lane l supplies A[l % 16][8 (l / 16) ..
// + 7] and B[8 (l / 16) .. + 7][l % 16]; it receives D[4 (l / 16) + i][l % 16],
// i = 0..3. Any K permutation shared by A and B gives the same product.
#include <hip/hip_runtime.h>

#include <algorithm>
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
typedef unsigned short v8u __attribute__((ext_vector_type(8)));

namespace {

#define HIPOK(expr)                                                                                   \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " (" #expr ")"); \
  } while (0)

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float silu(float z) { return z / (1.f + expf(-z)); }

enum { BM = 64, BN = 64, BK = 32, KPAD = BK + 8 };

struct Epi {
  const float* bias;            // [N] or null
  int act;                      // 1: SiLU
  float* Z;                     // pre-activation fp32 [M][ldc] or null
  uint16_t* Y;                  // activation bf16 [M][ldc] or null
  float* Yf;                    // activation fp32 [M][ldc] or null
  int ldc, nstore;              // row stride of the outputs; columns c < nstore are written
  const float *oscale, *oshift; // Yf = y * oscale[c] + oshift[c] when oscale is set
  const float* base;            // ... + base[r * ldbase + c] when set
  int ldbase;
};

// Stages rows [r0, r0 + 64) x k [k0, k0 + 32) of an operand into S[row][k].
// KCONTIG: stored [row][k] (leading dim ld); else stored [k][row]. Rows at or
// beyond R are zero; in the [k][row] form R is a multiple of 8.
template <bool KCONTIG>
__device__ __forceinline__ void stage(uint16_t (*S)[KPAD], const uint16_t* __restrict__ X, int ld, int r0, int k0,
                                      int R, int tid) {
  v8u v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (KCONTIG) {
    int row = tid >> 2, kc = (tid & 3) * 8;
    if (r0 + row < R) v = *reinterpret_cast<const v8u*>(X + (size_t)(r0 + row) * ld + k0 + kc);
    *reinterpret_cast<v8u*>(&S[row][kc]) = v;
  } else {
    int k = tid >> 3, rc = (tid & 7) * 8;
    if (r0 + rc < R) v = *reinterpret_cast<const v8u*>(X + (size_t)(k0 + k) * ld + r0 + rc);
    for (int j = 0; j < 8; j++) S[rc + j][k] = v[j];
  }
}

// C = op(A) op(B) with the epilogue. op(A) is M x K (TA: A stored K x M),
// op(B) is K x N (TB: B stored N x K). K is a multiple of 32.
// Split K (gridDim.z > 1): block z covers k in [z kchunk, (z + 1) kchunk) and
// stores its raw tile in part[z][M][N]; k_ksum adds the slices.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, Epi ep, int kchunk,
                                              float* __restrict__ part) {
  __shared__ __align__(16) uint16_t As[BM][KPAD];
  __shared__ __align__(16) uint16_t Bs[BN][KPAD];
  int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int wm = wave >> 1, wn = wave & 1;
  int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  int kb = blockIdx.z * kchunk, ke = kb + kchunk < K ? kb + kchunk : K;
  v4f acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += BK) {
    stage<!TA>(As, A, lda, m0, k0, M, tid);
    stage<TB>(Bs, B, ldb, n0, k0, N, tid);
    __syncthreads();
    v8bf a[2], b[2];
    for (int i = 0; i < 2; i++) a[i] = *reinterpret_cast<const v8bf*>(&As[32 * wm + 16 * i + (lane & 15)][8 * (lane >> 4)]);
    for (int j = 0; j < 2; j++) b[j] = *reinterpret_cast<const v8bf*>(&Bs[32 * wn + 16 * j + (lane & 15)][8 * (lane >> 4)]);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  if (part) {
    float* P = part + (size_t)blockIdx.z * M * N;
    for (int j = 0; j < 2; j++) {
      int c = n0 + 32 * wn + 16 * j + (lane & 15);
      if (c >= N) continue;
      for (int i = 0; i < 2; i++)
        for (int t = 0; t < 4; t++) {
          int r = m0 + 32 * wm + 16 * i + 4 * (lane >> 4) + t;
          if (r < M) P[(size_t)r * N + c] = acc[i][j][t];
        }
    }
    return;
  }
  for (int j = 0; j < 2; j++) {
    int c = n0 + 32 * wn + 16 * j + (lane & 15);
    if (c >= N || c >= ep.nstore) continue;
    float bb = ep.bias ? ep.bias[c] : 0.f;
    for (int i = 0; i < 2; i++)
      for (int t = 0; t < 4; t++) {
        int r = m0 + 32 * wm + 16 * i + 4 * (lane >> 4) + t;
        if (r >= M) continue;
        float z = acc[i][j][t] + bb;
        size_t o = (size_t)r * ep.ldc + c;
        if (ep.Z) ep.Z[o] = z;
        float y = ep.act ? silu(z) : z;
        if (ep.Y) ep.Y[o] = f2bf(y);
        if (ep.Yf) {
          float yo = ep.oscale ? y * ep.oscale[c] + ep.oshift[c] : y;
          if (ep.base) yo += ep.base[(size_t)r * ep.ldbase + c];
          ep.Yf[o] = yo;
        }
      }
  }
}

// dZ = dY * silu'(Z), as bf16 (next GEMMs) and fp32 (bias gradient)
__global__ void k_silu_bwd(int n, const float* __restrict__ dY, const float* __restrict__ Z, uint16_t* __restrict__ dZb,
                           float* __restrict__ dZf) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float z = Z[i], s = 1.f / (1.f + expf(-z));
  float g = dY[i] * s * (1.f + z * (1.f - s));
  dZf[i] = g;
  dZb[i] = f2bf(g);
}

// Y[r][c] = sum over the nz split-K slices of part[z][r][c], c < nstore
__global__ void k_ksum(int M, int N, int nz, const float* __restrict__ part, float* __restrict__ Y, int ldc, int nstore) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  int r = i / N, c = i - r * N;
  if (c >= nstore) return;
  float s = 0.f;
  for (int z = 0; z < nz; z++) s += part[(size_t)z * M * N + i];
  Y[(size_t)r * ldc + c] = s;
}

// column sums of an R x C row-major matrix: 64 columns x 4 row groups per block
__global__ void __launch_bounds__(256) k_colsum(int R, int C, const float* __restrict__ X, float* __restrict__ out) {
  __shared__ float part[4][64];
  int cx = threadIdx.x & 63, ry = threadIdx.x >> 6;
  int c = blockIdx.x * 64 + cx;
  float s = 0.f;
  if (c < C)
    for (int r = ry; r < R; r += 4) s += X[(size_t)r * C + c];
  part[ry][cx] = s;
  __syncthreads();
  if (ry == 0 && c < C) out[c] = (part[0][cx] + part[1][cx]) + (part[2][cx] + part[3][cx]);
}

// partial column sums over 128-row slices: part[slice][c]
__global__ void __launch_bounds__(256) k_colsum_part(int R, int C, const float* __restrict__ X, float* __restrict__ part) {
  __shared__ float ps[4][64];
  int cx = threadIdx.x & 63, ry = threadIdx.x >> 6;
  int c = blockIdx.x * 64 + cx;
  int r0 = blockIdx.y * 128, r1 = r0 + 128 < R ? r0 + 128 : R;
  float s = 0.f;
  if (c < C)
    for (int r = r0 + ry; r < r1; r += 4) s += X[(size_t)r * C + c];
  ps[ry][cx] = s;
  __syncthreads();
  if (ry == 0 && c < C) part[(size_t)blockIdx.y * C + c] = (ps[0][cx] + ps[1][cx]) + (ps[2][cx] + ps[3][cx]);
}

// squared error of the first nvalid columns of the first rvalid rows; gradient
// of scale * sum((p - t)^2) (fp32 + bf16, zero elsewhere) and per-row loss
__global__ void k_mse(int R, int C, int rvalid, int nvalid, const float* __restrict__ pred, const float* __restrict__ tgt,
                      float scale, float* __restrict__ dP, uint16_t* __restrict__ dPb, float* __restrict__ rowloss) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  float l = 0.f;
  for (int c = 0; c < C; c++) {
    size_t o = (size_t)r * C + c;
    float g = 0.f;
    if (r < rvalid && c < nvalid) {
      float d = pred[o] - tgt[o];
      l += d * d;
      g = 2.f * scale * d;
    }
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
  pb[i] = f2bf(pi);
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// normalised bf16 network input [rows][ip] from raw features [rows][nf] with
// the ctrl columns replaced
__global__ void k_pack(int rows, int nf, int ip, const float* __restrict__ raw, const float* __restrict__ ctrl,
                       int ctrl_off, int nu, const float* __restrict__ mean, const float* __restrict__ istd,
                       uint16_t* __restrict__ xb) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)rows * ip) return;
  int r = (int)(i / ip), c = (int)(i % ip);
  float x = 0.f;
  if (c < nf) {
    if (c >= ctrl_off && c < ctrl_off + nu) x = ctrl ? ctrl[(size_t)r * nu + c - ctrl_off] : 0.f;
    else x = raw[(size_t)r * nf + c];
    x = (x - mean[c]) * istd[c];
  }
  xb[i] = f2bf(x);
}

// training batch gathered from the data set: normalised bf16 inputs and fp32 targets
__global__ void k_gather(int rows, int nf, int ip, int nout, int op, const int* __restrict__ idx,
                         const float* __restrict__ X, const float* __restrict__ Y, const float* __restrict__ mean,
                         const float* __restrict__ istd, const float* __restrict__ omean, const float* __restrict__ ostd,
                         uint16_t* __restrict__ xb, float* __restrict__ tgt) {
  int r = blockIdx.x;
  if (r >= rows) return;
  size_t src = (size_t)idx[r];
  for (int c = threadIdx.x; c < ip; c += blockDim.x)
    xb[(size_t)r * ip + c] = c < nf ? f2bf((X[src * nf + c] - mean[c]) * istd[c]) : (uint16_t)0;
  for (int c = threadIdx.x; c < op; c += blockDim.x)
    tgt[(size_t)r * op + c] = c < nout ? (Y[src * nout + c] - omean[c]) / ostd[c] : 0.f;
}

// appends raw rows (ctrl columns inserted) and targets to the data set
__global__ void k_record(int rows, int nf, const float* __restrict__ raw, const float* __restrict__ ctrl, int ctrl_off,
                         int nu, const float* __restrict__ target, int nout, const float* __restrict__ base,
                         int base_ld, float* __restrict__ dsx, float* __restrict__ dsy) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t nx = (size_t)rows * nf;
  if (i < nx) {
    int r = (int)(i / nf), c = (int)(i % nf);
    float x = raw[i];
    if (c >= ctrl_off && c < ctrl_off + nu) x = ctrl ? ctrl[(size_t)r * nu + c - ctrl_off] : 0.f;
    dsx[i] = x;
  }
  if (i < (size_t)rows * nout) {
    int r = (int)(i / nout), c = (int)(i % nout);
    dsy[i] = target[i] - (base ? base[(size_t)r * base_ld + c] : 0.f);
  }
}

// per-chunk column statistics about a shift (the chunk's first row):
// part[(chunk * C + c) * 3 + {0,1,2}] = {shift, sum d, sum d^2}
__global__ void k_colstats(long R, int C, const float* __restrict__ X, long chunk, float* __restrict__ part) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  long r0 = (long)blockIdx.y * chunk, r1 = r0 + chunk < R ? r0 + chunk : R;
  float shift = r0 < R ? X[r0 * C + c] : 0.f, s1 = 0.f, s2 = 0.f;
  for (long r = r0; r < r1; r++) {
    float d = X[r * C + c] - shift;
    s1 += d;
    s2 += d * d;
  }
  float* o = part + ((size_t)blockIdx.y * C + c) * 3;
  o[0] = shift;
  o[1] = s1;
  o[2] = s2;
}

int rup(int x, int a) { return (x + a - 1) / a * a; }

void gemm(hipStream_t s, bool ta, bool tb, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
          const Epi& ep, int nz = 1, float* part = nullptr) {
  if (K % BK) throw std::runtime_error("gemm: K must be a multiple of 32");
  if ((ta && M % 8) || (!tb && N % 8)) throw std::runtime_error("gemm: row-contiguous operands need a multiple of 8");
  int kchunk = nz > 1 ? rup((K + nz - 1) / nz, BK) : K;
  if (nz > 1) nz = (K + kchunk - 1) / kchunk;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, nz);
  float* pp = nz > 1 ? part : nullptr;
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, ep, kchunk, pp);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, ep, kchunk, pp);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, ep, kchunk, pp);
  else hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, ep, kchunk, pp);
  if (nz > 1)
    hipLaunchKernelGGL(k_ksum, dim3((M * N + 255) / 256), dim3(256), 0, s, M, N, nz, (const float*)part, ep.Yf, ep.ldc,
                       ep.nstore);
  HIPOK(hipGetLastError());
}

Epi epi(const float* bias, int act, float* Z, uint16_t* Y, float* Yf, int ldc, int nstore) {
  Epi e{};
  e.bias = bias; e.act = act; e.Z = Z; e.Y = Y; e.Yf = Yf; e.ldc = ldc; e.nstore = nstore;
  return e;
}

// scratch of at least `floats` in n->part
float* part_scratch(MjmlxMLP* n, size_t floats) {
  if (floats > n->part_cap) {
    if (n->part) HIPOK(hipFree(n->part));
    HIPOK(hipMalloc(&n->part, floats * 4));
    n->part_cap = floats;
  }
  return n->part;
}

// weight gradient Yf = A^T B over the batch (K = rows): few output tiles and a
// long K, so K is split across enough blocks to cover the chip, the slices
// summed by k_ksum
void gemm_wgrad(MjmlxMLP* n, hipStream_t s, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
                int ldb, float* Yf, int ldc) {
  int tiles = ((N + BN - 1) / BN) * ((M + BM - 1) / BM);
  int nz = std::max(1, std::min(K / (4 * BK), 512 / std::max(tiles, 1)));
  float* part = nz > 1 ? part_scratch(n, (size_t)nz * M * N) : nullptr;
  gemm(s, true, false, M, N, K, A, lda, B, ldb, epi(nullptr, 0, nullptr, nullptr, Yf, ldc, ldc), nz, part);
}

// column sums of the first R rows of X [R][C] into out: 128-row slices in
// parallel, then one pass over the slices
void colsum(MjmlxMLP* n, hipStream_t s, int R, int C, const float* X, float* out) {
  int nch = (R + 127) / 128;
  if (nch <= 1) {
    hipLaunchKernelGGL(k_colsum, dim3((C + 63) / 64), dim3(256), 0, s, R, C, X, out);
    return;
  }
  float* part = part_scratch(n, (size_t)nch * C);
  hipLaunchKernelGGL(k_colsum_part, dim3((C + 63) / 64, nch), dim3(256), 0, s, R, C, X, part);
  hipLaunchKernelGGL(k_colsum, dim3((C + 63) / 64), dim3(256), 0, s, nch, C, (const float*)part, out);
}


// hidden layers for `rows` packed inputs x [rows][ip]
void hidden(MjmlxMLP* n, const uint16_t* x, int rows, hipStream_t s, bool keep_z) {
  gemm(s, false, false, rows, n->nh, n->ip, x, n->ip, n->Wb(0), n->nh,
       epi(n->bias(0), 1, keep_z ? n->z1 : nullptr, n->h1b, nullptr, n->nh, n->nh));
  gemm(s, false, false, rows, n->nh, n->nh, n->h1b, n->nh, n->Wb(1), n->nh,
       epi(n->bias(1), 1, keep_z ? n->z2 : nullptr, n->h2b, nullptr, n->nh, n->nh));
}

// forward in normalised units into n->out [rows][op]
void forward(MjmlxMLP* n, const uint16_t* x, int rows, hipStream_t s, bool keep_z) {
  hidden(n, x, rows, s, keep_z);
  gemm(s, false, false, rows, n->op, n->nh, n->h2b, n->nh, n->Wb(2), n->op,
       epi(n->bias(2), 0, nullptr, nullptr, n->out, n->op, n->op));
}

// gradient of mean over the first rvalid rows and nout columns of (out - tgt)^2 into n->g
void backward(MjmlxMLP* n, const uint16_t* x, int rows, int rvalid, hipStream_t s) {
  int B = rows, H = n->nh, O = n->op, I = n->ip;
  float scale = 1.f / (float)((size_t)rvalid * n->nout);
  hipLaunchKernelGGL(k_mse, dim3((B + 255) / 256), dim3(256), 0, s, B, O, rvalid, n->nout, n->out, n->tgt, scale,
                     n->dout, n->d3b, n->rowloss);
  gemm_wgrad(n, s, H, O, B, n->h2b, H, n->d3b, O, n->gW(2), O);
  colsum(n, s, B, O, n->dout, n->gb(2));
  gemm(s, false, true, B, H, O, n->d3b, O, n->Wb(2), O, epi(nullptr, 0, nullptr, nullptr, n->dh2, H, H));
  hipLaunchKernelGGL(k_silu_bwd, dim3((B * H + 255) / 256), dim3(256), 0, s, B * H, n->dh2, n->z2, n->d2b, n->d2f);
  gemm_wgrad(n, s, H, H, B, n->h1b, H, n->d2b, H, n->gW(1), H);
  colsum(n, s, B, H, n->d2f, n->gb(1));
  gemm(s, false, true, B, H, H, n->d2b, H, n->Wb(1), H, epi(nullptr, 0, nullptr, nullptr, n->dh1, H, H));
  hipLaunchKernelGGL(k_silu_bwd, dim3((B * H + 255) / 256), dim3(256), 0, s, B * H, n->dh1, n->z1, n->d1b, n->d1f);
  gemm_wgrad(n, s, I, H, B, x, I, n->d1b, H, n->gW(0), H);
  colsum(n, s, B, H, n->d1f, n->gb(0));
  HIPOK(hipGetLastError());
}

void adamw(MjmlxMLP* n, float lr, float wd, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n->step++;
  float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
  hipLaunchKernelGGL(k_adamw, dim3((unsigned)((n->nparam + 255) / 256)), dim3(256), 0, s, (int)n->nparam, n->p, n->g,
                     n->m, n->v, n->pb, lr, b1, b2, eps, wd, c1, c2);
  HIPOK(hipGetLastError());
}

// packed (unpadded) parameter blocks: W1 [nin x nh], b1, W2, b2, W3 [nh x nout], b3 [nout]
struct Block { size_t off; int rows, cols, ld; };
std::vector<Block> packed_blocks(const MjmlxMLP* n) {
  return {{n->off[0], n->nin, n->nh, n->nh}, {n->off[1], 1, n->nh, n->nh}, {n->off[2], n->nh, n->nh, n->nh},
          {n->off[3], 1, n->nh, n->nh},      {n->off[4], n->nh, n->nout, n->op}, {n->off[5], 1, n->nout, n->op}};
}

int copy_packed(const MjmlxMLP* n, float* dev, float* host, bool to_host) {
  size_t o = 0;
  for (auto& b : packed_blocks(n)) {
    size_t w = (size_t)b.cols * 4, pitch = (size_t)b.ld * 4;
    if (to_host) HIPOK(hipMemcpy2D(host + o, w, dev + b.off, pitch, w, b.rows, hipMemcpyDeviceToHost));
    else HIPOK(hipMemcpy2D(dev + b.off, pitch, host + o, w, w, b.rows, hipMemcpyHostToDevice));
    o += (size_t)b.rows * b.cols;
  }
  return (int)o;
}

// copies host rows [rows][nin] into the padded bf16 input xb (rows padded to rp with zeros)
void upload_inputs(MjmlxMLP* n, const float* x, int rows, int rp) {
  float* dx = nullptr;
  HIPOK(hipMalloc(&dx, (size_t)rp * n->ip * 4));
  std::vector<float> h((size_t)rp * n->ip, 0.f);
  for (int r = 0; r < rows; r++) std::memcpy(&h[(size_t)r * n->ip], x + (size_t)r * n->nin, n->nin * 4);
  HIPOK(hipMemcpy(dx, h.data(), h.size() * 4, hipMemcpyHostToDevice));
  size_t tot = h.size();
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, n->stream, tot, dx, n->xb);
  HIPOK(hipStreamSynchronize(n->stream));
  (void)hipFree(dx);
}

void upload_targets(MjmlxMLP* n, const float* y, int rows, int rp) {
  std::vector<float> h((size_t)rp * n->op, 0.f);
  for (int r = 0; r < rows; r++) std::memcpy(&h[(size_t)r * n->op], y + (size_t)r * n->nout, n->nout * 4);
  HIPOK(hipMemcpy(n->tgt, h.data(), h.size() * 4, hipMemcpyHostToDevice));
}

float mean_rowloss(MjmlxMLP* n, int rvalid) {
  std::vector<float> rl(rvalid);
  HIPOK(hipMemcpyAsync(rl.data(), n->rowloss, (size_t)rvalid * 4, hipMemcpyDeviceToHost, n->stream));
  HIPOK(hipStreamSynchronize(n->stream));
  double s = 0;
  for (float v : rl) s += v;
  return (float)(s / ((double)rvalid * n->nout));
}

void upload_norm(MjmlxMLP* n) {
  HIPOK(hipMemcpy(n->in_mean, n->h_in_mean.data(), n->ip * 4, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(n->in_istd, n->h_in_istd.data(), n->ip * 4, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(n->out_mean, n->h_out_mean.data(), n->op * 4, hipMemcpyHostToDevice));
  HIPOK(hipMemcpy(n->out_std, n->h_out_std.data(), n->op * 4, hipMemcpyHostToDevice));
}

}  // namespace

void MjmlxMLP::ensure(int rows) {
  if (rows <= cap) return;
  void* old[] = {xb, h1b, h2b, d3b, d2b, d1b, z1, z2, out, tgt, dout, dh2, dh1, d2f, d1f, rowloss};
  for (void* q : old) if (q) (void)hipFree(q);
  cap = rows;
  size_t R = rows;
  HIPOK(hipMalloc(&xb, R * ip * 2));
  HIPOK(hipMalloc(&h1b, R * nh * 2));
  HIPOK(hipMalloc(&h2b, R * nh * 2));
  HIPOK(hipMalloc(&d3b, R * op * 2));
  HIPOK(hipMalloc(&d2b, R * nh * 2));
  HIPOK(hipMalloc(&d1b, R * nh * 2));
  HIPOK(hipMalloc(&z1, R * nh * 4));
  HIPOK(hipMalloc(&z2, R * nh * 4));
  HIPOK(hipMalloc(&out, R * op * 4));
  HIPOK(hipMalloc(&tgt, R * op * 4));
  HIPOK(hipMalloc(&dout, R * op * 4));
  HIPOK(hipMalloc(&dh2, R * nh * 4));
  HIPOK(hipMalloc(&dh1, R * nh * 4));
  HIPOK(hipMalloc(&d2f, R * nh * 4));
  HIPOK(hipMalloc(&d1f, R * nh * 4));
  HIPOK(hipMalloc(&rowloss, R * 4));
}

namespace mjmlp {

void pack_features(MjmlxMLP* n, const float* raw, const float* ctrl, int ctrl_off, int nu, int rows, hipStream_t s) {
  n->ensure(rows);
  size_t tot = (size_t)rows * n->ip;
  hipLaunchKernelGGL(k_pack, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, rows, n->nin, n->ip, raw, ctrl,
                     ctrl_off, nu, n->in_mean, n->in_istd, n->xb);
  HIPOK(hipGetLastError());
}

void predict(MjmlxMLP* n, int rows, float* pred, hipStream_t s, const float* base, int base_ld) {
  hidden(n, n->xb, rows, s, false);
  Epi e = epi(n->bias(2), 0, nullptr, nullptr, pred, n->nout, n->nout);
  e.oscale = n->out_std;
  e.oshift = n->out_mean;
  e.base = base;
  e.ldbase = base_ld;
  gemm(s, false, false, rows, n->op, n->nh, n->h2b, n->nh, n->Wb(2), n->op, e);
}

long record(MjmlxMLP* n, const float* raw, const float* ctrl, int ctrl_off, int nu, const float* target, int rows,
            hipStream_t s, const float* base, int base_ld) {
  long k = std::min<long>(rows, n->ds_cap - n->ds_n);
  if (k <= 0) return 0;
  size_t tot = (size_t)k * std::max(n->nin, n->nout);
  hipLaunchKernelGGL(k_record, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, (int)k, n->nin, raw, ctrl,
                     ctrl_off, nu, target, n->nout, base, base_ld, n->ds_x + (size_t)n->ds_n * n->nin,
                     n->ds_y + (size_t)n->ds_n * n->nout);
  HIPOK(hipGetLastError());
  n->ds_n += k;
  return k;
}

}  // namespace mjmlp

// ------------------------------------------------------------------ C API
extern "C" {

MJMLX_API MjmlxMLP* mjmlx_mlp_create(int nin, int nhidden, int nout, int seed) {
  if (nin <= 0 || nhidden <= 0 || nout <= 0 || nhidden % 32) return nullptr;
  MjmlxMLP* n = nullptr;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("no HIP device available");
    n = new MjmlxMLP();
    n->nin = nin; n->nh = nhidden; n->nout = nout;
    n->ip = rup(nin, 32); n->op = rup(nout, 32);
    size_t shapes[6] = {(size_t)n->ip * nhidden, (size_t)nhidden, (size_t)nhidden * nhidden, (size_t)nhidden,
                        (size_t)nhidden * n->op, (size_t)n->op};
    size_t o = 0;
    for (int i = 0; i < 6; i++) { n->off[i] = o; n->len[i] = shapes[i]; o += (shapes[i] + 7) & ~size_t(7); }
    n->nparam = o;
    // He-uniform for the SiLU layers, a 10x smaller last layer, zero biases and padding
    std::vector<float> h(o, 0.f);
    std::mt19937 rng(seed);
    int fan[3] = {nin, nhidden, nhidden};
    int rows[3] = {nin, nhidden, nhidden}, cols[3] = {nhidden, nhidden, nout}, ld[3] = {nhidden, nhidden, n->op};
    for (int l = 0; l < 3; l++) {
      float lim = std::sqrt(6.f / (float)fan[l]) * (l == 2 ? 0.1f : 1.f);
      std::uniform_real_distribution<float> U(-lim, lim);
      for (int r = 0; r < rows[l]; r++)
        for (int c = 0; c < cols[l]; c++) h[n->off[2 * l] + (size_t)r * ld[l] + c] = U(rng);
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
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((o + 255) / 256)), dim3(256), 0, n->stream, o, n->p, n->pb);
    HIPOK(hipGetLastError());
    n->h_in_mean.assign(n->ip, 0.f); n->h_in_istd.assign(n->ip, 1.f);
    n->h_out_mean.assign(n->op, 0.f); n->h_out_std.assign(n->op, 1.f);
    HIPOK(hipMalloc(&n->in_mean, n->ip * 4));
    HIPOK(hipMalloc(&n->in_istd, n->ip * 4));
    HIPOK(hipMalloc(&n->out_mean, n->op * 4));
    HIPOK(hipMalloc(&n->out_std, n->op * 4));
    upload_norm(n);
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
                  n->dh1, n->d2f, n->d1f, n->rowloss, n->d_idx, n->ds_x, n->ds_y, n->part};
  for (void* q : ptrs) if (q) (void)hipFree(q);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

MJMLX_API int mjmlx_mlp_dims(const MjmlxMLP* n, int* nin, int* nh, int* nout, int* nparam) {
  if (!n) return -1;
  if (nin) *nin = n->nin;
  if (nh) *nh = n->nh;
  if (nout) *nout = n->nout;
  if (nparam) *nparam = n->nin * n->nh + n->nh + n->nh * n->nh + n->nh + n->nh * n->nout + n->nout;
  return 0;
}

MJMLX_API int mjmlx_mlp_set_normalization(MjmlxMLP* n, const float* in_mean, const float* in_std, const float* out_mean,
                                          const float* out_std) {
  if (!n) return -1;
  try {
    for (int i = 0; i < n->nin; i++) {
      n->h_in_mean[i] = in_mean ? in_mean[i] : 0.f;
      n->h_in_istd[i] = in_std ? (in_std[i] > 1e-6f ? 1.f / in_std[i] : 1.f) : 1.f;
    }
    for (int i = 0; i < n->nout; i++) {
      n->h_out_mean[i] = out_mean ? out_mean[i] : 0.f;
      n->h_out_std[i] = out_std ? (out_std[i] > 1e-6f ? out_std[i] : 1.f) : 1.f;
    }
    upload_norm(n);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_set_normalization error: %s\n", e.what());
    return -1;
  }
}

MJMLX_API int mjmlx_mlp_get_normalization(const MjmlxMLP* n, float* in_mean, float* in_std, float* out_mean,
                                          float* out_std) {
  if (!n) return -1;
  for (int i = 0; i < n->nin; i++) {
    if (in_mean) in_mean[i] = n->h_in_mean[i];
    if (in_std) in_std[i] = 1.f / n->h_in_istd[i];
  }
  for (int i = 0; i < n->nout; i++) {
    if (out_mean) out_mean[i] = n->h_out_mean[i];
    if (out_std) out_std[i] = n->h_out_std[i];
  }
  return 0;
}

MJMLX_API int mjmlx_mlp_get_params(const MjmlxMLP* n, float* out) {
  if (!n || !out) return -1;
  try { return copy_packed(n, n->p, out, true); }
  catch (const std::exception& e) { std::fprintf(stderr, "mjmlx_mlp_get_params error: %s\n", e.what()); return -1; }
}

MJMLX_API int mjmlx_mlp_set_params(MjmlxMLP* n, const float* in) {
  if (!n || !in) return -1;
  try {
    int k = copy_packed(n, n->p, const_cast<float*>(in), false);
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n->nparam + 255) / 256)), dim3(256), 0, n->stream, n->nparam, n->p,
                       n->pb);
    HIPOK(hipStreamSynchronize(n->stream));
    return k;
  } catch (const std::exception& e) { std::fprintf(stderr, "mjmlx_mlp_set_params error: %s\n", e.what()); return -1; }
}

MJMLX_API int mjmlx_mlp_forward(MjmlxMLP* n, const float* x, int rows, float* y) {
  if (!n || !x || !y || rows <= 0) return -1;
  try {
    n->ensure(rows);
    upload_inputs(n, x, rows, rows);
    forward(n, n->xb, rows, n->stream, false);
    HIPOK(hipMemcpy2DAsync(y, (size_t)n->nout * 4, n->out, (size_t)n->op * 4, (size_t)n->nout * 4, rows,
                           hipMemcpyDeviceToHost, n->stream));
    HIPOK(hipStreamSynchronize(n->stream));
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_forward error: %s\n", e.what());
    return -1;
  }
}

MJMLX_API float mjmlx_mlp_grad(MjmlxMLP* n, const float* x, const float* y, int rows, float* grad_out) {
  if (!n || !x || !y || rows <= 0) return -1.f;
  try {
    int rp = rup(rows, 32);
    n->ensure(rp);
    upload_inputs(n, x, rows, rp);
    upload_targets(n, y, rows, rp);
    forward(n, n->xb, rp, n->stream, true);
    backward(n, n->xb, rp, rows, n->stream);
    float loss = mean_rowloss(n, rows);
    if (grad_out) copy_packed(n, n->g, grad_out, true);
    return loss;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_grad error: %s\n", e.what());
    return -1.f;
  }
}

MJMLX_API float mjmlx_mlp_train_step(MjmlxMLP* n, const float* x, const float* y, int rows, float lr,
                                     float weight_decay) {
  if (!n || !x || !y || rows <= 0) return -1.f;
  try {
    int rp = rup(rows, 32);
    n->ensure(rp);
    upload_inputs(n, x, rows, rp);
    upload_targets(n, y, rows, rp);
    forward(n, n->xb, rp, n->stream, true);
    backward(n, n->xb, rp, rows, n->stream);
    adamw(n, lr, weight_decay, n->stream);
    return mean_rowloss(n, rows);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_train_step error: %s\n", e.what());
    return -1.f;
  }
}

MJMLX_API int mjmlx_mlp_dataset_reserve(MjmlxMLP* n, long rows) {
  if (!n || rows <= 0) return -1;
  try {
    HIPOK(hipStreamSynchronize(n->stream));
    if (n->ds_x) (void)hipFree(n->ds_x);
    if (n->ds_y) (void)hipFree(n->ds_y);
    n->ds_x = n->ds_y = nullptr;
    n->ds_cap = n->ds_n = 0;
    HIPOK(hipMalloc(&n->ds_x, (size_t)rows * n->nin * 4));
    HIPOK(hipMalloc(&n->ds_y, (size_t)rows * n->nout * 4));
    n->ds_cap = rows;
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_dataset_reserve error: %s\n", e.what());
    return -1;
  }
}

MJMLX_API long mjmlx_mlp_dataset_size(const MjmlxMLP* n) { return n ? n->ds_n : -1; }

MJMLX_API long mjmlx_mlp_dataset_append(MjmlxMLP* n, const float* x, const float* y, long rows) {
  if (!n || !x || !y || rows < 0) return -1;
  try {
    long k = std::min(rows, n->ds_cap - n->ds_n);
    if (k <= 0) return 0;
    HIPOK(hipMemcpy(n->ds_x + (size_t)n->ds_n * n->nin, x, (size_t)k * n->nin * 4, hipMemcpyHostToDevice));
    HIPOK(hipMemcpy(n->ds_y + (size_t)n->ds_n * n->nout, y, (size_t)k * n->nout * 4, hipMemcpyHostToDevice));
    n->ds_n += k;
    return k;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_dataset_append error: %s\n", e.what());
    return -1;
  }
}

MJMLX_API int mjmlx_mlp_dataset_get(const MjmlxMLP* n, long first, long rows, float* x, float* y) {
  if (!n || first < 0 || rows < 0 || first + rows > n->ds_n) return -1;
  try {
    HIPOK(hipStreamSynchronize(n->stream));
    HIPOK(hipDeviceSynchronize());
    if (x) HIPOK(hipMemcpy(x, n->ds_x + (size_t)first * n->nin, (size_t)rows * n->nin * 4, hipMemcpyDeviceToHost));
    if (y) HIPOK(hipMemcpy(y, n->ds_y + (size_t)first * n->nout, (size_t)rows * n->nout * 4, hipMemcpyDeviceToHost));
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_dataset_get error: %s\n", e.what());
    return -1;
  }
}

MJMLX_API int mjmlx_mlp_fit_normalization(MjmlxMLP* n) {
  if (!n || n->ds_n <= 1) return -1;
  try {
    HIPOK(hipDeviceSynchronize());
    const int nchunk = 256;
    long chunk = (n->ds_n + nchunk - 1) / nchunk;
    std::vector<float> mu_in, sd_in, mu_out, sd_out;
    for (int which = 0; which < 2; which++) {
      int C = which == 0 ? n->nin : n->nout;
      const float* X = which == 0 ? n->ds_x : n->ds_y;
      float* part = nullptr;
      HIPOK(hipMalloc(&part, (size_t)nchunk * C * 3 * 4));
      hipLaunchKernelGGL(k_colstats, dim3((C + 63) / 64, nchunk), dim3(64), 0, n->stream, n->ds_n, C, X, chunk, part);
      HIPOK(hipGetLastError());
      std::vector<float> h((size_t)nchunk * C * 3);
      HIPOK(hipMemcpyAsync(h.data(), part, h.size() * 4, hipMemcpyDeviceToHost, n->stream));
      HIPOK(hipStreamSynchronize(n->stream));
      (void)hipFree(part);
      std::vector<float> mu(C), sd(C);
      for (int c = 0; c < C; c++) {
        // parallel combination of per-chunk (count, mean, M2)
        double N = 0, M = 0, M2 = 0;
        for (int k = 0; k < nchunk; k++) {
          long r0 = (long)k * chunk, r1 = std::min(n->ds_n, r0 + chunk);
          if (r1 <= r0) continue;
          double nk = (double)(r1 - r0);
          const float* q = &h[((size_t)k * C + c) * 3];
          double mk = q[0] + q[1] / nk, m2k = std::max(0.0, (double)q[2] - (double)q[1] * q[1] / nk);
          double d = mk - M, Nn = N + nk;
          M += d * nk / Nn;
          M2 += m2k + d * d * N * nk / Nn;
          N = Nn;
        }
        mu[c] = (float)M;
        sd[c] = (float)std::sqrt(M2 / std::max(1.0, N - 1));
      }
      if (which == 0) { mu_in = mu; sd_in = sd; } else { mu_out = mu; sd_out = sd; }
    }
    return mjmlx_mlp_set_normalization(n, mu_in.data(), sd_in.data(), mu_out.data(), sd_out.data());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_fit_normalization error: %s\n", e.what());
    return -1;
  }
}

MJMLX_API float mjmlx_mlp_fit(MjmlxMLP* n, int epochs, int batch, float lr, float weight_decay, int seed) {
  if (!n || epochs <= 0 || batch <= 0 || batch % 32 || n->ds_n < batch) return -1.f;
  try {
    HIPOK(hipDeviceSynchronize());
    n->ensure(batch);
    long nb = n->ds_n / batch, used = nb * batch;
    if (n->idx_cap < n->ds_n) {
      if (n->d_idx) (void)hipFree(n->d_idx);
      n->d_idx = nullptr;
      HIPOK(hipMalloc(&n->d_idx, (size_t)n->ds_n * sizeof(int)));
      n->idx_cap = n->ds_n;
    }
    std::vector<int> perm(n->ds_n);
    for (long i = 0; i < n->ds_n; i++) perm[i] = (int)i;
    std::mt19937_64 rng((uint64_t)seed);
    double last = 0;
    for (int ep = 0; ep < epochs; ep++) {
      std::shuffle(perm.begin(), perm.end(), rng);
      HIPOK(hipMemcpyAsync(n->d_idx, perm.data(), (size_t)used * sizeof(int), hipMemcpyHostToDevice, n->stream));
      double tot = 0;
      long counted = 0;
      for (long bi = 0; bi < nb; bi++) {
        hipLaunchKernelGGL(k_gather, dim3(batch), dim3(64), 0, n->stream, batch, n->nin, n->ip, n->nout, n->op,
                           n->d_idx + bi * batch, n->ds_x, n->ds_y, n->in_mean, n->in_istd, n->out_mean, n->out_std,
                           n->xb, n->tgt);
        forward(n, n->xb, batch, n->stream, true);
        backward(n, n->xb, batch, batch, n->stream);
        adamw(n, lr, weight_decay, n->stream);
        if (ep == epochs - 1 && (bi % 16 == 0 || bi == nb - 1)) {
          tot += mean_rowloss(n, batch);
          counted++;
        }
      }
      // the next epoch's index upload must not overwrite indices still in use
      HIPOK(hipStreamSynchronize(n->stream));
      if (counted) last = tot / counted;
    }
    return (float)last;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjmlx_mlp_fit error: %s\n", e.what());
    return -1.f;
  }
}

}  // extern "C"
