// Fused step kernels of one narrow-phase set (MJK_NP, set by the build):
// all substeps of a control step, one env per 64-lane workgroup.
//
// Replaces the reference's per-stage Vulkan dispatches and the host substep
// loop (reference src/batched.cpp:980-1228,1469-1504).
#include "step_launch.h"

#ifndef MJK_NP
#error "step_np.hip is compiled once per narrow-phase set with -DMJK_NP=<0|1|2|4|5|6>"
#endif
#ifndef MJK_BIG
#error "step_np.hip is compiled with -DMJK_BIG=0 (models of at most SPARSE_ROWS_NV = 32 dofs) and -DMJK_BIG=1 (any size)"
#endif

// env stepped by workgroup slot `slot` (the launch's order: the identity
// unless the envs are sorted by their previous work)
__device__ __forceinline__ int slot_env(const mjk::StepIO& io, int slot) { return io.order[slot]; }

#define MJK_CAT2(a, b) a##b
#define MJK_CAT(a, b) MJK_CAT2(a, b)
#if defined(MJK_HALF) && MJK_HALF
#define MJK_LAUNCH MJK_CAT(MJK_CAT(launch_step_np, MJK_NP), _half)
#elif MJK_BIG
#define MJK_LAUNCH MJK_CAT(MJK_CAT(launch_step_np, MJK_NP), _big)
#else
#define MJK_LAUNCH MJK_CAT(launch_step_np, MJK_NP)
#endif

// GWS: the per-env workspace lives in device memory (gws + blockIdx.x *
// L.total) instead of LDS, for models whose workspace exceeds the LDS of a
// workgroup. Capped at 256 registers per lane (VGPRs and AGPRs together) so
// that two waves fit on a SIMD: without the cap the compiler spills VGPRs
// into AGPRs beyond 256 and the kernel runs one wave per SIMD, four envs per CU.
template <int NPSET, int GWS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_step(const mjk::KParams* __restrict__ kp, mjk::StepIO io, int B, int env0, int nsub, float* gws, int ls_cap) {
  extern __shared__ __align__(16) float lds[];
  if ((int)blockIdx.x >= B) return;
  const mjk::KParams& P = mjk::kparams(kp);
  float* w = GWS ? gws + (size_t)blockIdx.x * P.L.total : lds;
  mjk::LsCache lc{GWS ? reinterpret_cast<mjk::LsRow4*>(lds) : nullptr, GWS ? ls_cap : 0};
  mjk::run_env<NPSET>(P.m, P.L, w, io, slot_env(io, env0 + (int)blockIdx.x), nsub, lc);
}

// The same step capped at 168 VGPRs (three waves per SIMD, a few hundred
// bytes of spills per lane), for workspaces small enough that LDS admits more
// than eight envs per CU, and (GWS 1, picked by the autotuner) for device-memory
// workspaces, whose occupancy only registers limit.
template <int NPSET, int GWS>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
k_step_occ(const mjk::KParams* __restrict__ kp, mjk::StepIO io, int B, int env0, int nsub, float* gws, int ls_cap) {
  extern __shared__ __align__(16) float lds[];
  if ((int)blockIdx.x >= B) return;
  const mjk::KParams& P = mjk::kparams(kp);
  float* w = GWS ? gws + (size_t)blockIdx.x * P.L.total : lds;
  mjk::LsCache lc{GWS ? reinterpret_cast<mjk::LsRow4*>(lds) : nullptr, GWS ? ls_cap : 0};
  mjk::run_env<NPSET | 8>(P.m, P.L, w, io, slot_env(io, env0 + (int)blockIdx.x), nsub, lc);
}

// The step with the whole register file of a SIMD lane (one wave per SIMD),
// for LDS workspaces large enough that at most four envs fit on a CU anyway.
template <int NPSET>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
k_step_big(const mjk::KParams* __restrict__ kp, mjk::StepIO io, int B, int env0, int nsub) {
  extern __shared__ __align__(16) float lds[];
  if ((int)blockIdx.x >= B) return;
  const mjk::KParams& P = mjk::kparams(kp);
  mjk::run_env<NPSET>(P.m, P.L, lds, io, slot_env(io, env0 + (int)blockIdx.x), nsub);
}

#if defined(MJK_HALF) && MJK_HALF
// 32 lanes per env: each 64-lane workgroup steps envs 2b (lanes 0-31) and
// 2b + 1 (lanes 32-63), each half with its own LDS workspace; the second half
// of the last workgroup is idle when B is odd. Register budgets as k_step and
// k_step_occ.
template <int NPSET>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
k_step_half(const mjk::KParams* __restrict__ kp, mjk::StepIO io, int B, int env0, int nsub) {
  extern __shared__ __align__(16) float lds[];
  int h = (int)(threadIdx.x >> 5), e = 2 * (int)blockIdx.x + h;
  if (e >= B) return;
  const mjk::KParams& P = mjk::kparams(kp);
  mjk::run_env<NPSET>(P.m, P.L, lds + h * P.L.total, io, slot_env(io, env0 + e), nsub);
}

template <int NPSET>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3)))
k_step_half_occ(const mjk::KParams* __restrict__ kp, mjk::StepIO io, int B, int env0, int nsub) {
  extern __shared__ __align__(16) float lds[];
  int h = (int)(threadIdx.x >> 5), e = 2 * (int)blockIdx.x + h;
  if (e >= B) return;
  const mjk::KParams& P = mjk::kparams(kp);
  mjk::run_env<NPSET | 8>(P.m, P.L, lds + h * P.L.total, io, slot_env(io, env0 + e), nsub);
}

namespace mjk {

void MJK_LAUNCH(const StepLaunch& l, const StepIO& io, int nsub) {
  // bit 32 of the set names the 32-lane kernels
  constexpr int NP = MJK_NP | 32;
  dim3 g((l.B + 1) / 2), b(64);
  size_t sh = 2 * l.lds_bytes;
  if (l.variant == VAR_LDS_OCC_HALF)
    hipLaunchKernelGGL((k_step_half_occ<NP>), g, b, sh, l.stream, l.kp, io, l.B, l.env0, nsub);
  else
    hipLaunchKernelGGL((k_step_half<NP>), g, b, sh, l.stream, l.kp, io, l.B, l.env0, nsub);
}

}  // namespace mjk
#else
namespace mjk {

void MJK_LAUNCH(const StepLaunch& l, const StepIO& io, int nsub) {
  // bit 16 of the set names the kernels built with every path (MJK_BIG 1)
  constexpr int NP = MJK_NP | (MJK_BIG ? 16 : 0);
  dim3 g(l.B), b(64);
  bool gws = l.variant == VAR_GWS || l.variant == VAR_GWS_OCC;
  // device-memory workspaces: the dynamic LDS holds the line search's row cache
  size_t sh = gws ? (size_t)l.ls_cap * sizeof(LsRow4) : l.lds_bytes;
  switch (l.variant) {
    case VAR_GWS_OCC:
      hipLaunchKernelGGL((k_step_occ<NP, 1>), g, b, sh, l.stream, l.kp, io, l.B, l.env0, nsub, l.gws, l.ls_cap);
      break;
    case VAR_GWS:
      hipLaunchKernelGGL((k_step<NP, 1>), g, b, sh, l.stream, l.kp, io, l.B, l.env0, nsub, l.gws, l.ls_cap);
      break;
    case VAR_LDS_OCC:
      hipLaunchKernelGGL((k_step_occ<NP, 0>), g, b, sh, l.stream, l.kp, io, l.B, l.env0, nsub, l.gws, 0);
      break;
    case VAR_LDS_BIG:
      hipLaunchKernelGGL((k_step_big<NP>), g, b, sh, l.stream, l.kp, io, l.B, l.env0, nsub);
      break;
    default:
      hipLaunchKernelGGL((k_step<NP, 0>), g, b, sh, l.stream, l.kp, io, l.B, l.env0, nsub, l.gws, 0);
      break;
  }
}

}  // namespace mjk
#endif
