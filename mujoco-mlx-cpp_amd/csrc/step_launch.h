// Launch interface of the fused step kernels.
//
// The step kernel is instantiated once per narrow-phase set (NPSET: 0 basic
// pairs, 1 + capsule-box / box-box, 2 + GJK/EPA, +4 with the sensor stage) and
// per workspace/register variant. Each narrow-phase set lives in its own
// translation units (step_np.hip compiled with -DMJK_NP=<set>, once for models
// of at most SPARSE_ROWS_NV dofs and once with every path), so the sets build
// in parallel; the engine picks one through launch_step_np().
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace mjk {

// kernel variant: where the per-env workspace lives and the register budget per lane
enum {
  VAR_LDS = 0,      // LDS, 256 registers (two waves per SIMD)
  VAR_GWS = 1,      // device memory, 256 registers
  VAR_LDS_OCC = 2,  // LDS, 168 VGPRs (three waves per SIMD)
  VAR_LDS_BIG = 3,  // LDS, the whole register file (one wave per SIMD)
  VAR_GWS_OCC = 4,  // device memory, 168 VGPRs
  VAR_LDS_HALF = 5,      // LDS, 32 lanes per env (two envs per wave), 256 registers
  VAR_LDS_OCC_HALF = 6,  // LDS, 32 lanes per env, 168 VGPRs
  VAR_N = 7
};

struct StepLaunch {
  const KParams* kp;   // the sim's model and layout in device memory
  int B;               // envs of this launch
  int env0;            // index of the launch's first env in the io arrays
  int variant;         // VAR_*
  size_t lds_bytes;    // dynamic LDS per workgroup (LDS variants)
  float* gws;          // device-memory workspaces [B][L.total] (VAR_GWS*)
  hipStream_t stream;
  int ls_cap;          // VAR_GWS*: line-search rows staged in LDS (0: none)
};

// rows of the line-search cache beside a device-memory workspace: only for
// models whose rows exceed the two per lane held in registers, at most 800
// (12.8 KB of LDS per workgroup, within the share of a CU's LDS that the
// register-limited occupancy of those kernels leaves to each workgroup)
inline int ls_cache_rows(int maxefc) { return maxefc > 128 ? (maxefc < 800 ? maxefc : 800) : 0; }

// kernels built for models of at most SPARSE_ROWS_NV dofs (MJK_BIG 0) ...
void launch_step_np0(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np1(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np2(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np4(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np5(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np6(const StepLaunch& l, const StepIO& io, int nsub);
// ... and with every path, for any model size (MJK_BIG 1)
void launch_step_np0_big(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np1_big(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np2_big(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np4_big(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np5_big(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np6_big(const StepLaunch& l, const StepIO& io, int nsub);

// ... and with 32 lanes per env (models of at most SPARSE_ROWS_NV dofs)
void launch_step_np0_half(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np1_half(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np2_half(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np4_half(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np5_half(const StepLaunch& l, const StepIO& io, int nsub);
void launch_step_np6_half(const StepLaunch& l, const StepIO& io, int nsub);

// npset: narrow-phase set (+4 with sensors); big: the model has more than
// SPARSE_ROWS_NV dofs; the 32-lane variants (VAR_*_HALF) need a small model
inline void launch_step_np(int npset, bool big, const StepLaunch& l, const StepIO& io, int nsub) {
  if (l.variant == VAR_LDS_HALF || l.variant == VAR_LDS_OCC_HALF) {
    switch (npset) {
      case 0: launch_step_np0_half(l, io, nsub); break;
      case 1: launch_step_np1_half(l, io, nsub); break;
      case 2: launch_step_np2_half(l, io, nsub); break;
      case 4: launch_step_np4_half(l, io, nsub); break;
      case 5: launch_step_np5_half(l, io, nsub); break;
      default: launch_step_np6_half(l, io, nsub); break;
    }
    return;
  }
  if (big) {
    switch (npset) {
      case 0: launch_step_np0_big(l, io, nsub); break;
      case 1: launch_step_np1_big(l, io, nsub); break;
      case 2: launch_step_np2_big(l, io, nsub); break;
      case 4: launch_step_np4_big(l, io, nsub); break;
      case 5: launch_step_np5_big(l, io, nsub); break;
      default: launch_step_np6_big(l, io, nsub); break;
    }
    return;
  }
  switch (npset) {
    case 0: launch_step_np0(l, io, nsub); break;
    case 1: launch_step_np1(l, io, nsub); break;
    case 2: launch_step_np2(l, io, nsub); break;
    case 4: launch_step_np4(l, io, nsub); break;
    case 5: launch_step_np5(l, io, nsub); break;
    default: launch_step_np6(l, io, nsub); break;
  }
}

}  // namespace mjk
