/* Additions of the HIP engine to the mjmlx C API.
 *
 * The reference's caller-allocated structs cannot grow without breaking the
 * ABI, so knobs the reference does not have (warm-start mode, solver
 * statistics, launch shape, timing, debug dumps) are new functions with no
 * counterpart in the reference.
 */
#ifndef MJMLX_HIP_H
#define MJMLX_HIP_H

#include "mjmlx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* number of HIP devices visible to the engine (0 when none) */
MJMLX_API int mjmlx_hip_device_count(void);

/* Raw model field by MuJoCo name (integers converted to double); copies up to
 * cap values and returns the field length, or -1 for an unknown name. Also
 * serves the compiled collision pair table as cpair_{g1,g2,b1,b2,condim,
 * friction,solref,solimp,margin,gap}. */
MJMLX_API int mjmlx_model_field(const MjmlxModel* model, const char* name, double* out, int cap);

/* Warm start of the constraint solver:
 *   0  previous step's solver qacc, kept if its total cost is not above the
 *      cost at qacc_smooth (MuJoCo's warm start; default)
 *   1  qacc_smooth (warm start disabled)
 *   2  learned: the MLP installed with mjmlx_batched_set_warmstart_mlp
 *      predicts qacc; the solver starts from the cheapest of {prediction,
 *      previous qacc, qacc_smooth} and iterates to the same tolerance.
 * Returns 0, or -1 (bad handle / mode 2 without weights). */
MJMLX_API int mjmlx_batched_set_warmstart_mode(MjmlxBatchedSim* sim, int mode);
MJMLX_API int mjmlx_batched_get_qacc_warmstart(const MjmlxBatchedSim* sim, float* out);
MJMLX_API int mjmlx_batched_set_qacc_warmstart(MjmlxBatchedSim* sim, const float* in);

/* Per-env statistics of the last step call (any pointer may be NULL):
 * niter = solver iterations summed over the call's substeps, status = OR of
 * status bits (1 contact overflow, 2 row overflow, 4 unsupported pair or cone,
 * 16 Hessian not positive definite, 32 mass matrix not positive definite,
 * 64 non-finite state), ncon / nefc of the last substep. */
MJMLX_API int mjmlx_batched_solver_stats(const MjmlxBatchedSim* sim, int* niter, int* status, int* ncon, int* nefc);

/* Launch shape: persistent != 0 runs a grid of `grid` workgroups (0 = chosen
 * from the CU count) that pull envs from a device work queue; 0 launches one
 * workgroup per env. Returns 0 or -1. */
MJMLX_API int mjmlx_batched_set_launch(MjmlxBatchedSim* sim, int persistent, int grid);

/* Time in milliseconds of the last step call's kernel, measured with HIP
 * events on the sim's stream (0 before the first step). */
MJMLX_API float mjmlx_batched_last_step_ms(const MjmlxBatchedSim* sim);

/* Capacity per env: contacts and constraint rows kept in the per-env LDS
 * workspace (beyond them the status word reports an overflow). */
MJMLX_API int mjmlx_batched_capacity(const MjmlxBatchedSim* sim, int* maxcon, int* maxefc, int* lds_bytes);

/* Debug: one physics step like mjmlx_batched_step that also copies each env's
 * workspace after the constraint build (dump1) and after the solver (dump2),
 * float[num_envs * total] each (total from mjmlx_batched_layout("total")). */
MJMLX_API int mjmlx_batched_debug_step(MjmlxBatchedSim* sim, const float* ctrl, float* dump1, float* dump2);
/* Profiling builds (libmjmlx_prof.so) only: one call like step_n that also
 * returns per-env cycle counts of 16 pipeline stages (kinematics, com,
 * CRB, factorisation, velocities, smooth forces, collision, constraints,
 * solver set-up, line search, cost/gradient, Newton direction, other solver
 * work, integration, state I/O), unsigned long long[num_envs * 16].
 * Returns -1 in the product build. */
MJMLX_API int mjmlx_batched_profile_step(MjmlxBatchedSim* sim, const float* ctrl, int nsub, unsigned long long* out);
/* offset / length (floats) / kind (0 float, 1 int) of a workspace field */
MJMLX_API int mjmlx_batched_layout(const MjmlxBatchedSim* sim, const char* name, int* off, int* len, int* kind);

#ifdef __cplusplus
}
#endif

#endif /* MJMLX_HIP_H */
