/* Additions of the HIP engine to the mjmlx C API.
 *
 * The reference's caller-allocated structs cannot grow without breaking the
 * ABI, so knobs the reference does not have (warm-start mode, solver
 * statistics, launch shape, timing, debug dumps) are new functions with no
 * counterpart in the reference.
 */
#ifndef MJMLX_HIP_H
#define MJMLX_HIP_H

#include <stddef.h>

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

/* Name of object `id` of MuJoCo type obj_type (1 body, 2 xbody, 3 joint,
 * 5 geom, 6 site, 19 actuator, 20 sensor), or NULL for an unnamed object, an
 * id out of range or another type; the inverse of mjmlx_name2id. No mjmlx.h
 * counterpart: it serves mjb_id2name, which the reference forwards to
 * MuJoCo's mj_id2name (src/mjb.cpp:241-245). */
MJMLX_API const char* mjmlx_id2name(const MjmlxModel* model, int obj_type, int id);

/* Warm start of the constraint solver:
 *   0  previous step's solver qacc, kept if its total cost is not above the
 *      cost at qacc_smooth (MuJoCo's warm start; default)
 *   1  qacc_smooth (warm start disabled)
 *   2  learned: the MLP installed with mjmlx_batched_set_warmstart_mlp
 *      predicts qacc; the solver starts from the cheapest of {prediction,
 *      previous qacc, qacc_smooth} and iterates to the same tolerance; one
 *      network pass and one launch per physics substep.
 *   3  learned, first substep only: one network pass per step call (on the
 *      features of the state the call starts from) and one fused launch of
 *      all substeps; the other substeps warm-start from the previous qacc.
 * Returns 0, or -1 (bad handle / modes 2 and 3 without weights). */
MJMLX_API int mjmlx_batched_set_warmstart_mode(MjmlxBatchedSim* sim, int mode);
MJMLX_API int mjmlx_batched_get_qacc_warmstart(const MjmlxBatchedSim* sim, float* out);
MJMLX_API int mjmlx_batched_set_qacc_warmstart(MjmlxBatchedSim* sim, const float* in);

/* Per-env statistics of the last step call (any pointer may be NULL):
 * niter = solver iterations summed over the call's substeps, status = OR of
 * status bits (1 contact overflow, 2 row overflow, 4 unsupported pair or cone,
 * 16 Hessian not positive definite, 32 mass matrix not positive definite,
 * 64 non-finite state), ncon / nefc of the last substep. */
MJMLX_API int mjmlx_batched_solver_stats(const MjmlxBatchedSim* sim, int* niter, int* status, int* ncon, int* nefc);

/* Launch shape: 0 launches one workgroup per env (returns 0). persistent != 0
 * would run a grid of `grid` workgroups pulling envs from a device work queue;
 * that variant is switched off (it hung once with no cause found) and returns
 * -1 with a message. */
MJMLX_API int mjmlx_batched_set_launch(MjmlxBatchedSim* sim, int persistent, int grid);

/* Time in milliseconds of the last step call's kernel, measured with HIP
 * events on the sim's stream (0 before the first step). */
MJMLX_API float mjmlx_batched_last_step_ms(const MjmlxBatchedSim* sim);

/* Capacity per env: contacts stored and constraint rows kept in the per-env
 * workspace, and the workspace size in bytes (LDS up to 144 KB, else device
 * memory). Overflows set status bits. */
MJMLX_API int mjmlx_batched_capacity(const MjmlxBatchedSim* sim, int* maxcon, int* maxefc, int* lds_bytes);

/* Autotuner: when the workspace fits in LDS, drains the device, times `steps`
 * control steps from the current state with five kernel variants: workspace
 * in LDS at 256 registers per lane (two waves per SIMD), in device memory, in
 * LDS at 168 VGPRs (three waves per SIMD, with spills), in LDS at 512
 * registers (one wave per SIMD) and in device memory at 168 VGPRs, and, for
 * models of at most 32 dofs whose two workspaces fit a workgroup's LDS, two
 * more with 32 lanes per env (two envs per wave) at 256 registers and at 168
 * VGPRs; best of two rounds after a warm-up round; keeps the fastest (choice:
 * 0 LDS, 1 device memory, 2 LDS at 168 VGPRs, 3 LDS at 512 registers,
 * 4 device memory at 168 VGPRs, 5 LDS with 32 lanes per env, 6 the same at
 * 168 VGPRs) and restores every buffer the timed steps wrote
 * (qpos, qvel, warm start, ctrl, solver statistics, sensor values and the
 * accelerometers' history, the aux outputs read by the get_xpos / cvel /
 * cfrc_ext getters), so the sim continues exactly as without the call; the
 * timed steps use the current ctrl. No reference counterpart: the
 * reference's kernels have one workspace placement per size class. */
MJMLX_API int mjmlx_batched_autotune(MjmlxBatchedSim* sim, int steps, int* choice);

/* Selects the step's kernel variant by the autotuner's numbering (0-6).
 * Returns -1 for a variant the sim cannot run: an LDS variant when the
 * workspace exceeds a workgroup's LDS, a 32-lane variant for a model of more
 * than 32 dofs or whose two workspaces exceed it. A device-memory variant
 * allocates the workspaces on first selection. No reference counterpart. */
MJMLX_API int mjmlx_batched_set_variant(MjmlxBatchedSim* sim, int choice);

/* Learned warm start (mode 2): replay each control step's launches (features,
 * network, one step launch per substep) as a captured HIP graph instead of
 * issuing them one by one (on != 0). The graph is captured on first use and
 * again whenever the substep count, network or kernel variant changes; steps
 * without a network or without features yet run eagerly. No reference
 * counterpart (the reference has no learned warm start). */
MJMLX_API int mjmlx_batched_set_graph(MjmlxBatchedSim* sim, int on);

/* Debug: one physics step like mjmlx_batched_step that also copies each env's
 * workspace after the constraint build (dump1) and after the solver (dump2),
 * float[num_envs * total] each (total from mjmlx_batched_layout("total")). */
MJMLX_API int mjmlx_batched_debug_step(MjmlxBatchedSim* sim, const float* ctrl, float* dump1, float* dump2);
/* Profiling builds (libmjmlx_prof.so) only: one call like step_n that also
 * returns per-env cycle counts of 17 pipeline stages (kinematics, com,
 * CRB, factorisation, velocities, smooth forces, collision, constraints,
 * solver set-up, line search, cost/gradient, search direction, other solver
 * work, integration, state I/O, Newton Hessian build, Newton substitution),
 * unsigned long long[num_envs * 17].
 * Returns -1 in the product build. */
MJMLX_API int mjmlx_batched_profile_step(MjmlxBatchedSim* sim, const float* ctrl, int nsub, unsigned long long* out);
/* offset / length (floats) / kind (0 float, 1 int) of a workspace field */
MJMLX_API int mjmlx_batched_layout(const MjmlxBatchedSim* sim, const char* name, int* off, int* len, int* kind);

/* ---- learned warm start (no counterpart in the reference) ----------------
 * A small MLP nin -> nh -> nh -> nout (SiLU) with FP32 master weights and BF16
 * copies, run and trained on the GPU by the engine's own MFMA kernels. For the
 * warm start, nin = nq + nv + nu + 16 * 8 (mjmlx_batched_feature_dim) and
 * nout = nv. Inputs: qpos, qvel, ctrl and the first 16 contacts as (frame
 * normal, penetration depth, one-hot limb class of the two geoms); output:
 * the solver's converged qacc. One MLP must not be trained and used by a sim
 * at the same time. */
typedef struct MjmlxMLP MjmlxMLP;

/* nh must be a multiple of 32; returns NULL without a HIP device */
MJMLX_API MjmlxMLP* mjmlx_mlp_create(int nin, int nh, int nout, int seed);
MJMLX_API void mjmlx_mlp_free(MjmlxMLP* mlp);
/* nparam counts the packed parameters: W1 [nin x nh], b1, W2 [nh x nh], b2, W3 [nh x nout], b3 */
MJMLX_API int mjmlx_mlp_dims(const MjmlxMLP* mlp, int* nin, int* nh, int* nout, int* nparam);
/* x_n = (x - in_mean) / in_std and y = y_n * out_std + out_mean (NULL: 0 / 1) */
MJMLX_API int mjmlx_mlp_set_normalization(MjmlxMLP* mlp, const float* in_mean, const float* in_std,
                                          const float* out_mean, const float* out_std);
MJMLX_API int mjmlx_mlp_get_normalization(const MjmlxMLP* mlp, float* in_mean, float* in_std, float* out_mean,
                                          float* out_std);
/* packed FP32 parameters in the order W1 b1 W2 b2 W3 b3 (row-major, y = x W + b); return the count */
MJMLX_API int mjmlx_mlp_get_params(const MjmlxMLP* mlp, float* out);
MJMLX_API int mjmlx_mlp_set_params(MjmlxMLP* mlp, const float* in);
/* network on host rows in normalised units: y [rows x nout] from x [rows x nin] */
MJMLX_API int mjmlx_mlp_forward(MjmlxMLP* mlp, const float* x, int rows, float* y);
/* gradient of mean((MLP(x) - y)^2) (normalised units) w.r.t. the packed parameters; returns the loss */
MJMLX_API float mjmlx_mlp_grad(MjmlxMLP* mlp, const float* x, const float* y, int rows, float* grad_out);
/* one AdamW step (beta 0.9 / 0.999, eps 1e-8) on a host batch; returns the loss before the step */
MJMLX_API float mjmlx_mlp_train_step(MjmlxMLP* mlp, const float* x, const float* y, int rows, float lr,
                                     float weight_decay);
/* device-resident training set of raw (unnormalised) rows; reserve discards previous rows */
MJMLX_API int mjmlx_mlp_dataset_reserve(MjmlxMLP* mlp, long rows);
MJMLX_API long mjmlx_mlp_dataset_size(const MjmlxMLP* mlp);
MJMLX_API long mjmlx_mlp_dataset_append(MjmlxMLP* mlp, const float* x, const float* y, long rows);
MJMLX_API int mjmlx_mlp_dataset_get(const MjmlxMLP* mlp, long first, long rows, float* x, float* y);
/* normalisation set to the per-column mean / standard deviation of the training set */
MJMLX_API int mjmlx_mlp_fit_normalization(MjmlxMLP* mlp);
/* AdamW over the training set in shuffled batches (batch a multiple of 32);
 * returns the mean loss (normalised units) of the last epoch */
MJMLX_API float mjmlx_mlp_fit(MjmlxMLP* mlp, int epochs, int batch, float lr, float weight_decay, int seed);

/* width of the warm-start features of this sim's model */
MJMLX_API int mjmlx_batched_feature_dim(const MjmlxBatchedSim* sim);
/* feature mode 1 appends the previous substep's converged qacc (nv values) to
 * the features and makes the network residual: it is trained on, and
 * predicts, the change of qacc from that previous value. Mode 0 (default) is
 * the plain feature set. Changing the mode removes the installed network. */
MJMLX_API int mjmlx_batched_set_feature_mode(MjmlxBatchedSim* sim, int mode);
/* installs the network used by warm-start mode 2 (NULL removes it and leaves mode 2);
 * returns -1 when its widths do not match the model. The network's activation
 * buffers serve one sim at a time: sims sharing a network must not step
 * concurrently on different streams. */
MJMLX_API int mjmlx_batched_set_warmstart_mlp(MjmlxBatchedSim* sim, MjmlxMLP* mlp);
/* steps like step_n and appends one (features, converged qacc) row per env and
 * substep to mlp's training set (the first substep after a state change has no
 * features and is skipped); returns the rows appended or -1 */
MJMLX_API long mjmlx_batched_record(MjmlxBatchedSim* sim, MjmlxMLP* mlp, const float* ctrl, int nsub);
/* raw features after the last substep [num_envs x feature_dim] (ctrl columns zero); returns the width or -1 */
MJMLX_API int mjmlx_batched_get_features(const MjmlxBatchedSim* sim, float* out);

/* Allocation-free step. No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Steps like mjmlx_batched_step_n
 * with the plain warm start (modes 0 and 1; the learned modes return -1), with
 * the per-env workspaces in a fixed device array of the library (one per GPU,
 * held until the process ends) instead of memory the library allocates, and
 * calls no allocating routine. Limit: num_envs x workspace bytes (see
 * mjmlx_batched_capacity) at most mjmlx_batched_noalloc_scratch_bytes()
 * (64 MiB), else returns -1 with a message. Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (sims on different streams or
 * host threads) race. The array needs no clearing: each call initialises the
 * workspaces it uses from the sim's state. Returns 0, or -1 on error. */
MJMLX_API int mjmlx_batched_step_n_noalloc(MjmlxBatchedSim* sim, const float* ctrl, int n);
MJMLX_API size_t mjmlx_batched_noalloc_scratch_bytes(void);

#ifdef __cplusplus
}
#endif

#endif /* MJMLX_HIP_H */
