/* mjmlx C API served by the HIP engine (libmjmlx.so, gfx950).
 *
 * Every entry point has the name, signature and argument meaning of the
 * declaration it replaces in the reference (reference include/mjmlx/mjmlx.h,
 * line numbers in the comments), so existing load_model()/step() callers link
 * unchanged. Behaviour notes:
 *  - Handles are opaque; create/free come in pairs.
 *  - Getters return sim-owned host pointers, valid until the next step, set
 *    or free, with *n_out = num_envs * dim.
 *  - State is env-major: qpos[env * nq + i].
 *  - Errors: creators return NULL and print to stderr; int functions return
 *    0 on success, -1 for a bad handle / unsupported request, -2 for an
 *    index out of range. C++ exceptions never cross this boundary.
 *  - The single-env pipeline was removed in the reference (reference
 *    src/forward.cpp:165-167,285-297); its entry points here print an error
 *    and do nothing / return NULL. Camera entry points behave like the
 *    reference's camera stub (reference src/camera_stub.cpp:9-28).
 */
#ifndef MJMLX_H
#define MJMLX_H

#include "mjmlx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library info (reference mjmlx.h:37) */
MJMLX_API const char* mjmlx_version(void);

/* ---- model I/O (reference mjmlx.h:44-65). The MJCF subset compiled by the
 * engine: compiler, option(+flag), default classes, worldbody, body,
 * joint/freejoint, geom (plane sphere capsule ellipsoid cylinder box),
 * inertial, site, actuator (position motor velocity general on joints),
 * contact (pair exclude). */
MJMLX_API MjmlxModel* mjmlx_load_model(const char* xml_path);                  /* :44 */
MJMLX_API MjmlxModel* mjmlx_load_model_from_string(const char* xml_string);    /* :47 */
typedef struct MjmlxModelOptions {                                              /* :53-56 */
  int self_collision;
  int max_condim;
} MjmlxModelOptions;
MJMLX_API MjmlxModelOptions mjmlx_default_model_options(void);                 /* :57 */
MJMLX_API MjmlxModel* mjmlx_load_model_ex(const char* xml_path, const MjmlxModelOptions* options);            /* :58 */
MJMLX_API MjmlxModel* mjmlx_load_model_from_string_ex(const char* xml_string, const MjmlxModelOptions* options); /* :59 */
MJMLX_API void mjmlx_free_model(MjmlxModel* model);                            /* :62 */
MJMLX_API MjmlxModelInfo mjmlx_model_info(const MjmlxModel* model);            /* :65 */

/* ---- model field accessors (reference mjmlx.h:72-95) */
MJMLX_API float mjmlx_model_opt_timestep(const MjmlxModel* model);             /* :72 */
MJMLX_API void mjmlx_model_set_opt_timestep(MjmlxModel* model, float dt);      /* :75 */
MJMLX_API float mjmlx_model_body_mass(const MjmlxModel* model, int body_id);   /* :78 */
/* obj_type: MuJoCo mjtObj (1 body, 2 joint, 5 geom, 19 actuator) (:83) */
MJMLX_API int mjmlx_name2id(const MjmlxModel* model, int obj_type, const char* name);
/* sensors of the MJCF <sensor> section (reference src/sensor.cpp:362-377):
 * jointpos/vel, actuatorpos/vel, jointactuatorfrc, ballquat/angvel, frame
 * pos/quat/x,y,z axis/linvel/angvel of a site, geom, body or xbody, gyro,
 * velocimeter, accelerometer, magnetometer, subtreecom; types are MuJoCo's
 * mjtSensor ids */
MJMLX_API int mjmlx_model_nsensor(const MjmlxModel* model);                    /* :86 */
MJMLX_API int mjmlx_model_nsensordata(const MjmlxModel* model);                /* :87 */
MJMLX_API const char* mjmlx_model_sensor_name(const MjmlxModel* model, int sensor_id); /* :88 */
MJMLX_API int mjmlx_model_sensor_type(const MjmlxModel* model, int sensor_id); /* :89 */
MJMLX_API int mjmlx_model_sensor_adr(const MjmlxModel* model, int sensor_id);  /* :90 */
MJMLX_API int mjmlx_model_sensor_dim(const MjmlxModel* model, int sensor_id);  /* :91 */
/* the reference returns MuJoCo's mjModel*; MuJoCo C is not part of this
 * engine, so this returns NULL (:95) */
MJMLX_API const void* mjmlx_get_mj_model(const MjmlxModel* model);

/* ---- single-env simulation (reference mjmlx.h:102-164; removed upstream) */
MJMLX_API MjmlxData* mjmlx_make_data(const MjmlxModel* model);
MJMLX_API void mjmlx_free_data(MjmlxData* data);
MJMLX_API void mjmlx_reset_data(const MjmlxModel* model, MjmlxData* data);
MJMLX_API void mjmlx_forward(const MjmlxModel* model, MjmlxData* data);
MJMLX_API void mjmlx_step(const MjmlxModel* model, MjmlxData* data);
MJMLX_API void mjmlx_step1(const MjmlxModel* model, MjmlxData* data);
MJMLX_API void mjmlx_step2(const MjmlxModel* model, MjmlxData* data);
MJMLX_API void mjmlx_kinematics(const MjmlxModel* model, MjmlxData* data);
MJMLX_API void mjmlx_set_qpos(MjmlxData* data, const float* qpos, int n);
MJMLX_API void mjmlx_set_qvel(MjmlxData* data, const float* qvel, int n);
MJMLX_API void mjmlx_set_ctrl(MjmlxData* data, const float* ctrl, int n);
MJMLX_API const float* mjmlx_get_qpos(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qvel(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_ctrl(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_xpos(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_xquat(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_xipos(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_cvel(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qfrc_bias(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qacc(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qfrc_constraint(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qfrc_actuator(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qfrc_passive(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qfrc_gravcomp(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qfrc_smooth(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_qacc_smooth(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_subtree_com(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_cinert(const MjmlxData* data, int* n_out);
MJMLX_API const float* mjmlx_get_cfrc_ext(const MjmlxData* data, int* n_out);
MJMLX_API int mjmlx_get_ncon(const MjmlxData* data);
MJMLX_API int mjmlx_get_nefc(const MjmlxData* data);
MJMLX_API void mjmlx_rne_post_constraint(const MjmlxModel* model, MjmlxData* data);

/* ---- batched simulation on the GPU (reference mjmlx.h:171-252) */
MJMLX_API MjmlxBatchedSim* mjmlx_batched_create(const MjmlxModel* model, const MjmlxBatchedConfig* config); /* :171 */
MJMLX_API void mjmlx_batched_free(MjmlxBatchedSim* sim);                                  /* :176 */
/* one physics step of every env; ctrl float[num_envs*nu], NULL = zero (:181) */
MJMLX_API void mjmlx_batched_step(MjmlxBatchedSim* sim, const float* ctrl);
/* n physics steps with the same ctrl, fused into one kernel launch (:186) */
MJMLX_API void mjmlx_batched_step_n(MjmlxBatchedSim* sim, const float* ctrl, int n);
/* reset envs with mask[i] != 0 to qpos0, zero velocity and warm start (:193) */
MJMLX_API void mjmlx_batched_reset(MjmlxBatchedSim* sim, const int* reset_mask);
MJMLX_API void mjmlx_batched_get_state(const MjmlxBatchedSim* sim, float* qpos_out, float* qvel_out,
                                       int* nq_out, int* nv_out);                           /* :198 */
MJMLX_API void mjmlx_batched_set_state(MjmlxBatchedSim* sim, const float* qpos, const float* qvel); /* :203 */
MJMLX_API const float* mjmlx_batched_get_qpos(const MjmlxBatchedSim* sim, int* n_out);          /* :208 */
MJMLX_API const float* mjmlx_batched_get_qvel(const MjmlxBatchedSim* sim, int* n_out);          /* :209 */
/* derived quantities of the last step's pre-integration state (:210-216) */
MJMLX_API const float* mjmlx_batched_get_xpos(const MjmlxBatchedSim* sim, int* n_out);
MJMLX_API const float* mjmlx_batched_get_xquat(const MjmlxBatchedSim* sim, int* n_out);
MJMLX_API const float* mjmlx_batched_get_subtree_com(const MjmlxBatchedSim* sim, int* n_out);
MJMLX_API const float* mjmlx_batched_get_cinert(const MjmlxBatchedSim* sim, int* n_out);
MJMLX_API const float* mjmlx_batched_get_cvel(const MjmlxBatchedSim* sim, int* n_out);
MJMLX_API const float* mjmlx_batched_get_qfrc_actuator(const MjmlxBatchedSim* sim, int* n_out);
MJMLX_API const float* mjmlx_batched_get_cfrc_ext(const MjmlxBatchedSim* sim, int* n_out);
MJMLX_API int mjmlx_batched_nsensordata(const MjmlxBatchedSim* sim);                            /* :220 */
/* [num_envs][nsensordata] of the state the last step's last substep started
 * from (as the reference's outputs); the accelerometer differences its frame's
 * velocity over one substep. NULL without sensors. */
MJMLX_API const float* mjmlx_batched_get_sensordata(const MjmlxBatchedSim* sim, int* n_out);     /* :221 */
MJMLX_API void mjmlx_batched_set_env_qpos(MjmlxBatchedSim* sim, int env_idx, const float* qpos, int nq); /* :224 */
MJMLX_API void mjmlx_batched_set_env_qvel(MjmlxBatchedSim* sim, int env_idx, const float* qvel, int nv); /* :225 */
/* per-env physics randomisation of a sim created with domain_rand = 1 (:233-238):
 * mass_scale [n][nbody] scales each body's mass and inertia, com_delta
 * [n][nbody][3] shifts its COM in the body frame, friction_scale [n] scales
 * every contact friction coefficient, gain_scale [n][nu] scales each
 * actuator's force, drag_k [n][num_free_joints] adds F = -k |v| v on each
 * free joint's translation. Rows go to env_idx[i] (env i when env_idx is
 * NULL and n = num_envs); NULL arrays are kept. 0 ok, -1 bad handle or no
 * domain_rand, -2 env index out of range, -3 (drag) no free joint. */
MJMLX_API int mjmlx_batched_set_env_params(MjmlxBatchedSim* sim, const int* env_idx, int n, const float* mass_scale,
                                           const float* com_delta, const float* friction_scale, const float* gain_scale);
MJMLX_API int mjmlx_batched_set_free_drag(MjmlxBatchedSim* sim, const int* env_idx, int n, const float* drag_k); /* :237 */
MJMLX_API int mjmlx_batched_num_free_joints(const MjmlxBatchedSim* sim);                                           /* :238 */
/* packed readback [qpos, qvel, |contact force| of chosen bodies, sum of
 * squares and of absolute values of qfrc_actuator over a dof range] (:241-242) */
MJMLX_API int mjmlx_batched_set_pack_spec(MjmlxBatchedSim* sim, const int* body_ids, int nbody_sel, int dof_lo, int ndof);
MJMLX_API int mjmlx_batched_get_pack(MjmlxBatchedSim* sim, float* out);
/* mj_ray per env and ray at each env's current qpos (:248): pnt / vec
 * [num_envs][nray][3]; dist_out (units of |vec|) and geomid_out [num_envs][nray],
 * -1 / -1 for no hit. Skips geoms of body bodyexclude, of groups off in
 * geomgroup[6] (NULL: all on) and, when flg_static is 0, of bodies welded to
 * the world. Planes are one-sided. Returns 0, or -1 for bad arguments. */
MJMLX_API int mjmlx_batched_raycast(MjmlxBatchedSim* sim, const float* pnt, const float* vec, int nray, int bodyexclude,
                                   const unsigned char* geomgroup, int flg_static, float* dist_out, int* geomid_out);
/* completes all queued work of the sim (:252) */
MJMLX_API void mjmlx_batched_eval_state(const MjmlxBatchedSim* sim);

/* ---- differentiable simulation (reference mjmlx.h:261-268; stubs upstream) */
MJMLX_API void mjmlx_grad_step(const MjmlxModel* model, const MjmlxData* data, float* grad_out);
MJMLX_API void mjmlx_batched_grad_step(MjmlxBatchedSim* sim, float* grad_out);

/* ---- camera (reference mjmlx.h:277-295; rendering is not part of this engine) */
MJMLX_API MjmlxCameraRenderer* mjmlx_camera_create(MjmlxBatchedSim* sim, int width, int height);
MJMLX_API void mjmlx_camera_free(MjmlxCameraRenderer* r);
MJMLX_API int mjmlx_camera_count(const MjmlxCameraRenderer* r);
MJMLX_API int mjmlx_camera_render(MjmlxCameraRenderer* r);
MJMLX_API int mjmlx_camera_draw_call_count(const MjmlxCameraRenderer* r);
MJMLX_API int mjmlx_camera_get_rgb(MjmlxCameraRenderer* r, uint8_t* out);
MJMLX_API int mjmlx_camera_get_depth(MjmlxCameraRenderer* r, float* out);
MJMLX_API int mjmlx_camera_gpu_supported(const MjmlxCameraRenderer* r);
MJMLX_API int mjmlx_camera_gpu_copy(MjmlxCameraRenderer* r);
MJMLX_API int mjmlx_camera_gpu_export_fd(MjmlxCameraRenderer* r, int which, uint64_t* alloc_size);

#ifdef __cplusplus
}
#endif

#endif /* MJMLX_H */
