/* mjb unified C API, HIP engine build (libmjb.so).
 *
 * Same entry points as the reference (reference include/mjmlx/mjb.h:38-240).
 * MJB_BACKEND_MLX (== MJB_BACKEND_HIP) dispatches to the HIP engine through
 * the mjmlx C API. MJB_BACKEND_CPU is MuJoCo C in the reference
 * (reference src/mjb.cpp:807-839); MuJoCo C is not part of this engine, so
 * loading a model on that backend fails with a message on stderr and NULL.
 * Single-env entry points follow the removed single-env pipeline of the
 * reference's GPU engine: they print an error and return NULL / do nothing.
 */
#ifndef MJB_H
#define MJB_H

#include "mjb_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* backend lifecycle (:38-40) */
MJB_API MjbBackend* mjb_create_backend(MjbBackendType type);
MJB_API void mjb_free_backend(MjbBackend* backend);
MJB_API MjbBackendType mjb_backend_type(const MjbBackend* backend);

/* model I/O (:46-48) */
MJB_API MjbModel* mjb_load_model(MjbBackend* b, const char* xml_path);
MJB_API MjbModel* mjb_load_model_from_string(MjbBackend* b, const char* xml_string);
MJB_API void mjb_free_model(MjbModel* model);

/* model accessors (:54-74) */
MJB_API MjbModelInfo mjb_model_info(const MjbModel* model);
MJB_API float mjb_model_opt_timestep(const MjbModel* model);
MJB_API void mjb_model_set_opt_timestep(MjbModel* model, float dt);
MJB_API float mjb_model_body_mass(const MjbModel* model, int body_id);
MJB_API int mjb_name2id(const MjbModel* model, int obj_type, const char* name);
MJB_API const char* mjb_id2name(const MjbModel* model, int obj_type, int id);
MJB_API int mjb_model_jnt_qposadr(const MjbModel* model, int jnt_id);
MJB_API int mjb_model_jnt_dofadr(const MjbModel* model, int jnt_id);
MJB_API int mjb_model_jnt_type(const MjbModel* model, int jnt_id);
MJB_API int mjb_model_nconmax(const MjbModel* model);
MJB_API int mjb_model_geom_type(const MjbModel* model, int geom_id);
MJB_API int mjb_model_sensor_adr(const MjbModel* model, int sensor_id);
MJB_API int mjb_model_body_mocapid(const MjbModel* model, int body_id);
MJB_API float mjb_model_tendon_width(const MjbModel* model, int tendon_id);
MJB_API int mjb_model_hfield_adr(const MjbModel* model, int hfield_id);
MJB_API const float* mjb_model_eq_data(const MjbModel* model, int* n_out);
MJB_API const float* mjb_model_hfield_data(const MjbModel* model, int* n_out);

/* single-env data and simulation (:80-187) */
MJB_API MjbData* mjb_make_data(MjbModel* model);
MJB_API void mjb_free_data(MjbData* data);
MJB_API void mjb_reset_data(MjbModel* model, MjbData* data);
MJB_API void mjb_step(MjbModel* model, MjbData* data);
MJB_API void mjb_forward(MjbModel* model, MjbData* data);
MJB_API void mjb_step1(MjbModel* model, MjbData* data);
MJB_API void mjb_step2(MjbModel* model, MjbData* data);
MJB_API void mjb_kinematics(MjbModel* model, MjbData* data);
MJB_API void mjb_rne_post_constraint(MjbModel* model, MjbData* data);
MJB_API void mjb_set_qpos(MjbData* data, const float* qpos, int n);
MJB_API void mjb_set_qvel(MjbData* data, const float* qvel, int n);
MJB_API void mjb_set_ctrl(MjbData* data, const float* ctrl, int n);
MJB_API const float* mjb_get_qpos(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_qvel(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_ctrl(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_xpos(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_xquat(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_xipos(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_cvel(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_qfrc_actuator(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_subtree_com(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_cinert(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_cfrc_ext(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_geom_xpos(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_geom_xmat(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_sensordata(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_xaxis(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_site_xpos(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_site_xmat(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_actuator_length(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_actuator_velocity(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_actuator_force(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_mocap_pos(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_mocap_quat(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_ten_length(const MjbData* data, int* n_out);
MJB_API const float* mjb_get_wrap_xpos(const MjbData* data, int* n_out);
MJB_API const int* mjb_get_ten_wrapadr(const MjbData* data, int* n_out);
MJB_API const int* mjb_get_ten_wrapnum(const MjbData* data, int* n_out);
MJB_API const int* mjb_get_wrap_obj(const MjbData* data, int* n_out);
MJB_API void mjb_set_mocap_pos(MjbData* data, const float* pos, int n);
MJB_API void mjb_set_mocap_quat(MjbData* data, const float* quat, int n);
MJB_API const float* mjb_model_geom_pos(const MjbModel* model, int* n_out);
MJB_API const float* mjb_model_geom_quat(const MjbModel* model, int* n_out);
MJB_API void mjb_set_qpos_at(MjbData* data, int index, float value);
MJB_API void mjb_set_qvel_at(MjbData* data, int index, float value);
MJB_API void mjb_set_ctrl_at(MjbData* data, int index, float value);
MJB_API const float* mjb_get_xfrc_applied(const MjbData* data, int* n_out);
MJB_API void mjb_set_xfrc_applied(MjbData* data, const float* values, int n);
MJB_API int mjb_get_warning_count(const MjbData* data, int index);
MJB_API int mjb_save_last_xml(const MjbModel* model, const char* path, char* error_buf, int error_buf_size);
MJB_API void mjb_object_velocity(const MjbModel* model, const MjbData* data, int objtype, int objid, int flg_local,
                                 float* result6);
MJB_API void mjb_load_plugin_library(const char* path);
MJB_API void mjb_model_set_hfield_data(MjbModel* model, int offset, const float* values, int n);

/* batched simulation (:197-233). One mjb_batched_step is one physics step. */
MJB_API MjbBatchedSim* mjb_batched_create(MjbModel* model, const MjbBatchedConfig* config);
MJB_API void mjb_batched_free(MjbBatchedSim* sim);
MJB_API int mjb_batched_is_gpu(const MjbBatchedSim* sim);
MJB_API void mjb_batched_step(MjbBatchedSim* sim, const float* ctrl);
MJB_API void mjb_batched_reset(MjbBatchedSim* sim, const int* reset_mask);
MJB_API const float* mjb_batched_get_qpos(const MjbBatchedSim* sim, int* n_out);
MJB_API void mjb_batched_get_state(const MjbBatchedSim* sim, float* qpos_out, float* qvel_out, int* nq_out, int* nv_out);
MJB_API void mjb_batched_set_state(MjbBatchedSim* sim, const float* qpos, const float* qvel);
MJB_API const float* mjb_batched_get_qvel(const MjbBatchedSim* sim, int* n_out);
MJB_API const float* mjb_batched_get_xpos(const MjbBatchedSim* sim, int* n_out);
MJB_API const float* mjb_batched_get_xquat(const MjbBatchedSim* sim, int* n_out);
MJB_API const float* mjb_batched_get_subtree_com(const MjbBatchedSim* sim, int* n_out);
MJB_API const float* mjb_batched_get_cinert(const MjbBatchedSim* sim, int* n_out);
MJB_API const float* mjb_batched_get_cvel(const MjbBatchedSim* sim, int* n_out);
MJB_API const float* mjb_batched_get_qfrc_actuator(const MjbBatchedSim* sim, int* n_out);
MJB_API const float* mjb_batched_get_cfrc_ext(const MjbBatchedSim* sim, int* n_out);
MJB_API void mjb_batched_set_env_qpos(MjbBatchedSim* sim, int env_idx, const float* qpos, int nq);
MJB_API void mjb_batched_set_env_qvel(MjbBatchedSim* sim, int env_idx, const float* qvel, int nv);
MJB_API void mjb_batched_eval_state(const MjbBatchedSim* sim);

/* differentiable simulation (:240): -1 = not supported */
MJB_API int mjb_grad_step(MjbModel* model, MjbData* data, float* grad_out);

#ifdef __cplusplus
}
#endif

#endif /* MJB_H */
