// mjb unified C API over the HIP engine (libmjb.so links libmjmlx.so).
//
// Reference src/mjb.cpp routes MJB_BACKEND_MLX to libmjmlx (mjb.cpp:744-793,
// 841) and MJB_BACKEND_CPU to MuJoCo C. Only the first exists here; the CPU
// backend reports that MuJoCo C is unavailable instead of silently stepping
// something else.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mjmlx/mjb.h"
#include "mjmlx/mjmlx_hip.h"

struct MjbBackend { MjbBackendType type; };
struct MjbModel { MjbBackendType type; MjmlxModel* mlx; std::vector<float> scratch; };
struct MjbData { int unused; };
struct MjbBatchedSim { MjbModel* model; MjmlxBatchedSim* mlx; };

namespace {
void cpu_unavailable(const char* fn) {
  std::fprintf(stderr, "%s error: MJB_BACKEND_CPU needs MuJoCo C (libmujoco 3.5.0), which is not part of the HIP engine\n", fn);
}
void single_env(const char* fn) {
  std::fprintf(stderr, "%s error: the single-env pipeline is not available; use the batched API\n", fn);
}
double field1(const MjbModel* m, const char* name, int idx, double def) {
  if (!m || !m->mlx || idx < 0) return def;
  int n = mjmlx_model_field(m->mlx, name, nullptr, 0);
  if (n <= idx) return def;
  std::vector<double> v(n);
  mjmlx_model_field(m->mlx, name, v.data(), n);
  return v[idx];
}
}  // namespace

extern "C" {

MJB_API MjbBackend* mjb_create_backend(MjbBackendType type) {
  if (type != MJB_BACKEND_CPU && type != MJB_BACKEND_MLX) return nullptr;
  return new MjbBackend{type};
}
MJB_API void mjb_free_backend(MjbBackend* b) { delete b; }
MJB_API MjbBackendType mjb_backend_type(const MjbBackend* b) { return b ? b->type : MJB_BACKEND_CPU; }

MJB_API MjbModel* mjb_load_model(MjbBackend* b, const char* path) {
  if (!b || !path) return nullptr;
  if (b->type == MJB_BACKEND_CPU) { cpu_unavailable("mjb_load_model"); return nullptr; }
  MjmlxModel* m = mjmlx_load_model(path);
  return m ? new MjbModel{b->type, m, {}} : nullptr;
}
MJB_API MjbModel* mjb_load_model_from_string(MjbBackend* b, const char* xml) {
  if (!b || !xml) return nullptr;
  if (b->type == MJB_BACKEND_CPU) { cpu_unavailable("mjb_load_model_from_string"); return nullptr; }
  MjmlxModel* m = mjmlx_load_model_from_string(xml);
  return m ? new MjbModel{b->type, m, {}} : nullptr;
}
MJB_API void mjb_free_model(MjbModel* m) {
  if (!m) return;
  mjmlx_free_model(m->mlx);
  delete m;
}

MJB_API MjbModelInfo mjb_model_info(const MjbModel* m) {
  MjbModelInfo i;
  std::memset(&i, 0, sizeof(i));
  if (!m) return i;
  MjmlxModelInfo x = mjmlx_model_info(m->mlx);
  i.nq = x.nq; i.nv = x.nv; i.nu = x.nu; i.nbody = x.nbody; i.njnt = x.njnt; i.ngeom = x.ngeom; i.nsite = x.nsite;
  return i;
}
MJB_API float mjb_model_opt_timestep(const MjbModel* m) { return m ? mjmlx_model_opt_timestep(m->mlx) : 0.f; }
MJB_API void mjb_model_set_opt_timestep(MjbModel* m, float dt) { if (m) mjmlx_model_set_opt_timestep(m->mlx, dt); }
MJB_API float mjb_model_body_mass(const MjbModel* m, int id) { return m ? mjmlx_model_body_mass(m->mlx, id) : 0.f; }
MJB_API int mjb_name2id(const MjbModel* m, int t, const char* name) { return m ? mjmlx_name2id(m->mlx, t, name) : -1; }
MJB_API const char* mjb_id2name(const MjbModel* m, int t, int id) {
  return m && m->mlx ? mjmlx_id2name(m->mlx, t, id) : nullptr;
}
MJB_API int mjb_model_jnt_qposadr(const MjbModel* m, int j) { return (int)field1(m, "jnt_qposadr", j, -1); }
MJB_API int mjb_model_jnt_dofadr(const MjbModel* m, int j) { return (int)field1(m, "jnt_dofadr", j, -1); }
MJB_API int mjb_model_jnt_type(const MjbModel* m, int j) { return (int)field1(m, "jnt_type", j, -1); }
MJB_API int mjb_model_nconmax(const MjbModel* m) { return m ? mjmlx_model_info(m->mlx).ncon : 0; }
MJB_API int mjb_model_geom_type(const MjbModel* m, int g) { return (int)field1(m, "geom_type", g, -1); }
MJB_API int mjb_model_sensor_adr(const MjbModel*, int) { return -1; }
MJB_API int mjb_model_body_mocapid(const MjbModel*, int) { return -1; }
MJB_API float mjb_model_tendon_width(const MjbModel*, int) { return 0.f; }
MJB_API int mjb_model_hfield_adr(const MjbModel*, int) { return -1; }
MJB_API const float* mjb_model_eq_data(const MjbModel*, int* n) { if (n) *n = 0; return nullptr; }
MJB_API const float* mjb_model_hfield_data(const MjbModel*, int* n) { if (n) *n = 0; return nullptr; }

MJB_API MjbData* mjb_make_data(MjbModel*) { single_env("mjb_make_data"); return nullptr; }
MJB_API void mjb_free_data(MjbData* d) { delete d; }
MJB_API void mjb_reset_data(MjbModel*, MjbData*) { single_env("mjb_reset_data"); }
MJB_API void mjb_step(MjbModel*, MjbData*) { single_env("mjb_step"); }
MJB_API void mjb_forward(MjbModel*, MjbData*) { single_env("mjb_forward"); }
MJB_API void mjb_step1(MjbModel*, MjbData*) { single_env("mjb_step1"); }
MJB_API void mjb_step2(MjbModel*, MjbData*) { single_env("mjb_step2"); }
MJB_API void mjb_kinematics(MjbModel*, MjbData*) { single_env("mjb_kinematics"); }
MJB_API void mjb_rne_post_constraint(MjbModel*, MjbData*) { single_env("mjb_rne_post_constraint"); }
MJB_API void mjb_set_qpos(MjbData*, const float*, int) { single_env("mjb_set_qpos"); }
MJB_API void mjb_set_qvel(MjbData*, const float*, int) { single_env("mjb_set_qvel"); }
MJB_API void mjb_set_ctrl(MjbData*, const float*, int) { single_env("mjb_set_ctrl"); }
#define MJB_NULL_GETTER(fn, T) \
  MJB_API const T* fn(const MjbData*, int* n_out) { if (n_out) *n_out = 0; single_env(#fn); return nullptr; }
MJB_NULL_GETTER(mjb_get_qpos, float)
MJB_NULL_GETTER(mjb_get_qvel, float)
MJB_NULL_GETTER(mjb_get_ctrl, float)
MJB_NULL_GETTER(mjb_get_xpos, float)
MJB_NULL_GETTER(mjb_get_xquat, float)
MJB_NULL_GETTER(mjb_get_xipos, float)
MJB_NULL_GETTER(mjb_get_cvel, float)
MJB_NULL_GETTER(mjb_get_qfrc_actuator, float)
MJB_NULL_GETTER(mjb_get_subtree_com, float)
MJB_NULL_GETTER(mjb_get_cinert, float)
MJB_NULL_GETTER(mjb_get_cfrc_ext, float)
MJB_NULL_GETTER(mjb_get_geom_xpos, float)
MJB_NULL_GETTER(mjb_get_geom_xmat, float)
MJB_NULL_GETTER(mjb_get_sensordata, float)
MJB_NULL_GETTER(mjb_get_xaxis, float)
MJB_NULL_GETTER(mjb_get_site_xpos, float)
MJB_NULL_GETTER(mjb_get_site_xmat, float)
MJB_NULL_GETTER(mjb_get_actuator_length, float)
MJB_NULL_GETTER(mjb_get_actuator_velocity, float)
MJB_NULL_GETTER(mjb_get_actuator_force, float)
MJB_NULL_GETTER(mjb_get_mocap_pos, float)
MJB_NULL_GETTER(mjb_get_mocap_quat, float)
MJB_NULL_GETTER(mjb_get_ten_length, float)
MJB_NULL_GETTER(mjb_get_wrap_xpos, float)
MJB_NULL_GETTER(mjb_get_ten_wrapadr, int)
MJB_NULL_GETTER(mjb_get_ten_wrapnum, int)
MJB_NULL_GETTER(mjb_get_wrap_obj, int)
MJB_NULL_GETTER(mjb_get_xfrc_applied, float)
#undef MJB_NULL_GETTER
MJB_API void mjb_set_mocap_pos(MjbData*, const float*, int) { single_env("mjb_set_mocap_pos"); }
MJB_API void mjb_set_mocap_quat(MjbData*, const float*, int) { single_env("mjb_set_mocap_quat"); }
MJB_API const float* mjb_model_geom_pos(const MjbModel* cm, int* n_out) {
  auto* m = const_cast<MjbModel*>(cm);
  if (!m) return nullptr;
  int n = mjmlx_model_field(m->mlx, "geom_pos", nullptr, 0);
  std::vector<double> v(n > 0 ? n : 1);
  mjmlx_model_field(m->mlx, "geom_pos", v.data(), n);
  m->scratch.assign(v.begin(), v.begin() + (n > 0 ? n : 0));
  if (n_out) *n_out = n;
  return m->scratch.data();
}
MJB_API const float* mjb_model_geom_quat(const MjbModel* cm, int* n_out) {
  auto* m = const_cast<MjbModel*>(cm);
  if (!m) return nullptr;
  int n = mjmlx_model_field(m->mlx, "geom_quat", nullptr, 0);
  std::vector<double> v(n > 0 ? n : 1);
  mjmlx_model_field(m->mlx, "geom_quat", v.data(), n);
  m->scratch.assign(v.begin(), v.begin() + (n > 0 ? n : 0));
  if (n_out) *n_out = n;
  return m->scratch.data();
}
MJB_API void mjb_set_qpos_at(MjbData*, int, float) { single_env("mjb_set_qpos_at"); }
MJB_API void mjb_set_qvel_at(MjbData*, int, float) { single_env("mjb_set_qvel_at"); }
MJB_API void mjb_set_ctrl_at(MjbData*, int, float) { single_env("mjb_set_ctrl_at"); }
MJB_API void mjb_set_xfrc_applied(MjbData*, const float*, int) { single_env("mjb_set_xfrc_applied"); }
MJB_API int mjb_get_warning_count(const MjbData*, int) { return 0; }
MJB_API int mjb_save_last_xml(const MjbModel*, const char*, char* err, int n) {
  if (err && n > 0) std::snprintf(err, n, "saving XML is not supported by the HIP engine");
  return -1;
}
MJB_API void mjb_object_velocity(const MjbModel*, const MjbData*, int, int, int, float* r) {
  if (r) for (int i = 0; i < 6; i++) r[i] = 0.f;
  single_env("mjb_object_velocity");
}
MJB_API void mjb_load_plugin_library(const char*) {}
MJB_API void mjb_model_set_hfield_data(MjbModel*, int, const float*, int) {}

MJB_API MjbBatchedSim* mjb_batched_create(MjbModel* m, const MjbBatchedConfig* cfg) {
  if (!m || !cfg || cfg->num_envs <= 0) return nullptr;
  if (m->type == MJB_BACKEND_CPU) { cpu_unavailable("mjb_batched_create"); return nullptr; }
  MjmlxBatchedConfig c;
  std::memset(&c, 0, sizeof(c));
  c.num_envs = cfg->num_envs;
  c.integrator = MJMLX_INTEGRATOR_EULER;  // reference src/mjb.cpp:762-766
  c.use_gpu = 1;
  c.solver_iterations = cfg->solver_iterations;
  MjmlxBatchedSim* s = mjmlx_batched_create(m->mlx, &c);
  if (!s) return nullptr;  // no silent CPU fallback: the CPU backend is unavailable
  return new MjbBatchedSim{m, s};
}
MJB_API void mjb_batched_free(MjbBatchedSim* s) {
  if (!s) return;
  mjmlx_batched_free(s->mlx);
  delete s;
}
MJB_API int mjb_batched_is_gpu(const MjbBatchedSim* s) { return s ? 1 : 0; }
MJB_API void mjb_batched_step(MjbBatchedSim* s, const float* ctrl) { if (s) mjmlx_batched_step(s->mlx, ctrl); }
MJB_API void mjb_batched_reset(MjbBatchedSim* s, const int* mask) { if (s) mjmlx_batched_reset(s->mlx, mask); }
MJB_API const float* mjb_batched_get_qpos(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_qpos(s->mlx, n) : nullptr; }
MJB_API void mjb_batched_get_state(const MjbBatchedSim* s, float* qp, float* qv, int* nq, int* nv) {
  if (s) mjmlx_batched_get_state(s->mlx, qp, qv, nq, nv);
}
MJB_API void mjb_batched_set_state(MjbBatchedSim* s, const float* qp, const float* qv) { if (s) mjmlx_batched_set_state(s->mlx, qp, qv); }
MJB_API const float* mjb_batched_get_qvel(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_qvel(s->mlx, n) : nullptr; }
MJB_API const float* mjb_batched_get_xpos(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_xpos(s->mlx, n) : nullptr; }
MJB_API const float* mjb_batched_get_xquat(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_xquat(s->mlx, n) : nullptr; }
MJB_API const float* mjb_batched_get_subtree_com(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_subtree_com(s->mlx, n) : nullptr; }
MJB_API const float* mjb_batched_get_cinert(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_cinert(s->mlx, n) : nullptr; }
MJB_API const float* mjb_batched_get_cvel(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_cvel(s->mlx, n) : nullptr; }
MJB_API const float* mjb_batched_get_qfrc_actuator(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_qfrc_actuator(s->mlx, n) : nullptr; }
MJB_API const float* mjb_batched_get_cfrc_ext(const MjbBatchedSim* s, int* n) { return s ? mjmlx_batched_get_cfrc_ext(s->mlx, n) : nullptr; }
MJB_API void mjb_batched_set_env_qpos(MjbBatchedSim* s, int e, const float* q, int nq) { if (s) mjmlx_batched_set_env_qpos(s->mlx, e, q, nq); }
MJB_API void mjb_batched_set_env_qvel(MjbBatchedSim* s, int e, const float* v, int nv) { if (s) mjmlx_batched_set_env_qvel(s->mlx, e, v, nv); }
MJB_API void mjb_batched_eval_state(const MjbBatchedSim* s) { if (s) mjmlx_batched_eval_state(s->mlx); }
MJB_API int mjb_grad_step(MjbModel*, MjbData*, float*) { return -1; }

}  // extern "C"
