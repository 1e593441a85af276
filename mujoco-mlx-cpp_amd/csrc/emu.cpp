// TEST-ONLY serial interpreter of the kernel source (kernels.h with WAVE == 1).
//
// It runs the exact stage sequence the gfx950 kernel runs, one work item at a
// time on the host, so the CPU test suite can check the kernel algorithm
// against the FP64 oracle without a GPU. It is built into its own library
// (libmjmlx_emu.so) that only tests load; the engine library libmjmlx.so never
// contains or calls it.
#define MJ_EMU 1
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hostmodel.h"
#include "kernels.h"

namespace {
struct EmuSim {
  mjh::Model model;
  mjh::Blob blob;
  mjk::DModel dm;
  mjk::Layout L;
  std::vector<mjh::LayoutField> fields;
  int B;
  std::vector<float> qpos, qvel, qacc_ws, ws, dump1, dump2, feat, pred, rand, sensordata, sens_prev;
  bool use_pred = false;
  mjh::Caps caps;
  int nrand = 0;
  std::vector<int> niter, status, ncon, nefc;
  int poison_mode = 0;
  unsigned poison = 0;
};
}  // namespace

extern "C" {

__attribute__((visibility("default"))) void* mjemu_create(const char* xml, int self_collision, int max_condim, int B,
                                                          int maxcon, int maxefc) {
  try {
    auto* s = new EmuSim();
    mjh::LoadOptions o;
    o.self_collision = self_collision;
    o.max_condim = max_condim;
    s->model = mjh::compile_mjcf(xml, o);
    std::vector<mjh::CollisionPair> pairs;
    s->blob = mjh::build_blob(s->model, &pairs);
    s->dm = s->blob.bind(s->blob.bytes.data());
    // maxcon / maxefc > 0 replace the contact cap and the row budget
    mjh::Caps caps = mjh::default_caps(s->model, pairs, maxcon, maxefc);
    s->caps = caps;
    auto& m = s->model;
    s->L = mjk::make_layout(m.nq, m.nv, m.nu, m.nbody, m.njnt, m.ngeom, caps.maxcon, caps.maxefc,
                            m.solver == mjh::SOL_NEWTON, 0, s->dm.jw, s->dm.nM);
    s->L.efc_cap = caps.efc_cap;
    s->L.con_cap = caps.con_cap;
    s->fields = mjh::layout_fields(m, s->L);
    s->B = B;
    s->qpos.resize((size_t)B * m.nq);
    for (int e = 0; e < B; e++) for (int i = 0; i < m.nq; i++) s->qpos[(size_t)e * m.nq + i] = (float)m.qpos0[i];
    s->qvel.assign((size_t)B * m.nv, 0.f);
    s->qacc_ws.assign((size_t)B * m.nv, 0.f);
    s->ws.assign(s->L.total, 0.f);
    s->niter.assign(B, 0); s->status.assign(B, 0); s->ncon.assign(B, 0); s->nefc.assign(B, 0);
    s->sensordata.assign((size_t)B * std::max(1, m.nsensordata), 0.f);
    s->sens_prev.assign((size_t)B * (3 * s->dm.nsens_acc + 1), 0.f);
    return s;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjemu_create: %s\n", e.what());
    return nullptr;
  }
}

// Domain randomisation: per-env parameter blocks [B][rand_size] in the HIP
// engine's order (mass scales, COM shifts, friction scale, force scales, free
// joint drag); NULL switches it off. Returns the block size.
__attribute__((visibility("default"))) int mjemu_set_rand(void* p, const float* blocks) {
  auto* s = static_cast<EmuSim*>(p);
  auto& m = s->model;
  int nfree = 0;
  for (int j = 0; j < m.njnt; j++) nfree += m.jnt_type[j] == mjk::JFREE;
  s->nrand = blocks ? mjk::rand_size(m.nbody, m.nu, nfree) : 0;
  s->rand.assign(blocks, blocks ? blocks + (size_t)s->B * s->nrand : blocks);
  s->L = mjk::make_layout(m.nq, m.nv, m.nu, m.nbody, m.njnt, m.ngeom, s->caps.maxcon, s->caps.maxefc,
                          m.solver == mjh::SOL_NEWTON, s->nrand, s->dm.jw, s->dm.nM);
  s->L.efc_cap = s->caps.efc_cap;
  s->L.con_cap = s->caps.con_cap;
  s->fields = mjh::layout_fields(m, s->L);
  s->ws.assign(s->L.total, 0.f);
  return s->nrand;
}

__attribute__((visibility("default"))) void mjemu_free(void* p) { delete static_cast<EmuSim*>(p); }

__attribute__((visibility("default"))) void mjemu_set_state(void* p, const float* qpos, const float* qvel, const float* qacc_ws) {
  auto* s = static_cast<EmuSim*>(p);
  if (qpos) std::memcpy(s->qpos.data(), qpos, s->qpos.size() * sizeof(float));
  if (qvel) std::memcpy(s->qvel.data(), qvel, s->qvel.size() * sizeof(float));
  if (qacc_ws) std::memcpy(s->qacc_ws.data(), qacc_ws, s->qacc_ws.size() * sizeof(float));
}

__attribute__((visibility("default"))) void mjemu_get_state(void* p, float* qpos, float* qvel, float* qacc_ws) {
  auto* s = static_cast<EmuSim*>(p);
  if (qpos) std::memcpy(qpos, s->qpos.data(), s->qpos.size() * sizeof(float));
  if (qvel) std::memcpy(qvel, s->qvel.data(), s->qvel.size() * sizeof(float));
  if (qacc_ws) std::memcpy(qacc_ws, s->qacc_ws.data(), s->qacc_ws.size() * sizeof(float));
}

__attribute__((visibility("default"))) int mjemu_step(void* p, const float* ctrl, int nsub, int dump) {
  auto* s = static_cast<EmuSim*>(p);
  mjk::StepIO io{};
  io.qpos = s->qpos.data(); io.qvel = s->qvel.data(); io.qacc_ws = s->qacc_ws.data(); io.ctrl = ctrl;
  io.niter = s->niter.data(); io.status = s->status.data(); io.ncon = s->ncon.data(); io.nefc = s->nefc.data();
  s->feat.assign((size_t)s->B * s->dm.nfeat, 0.f);
  io.feat = s->feat.data();
  if (s->nrand > 0) { io.rand = s->rand.data(); io.nrand = s->nrand; }
  io.sensordata = s->sensordata.data(); io.sens_prev = s->sens_prev.data();
  if (s->use_pred) io.qacc_pred = s->pred.data();
  if (dump) {
    s->dump1.assign((size_t)s->B * s->L.total, 0.f);
    s->dump2.assign((size_t)s->B * s->L.total, 0.f);
    io.dump1 = s->dump1.data(); io.dump2 = s->dump2.data();
  }
  int bad = 0;
  for (int e = 0; e < s->B; e++) {
    // poison 0: zeroed workspace per env; 1: the bit pattern s->poison per
    // env; 2: the previous env's workspace left as it is (one workspace
    // serving successive envs, as a persistent launch does)
    if (s->poison_mode == 0) std::fill(s->ws.begin(), s->ws.end(), 0.f);
    else if (s->poison_mode == 1 || e == 0) {
      float f;
      std::memcpy(&f, &s->poison, sizeof(f));
      std::fill(s->ws.begin(), s->ws.end(), f);
    }
    mjk::run_env<6>(s->dm, s->L, s->ws.data(), io, e, nsub);
    if (s->status[e]) bad++;
  }
  return bad;
}

// workspace initialisation before each env (0 zeros, 1 the given bit
// pattern, 2 the pattern once, then each env reuses the previous env's
// workspace): results must not depend on it
__attribute__((visibility("default"))) void mjemu_set_poison(void* p, int mode, unsigned pattern) {
  auto* s = static_cast<EmuSim*>(p);
  s->poison_mode = mode;
  s->poison = pattern;
}

// learned warm-start features written by the last step, float[B * nfeat]; returns nfeat
__attribute__((visibility("default"))) int mjemu_features(void* p, float* out) {
  auto* s = static_cast<EmuSim*>(p);
  if (out && !s->feat.empty()) std::memcpy(out, s->feat.data(), s->feat.size() * sizeof(float));
  return s->dm.nfeat;
}

// predicted qacc offered to the next steps' solver as a warm-start candidate (null: none)
__attribute__((visibility("default"))) void mjemu_set_pred(void* p, const float* pred) {
  auto* s = static_cast<EmuSim*>(p);
  s->use_pred = pred != nullptr;
  if (pred) s->pred.assign(pred, pred + (size_t)s->B * s->model.nv);
}

__attribute__((visibility("default"))) int mjemu_layout(void* p, const char* name, int* off, int* len, int* kind) {
  auto* s = static_cast<EmuSim*>(p);
  if (std::strcmp(name, "total") == 0) { *off = 0; *len = s->L.total; *kind = 0; return 0; }
  for (auto& f : s->fields)
    if (std::strcmp(f.name, name) == 0) { *off = f.off; *len = f.len; *kind = f.kind; return 0; }
  return -1;
}

__attribute__((visibility("default"))) const float* mjemu_dump(void* p, int which) {
  auto* s = static_cast<EmuSim*>(p);
  return which == 1 ? s->dump1.data() : s->dump2.data();
}

__attribute__((visibility("default"))) void mjemu_stats(void* p, int* niter, int* status, int* ncon, int* nefc) {
  auto* s = static_cast<EmuSim*>(p);
  if (niter) std::memcpy(niter, s->niter.data(), s->B * sizeof(int));
  if (status) std::memcpy(status, s->status.data(), s->B * sizeof(int));
  if (ncon) std::memcpy(ncon, s->ncon.data(), s->B * sizeof(int));
  if (nefc) std::memcpy(nefc, s->nefc.data(), s->B * sizeof(int));
}

// The kernel's narrow phase for one geom pair (geom types as in the model,
// the lower type first); out holds up to 8 contacts of (dist, pos[3], n[3]).
__attribute__((visibility("default"))) int mjemu_narrow(int t1, int t2, const float* s1, const float* s2,
                                                       const float* x1, const float* R1, const float* x2,
                                                       const float* R2, float margin, float* out) {
  mjk::RawC rc[8];
  int n = mjk::k_narrow<2>(t1, t2, s1, s2, x1, R1, x2, R2, margin, rc);
  for (int k = 0; k < n && k < 8; k++) {
    out[7 * k] = rc[k].dist;
    for (int q = 0; q < 3; q++) { out[7 * k + 1 + q] = rc[k].pos[q]; out[7 * k + 4 + q] = rc[k].n[q]; }
  }
  return n;
}

// sensordata of the last step [B][nsensordata]; returns nsensordata
__attribute__((visibility("default"))) int mjemu_sensordata(void* p, float* out) {
  auto* s = static_cast<EmuSim*>(p);
  int nsd = s->model.nsensordata;
  if (out && nsd > 0) std::memcpy(out, s->sensordata.data(), (size_t)s->B * nsd * sizeof(float));
  return nsd;
}

// the kernel's ray-primitive intersection (ray_geom) for one ray in the geom frame
__attribute__((visibility("default"))) float mjemu_ray_geom(int type, const float* size, const float* p, const float* v) {
  return mjk::ray_geom(type, size, p, v);
}

// the kernel's per-env ray routine: rays [B][nray][6] at the sim's current
// qpos, geom filter gmask [ngeom]; dist / gid [B][nray]
__attribute__((visibility("default"))) void mjemu_rays(void* p, const float* rays, int nray, const int* gmask,
                                                       float* dist, int* gid) {
  auto* s = static_cast<EmuSim*>(p);
  for (int e = 0; e < s->B; e++) {
    std::fill(s->ws.begin(), s->ws.end(), 0.f);
    mjk::run_rays(s->dm, s->L, s->ws.data(), s->qpos.data(), s->rand.data(), s->nrand, rays, nray, gmask, dist, gid, e);
  }
}

}  // extern "C"

extern "C" __attribute__((visibility("default"))) int mjemu_model_field(void* p, const char* name, double* out, int cap) {
  return mjh::model_field(static_cast<EmuSim*>(p)->model, name, out, cap);
}

// compile only (no kernel tables): model fields of any MJCF the compiler accepts
extern "C" __attribute__((visibility("default"))) void* mjemu_compile(const char* xml, int self_collision, int max_condim) {
  try {
    auto* s = new EmuSim();
    mjh::LoadOptions o;
    o.self_collision = self_collision;
    o.max_condim = max_condim;
    s->model = mjh::compile_mjcf(xml, o);
    s->B = 0;
    return s;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "mjemu_compile: %s\n", e.what());
    return nullptr;
  }
}
