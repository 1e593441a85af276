// Kernel tables from a compiled model. Static derived data mirrors what the
// reference precomputes on the host: the collision pair table
// (reference src/batched.cpp:317-367), per-pair solver properties (:529-548),
// limit/friction tables (:378-409,458-478) and per-body dof chains (:551-561).
#include "hostmodel.h"

#include <cstdlib>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <stdexcept>

namespace mjh {

namespace {

struct Builder {
  std::vector<char> bytes;
  template <class T>
  size_t add(const std::vector<T>& v) {
    size_t o = (bytes.size() + 15) & ~size_t(15);
    bytes.resize(o + std::max<size_t>(1, v.size()) * sizeof(T), 0);
    if (!v.empty()) std::memcpy(bytes.data() + o, v.data(), v.size() * sizeof(T));
    return o;
  }
};

std::vector<float> tof(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }

bool pair_supported(int t1, int t2) {
  using namespace mjk;
  bool cvx1 = t1 == GSPHERE || t1 == GCAPSULE || t1 == GELLIPSOID || t1 == GCYLINDER || t1 == GBOX;
  bool cvx2 = t2 == GSPHERE || t2 == GCAPSULE || t2 == GELLIPSOID || t2 == GCYLINDER || t2 == GBOX;
  if (t1 == GPLANE && (t2 == GELLIPSOID || t2 == GCYLINDER)) return true;
  if ((t1 == GELLIPSOID || t1 == GCYLINDER || t2 == GELLIPSOID || t2 == GCYLINDER) && cvx1 && cvx2) return true;
  if (t1 == GPLANE) return t2 == GSPHERE || t2 == GCAPSULE || t2 == GBOX;
  if (t1 == GSPHERE) return t2 == GSPHERE || t2 == GCAPSULE || t2 == GBOX;
  if (t1 == GCAPSULE) return t2 == GCAPSULE || t2 == GBOX;
  if (t1 == GBOX) return t2 == GBOX;
  return false;
}

int pair_maxc(int t1, int t2) {
  using namespace mjk;
  if (t1 == GPLANE && t2 == GCAPSULE) return 2;
  if (t1 == GPLANE && t2 == GBOX) return 4;
  if (t1 == GCAPSULE && t2 == GBOX) return 2;
  if (t1 == GBOX && t2 == GBOX) return 4;
  return 1;
}

}  // namespace

mjk::DModel Blob::bind(const char* base) const {
  mjk::DModel d = proto;
  char* p0 = reinterpret_cast<char*>(&d) + offsetof(mjk::DModel, body_parentid);
  char* p1 = reinterpret_cast<char*>(&d) + sizeof(mjk::DModel);
  for (char* p = p0; p + sizeof(uintptr_t) <= p1; p += sizeof(uintptr_t)) {
    uintptr_t v;
    std::memcpy(&v, p, sizeof(v));
    v += reinterpret_cast<uintptr_t>(base);
    std::memcpy(p, &v, sizeof(v));
  }
  return d;
}

int max_contacts(const Model& m, const std::vector<CollisionPair>& pairs) {
  int s = 0;
  for (auto& p : pairs) s += pair_maxc(m.geom_type[p.g1], m.geom_type[p.g2]);
  return s;
}

int max_rows(const Model& m, const std::vector<CollisionPair>& pairs) {
  int s = 0;
  for (int k = 0; k < m.nv; k++) if (m.dof_frictionloss[k] > 0) s++;
  for (int j = 0; j < m.njnt; j++) if (m.jnt_limited[j]) s += 2;
  for (auto& p : pairs) {
    int per = p.condim == 1 ? 1 : (m.cone == CONE_PYRAMIDAL ? 2 * (p.condim - 1) : p.condim);
    s += per * pair_maxc(m.geom_type[p.g1], m.geom_type[p.g2]);
  }
  return s;
}

Caps default_caps(const Model& m, const std::vector<CollisionPair>& pairs, int maxcon_override,
                  int efc_cap_override) {
  Caps c;
  bool budget_env = false;
  if (const char* v = std::getenv("MJMLX_MAX_EFC")) budget_env = std::atoi(v) > 0;
  bool ref_caps = false;
  if (const char* v = std::getenv("MJMLX_HIP_REF_CAPS")) ref_caps = std::atoi(v) != 0;
  int nfree = 0;
  for (int j = 0; j < m.njnt; j++) nfree += m.jnt_type[j] == mjk::JFREE;
  if ((m.nv > 32 || nfree > 1) && efc_cap_override <= 0 && !budget_env && !ref_caps) {
    // Models of more than 32 dofs or of several free bodies (contacts grow
    // with the bodies): MuJoCo's semantics, no row budget. Every stored
    // contact gets its rows; up to min(512, the pairs' contact maxima)
    // contacts are stored and counted. (The reference caps these at 128
    // contacts and a 72 / 256-row budget, reference src/batched.cpp:290,
    // src/compat/mkx_kernels/solver.hpp:29-33, which leaves contacts of
    // landed box piles without force.)
    int mc = max_contacts(m, pairs);
    c.con_cap = maxcon_override > 0 ? maxcon_override : std::max(1, std::min(mc, 512));
    c.maxcon = std::max(1, std::min(mc, c.con_cap));
    int rmax = 1, fixed = 0;
    for (auto& p : pairs) {
      int per = p.condim == 1 ? 1 : (m.cone == CONE_PYRAMIDAL ? 2 * (p.condim - 1) : p.condim);
      rmax = std::max(rmax, per);
    }
    for (int k = 0; k < m.nv; k++) if (m.dof_frictionloss[k] > 0) fixed++;
    for (int j = 0; j < m.njnt; j++) if (m.jnt_limited[j]) fixed += 2;
    c.maxefc = std::max(1, std::min(max_rows(m, pairs), fixed + rmax * c.maxcon));
    c.efc_cap = c.maxefc + 10;  // no admission test binds before the buffer ends
    return c;
  }
  // The reference's capacities (reference src/batched.cpp:290 and
  // src/compat/mkx_kernels/solver.hpp:29-33,299,349,403): 128 contacts counted,
  // a row budget of 72 rows for nv <= 32 and 256 beyond
  int con_cap = maxcon_override > 0 ? maxcon_override : 128;
  c.efc_cap = efc_cap_override > 0 ? efc_cap_override : (m.nv <= 32 ? 72 : 256);
  if (const char* v = std::getenv("MJMLX_MAX_EFC")) if (efc_cap_override <= 0 && std::atoi(v) > 0) c.efc_cap = std::atoi(v);
  c.con_cap = con_cap;
  c.maxcon = std::max(1, std::min(max_contacts(m, pairs), con_cap));
  // admitted rows never exceed efc_cap - 1 (a contact starts below efc_cap - 10
  // and adds at most 10 rows)
  c.maxefc = std::max(1, std::min(max_rows(m, pairs), c.efc_cap));
  // when every stored contact is active (no pair with a gap), contacts get
  // rows in list order until efc_cap - 10 rows exist: with r rows per contact
  // at least, no contact past the first ceil((efc_cap - 10) / r) gets rows
  bool gap = false;
  int rmin = 10;
  for (auto& p : pairs) {
    gap = gap || p.gap > 0;
    int per = p.condim == 1 ? 1 : (m.cone == CONE_PYRAMIDAL ? 2 * (p.condim - 1) : p.condim);
    rmin = std::min(rmin, per);
  }
  if (!gap && rmin > 0) c.maxcon = std::max(1, std::min(c.maxcon, (std::max(c.efc_cap - 10, 0) + rmin - 1) / rmin));
  return c;
}

Blob build_blob(const Model& m, std::vector<CollisionPair>* pairs_out) {
  using namespace mjk;
  if (m.nv > 512) throw std::runtime_error("HIP engine: models with nv > 512 are not supported by the dense-row kernel");
  if (m.solver == SOL_PGS) throw std::runtime_error("HIP engine: the PGS solver is not implemented (use Newton or CG)");
  if (m.integrator == INT_RK4) throw std::runtime_error("HIP engine: the RK4 integrator is not implemented");
  std::vector<CollisionPair> pairs = build_pairs(m);
  for (auto& p : pairs)
    if (!pair_supported(m.geom_type[p.g1], m.geom_type[p.g2]))
      throw std::runtime_error("HIP engine: geom pair types " + std::to_string(m.geom_type[p.g1]) + "/" +
                               std::to_string(m.geom_type[p.g2]) + " have no narrow phase yet");
  for (int j = 0; j < m.njnt; j++)
    if (m.jnt_limited[j] && (m.jnt_type[j] == JNT_BALL)) throw std::runtime_error("HIP engine: ball joint limits are not implemented");
  int nb = m.nbody, nv = m.nv;
  // subtree ranges (depth-first numbering)
  std::vector<int> send(nb);
  for (int b = 0; b < nb; b++) send[b] = b + 1;
  for (int b = nb - 1; b > 0; b--) send[m.body_parentid[b]] = std::max(send[m.body_parentid[b]], send[b]);
  for (int b = 1; b < nb; b++)
    if (m.body_parentid[b] >= b) throw std::runtime_error("HIP engine: bodies must be numbered parent-first");
  // tree levels (depth 1..)
  std::vector<int> depth(nb, 0);
  int maxd = 0;
  for (int b = 1; b < nb; b++) { depth[b] = depth[m.body_parentid[b]] + 1; maxd = std::max(maxd, depth[b]); }
  std::vector<int> level_adr(maxd + 1, 0), level_body;
  for (int lv = 1; lv <= maxd; lv++) {
    level_adr[lv - 1] = (int)level_body.size();
    for (int b = 1; b < nb; b++) if (depth[b] == lv) level_body.push_back(b);
  }
  level_adr[maxd] = (int)level_body.size();
  // last dof of each body's kinematic chain (-1: none)
  std::vector<int> lastdof(nb, -1);
  for (int b = 1; b < nb; b++) {
    int bb = b;
    while (bb > 0 && m.body_dofnum[bb] == 0) bb = m.body_parentid[bb];
    if (bb > 0) lastdof[b] = m.body_dofadr[bb] + m.body_dofnum[bb] - 1;
  }
  // dof tree
  std::vector<int> ddepth(nv, 0), dsend(nv), nchain(nv), chain((size_t)std::max(nv, 1) * MAX_DOF_DEPTH, -1);
  int maxdd = 0;
  for (int k = 0; k < nv; k++) {
    int p = m.dof_parentid[k];
    if (p >= k) throw std::runtime_error("HIP engine: dof parents must precede their children");
    ddepth[k] = p < 0 ? 0 : ddepth[p] + 1;
    maxdd = std::max(maxdd, ddepth[k]);
    int n = 0;
    for (int a = k; a >= 0; a = m.dof_parentid[a]) {
      if (n >= MAX_DOF_DEPTH) throw std::runtime_error("HIP engine: dof tree deeper than 64");
      chain[(size_t)k * MAX_DOF_DEPTH + n++] = a;
    }
    nchain[k] = n;
  }
  // qM storage: dense up to COMPACT_M_NV dofs, else compact rows
  std::vector<int> madr(std::max(nv, 1), 0);
  int nM = nv * nv;
  if (nv > COMPACT_M_NV) {
    nM = 0;
    for (int k = 0; k < nv; k++) { madr[k] = nM; nM += nchain[k]; }
  } else {
    for (int k = 0; k < nv; k++) madr[k] = k * nv;
  }
  for (int k = 0; k < nv; k++) dsend[k] = k + 1;
  for (int k = nv - 1; k >= 0; k--) { int p = m.dof_parentid[k]; if (p >= 0) dsend[p] = std::max(dsend[p], dsend[k]); }
  // dof trees (a root dof's subtree, a contiguous range) that are all paths of
  // at most PATH_TREE_MAX dofs
  std::vector<int> tree_adr;
  {
    bool paths = nv > 0;
    for (int k = 0; k < nv && paths; k++) {
      int p = m.dof_parentid[k];
      if (p < 0) {
        tree_adr.push_back(k);
        paths = dsend[k] - k <= PATH_TREE_MAX;
      } else {
        paths = p == k - 1;
      }
    }
    if (paths) tree_adr.push_back(nv);
    else tree_adr.assign(1, 0);
  }
  for (int k = 0; k < nv; k++)
    for (int q = k + 1; q < dsend[k]; q++) {
      bool desc = false;
      for (int a = q; a >= 0; a = m.dof_parentid[a]) if (a == k) { desc = true; break; }
      if (!desc) throw std::runtime_error("HIP engine: dof subtree is not contiguous");
    }
  // compressed constraint rows beyond SPARSE_ROWS_NV dofs: a contact row's
  // nonzeros are body2's dof chain then body1's up to the first shared dof
  int jw = 0;
  bool dense_rows = false;
  if (const char* v = std::getenv("MJMLX_HIP_DENSE_ROWS")) dense_rows = std::atoi(v) != 0;
  if (nv > SPARSE_ROWS_NV && !dense_rows) {
    jw = 1;
    auto on_chain = [&](int d, int last) { return last >= d && last < dsend[d]; };
    for (auto& p : pairs) {
      int l2 = lastdof[p.b2], l1 = lastdof[p.b1];
      int n = l2 >= 0 ? nchain[l2] : 0;
      for (int d = l1; d >= 0 && !(l2 >= 0 && on_chain(d, l2)); d = m.dof_parentid[d]) n++;
      jw = std::max(jw, n);
    }
  }
  int ndl = nv > 0 ? maxdd + 1 : 0;
  // level-parallel LDL of qM: for each dof level (deepest first, the root
  // level has no ancestors to update) the entries (i, j), i and j strict
  // ancestors of a dof k of the level with j in the chain of i, each with the
  // list of such k
  std::vector<int> flev_adr, flev_i, flev_j, flev_kadr, flev_kcnt, flev_k;
  for (int lv = ndl - 1; lv >= 1; lv--) {
    flev_adr.push_back((int)flev_i.size());
    std::vector<std::vector<int>> ks((size_t)nv * nv);
    for (int k = 0; k < nv; k++) {
      if (ddepth[k] != lv) continue;
      for (int i = m.dof_parentid[k]; i >= 0; i = m.dof_parentid[i])
        for (int j = i; j >= 0; j = m.dof_parentid[j]) ks[(size_t)i * nv + j].push_back(k);
    }
    for (int i = 0; i < nv; i++)
      for (int j = 0; j <= i; j++) {
        auto& v = ks[(size_t)i * nv + j];
        if (v.empty()) continue;
        flev_i.push_back(i); flev_j.push_back(j);
        flev_kadr.push_back((int)flev_k.size()); flev_kcnt.push_back((int)v.size());
        for (int k : v) flev_k.push_back(k);
      }
  }
  flev_adr.push_back((int)flev_i.size());
  std::vector<int> dl_adr(ndl + 1, 0), dl_dof;
  for (int lv = 0; lv < ndl; lv++) {
    dl_adr[lv] = (int)dl_dof.size();
    for (int k = 0; k < nv; k++) if (ddepth[k] == lv) dl_dof.push_back(k);
  }
  dl_adr[ndl] = (int)dl_dof.size();
  // actuators per dof
  std::vector<int> act_dof(m.nu), act_qadr(m.nu), dof_act_adr(nv, 0), dof_act_num(nv, 0), dof_act_list;
  for (int u = 0; u < m.nu; u++) { act_dof[u] = m.jnt_dofadr[m.actuator_trnid[u]]; act_qadr[u] = m.jnt_qposadr[m.actuator_trnid[u]]; }
  for (int k = 0; k < nv; k++) {
    dof_act_adr[k] = (int)dof_act_list.size();
    for (int u = 0; u < m.nu; u++) if (act_dof[u] == k) dof_act_list.push_back(u);
    dof_act_num[k] = (int)dof_act_list.size() - dof_act_adr[k];
  }
  std::vector<int> floss_dof, lim_jnt;
  for (int k = 0; k < nv; k++) if (m.dof_frictionloss[k] > 0) floss_dof.push_back(k);
  for (int j = 0; j < m.njnt; j++)
    if (m.jnt_limited[j] && (m.jnt_type[j] == JNT_HINGE || m.jnt_type[j] == JNT_SLIDE)) lim_jnt.push_back(j);
  // pairs
  int np = (int)pairs.size();
  std::vector<int> pg1(np), pg2(np), pb1(np), pb2(np), pcd(np);
  std::vector<float> pfr(5 * np), psr(2 * np), psi(5 * np), pmg(np), pgp(np);
  for (int i = 0; i < np; i++) {
    auto& p = pairs[i];
    pg1[i] = p.g1; pg2[i] = p.g2; pb1[i] = p.b1; pb2[i] = p.b2; pcd[i] = p.condim;
    for (int k = 0; k < 5; k++) { pfr[5 * i + k] = (float)p.friction[k]; psi[5 * i + k] = (float)p.solimp[k]; }
    for (int k = 0; k < 2; k++) psr[2 * i + k] = (float)p.solref[k];
    pmg[i] = (float)p.margin; pgp[i] = (float)p.gap;
  }
  // limb class of each geom for the learned warm-start features: the index
  // (mod FEAT_NCLASS) of the root body's child whose subtree holds the geom;
  // -1 on the world and on root bodies
  std::vector<int> fclass(m.ngeom, -1);
  {
    std::vector<int> limb(nb, -1);
    for (int b = 1; b < nb; b++) {
      int par = m.body_parentid[b];
      if (par == 0) continue;
      if (m.body_parentid[par] == 0) {
        int k = 0;
        for (int c = 1; c < b; c++) if (m.body_parentid[c] == par) k++;
        limb[b] = k % FEAT_NCLASS;
      } else {
        limb[b] = limb[par];
      }
    }
    for (int g = 0; g < m.ngeom; g++) fclass[g] = limb[m.geom_bodyid[g]];
  }
  // sensors: frame body and local pose (reference src/sensor.cpp:112-124:
  // site and geom frames, a body's inertial frame, an xbody's own frame)
  int ns = m.nsensor, nacc = 0;
  std::vector<int> s_body(ns, -1), s_qadr(ns, 0), s_dadr(ns, 0), s_acc(ns, -1);
  std::vector<float> s_pos(3 * ns, 0.f), s_quat(4 * ns, 0.f), s_gear(ns, 1.f);
  for (int i = 0; i < ns; i++) {
    int t = m.sensor_type[i], ot = m.sensor_objtype[i], id = m.sensor_objid[i];
    s_quat[4 * i] = 1.f;
    auto frame = [&](int body, const double* p, const double* q) {
      s_body[i] = body;
      for (int k = 0; k < 3; k++) s_pos[3 * i + k] = p ? (float)p[k] : 0.f;
      for (int k = 0; k < 4; k++) s_quat[4 * i + k] = q ? (float)q[k] : (k == 0 ? 1.f : 0.f);
    };
    if (ot == 6) frame(m.site_bodyid[id], &m.site_pos[3 * id], &m.site_quat[4 * id]);
    else if (ot == 5) frame(m.geom_bodyid[id], &m.geom_pos[3 * id], &m.geom_quat[4 * id]);
    else if (ot == 1 && t != SENS_SUBTREECOM) frame(id, &m.body_ipos[3 * id], &m.body_iquat[4 * id]);
    else if (ot == 1 || ot == 2) frame(id, nullptr, nullptr);
    else if (ot == 3) { s_qadr[i] = m.jnt_qposadr[id]; s_dadr[i] = m.jnt_dofadr[id]; }
    else if (ot == 19) {
      int j = m.actuator_trnid[id];
      s_qadr[i] = m.jnt_qposadr[j]; s_dadr[i] = m.jnt_dofadr[j]; s_gear[i] = (float)m.actuator_gear[id];
    }
    if (t == SENS_ACCELEROMETER) s_acc[i] = nacc++;
  }
  std::vector<float> gain(3 * m.nu), bias(3 * m.nu);
  for (int i = 0; i < 3 * m.nu; i++) { gain[i] = (float)m.actuator_gainprm[i]; bias[i] = (float)m.actuator_biasprm[i]; }

  Blob blob;
  Builder B;
  DModel& d = blob.proto;
  std::memset(&d, 0, sizeof(d));
  d.nM = nM;
  d.ntree = (int)tree_adr.size() - 1;
  d.nq = m.nq; d.nv = nv; d.nu = m.nu; d.nbody = nb; d.njnt = m.njnt; d.ngeom = m.ngeom; d.npair = np;
  d.nlevel = maxd; d.ndoflevel = ndl; d.nflev = (int)flev_adr.size() - 1; d.maxdepth = maxdd; d.nfloss = (int)floss_dof.size(); d.nlim = (int)lim_jnt.size();
  d.iterations = m.iterations; d.ls_iterations = m.ls_iterations; d.solver = m.solver; d.cone = m.cone;
  d.integrator = m.integrator; d.disableflags = m.disableflags;
  d.has_damping = 0;
  for (int k = 0; k < nv; k++) if (m.dof_damping[k] > 0) d.has_damping = 1;
  d.timestep = (float)m.timestep;
  for (int i = 0; i < 3; i++) d.gravity[i] = (float)m.gravity[i];
  d.tolerance = (float)m.tolerance; d.ls_tolerance = (float)m.ls_tolerance; d.impratio = (float)m.impratio;
  d.meaninertia = (float)m.meaninertia;
  d.nfeat = m.nq + nv + m.nu + FEAT_NCON * FEAT_PER_CON;
  d.nsensor = ns; d.nsensordata = m.nsensordata; d.nsens_acc = nacc;
  d.jw = jw;
  for (int i = 0; i < 3; i++) d.magnetic[i] = (float)m.magnetic[i];
  std::vector<float> subtreemass = tof(m.body_subtreemass);
#define ADDP(field, vec) d.field = reinterpret_cast<decltype(d.field)>(B.add(vec))
  ADDP(body_parentid, m.body_parentid); ADDP(body_rootid, m.body_rootid); ADDP(body_jntnum, m.body_jntnum);
  ADDP(body_jntadr, m.body_jntadr); ADDP(body_dofnum, m.body_dofnum); ADDP(body_dofadr, m.body_dofadr);
  ADDP(body_subtree_end, send); ADDP(level_adr, level_adr); ADDP(level_body, level_body); ADDP(body_lastdof, lastdof);
  ADDP(body_pos, tof(m.body_pos)); ADDP(body_quat, tof(m.body_quat)); ADDP(body_ipos, tof(m.body_ipos));
  ADDP(body_iquat, tof(m.body_iquat)); ADDP(body_mass, tof(m.body_mass)); ADDP(body_inertia, tof(m.body_inertia));
  ADDP(body_subtreemass, subtreemass); ADDP(body_invweight0, tof(m.body_invweight0));
  ADDP(jnt_type, m.jnt_type); ADDP(jnt_qposadr, m.jnt_qposadr); ADDP(jnt_dofadr, m.jnt_dofadr); ADDP(jnt_bodyid, m.jnt_bodyid);
  ADDP(jnt_pos, tof(m.jnt_pos)); ADDP(jnt_axis, tof(m.jnt_axis)); ADDP(jnt_range, tof(m.jnt_range));
  ADDP(jnt_margin, tof(m.jnt_margin)); ADDP(jnt_solref, tof(m.jnt_solref)); ADDP(jnt_solimp, tof(m.jnt_solimp));
  ADDP(jnt_stiffness, tof(m.jnt_stiffness)); ADDP(qpos0, tof(m.qpos0)); ADDP(qpos_spring, tof(m.qpos_spring));
  ADDP(dof_bodyid, m.dof_bodyid); ADDP(dof_jntid, m.dof_jntid); ADDP(dof_parentid, m.dof_parentid);
  ADDP(dof_depth, ddepth); ADDP(dof_subtree_end, dsend); ADDP(doflevel_adr, dl_adr); ADDP(doflevel_dof, dl_dof);
  ADDP(dof_chain, chain); ADDP(dof_nchain, nchain); ADDP(dof_act_adr, dof_act_adr); ADDP(dof_act_num, dof_act_num);
  ADDP(dof_act_list, dof_act_list); ADDP(dof_madr, madr); ADDP(tree_adr, tree_adr);
  ADDP(flev_adr, flev_adr); ADDP(flev_i, flev_i); ADDP(flev_j, flev_j); ADDP(flev_kadr, flev_kadr);
  ADDP(flev_kcnt, flev_kcnt); ADDP(flev_k, flev_k);
  ADDP(dof_armature, tof(m.dof_armature)); ADDP(dof_damping, tof(m.dof_damping)); ADDP(dof_frictionloss, tof(m.dof_frictionloss));
  ADDP(dof_invweight0, tof(m.dof_invweight0)); ADDP(dof_solref, tof(m.dof_solref)); ADDP(dof_solimp, tof(m.dof_solimp));
  ADDP(geom_type, m.geom_type); ADDP(geom_bodyid, m.geom_bodyid); ADDP(geom_fclass, fclass);
  ADDP(geom_size, tof(m.geom_size)); ADDP(geom_pos, tof(m.geom_pos)); ADDP(geom_quat, tof(m.geom_quat));
  ADDP(geom_rbound, tof(m.geom_rbound));
  ADDP(pair_g1, pg1); ADDP(pair_g2, pg2); ADDP(pair_b1, pb1); ADDP(pair_b2, pb2); ADDP(pair_condim, pcd);
  ADDP(pair_friction, pfr); ADDP(pair_solref, psr); ADDP(pair_solimp, psi); ADDP(pair_margin, pmg); ADDP(pair_gap, pgp);
  ADDP(act_dof, act_dof); ADDP(act_qadr, act_qadr); ADDP(act_ctrllimited, m.actuator_ctrllimited);
  ADDP(act_forcelimited, m.actuator_forcelimited);
  ADDP(act_gear, tof(m.actuator_gear)); ADDP(act_gain, gain); ADDP(act_bias, bias);
  ADDP(act_ctrlrange, tof(m.actuator_ctrlrange)); ADDP(act_forcerange, tof(m.actuator_forcerange));
  ADDP(floss_dof, floss_dof); ADDP(lim_jnt, lim_jnt);
  ADDP(sens_type, m.sensor_type); ADDP(sens_adr, m.sensor_adr); ADDP(sens_dim, m.sensor_dim); ADDP(sens_body, s_body);
  ADDP(sens_qadr, s_qadr); ADDP(sens_dadr, s_dadr); ADDP(sens_acc, s_acc);
  ADDP(sens_pos, s_pos); ADDP(sens_quat, s_quat); ADDP(sens_gear, s_gear);
#undef ADDP
  blob.bytes = std::move(B.bytes);
  if (pairs_out) *pairs_out = pairs;
  return blob;
}

std::vector<LayoutField> layout_fields(const Model& m, const mjk::Layout& L) {
  int nb = m.nbody, nv = m.nv, nj = m.njnt, ng = m.ngeom;
  std::vector<LayoutField> f = {
      {"qpos", L.qpos, m.nq, 0}, {"qvel", L.qvel, nv, 0}, {"ctrl", L.ctrl, m.nu, 0},
      {"qacc_warmstart", L.qacc_ws, nv, 0}, {"qacc", L.qacc, nv, 0}, {"qM", L.qM, L.nM, 0},
      {"qLD", L.qLD, L.nM, 0}, {"qfrc_smooth", L.qfrc_smooth, nv, 0}, {"qacc_smooth", L.qacc_smooth, nv, 0},
      {"qfrc_constraint", L.qfrc_constraint, nv, 0}, {"actuator_force", L.act_force, m.nu, 0},
      {"misc", L.misc, mjk::MISC_SIZE, 1},
      {"xpos", L.xpos, 3 * nb, 0}, {"xquat", L.xquat, 4 * nb, 0}, {"xmat", L.xmat, 9 * nb, 0},
      {"xipos", L.xipos, 3 * nb, 0}, {"xanchor", L.xanchor, 3 * nj, 0}, {"xaxis", L.xaxis, 3 * nj, 0},
      {"geom_xpos", L.gxpos, 3 * ng, 0}, {"geom_xmat", L.gxmat, 9 * ng, 0}, {"subtree_com", L.scom, 3 * nb, 0},
      {"cinert", L.cinert, 10 * nb, 0}, {"cdof", L.cdof, 6 * nv, 0}, {"cvel", L.cvel, 6 * nb, 0},
      {"cdof_dot", L.cdofdot, 6 * nv, 0}, {"crb", L.crb, 10 * nb, 0},
      {"contact", L.con, mjk::CON_STRIDE * L.maxcon, 0},
      {"efc_J", L.efc_J, L.maxefc * (L.jw > 0 ? L.jw : nv), 0}, {"efc_D", L.efc_D, L.maxefc, 0},
      {"efc_aref", L.efc_aref, L.maxefc, 0}, {"efc_force", L.efc_force, L.maxefc, 0},
      {"efc_pos", L.efc_pos, L.maxefc, 0}, {"efc_margin", L.efc_margin, L.maxefc, 0},
      {"efc_type", L.efc_type, L.maxefc, 1}, {"efc_id", L.efc_id, L.maxefc, 1},
  };
  if (L.rand >= 0) {  // domain randomisation: the env's parameter block
    int nfree = 0;
    for (int j = 0; j < nj; j++) nfree += m.jnt_type[j] == mjk::JFREE;
    f.push_back({"rand", L.rand, mjk::rand_size(nb, m.nu, nfree), 0});
  }
  return f;
}

}  // namespace mjh

namespace mjh {

// Raw model field by MuJoCo name (integers are returned as doubles). Scalars:
// nq nv nu nbody njnt ngeom timestep gravity tolerance ls_tolerance impratio
// iterations ls_iterations solver cone integrator disableflags meaninertia
// npair nexclude. Compiled pair table: cpair_* (g1 g2 b1 b2 condim friction
// solref solimp margin gap). Returns the length, or -1 for an unknown name.
int model_field(const Model& m, const char* name, double* out, int cap) {
  std::string n(name);
  auto put = [&](const std::vector<double>& v) {
    for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = v[i];
    return (int)v.size();
  };
  auto puti = [&](const std::vector<int>& v) {
    for (int i = 0; i < (int)v.size() && i < cap; i++) out[i] = v[i];
    return (int)v.size();
  };
  auto scal = [&](double x) { if (cap > 0) out[0] = x; return 1; };
  if (n == "nq") return scal(m.nq);
  if (n == "nv") return scal(m.nv);
  if (n == "nu") return scal(m.nu);
  if (n == "nbody") return scal(m.nbody);
  if (n == "njnt") return scal(m.njnt);
  if (n == "ngeom") return scal(m.ngeom);
  if (n == "timestep") return scal(m.timestep);
  if (n == "gravity") return put({m.gravity[0], m.gravity[1], m.gravity[2]});
  if (n == "tolerance") return scal(m.tolerance);
  if (n == "ls_tolerance") return scal(m.ls_tolerance);
  if (n == "impratio") return scal(m.impratio);
  if (n == "iterations") return scal(m.iterations);
  if (n == "ls_iterations") return scal(m.ls_iterations);
  if (n == "solver") return scal(m.solver);
  if (n == "cone") return scal(m.cone);
  if (n == "integrator") return scal(m.integrator);
  if (n == "disableflags") return scal(m.disableflags);
  if (n == "meaninertia") return scal(m.meaninertia);
  if (n == "npair") return scal((double)m.pair_geom1.size());
  if (n == "nexclude") return scal((double)m.exclude_signature.size());
  if (n == "nsite") return scal(m.nsite);
  if (n == "nsensor") return scal(m.nsensor);
  if (n == "nsensordata") return scal(m.nsensordata);
  if (n == "magnetic") return put({m.magnetic[0], m.magnetic[1], m.magnetic[2]});
#define F(x) if (n == #x) return put(m.x);
#define I(x) if (n == #x) return puti(m.x);
  I(body_parentid) I(body_rootid) I(body_weldid) I(body_jntnum) I(body_jntadr) I(body_dofnum) I(body_dofadr)
  I(body_geomnum) I(body_geomadr)
  F(body_pos) F(body_quat) F(body_ipos) F(body_iquat) F(body_mass) F(body_inertia) F(body_invweight0) F(body_subtreemass)
  I(jnt_type) I(jnt_qposadr) I(jnt_dofadr) I(jnt_bodyid) I(jnt_limited)
  F(jnt_pos) F(jnt_axis) F(jnt_range) F(jnt_margin) F(jnt_solref) F(jnt_solimp) F(jnt_stiffness)
  I(dof_bodyid) I(dof_jntid) I(dof_parentid)
  F(dof_armature) F(dof_damping) F(dof_frictionloss) F(dof_invweight0) F(dof_solref) F(dof_solimp)
  I(geom_type) I(geom_contype) I(geom_conaffinity) I(geom_condim) I(geom_bodyid) I(geom_priority) I(geom_group)
  F(geom_size) F(geom_pos) F(geom_quat) F(geom_friction) F(geom_solref) F(geom_solimp) F(geom_margin) F(geom_gap)
  F(geom_solmix) F(geom_rbound)
  I(actuator_trnid) I(actuator_ctrllimited) I(actuator_forcelimited)
  F(actuator_gear) F(actuator_gainprm) F(actuator_biasprm) F(actuator_ctrlrange) F(actuator_forcerange)
  I(pair_geom1) I(pair_geom2) I(pair_dim) F(pair_friction) F(pair_solref) F(pair_solimp) F(pair_margin) F(pair_gap)
  I(exclude_signature) F(qpos0) F(qpos_spring)
  I(site_bodyid) F(site_pos) F(site_quat)
  I(sensor_type) I(sensor_objtype) I(sensor_objid) I(sensor_adr) I(sensor_dim)
#undef F
#undef I
  if (n.rfind("cpair_", 0) == 0) {
    std::vector<CollisionPair> ps = build_pairs(m);
    std::vector<double> v;
    std::string f = n.substr(6);
    for (auto& p : ps) {
      if (f == "g1") v.push_back(p.g1);
      else if (f == "g2") v.push_back(p.g2);
      else if (f == "b1") v.push_back(p.b1);
      else if (f == "b2") v.push_back(p.b2);
      else if (f == "condim") v.push_back(p.condim);
      else if (f == "friction") v.insert(v.end(), p.friction, p.friction + 5);
      else if (f == "solref") v.insert(v.end(), p.solref, p.solref + 2);
      else if (f == "solimp") v.insert(v.end(), p.solimp, p.solimp + 5);
      else if (f == "margin") v.push_back(p.margin);
      else if (f == "gap") v.push_back(p.gap);
      else return -1;
    }
    return put(v);
  }
  return -1;
}

}  // namespace mjh
