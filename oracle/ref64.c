/* ref64.c -- FP64 CPU oracle for the batched MuJoCo step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see ref64.h). Restates the MuJoCo C
 * 3.5.0 step that the reference treats as ground truth (reference
 * src/mjb.cpp:807-839, tests/test_gpu_vs_cpu.cpp:177-267) in plain C double
 * precision. Where the reference's own kernels restate a MuJoCo formula the
 * citation points there; where the reference departs from MuJoCo (single
 * contact per pair, no warm start, its own solver tolerances: reference
 * src/compat/mkx_kernels/collision.hpp:493-500, solver.hpp:41-45,592) this file
 * follows MuJoCo instead.
 *
 * How each part is pinned (closed forms, brute force, optimality, the
 * reference's known-answer tests) is checked by tests/test_oracle.py and the
 * per-feature tests beside it; the parts no check pins are listed in ref64.h.
 */
#include "ref64.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define MINVAL 1e-15
#define MINIMP 0.0001
#define MAXIMP 0.9999

enum { JFREE = 0, JBALL = 1, JSLIDE = 2, JHINGE = 3 };
enum { GPLANE = 0, GSPHERE = 2, GCAPSULE = 3, GELLIPSOID = 4, GCYLINDER = 5, GBOX = 6 };
enum { EFC_EQUALITY = 0, EFC_FRICTION_DOF = 1, EFC_LIMIT_JOINT = 3, EFC_CONTACT_FRICTIONLESS = 5,
       EFC_CONTACT_PYRAMIDAL = 6, EFC_CONTACT_ELLIPTIC = 7 };
#define DS_CONSTRAINT (1 << 0)
#define DS_FRICTIONLOSS (1 << 2)
#define DS_LIMIT (1 << 3)
#define DS_CONTACT (1 << 4)
#define DS_PASSIVE (1 << 5)
#define DS_GRAVITY (1 << 6)
#define DS_CLAMPCTRL (1 << 7)
#define DS_WARMSTART (1 << 8)
#define DS_FILTERPARENT (1 << 9)
#define DS_ACTUATION (1 << 10)
#define DS_REFSAFE (1 << 11)
#define DS_EULERDAMP (1 << 15)

struct orc_prep {
  orc_model_t m;
  int npair;
  orc_pair_t* pairs;
  double* body_invweight0;
  double* dof_invweight0;
  double* body_subtreemass;
  double meaninertia;
  int con_cap, efc_cap; /* 0: unbounded (ORC_MAXCON / ORC_MAXEFC) */
};

struct orc_data {
  int nq, nv, nu, nbody, njnt, ngeom;
  double *qpos, *qvel, *ctrl, *qacc_ws, *qacc, *qacc_smooth;
  double *xpos, *xquat, *xmat, *xipos, *ximat, *xanchor, *xaxis, *geom_xpos, *geom_xmat;
  double *subtree_com, *cinert, *cdof, *cvel, *cdof_dot, *crb, *cacc, *cfrc;
  double *qM, *qLD;
  double *qfrc_bias, *qfrc_passive, *qfrc_actuator, *qfrc_smooth, *qfrc_constraint, *actuator_force;
  int ncon;
  orc_contact_t con[ORC_MAXCON];
  int nefc;
  int efc_type[ORC_MAXEFC], efc_id[ORC_MAXEFC];
  double *efc_J;
  double efc_pos[ORC_MAXEFC], efc_margin[ORC_MAXEFC], efc_D[ORC_MAXEFC], efc_R[ORC_MAXEFC],
      efc_aref[ORC_MAXEFC], efc_force[ORC_MAXEFC], efc_floss[ORC_MAXEFC], efc_vel[ORC_MAXEFC],
      efc_diagApprox[ORC_MAXEFC], efc_Jaref[ORC_MAXEFC], efc_Jv[ORC_MAXEFC], efc_imp[ORC_MAXEFC];
  int efc_active[ORC_MAXEFC];
  int niter;
  int status;
  /* scratch */
  double *Ma, *grad, *Mgrad, *search, *Mv, *H, *tmp1, *tmp2;
  double cost;
  /* domain randomisation (reference src/batched.cpp:1508-1538): per-body mass
   * scale and body-frame COM shift, friction scale, per-actuator force scale,
   * per-free-joint quadratic drag; off until orc_set_env_params is called */
  int has_rand, nfree;
  double *rand_mass, *rand_com, *rand_gain, *rand_drag, rand_fric;
  /* injected contacts (orc_inject_contacts): when ninject >= 0, collision
   * uses these instead of the narrow phase */
  int ninject;
  orc_contact_t* inject;
};

/* ------------------------------------------------------------------ math */
static void quat2mat(const double q[4], double R[9]) {
  double w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = w * w + x * x - y * y - z * z; R[1] = 2 * (x * y - w * z); R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z); R[4] = w * w - x * x + y * y - z * z; R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y); R[7] = 2 * (y * z + w * x); R[8] = w * w - x * x - y * y + z * z;
}
static void mulquat(const double a[4], const double b[4], double o[4]) {
  double r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  double r1 = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  double r2 = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  double r3 = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
  o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3;
}
static double normalize4(double q[4]) {
  double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (n < MINVAL) { q[0] = 1; q[1] = q[2] = q[3] = 0; return 0; }
  q[0] /= n; q[1] /= n; q[2] /= n; q[3] /= n;
  return n;
}
static double normalize3(double v[3]) {
  double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  if (n < MINVAL) { v[0] = 1; v[1] = v[2] = 0; return 0; }
  v[0] /= n; v[1] /= n; v[2] /= n;
  return n;
}
static void mulmatvec3(const double R[9], const double v[3], double o[3]) {
  double a = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  double b = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  double c = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
  o[0] = a; o[1] = b; o[2] = c;
}
static void mulmatTvec3(const double R[9], const double v[3], double o[3]) {
  double a = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
  double b = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
  double c = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
  o[0] = a; o[1] = b; o[2] = c;
}
static void cross3(const double a[3], const double b[3], double o[3]) {
  double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
  o[0] = x; o[1] = y; o[2] = z;
}
static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void axisangle2quat(const double ax[3], double a, double q[4]) {
  double s = sin(0.5 * a);
  q[0] = cos(0.5 * a); q[1] = ax[0] * s; q[2] = ax[1] * s; q[3] = ax[2] * s;
}
static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* spatial inertia (10-vector about a point) times motion vector [ang; lin] */
static void mul_inert_vec(const double* i, const double* v, double* r) {
  r[0] = i[0] * v[0] + i[3] * v[1] + i[4] * v[2] - i[8] * v[4] + i[7] * v[5];
  r[1] = i[3] * v[0] + i[1] * v[1] + i[5] * v[2] + i[8] * v[3] - i[6] * v[5];
  r[2] = i[4] * v[0] + i[5] * v[1] + i[2] * v[2] - i[7] * v[3] + i[6] * v[4];
  r[3] = i[8] * v[1] - i[7] * v[2] + i[9] * v[3];
  r[4] = i[6] * v[2] - i[8] * v[0] + i[9] * v[4];
  r[5] = i[7] * v[0] - i[6] * v[1] + i[9] * v[5];
}
/* motion cross motion: v x u */
static void cross_motion(const double* v, const double* u, double* r) {
  double a[3], b[3], c[3];
  cross3(v, u, a);
  cross3(v, u + 3, b);
  cross3(v + 3, u, c);
  r[0] = a[0]; r[1] = a[1]; r[2] = a[2];
  r[3] = b[0] + c[0]; r[4] = b[1] + c[1]; r[5] = b[2] + c[2];
}
/* motion cross force: v x* f */
static void cross_force(const double* v, const double* f, double* r) {
  double a[3], b[3], c[3];
  cross3(v, f, a);
  cross3(v + 3, f + 3, b);
  cross3(v, f + 3, c);
  r[0] = a[0] + b[0]; r[1] = a[1] + b[1]; r[2] = a[2] + b[2];
  r[3] = c[0]; r[4] = c[1]; r[5] = c[2];
}

/* ------------------------------------------------------------------ prepare */
static int excluded(const orc_model_t* m, int b1, int b2) {
  int lo = b1 < b2 ? b1 : b2, hi = b1 < b2 ? b2 : b1;
  int sig = (lo << 16) | hi;
  for (int i = 0; i < m->nexclude; i++) if (m->exclude_signature[i] == sig) return 1;
  return 0;
}

/* Static pair table: contype/conaffinity mask, weld, filterparent, exclude,
 * then MuJoCo's parameter mixing (priority, solmix, max friction/condim).
 * Follows reference src/io.cpp:922-959 (mixing) and :961-1070 (filter). */
static void build_pairs(orc_prep_t* p) {
  const orc_model_t* m = &p->m;
  int cap = m->ngeom * (m->ngeom - 1) / 2 + m->npair + 1;
  p->pairs = (orc_pair_t*)calloc(cap, sizeof(orc_pair_t));
  p->npair = 0;
  if (m->disableflags & (DS_CONTACT | DS_CONSTRAINT)) return;
  for (int i = 0; i < m->ngeom; i++) {
    for (int j = i + 1; j < m->ngeom; j++) {
      int g1 = i, g2 = j;
      if (m->geom_type[g1] > m->geom_type[g2]) { g1 = j; g2 = i; }
      int mask = (m->geom_contype[g1] & m->geom_conaffinity[g2]) | (m->geom_contype[g2] & m->geom_conaffinity[g1]);
      if (!mask) continue;
      int b1 = m->geom_bodyid[g1], b2 = m->geom_bodyid[g2];
      int w1 = m->body_weldid[b1], w2 = m->body_weldid[b2];
      if (w1 == w2) continue;
      if (!(m->disableflags & DS_FILTERPARENT)) {
        int w1p = m->body_weldid[m->body_parentid[w1]], w2p = m->body_weldid[m->body_parentid[w2]];
        if (w1 != 0 && w2 != 0 && (w1 == w2p || w2 == w1p)) continue;
      }
      if (excluded(m, b1, b2)) continue;
      if (m->geom_type[g1] == GPLANE && m->geom_type[g2] == GPLANE) continue;
      orc_pair_t* q = &p->pairs[p->npair++];
      q->g1 = g1; q->g2 = g2; q->b1 = b1; q->b2 = b2;
      q->margin = fmax(m->geom_margin[g1], m->geom_margin[g2]);
      q->gap = fmax(m->geom_gap[g1], m->geom_gap[g2]);
      double fr[3];
      int p1 = m->geom_priority[g1], p2 = m->geom_priority[g2];
      if (p1 != p2) {
        int gp = p1 > p2 ? g1 : g2;
        q->condim = m->geom_condim[gp];
        for (int k = 0; k < 3; k++) fr[k] = m->geom_friction[3 * gp + k];
        for (int k = 0; k < 2; k++) q->solref[k] = m->geom_solref[2 * gp + k];
        for (int k = 0; k < 5; k++) q->solimp[k] = m->geom_solimp[5 * gp + k];
      } else {
        q->condim = m->geom_condim[g1] > m->geom_condim[g2] ? m->geom_condim[g1] : m->geom_condim[g2];
        double s1 = m->geom_solmix[g1], s2 = m->geom_solmix[g2], mix;
        if (s1 >= MINVAL && s2 >= MINVAL) mix = s1 / (s1 + s2);
        else if (s1 < MINVAL && s2 < MINVAL) mix = 0.5;
        else mix = s1 < MINVAL ? 0.0 : 1.0;
        const double *r1 = m->geom_solref + 2 * g1, *r2 = m->geom_solref + 2 * g2;
        if (r1[0] > 0 && r2[0] > 0) for (int k = 0; k < 2; k++) q->solref[k] = mix * r1[k] + (1 - mix) * r2[k];
        else for (int k = 0; k < 2; k++) q->solref[k] = fmin(r1[k], r2[k]);
        for (int k = 0; k < 5; k++) q->solimp[k] = mix * m->geom_solimp[5 * g1 + k] + (1 - mix) * m->geom_solimp[5 * g2 + k];
        for (int k = 0; k < 3; k++) fr[k] = fmax(m->geom_friction[3 * g1 + k], m->geom_friction[3 * g2 + k]);
      }
      q->friction[0] = fr[0]; q->friction[1] = fr[0]; q->friction[2] = fr[1]; q->friction[3] = fr[2]; q->friction[4] = fr[2];
    }
  }
  for (int k = 0; k < m->npair; k++) {
    int g1 = m->pair_geom1[k], g2 = m->pair_geom2[k];
    if (m->geom_type[g1] > m->geom_type[g2]) { int t = g1; g1 = g2; g2 = t; }
    orc_pair_t q;
    memset(&q, 0, sizeof(q));
    q.g1 = g1; q.g2 = g2; q.b1 = m->geom_bodyid[g1]; q.b2 = m->geom_bodyid[g2];
    q.condim = m->pair_dim[k]; q.margin = m->pair_margin[k]; q.gap = m->pair_gap[k];
    for (int i = 0; i < 5; i++) { q.friction[i] = m->pair_friction[5 * k + i]; q.solimp[i] = m->pair_solimp[5 * k + i]; }
    for (int i = 0; i < 2; i++) q.solref[i] = m->pair_solref[2 * k + i];
    int found = -1;
    for (int e = 0; e < p->npair; e++) if (p->pairs[e].g1 == g1 && p->pairs[e].g2 == g2) found = e;
    if (found >= 0) p->pairs[found] = q;
    else p->pairs[p->npair++] = q;
  }
}

static void data_alloc(orc_data_t* d, const orc_model_t* m) {
  d->nq = m->nq; d->nv = m->nv; d->nu = m->nu; d->nbody = m->nbody; d->njnt = m->njnt; d->ngeom = m->ngeom;
  int nv = m->nv > 0 ? m->nv : 1, nq = m->nq > 0 ? m->nq : 1, nu = m->nu > 0 ? m->nu : 1;
  int nb = m->nbody, nj = m->njnt > 0 ? m->njnt : 1, ng = m->ngeom > 0 ? m->ngeom : 1;
#define AL(f, n) d->f = (double*)calloc((size_t)(n), sizeof(double))
  AL(qpos, nq); AL(qvel, nv); AL(ctrl, nu); AL(qacc_ws, nv); AL(qacc, nv); AL(qacc_smooth, nv);
  AL(xpos, 3 * nb); AL(xquat, 4 * nb); AL(xmat, 9 * nb); AL(xipos, 3 * nb); AL(ximat, 9 * nb);
  AL(xanchor, 3 * nj); AL(xaxis, 3 * nj); AL(geom_xpos, 3 * ng); AL(geom_xmat, 9 * ng);
  AL(subtree_com, 3 * nb); AL(cinert, 10 * nb); AL(cdof, 6 * nv); AL(cvel, 6 * nb); AL(cdof_dot, 6 * nv);
  AL(crb, 10 * nb); AL(cacc, 6 * nb); AL(cfrc, 6 * nb);
  AL(qM, nv * nv); AL(qLD, nv * nv);
  AL(qfrc_bias, nv); AL(qfrc_passive, nv); AL(qfrc_actuator, nv); AL(qfrc_smooth, nv); AL(qfrc_constraint, nv);
  AL(actuator_force, nu);
  AL(efc_J, (size_t)ORC_MAXEFC * nv);
  AL(Ma, nv); AL(grad, nv); AL(Mgrad, nv); AL(search, nv); AL(Mv, nv); AL(H, nv * nv); AL(tmp1, nv); AL(tmp2, nv);
  d->nfree = 0;
  for (int j = 0; j < m->njnt; j++) d->nfree += m->jnt_type[j] == JFREE;
  AL(rand_mass, nb); AL(rand_com, 3 * nb); AL(rand_gain, nu); AL(rand_drag, d->nfree > 0 ? d->nfree : 1);
  for (int b = 0; b < nb; b++) d->rand_mass[b] = 1;
  for (int u = 0; u < nu; u++) d->rand_gain[u] = 1;
  d->rand_fric = 1;
  d->has_rand = 0;
#undef AL
}

static void free_data_arrays(orc_data_t* d) {
  double* ptrs[] = {d->qpos, d->qvel, d->ctrl, d->qacc_ws, d->qacc, d->qacc_smooth, d->xpos, d->xquat, d->xmat,
                    d->xipos, d->ximat, d->xanchor, d->xaxis, d->geom_xpos, d->geom_xmat, d->subtree_com,
                    d->cinert, d->cdof, d->cvel, d->cdof_dot, d->crb, d->cacc, d->cfrc, d->qM, d->qLD,
                    d->qfrc_bias, d->qfrc_passive, d->qfrc_actuator, d->qfrc_smooth, d->qfrc_constraint,
                    d->actuator_force, d->efc_J, d->Ma, d->grad, d->Mgrad, d->search, d->Mv, d->H, d->tmp1, d->tmp2,
                    d->rand_mass, d->rand_com, d->rand_gain, d->rand_drag};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
}

/* forward declarations of the pipeline pieces used by orc_prepare */
static void kinematics(const orc_prep_t* p, orc_data_t* d);
static void com_pos(const orc_prep_t* p, orc_data_t* d);
static void crb_mass(const orc_prep_t* p, orc_data_t* d);
static int factor_m(const orc_prep_t* p, orc_data_t* d);
static void solve_m(const orc_prep_t* p, const orc_data_t* d, double* x);

orc_prep_t* orc_prepare(const orc_model_t* m) {
  orc_prep_t* p = (orc_prep_t*)calloc(1, sizeof(orc_prep_t));
  p->m = *m;
  int nb = m->nbody, nv = m->nv;
  p->body_subtreemass = (double*)calloc(nb, sizeof(double));
  for (int b = 0; b < nb; b++) p->body_subtreemass[b] = m->body_mass[b];
  for (int b = nb - 1; b > 0; b--) p->body_subtreemass[m->body_parentid[b]] += p->body_subtreemass[b];
  build_pairs(p);
  p->body_invweight0 = (double*)calloc(2 * nb, sizeof(double));
  p->dof_invweight0 = (double*)calloc(nv > 0 ? nv : 1, sizeof(double));
  p->meaninertia = 1.0;
  if (nv == 0) return p;
  /* mj_setConst subset: kinematics and M at qpos0 */
  orc_data_t* d = orc_data_new(p);
  memcpy(d->qpos, m->qpos0, sizeof(double) * m->nq);
  kinematics(p, d);
  com_pos(p, d);
  crb_mass(p, d);
  if (factor_m(p, d) != 0) { orc_data_free(d); orc_free(p); return NULL; }
  double tr = 0;
  for (int i = 0; i < nv; i++) tr += d->qM[i * nv + i];
  p->meaninertia = tr / nv;
  if (p->meaninertia < MINVAL) p->meaninertia = 1.0;
  /* columns of M^-1 */
  double* Minv = (double*)calloc((size_t)nv * nv, sizeof(double));
  double* e = (double*)calloc(nv, sizeof(double));
  for (int c = 0; c < nv; c++) {
    memset(e, 0, sizeof(double) * nv);
    e[c] = 1;
    solve_m(p, d, e);
    for (int r = 0; r < nv; r++) Minv[r * nv + c] = e[r];
  }
  /* body invweight0: mean translational / rotational diagonal of J M^-1 J^T at the body COM */
  double* J = (double*)calloc((size_t)6 * nv, sizeof(double));
  for (int b = 1; b < nb; b++) {
    if (m->body_weldid[b] == 0) continue;
    memset(J, 0, sizeof(double) * 6 * nv);
    int bb = b;
    while (bb > 0 && m->body_dofnum[bb] == 0) bb = m->body_parentid[bb];
    if (bb == 0) continue;
    const double* c = d->subtree_com + 3 * m->body_rootid[b];
    double off[3] = {d->xipos[3 * b] - c[0], d->xipos[3 * b + 1] - c[1], d->xipos[3 * b + 2] - c[2]};
    for (int dd = m->body_dofadr[bb] + m->body_dofnum[bb] - 1; dd >= 0; dd = m->dof_parentid[dd]) {
      const double* cd = d->cdof + 6 * dd;
      double t[3];
      cross3(cd, off, t);
      J[0 * nv + dd] = cd[3] + t[0]; J[1 * nv + dd] = cd[4] + t[1]; J[2 * nv + dd] = cd[5] + t[2];
      J[3 * nv + dd] = cd[0]; J[4 * nv + dd] = cd[1]; J[5 * nv + dd] = cd[2];
    }
    double diag[6];
    for (int r = 0; r < 6; r++) {
      double s = 0;
      for (int i = 0; i < nv; i++) for (int k = 0; k < nv; k++) s += J[r * nv + i] * Minv[i * nv + k] * J[r * nv + k];
      diag[r] = s;
    }
    p->body_invweight0[2 * b] = (diag[0] + diag[1] + diag[2]) / 3.0;
    p->body_invweight0[2 * b + 1] = (diag[3] + diag[4] + diag[5]) / 3.0;
  }
  for (int j = 0; j < m->njnt; j++) {
    int a = m->jnt_dofadr[j];
    if (m->jnt_type[j] == JFREE) {
      double t = (Minv[a * nv + a] + Minv[(a + 1) * nv + a + 1] + Minv[(a + 2) * nv + a + 2]) / 3.0;
      double r = (Minv[(a + 3) * nv + a + 3] + Minv[(a + 4) * nv + a + 4] + Minv[(a + 5) * nv + a + 5]) / 3.0;
      for (int i = 0; i < 3; i++) { p->dof_invweight0[a + i] = t; p->dof_invweight0[a + 3 + i] = r; }
    } else if (m->jnt_type[j] == JBALL) {
      double r = (Minv[a * nv + a] + Minv[(a + 1) * nv + a + 1] + Minv[(a + 2) * nv + a + 2]) / 3.0;
      for (int i = 0; i < 3; i++) p->dof_invweight0[a + i] = r;
    } else {
      p->dof_invweight0[a] = Minv[a * nv + a];
    }
  }
  free(J); free(e); free(Minv);
  orc_data_free(d);
  return p;
}

/* Per-env capacities of the reference (src/batched.cpp:290 contacts;
 * src/compat/mkx_kernels/solver.hpp:299,349,403 row budget: friction and limit
 * rows while fewer than efc_cap - 4 rows exist, a contact while fewer than
 * efc_cap - 10). 0 leaves a cap unbounded. */
void orc_set_caps(orc_prep_t* p, int con_cap, int efc_cap) {
  p->con_cap = con_cap;
  p->efc_cap = efc_cap;
}

void orc_free(orc_prep_t* p) {
  if (!p) return;
  free(p->pairs); free(p->body_invweight0); free(p->dof_invweight0); free(p->body_subtreemass);
  free(p);
}
int orc_npair(const orc_prep_t* p) { return p->npair; }
const orc_pair_t* orc_pairs(const orc_prep_t* p) { return p->pairs; }
const double* orc_body_invweight0(const orc_prep_t* p) { return p->body_invweight0; }
const double* orc_dof_invweight0(const orc_prep_t* p) { return p->dof_invweight0; }
double orc_meaninertia(const orc_prep_t* p) { return p->meaninertia; }

orc_data_t* orc_data_new(const orc_prep_t* p) {
  orc_data_t* d = (orc_data_t*)calloc(1, sizeof(orc_data_t));
  d->ninject = -1;
  data_alloc(d, &p->m);
  memcpy(d->qpos, p->m.qpos0, sizeof(double) * p->m.nq);
  return d;
}
void orc_data_free(orc_data_t* d) {
  if (!d) return;
  free(d->inject);
  free_data_arrays(d);
  free(d);
}
void orc_set_state(orc_data_t* d, const double* qpos, const double* qvel, const double* qacc_ws, const double* ctrl) {
  if (qpos) memcpy(d->qpos, qpos, sizeof(double) * d->nq);
  if (qvel) memcpy(d->qvel, qvel, sizeof(double) * d->nv);
  if (qacc_ws) memcpy(d->qacc_ws, qacc_ws, sizeof(double) * d->nv);
  if (ctrl) memcpy(d->ctrl, ctrl, sizeof(double) * d->nu);
}
void orc_set_env_params(orc_data_t* d, const double* mass_scale, const double* com_delta, const double* friction_scale,
                        const double* gain_scale, const double* drag) {
  d->has_rand = 1;
  if (mass_scale) memcpy(d->rand_mass, mass_scale, sizeof(double) * d->nbody);
  if (com_delta) memcpy(d->rand_com, com_delta, sizeof(double) * 3 * d->nbody);
  if (friction_scale) d->rand_fric = friction_scale[0];
  if (gain_scale) memcpy(d->rand_gain, gain_scale, sizeof(double) * d->nu);
  if (drag) memcpy(d->rand_drag, drag, sizeof(double) * d->nfree);
}
int orc_nfree(const orc_data_t* d) { return d->nfree; }
void orc_get_state(const orc_data_t* d, double* qpos, double* qvel, double* qacc_ws) {
  if (qpos) memcpy(qpos, d->qpos, sizeof(double) * d->nq);
  if (qvel) memcpy(qvel, d->qvel, sizeof(double) * d->nv);
  if (qacc_ws) memcpy(qacc_ws, d->qacc_ws, sizeof(double) * d->nv);
}

/* ------------------------------------------------------------------ smooth dynamics */
/* mj_kinematics: body chain, joint anchors/axes, inertial and geom frames.
 * Same composition as reference src/compat/mkx_kernels/kinematics.hpp:79-144. */
static void kinematics(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  d->xpos[0] = d->xpos[1] = d->xpos[2] = 0;
  d->xquat[0] = 1; d->xquat[1] = d->xquat[2] = d->xquat[3] = 0;
  quat2mat(d->xquat, d->xmat);
  memcpy(d->xipos, d->xpos, 3 * sizeof(double));
  memcpy(d->ximat, d->xmat, 9 * sizeof(double));
  for (int b = 1; b < m->nbody; b++) {
    int par = m->body_parentid[b];
    int ja = m->body_jntadr[b], jn = m->body_jntnum[b];
    double pos[3], q[4];
    if (jn > 0 && m->jnt_type[ja] == JFREE) {
      int a = m->jnt_qposadr[ja];
      for (int i = 0; i < 3; i++) pos[i] = d->qpos[a + i];
      for (int i = 0; i < 4; i++) q[i] = d->qpos[a + 3 + i];
      normalize4(q);
      for (int i = 0; i < 3; i++) { d->xanchor[3 * ja + i] = pos[i]; d->xaxis[3 * ja + i] = m->jnt_axis[3 * ja + i]; }
    } else {
      double t[3];
      mulmatvec3(d->xmat + 9 * par, m->body_pos + 3 * b, t);
      for (int i = 0; i < 3; i++) pos[i] = d->xpos[3 * par + i] + t[i];
      mulquat(d->xquat + 4 * par, m->body_quat + 4 * b, q);
      for (int j = ja; j < ja + jn; j++) {
        int a = m->jnt_qposadr[j];
        double R[9], anc[3], ax[3];
        quat2mat(q, R);
        mulmatvec3(R, m->jnt_pos + 3 * j, anc);
        for (int i = 0; i < 3; i++) anc[i] += pos[i];
        mulmatvec3(R, m->jnt_axis + 3 * j, ax);
        memcpy(d->xanchor + 3 * j, anc, 3 * sizeof(double));
        memcpy(d->xaxis + 3 * j, ax, 3 * sizeof(double));
        if (m->jnt_type[j] == JSLIDE) {
          double s = d->qpos[a] - m->qpos0[a];
          for (int i = 0; i < 3; i++) pos[i] += ax[i] * s;
        } else {
          double ql[4];
          if (m->jnt_type[j] == JHINGE) axisangle2quat(m->jnt_axis + 3 * j, d->qpos[a] - m->qpos0[a], ql);
          else { for (int i = 0; i < 4; i++) ql[i] = d->qpos[a + i]; normalize4(ql); }
          mulquat(q, ql, q);
          double R2[9], r[3];
          quat2mat(q, R2);
          mulmatvec3(R2, m->jnt_pos + 3 * j, r);
          for (int i = 0; i < 3; i++) pos[i] = anc[i] - r[i];
        }
      }
      normalize4(q);
    }
    memcpy(d->xpos + 3 * b, pos, 3 * sizeof(double));
    memcpy(d->xquat + 4 * b, q, 4 * sizeof(double));
    quat2mat(q, d->xmat + 9 * b);
    double t[3], qi[4], ip[3];
    /* randomised COM: reference forward.hpp:113-114 adds xmat * com_delta to xipos */
    for (int i = 0; i < 3; i++) ip[i] = m->body_ipos[3 * b + i] + (d->has_rand ? d->rand_com[3 * b + i] : 0.0);
    mulmatvec3(d->xmat + 9 * b, ip, t);
    for (int i = 0; i < 3; i++) d->xipos[3 * b + i] = pos[i] + t[i];
    mulquat(q, m->body_iquat + 4 * b, qi);
    quat2mat(qi, d->ximat + 9 * b);
  }
  for (int g = 0; g < m->ngeom; g++) {
    int b = m->geom_bodyid[g];
    double t[3], qg[4];
    mulmatvec3(d->xmat + 9 * b, m->geom_pos + 3 * g, t);
    for (int i = 0; i < 3; i++) d->geom_xpos[3 * g + i] = d->xpos[3 * b + i] + t[i];
    mulquat(d->xquat + 4 * b, m->geom_quat + 4 * g, qg);
    quat2mat(qg, d->geom_xmat + 9 * g);
  }
}

/* mj_comPos: subtree COM, body inertia about the root's subtree COM (cinert)
 * and dof motion subspaces (cdof). Reference forward.hpp:146-233. */
static void com_pos(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  int nb = m->nbody;
  /* randomised masses scale mass and inertia (reference forward.hpp:147,167,173) */
  double smass[nb > 0 ? nb : 1];
  for (int b = 0; b < nb; b++) {
    smass[b] = m->body_mass[b] * (d->has_rand ? d->rand_mass[b] : 1.0);
    for (int i = 0; i < 3; i++) d->subtree_com[3 * b + i] = smass[b] * d->xipos[3 * b + i];
  }
  for (int b = nb - 1; b > 0; b--) {
    int par = m->body_parentid[b];
    for (int i = 0; i < 3; i++) d->subtree_com[3 * par + i] += d->subtree_com[3 * b + i];
    smass[par] += smass[b];
  }
  for (int b = 0; b < nb; b++) {
    double sm = d->has_rand ? smass[b] : p->body_subtreemass[b];
    for (int i = 0; i < 3; i++) d->subtree_com[3 * b + i] = sm < MINVAL ? d->xipos[3 * b + i] : d->subtree_com[3 * b + i] / sm;
  }
  memset(d->cinert, 0, 10 * sizeof(double));
  for (int b = 1; b < nb; b++) {
    const double* R = d->ximat + 9 * b;
    double sc = d->has_rand ? d->rand_mass[b] : 1.0;
    double I[3] = {m->body_inertia[3 * b] * sc, m->body_inertia[3 * b + 1] * sc, m->body_inertia[3 * b + 2] * sc};
    double mass = m->body_mass[b] * sc;
    const double* c = d->subtree_com + 3 * m->body_rootid[b];
    double dd[3] = {d->xipos[3 * b] - c[0], d->xipos[3 * b + 1] - c[1], d->xipos[3 * b + 2] - c[2]};
    double A[9];
    for (int r = 0; r < 3; r++)
      for (int k = 0; k < 3; k++) A[3 * r + k] = R[3 * r] * I[0] * R[3 * k] + R[3 * r + 1] * I[1] * R[3 * k + 1] + R[3 * r + 2] * I[2] * R[3 * k + 2];
    double* o = d->cinert + 10 * b;
    double sq = dot3(dd, dd);
    o[0] = A[0] + mass * (sq - dd[0] * dd[0]);
    o[1] = A[4] + mass * (sq - dd[1] * dd[1]);
    o[2] = A[8] + mass * (sq - dd[2] * dd[2]);
    o[3] = A[1] - mass * dd[0] * dd[1];
    o[4] = A[2] - mass * dd[0] * dd[2];
    o[5] = A[5] - mass * dd[1] * dd[2];
    o[6] = mass * dd[0]; o[7] = mass * dd[1]; o[8] = mass * dd[2]; o[9] = mass;
  }
  for (int j = 0; j < m->njnt; j++) {
    int b = m->jnt_bodyid[j], da = m->jnt_dofadr[j];
    const double* c = d->subtree_com + 3 * m->body_rootid[b];
    const double* anc = d->xanchor + 3 * j;
    double off[3] = {c[0] - anc[0], c[1] - anc[1], c[2] - anc[2]};
    const double* R = d->xmat + 9 * b;
    int t = m->jnt_type[j];
    if (t == JFREE || t == JBALL) {
      int r0 = da;
      if (t == JFREE) {
        for (int i = 0; i < 3; i++) {
          double* cd = d->cdof + 6 * (da + i);
          memset(cd, 0, 6 * sizeof(double));
          cd[3 + i] = 1;
        }
        r0 = da + 3;
      }
      for (int i = 0; i < 3; i++) {
        double ax[3] = {R[i], R[3 + i], R[6 + i]};
        double* cd = d->cdof + 6 * (r0 + i);
        cd[0] = ax[0]; cd[1] = ax[1]; cd[2] = ax[2];
        cross3(ax, off, cd + 3);
      }
    } else if (t == JSLIDE) {
      double* cd = d->cdof + 6 * da;
      cd[0] = cd[1] = cd[2] = 0;
      memcpy(cd + 3, d->xaxis + 3 * j, 3 * sizeof(double));
    } else {
      double* cd = d->cdof + 6 * da;
      memcpy(cd, d->xaxis + 3 * j, 3 * sizeof(double));
      cross3(d->xaxis + 3 * j, off, cd + 3);
    }
  }
}

/* mj_crb: composite inertia and the joint-space inertia matrix (dense storage
 * of the tree-sparse qM). Reference forward.hpp:240-279. */
static void crb_mass(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  int nv = m->nv;
  memcpy(d->crb, d->cinert, sizeof(double) * 10 * m->nbody);
  for (int b = m->nbody - 1; b > 0; b--) {
    int par = m->body_parentid[b];
    if (par > 0) for (int k = 0; k < 10; k++) d->crb[10 * par + k] += d->crb[10 * b + k];
  }
  memset(d->qM, 0, sizeof(double) * nv * nv);
  for (int i = 0; i < nv; i++) {
    double buf[6];
    mul_inert_vec(d->crb + 10 * m->dof_bodyid[i], d->cdof + 6 * i, buf);
    for (int j = i; j >= 0; j = m->dof_parentid[j]) {
      double s = 0;
      for (int k = 0; k < 6; k++) s += d->cdof[6 * j + k] * buf[k];
      d->qM[i * nv + j] = s;
      d->qM[j * nv + i] = s;
    }
    d->qM[i * nv + i] += m->dof_armature[i];
  }
}

/* mj_factorM: tree-sparse L^T D L factorisation of qM (dense storage; entries
 * outside the dof-ancestor pattern stay zero). L is unit lower triangular, D is
 * kept on the diagonal of qLD. */
static int factor_m_into(const orc_model_t* m, const double* M, double* LD) {
  int nv = m->nv;
  memcpy(LD, M, sizeof(double) * nv * nv);
  for (int k = nv - 1; k >= 0; k--) {
    double dk = LD[k * nv + k];
    if (!(dk > MINVAL)) return -1;
    for (int i = m->dof_parentid[k]; i >= 0; i = m->dof_parentid[i]) {
      double a = LD[k * nv + i] / dk;
      for (int j = i; j >= 0; j = m->dof_parentid[j]) LD[i * nv + j] -= a * LD[k * nv + j];
      LD[k * nv + i] = a;
    }
  }
  return 0;
}
static void solve_ld(const orc_model_t* m, const double* LD, double* x) {
  int nv = m->nv;
  for (int k = nv - 1; k >= 0; k--)
    for (int i = m->dof_parentid[k]; i >= 0; i = m->dof_parentid[i]) x[i] -= LD[k * nv + i] * x[k];
  for (int k = 0; k < nv; k++) x[k] /= LD[k * nv + k];
  for (int k = 0; k < nv; k++)
    for (int i = m->dof_parentid[k]; i >= 0; i = m->dof_parentid[i]) x[k] -= LD[k * nv + i] * x[i];
}
static int factor_m(const orc_prep_t* p, orc_data_t* d) { return factor_m_into(&p->m, d->qM, d->qLD); }
static void solve_m(const orc_prep_t* p, const orc_data_t* d, double* x) { solve_ld(&p->m, d->qLD, x); }

/* mj_comVel: body velocities and cdof_dot with MuJoCo's grouping of free and
 * ball dofs (reference forward.hpp:287-332). */
static void com_vel(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  memset(d->cvel, 0, 6 * sizeof(double));
  for (int b = 1; b < m->nbody; b++) {
    double cvel[6];
    memcpy(cvel, d->cvel + 6 * m->body_parentid[b], 6 * sizeof(double));
    int ja = m->body_jntadr[b], jn = m->body_jntnum[b];
    for (int j = ja; j < ja + jn; j++) {
      int da = m->jnt_dofadr[j], t = m->jnt_type[j];
      if (t == JFREE) {
        for (int k = 0; k < 3; k++) memset(d->cdof_dot + 6 * (da + k), 0, 6 * sizeof(double));
        for (int k = 0; k < 3; k++) for (int c = 0; c < 6; c++) cvel[c] += d->cdof[6 * (da + k) + c] * d->qvel[da + k];
        da += 3;
      }
      if (t == JFREE || t == JBALL) {
        for (int k = 0; k < 3; k++) cross_motion(cvel, d->cdof + 6 * (da + k), d->cdof_dot + 6 * (da + k));
        for (int k = 0; k < 3; k++) for (int c = 0; c < 6; c++) cvel[c] += d->cdof[6 * (da + k) + c] * d->qvel[da + k];
      } else {
        cross_motion(cvel, d->cdof + 6 * da, d->cdof_dot + 6 * da);
        for (int c = 0; c < 6; c++) cvel[c] += d->cdof[6 * da + c] * d->qvel[da];
      }
    }
    memcpy(d->cvel + 6 * b, cvel, 6 * sizeof(double));
  }
}

/* mj_rne with zero acceleration: qfrc_bias = C(q, qdot) qdot + g(q).
 * Reference forward.hpp:338-397. */
static void rne_bias(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  memset(d->cacc, 0, 6 * sizeof(double));
  if (!(m->disableflags & DS_GRAVITY)) for (int i = 0; i < 3; i++) d->cacc[3 + i] = -m->gravity[i];
  for (int b = 1; b < m->nbody; b++) {
    double* a = d->cacc + 6 * b;
    memcpy(a, d->cacc + 6 * m->body_parentid[b], 6 * sizeof(double));
    int da = m->body_dofadr[b], dn = m->body_dofnum[b];
    for (int k = da; k < da + dn; k++) for (int c = 0; c < 6; c++) a[c] += d->cdof_dot[6 * k + c] * d->qvel[k];
    double f1[6], mv[6], f2[6];
    mul_inert_vec(d->cinert + 10 * b, a, f1);
    mul_inert_vec(d->cinert + 10 * b, d->cvel + 6 * b, mv);
    cross_force(d->cvel + 6 * b, mv, f2);
    for (int c = 0; c < 6; c++) d->cfrc[6 * b + c] = f1[c] + f2[c];
  }
  memset(d->cfrc, 0, 6 * sizeof(double));
  for (int b = m->nbody - 1; b > 0; b--) {
    int par = m->body_parentid[b];
    for (int c = 0; c < 6; c++) d->cfrc[6 * par + c] += d->cfrc[6 * b + c];
  }
  for (int k = 0; k < m->nv; k++) {
    double s = 0;
    for (int c = 0; c < 6; c++) s += d->cdof[6 * k + c] * d->cfrc[6 * m->dof_bodyid[k] + c];
    d->qfrc_bias[k] = s;
  }
}

/* mj_passive subset: joint springs (hinge/slide) and dof damping. */
static void passive(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  memset(d->qfrc_passive, 0, sizeof(double) * m->nv);
  if (m->disableflags & DS_PASSIVE) return;
  for (int j = 0; j < m->njnt; j++) {
    int t = m->jnt_type[j];
    if ((t == JHINGE || t == JSLIDE) && m->jnt_stiffness[j] != 0) {
      int a = m->jnt_qposadr[j];
      d->qfrc_passive[m->jnt_dofadr[j]] -= m->jnt_stiffness[j] * (d->qpos[a] - m->qpos_spring[a]);
    }
  }
  for (int k = 0; k < m->nv; k++) d->qfrc_passive[k] -= m->dof_damping[k] * d->qvel[k];
}

/* quadratic drag F = -k |v| v on each free joint's translational dofs, k per
 * free joint in joint order (reference forward.hpp:455-465) */
static void free_drag(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  int f = 0;
  for (int j = 0; j < m->njnt; j++) {
    if (m->jnt_type[j] != JFREE) continue;
    const double* v = d->qvel + m->jnt_dofadr[j];
    double kd = -d->rand_drag[f++] * sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    for (int i = 0; i < 3; i++) d->qfrc_smooth[m->jnt_dofadr[j] + i] += kd * v[i];
  }
}

/* mj_fwdActuation for joint transmissions: force = gain*ctrl + b0 + b1*length
 * + b2*velocity, ctrl and force clamping, qfrc_actuator = moment * force.
 * Reference forward.hpp:411-435. */
static void actuation(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  memset(d->qfrc_actuator, 0, sizeof(double) * m->nv);
  for (int u = 0; u < m->nu; u++) {
    d->actuator_force[u] = 0;
    if (m->disableflags & DS_ACTUATION) continue;
    int j = m->actuator_trnid[u];
    int da = m->jnt_dofadr[j], qa = m->jnt_qposadr[j];
    double gear = m->actuator_gear[u];
    double c = d->ctrl[u];
    if (m->actuator_ctrllimited[u] && !(m->disableflags & DS_CLAMPCTRL)) c = clampd(c, m->actuator_ctrlrange[2 * u], m->actuator_ctrlrange[2 * u + 1]);
    double len = gear * d->qpos[qa], vel = gear * d->qvel[da];
    const double *g = m->actuator_gainprm + 3 * u, *bi = m->actuator_biasprm + 3 * u;
    double f = g[0] * c + bi[0] + bi[1] * len + bi[2] * vel;
    if (d->has_rand) f *= d->rand_gain[u]; /* reference forward.hpp:427 */
    if (m->actuator_forcelimited[u]) f = clampd(f, m->actuator_forcerange[2 * u], m->actuator_forcerange[2 * u + 1]);
    d->actuator_force[u] = f;
    d->qfrc_actuator[da] += gear * f;
  }
}

/* ------------------------------------------------------------------ collision */
/* Contact frame from the normal: tangent 1 is the projection of (0,1,0) --
 * or (0,0,1) when the normal is close to y -- orthogonal to the normal, and
 * tangent 2 = normal x tangent 1 (MuJoCo mju_makeFrame). */
static void make_frame(const double n_in[3], double f[9]) {
  double n[3] = {n_in[0], n_in[1], n_in[2]};
  normalize3(n);
  double y[3];
  if (fabs(n[1]) < 0.5) { y[0] = 0; y[1] = 1; y[2] = 0; }
  else { y[0] = 0; y[1] = 0; y[2] = 1; }
  double s = dot3(n, y);
  for (int i = 0; i < 3; i++) y[i] -= s * n[i];
  normalize3(y);
  double z[3];
  cross3(n, y, z);
  f[0] = n[0]; f[1] = n[1]; f[2] = n[2];
  f[3] = y[0]; f[4] = y[1]; f[5] = y[2];
  f[6] = z[0]; f[7] = z[1]; f[8] = z[2];
}

typedef struct { double dist, pos[3], n[3]; } rawc_t;

static int sphere_sphere(const double c1[3], double r1, const double c2[3], double r2, double margin, rawc_t* out) {
  double dv[3] = {c2[0] - c1[0], c2[1] - c1[1], c2[2] - c1[2]};
  double L = sqrt(dot3(dv, dv));
  double dist = L - r1 - r2;
  if (dist > margin) return 0;
  double n[3];
  if (L < MINVAL) { n[0] = 1; n[1] = 0; n[2] = 0; }
  else { n[0] = dv[0] / L; n[1] = dv[1] / L; n[2] = dv[2] / L; }
  out->dist = dist;
  for (int i = 0; i < 3; i++) { out->n[i] = n[i]; out->pos[i] = c1[i] + n[i] * (r1 + 0.5 * dist); }
  return 1;
}

/* plane (z axis of its frame) against a sphere centre */
static int plane_point(const double pp[3], const double pn[3], const double c[3], double r, double margin, rawc_t* out) {
  double dv[3] = {c[0] - pp[0], c[1] - pp[1], c[2] - pp[2]};
  double dist = dot3(pn, dv) - r;
  if (dist > margin) return 0;
  out->dist = dist;
  for (int i = 0; i < 3; i++) { out->n[i] = pn[i]; out->pos[i] = c[i] - pn[i] * (r + 0.5 * dist); }
  return 1;
}

/* closest points of segments p1 + s*d1 and p2 + t*d2, s,t in [0,1]
 * (the clamped-parameter method of Ericson, Real-Time Collision Detection 5.1.9) */
static void seg_seg(const double p1[3], const double d1[3], const double p2[3], const double d2[3], double* s_out, double* t_out) {
  double r[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  double a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r);
  double s, t;
  if (a <= MINVAL && e <= MINVAL) { s = t = 0; }
  else if (a <= MINVAL) { s = 0; t = clampd(f / e, 0, 1); }
  else {
    double c = dot3(d1, r);
    if (e <= MINVAL) { t = 0; s = clampd(-c / a, 0, 1); }
    else {
      double b = dot3(d1, d2);
      double denom = a * e - b * b;
      s = (denom > 1e-10 * a * e) ? clampd((b * f - c * e) / denom, 0, 1) : 0;
      t = (b * s + f) / e;
      if (t < 0) { t = 0; s = clampd(-c / a, 0, 1); }
      else if (t > 1) { t = 1; s = clampd((b - c) / a, 0, 1); }
    }
  }
  *s_out = s; *t_out = t;
}

static void capsule_ends(const double* x, const double* R, double h, double p0[3], double d[3]) {
  double ax[3] = {R[2], R[5], R[8]};
  for (int i = 0; i < 3; i++) { p0[i] = x[i] - ax[i] * h; d[i] = 2 * h * ax[i]; }
}

/* sphere centre vs box (box frame x/R, half sizes s): one contact; normal points
 * from the sphere towards the box */
static int sphere_box(const double c[3], double r, const double bx[3], const double R[9], const double s[3], double margin, rawc_t* out) {
  double rel[3] = {c[0] - bx[0], c[1] - bx[1], c[2] - bx[2]}, pl[3];
  mulmatTvec3(R, rel, pl);
  double q[3];
  int inside = 1;
  for (int i = 0; i < 3; i++) {
    q[i] = clampd(pl[i], -s[i], s[i]);
    if (pl[i] < -s[i] || pl[i] > s[i]) inside = 0;
  }
  double nl[3], dist;
  if (!inside) {
    double dv[3] = {pl[0] - q[0], pl[1] - q[1], pl[2] - q[2]};
    double L = sqrt(dot3(dv, dv));
    dist = L - r;
    if (dist > margin) return 0;
    for (int i = 0; i < 3; i++) nl[i] = -dv[i] / L;  /* towards the box */
  } else {
    int k = 0;
    double best = s[0] - fabs(pl[0]);
    for (int i = 1; i < 3; i++) { double dp = s[i] - fabs(pl[i]); if (dp < best) { best = dp; k = i; } }
    dist = -best - r;
    if (dist > margin) return 0;
    nl[0] = nl[1] = nl[2] = 0;
    nl[k] = pl[k] >= 0 ? -1.0 : 1.0;
  }
  double n[3];
  mulmatvec3(R, nl, n);
  out->dist = dist;
  for (int i = 0; i < 3; i++) { out->n[i] = n[i]; out->pos[i] = c[i] + n[i] * (r + 0.5 * dist); }
  return 1;
}

/* signed distance of a point in the box frame (negative inside) */
static double box_sdf(const double p[3], const double h[3]) {
  double q[3], o[3];
  for (int i = 0; i < 3; i++) { q[i] = fabs(p[i]) - h[i]; o[i] = q[i] > 0 ? q[i] : 0; }
  double in = q[0];
  if (q[1] > in) in = q[1];
  if (q[2] > in) in = q[2];
  return sqrt(dot3(o, o)) + (in < 0 ? in : 0);
}

/* capsule (geom1) vs box (geom2), up to 2 contacts: spheres on the capsule
 * axis at the minimiser of the (convex) box signed distance along the axis,
 * and at the ends when they are within the margin and apart from it; both
 * ends when both touch and the interior is no deeper. The minimiser is found
 * by golden-section search. */
static int capsule_box(const double* xc, const double* Rc, const double* sc, const double* xb, const double* Rb,
                       const double* sb, double margin, rawc_t* out) {
  double p0[3], d[3], l0[3], ld[3], rel[3];
  capsule_ends(xc, Rc, sc[1], p0, d);
  for (int i = 0; i < 3; i++) rel[i] = p0[i] - xb[i];
  mulmatTvec3(Rb, rel, l0);
  mulmatTvec3(Rb, d, ld);
  double r = sc[0];
#define ALONG(tt, q) do { for (int i_ = 0; i_ < 3; i_++) (q)[i_] = l0[i_] + (tt) * ld[i_]; } while (0)
  const double gr = 0.5 * (3.0 - sqrt(5.0));
  double a = 0, b = 1, t1 = a + gr * (b - a), t2 = b - gr * (b - a), q[3];
  ALONG(t1, q); double f1 = box_sdf(q, sb);
  ALONG(t2, q); double f2 = box_sdf(q, sb);
  for (int it = 0; it < 80; it++) {
    if (f1 <= f2) { b = t2; t2 = t1; f2 = f1; t1 = a + gr * (b - a); ALONG(t1, q); f1 = box_sdf(q, sb); }
    else { a = t1; t1 = t2; f1 = f2; t2 = b - gr * (b - a); ALONG(t2, q); f2 = box_sdf(q, sb); }
  }
  double ts = 0.5 * (a + b), fs, fa, fb;
  ALONG(ts, q); fs = box_sdf(q, sb);
  ALONG(0.0, q); fa = box_sdf(q, sb);
  ALONG(1.0, q); fb = box_sdf(q, sb);
#undef ALONG
  if (fs - r > margin) return 0;
  double tc[2];
  int nc = 0;
  double fe = fa <= fb ? fa : fb;
  if (fa - r <= margin && fb - r <= margin && fs >= fe - 1e-6) { tc[0] = 0; tc[1] = 1; nc = 2; }
  else {
    tc[nc++] = ts;
    double te = fa <= fb ? 0.0 : 1.0;
    if (fe - r <= margin && fabs(te - ts) > 0.1) tc[nc++] = te;
  }
  int n = 0;
  for (int k = 0; k < nc; k++) {
    double c[3] = {p0[0] + tc[k] * d[0], p0[1] + tc[k] * d[1], p0[2] + tc[k] * d[2]};
    n += sphere_box(c, r, xb, Rb, sb, margin, out + n);
  }
  return n;
}

/* box (geom1) vs box (geom2) by separating axes, up to 4 contacts.
 * Face axis (preferred unless an edge axis separates more by the tolerance
 * sep_e > 0.95 sep_f + 0.01 h_min): clip the incident face against the
 * reference face's side planes, keep points within the margin below it;
 * more than 4 are reduced to the deepest, the farthest from it, and the
 * points of largest positive and negative area across that chord. Edge
 * axis: one contact at the closest points of the supporting edges. */
static int box_box(const double* x1, const double* R1, const double* h1, const double* x2, const double* R2,
                   const double* h2, double margin, rawc_t* out) {
  double C[3][3], t[3], dw[3];
  for (int i = 0; i < 3; i++) dw[i] = x2[i] - x1[i];
  mulmatTvec3(R1, dw, t);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i][j] = R1[i] * R2[j] + R1[3 + i] * R2[3 + j] + R1[6 + i] * R2[6 + j];
  double bf = -1e300, be = -1e300, eax[3] = {0, 0, 0};
  int face = -1, edge = -1;
  for (int i = 0; i < 3; i++) {
    double sep = fabs(t[i]) - h1[i] - (h2[0] * fabs(C[i][0]) + h2[1] * fabs(C[i][1]) + h2[2] * fabs(C[i][2]));
    if (sep > margin) return 0;
    if (sep > bf) { bf = sep; face = i; }
  }
  for (int j = 0; j < 3; j++) {
    double tb = t[0] * C[0][j] + t[1] * C[1][j] + t[2] * C[2][j];
    double sep = fabs(tb) - h2[j] - (h1[0] * fabs(C[0][j]) + h1[1] * fabs(C[1][j]) + h1[2] * fabs(C[2][j]));
    if (sep > margin) return 0;
    if (sep > bf) { bf = sep; face = 3 + j; }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double ai[3] = {0, 0, 0}, bj[3] = {C[0][j], C[1][j], C[2][j]}, L[3];
      ai[i] = 1;
      cross3(ai, bj, L);
      double ln = sqrt(dot3(L, L));
      if (ln < 1e-5) continue;
      for (int k = 0; k < 3; k++) L[k] /= ln;
      double lb[3];
      for (int k = 0; k < 3; k++) lb[k] = L[0] * C[0][k] + L[1] * C[1][k] + L[2] * C[2][k];
      double sep = fabs(dot3(t, L)) - (h1[0] * fabs(L[0]) + h1[1] * fabs(L[1]) + h1[2] * fabs(L[2]))
                   - (h2[0] * fabs(lb[0]) + h2[1] * fabs(lb[1]) + h2[2] * fabs(lb[2]));
      if (sep > margin) return 0;
      if (sep > be) {
        be = sep; edge = 3 * i + j;
        double sg = dot3(t, L) < 0 ? -1 : 1;
        for (int k = 0; k < 3; k++) eax[k] = sg * L[k];
      }
    }
  double hmin = h1[0];
  for (int k = 0; k < 3; k++) { if (h1[k] < hmin) hmin = h1[k]; if (h2[k] < hmin) hmin = h2[k]; }
  if (edge >= 0 && be > 0.95 * bf + 0.01 * hmin) {
    int i = edge / 3, j = edge % 3;
    double p1[3], e1[3] = {0, 0, 0}, p2[3], e2[3], c2[3] = {t[0], t[1], t[2]};
    for (int k = 0; k < 3; k++) p1[k] = eax[k] >= 0 ? h1[k] : -h1[k];
    p1[i] = -h1[i];
    e1[i] = 2 * h1[i];
    for (int k = 0; k < 3; k++) {
      if (k == j) continue;
      double bk[3] = {C[0][k], C[1][k], C[2][k]};
      double sg = dot3(bk, eax) > 0 ? -1 : 1;
      for (int q = 0; q < 3; q++) c2[q] += sg * h2[k] * bk[q];
    }
    for (int q = 0; q < 3; q++) { p2[q] = c2[q] - h2[j] * C[q][j]; e2[q] = 2 * h2[j] * C[q][j]; }
    double s, u, m1[3], mw[3], nw[3];
    seg_seg(p1, e1, p2, e2, &s, &u);
    for (int q = 0; q < 3; q++) m1[q] = 0.5 * (p1[q] + s * e1[q] + p2[q] + u * e2[q]);
    mulmatvec3(R1, m1, mw);
    mulmatvec3(R1, eax, nw);
    out->dist = be;
    for (int q = 0; q < 3; q++) { out->pos[q] = mw[q] + x1[q]; out->n[q] = nw[q]; }
    return 1;
  }
  int ref1 = face < 3, ax = ref1 ? face : face - 3;
  const double *xr = ref1 ? x1 : x2, *Rr = ref1 ? R1 : R2, *hr = ref1 ? h1 : h2;
  const double *xi = ref1 ? x2 : x1, *Ri = ref1 ? R2 : R1, *hi = ref1 ? h2 : h1;
  double nr[3] = {Rr[ax], Rr[3 + ax], Rr[6 + ax]}, dri[3];
  for (int q = 0; q < 3; q++) dri[q] = xi[q] - xr[q];
  if (dot3(nr, dri) < 0) for (int q = 0; q < 3; q++) nr[q] = -nr[q];
  int fi = 0;
  double best = -1, fsg = 1;
  for (int k = 0; k < 3; k++) {
    double dk = Ri[k] * nr[0] + Ri[3 + k] * nr[1] + Ri[6 + k] * nr[2];
    if (fabs(dk) > best) { best = fabs(dk); fi = k; fsg = dk > 0 ? -1 : 1; }
  }
  double poly[8][3], tmp[8][3];
  int np = 4;
  int u = (fi + 1) % 3, v = (fi + 2) % 3;
  const double su[4] = {-1, 1, 1, -1}, sv[4] = {-1, -1, 1, 1};
  for (int c = 0; c < 4; c++) {
    double pl[3] = {0, 0, 0}, pw[3];
    pl[fi] = fsg * hi[fi]; pl[u] = su[c] * hi[u]; pl[v] = sv[c] * hi[v];
    mulmatvec3(Ri, pl, pw);
    for (int q = 0; q < 3; q++) pw[q] += xi[q] - xr[q];
    mulmatTvec3(Rr, pw, poly[c]);
  }
  for (int k = 0; k < 3; k++) {
    if (k == ax) continue;
    for (int sg = -1; sg <= 1; sg += 2) {
      int nn = 0;
      for (int c = 0; c < np; c++) {
        const double *A = poly[c], *Bq = poly[(c + 1) % np];
        double da = sg * A[k] - hr[k], db = sg * Bq[k] - hr[k];
        if (da <= 0 && nn < 8) { memcpy(tmp[nn], A, sizeof(tmp[nn])); nn++; }
        if (((da < 0) != (db < 0)) && nn < 8) {
          double tt = da / (da - db);
          for (int q = 0; q < 3; q++) tmp[nn][q] = A[q] + tt * (Bq[q] - A[q]);
          nn++;
        }
      }
      np = nn;
      memcpy(poly, tmp, sizeof(double) * 3 * (size_t)np);
    }
  }
  double nl[3];
  mulmatTvec3(Rr, nr, nl);
  double sgn = nl[ax] >= 0 ? 1 : -1;
  double dep[8];
  int keep[8], nk = 0;
  for (int c = 0; c < np; c++) {
    dep[c] = sgn * poly[c][ax] - hr[ax];
    if (dep[c] <= margin) keep[nk++] = c;
  }
  if (!nk) return 0;
  int sel[4], ns = 0;
  if (nk <= 4) for (int c = 0; c < nk; c++) sel[ns++] = keep[c];
  else {
    int i0 = keep[0], i1 = -1, i2 = -1, i3 = -1;
    for (int c = 1; c < nk; c++) if (dep[keep[c]] < dep[i0]) i0 = keep[c];
    double dbest = -1, amax = 0, amin = 0, e[3];
    for (int c = 0; c < nk; c++) {
      double dd[3];
      for (int q = 0; q < 3; q++) dd[q] = poly[keep[c]][q] - poly[i0][q];
      if (dot3(dd, dd) > dbest) { dbest = dot3(dd, dd); i1 = keep[c]; }
    }
    for (int q = 0; q < 3; q++) e[q] = poly[i1][q] - poly[i0][q];
    for (int c = 0; c < nk; c++) {
      double f[3], cr[3];
      for (int q = 0; q < 3; q++) f[q] = poly[keep[c]][q] - poly[i0][q];
      cross3(e, f, cr);
      double area = cr[ax] * sgn;
      if (area > amax) { amax = area; i2 = keep[c]; }
      if (area < amin) { amin = area; i3 = keep[c]; }
    }
    sel[ns++] = i0; sel[ns++] = i1;
    if (i2 >= 0) sel[ns++] = i2;
    if (i3 >= 0) sel[ns++] = i3;
  }
  for (int c = 0; c < ns; c++) {
    double mid[3] = {poly[sel[c]][0], poly[sel[c]][1], poly[sel[c]][2]}, mw[3];
    mid[ax] -= 0.5 * sgn * dep[sel[c]];
    mulmatvec3(Rr, mid, mw);
    out[c].dist = dep[sel[c]];
    for (int q = 0; q < 3; q++) { out[c].pos[q] = mw[q] + xr[q]; out[c].n[q] = ref1 ? nr[q] : -nr[q]; }
  }
  return ns;
}

/* ------------------------------------------------------------------ convex primitives: GJK / EPA
 * Pairs involving an ellipsoid or a cylinder (and no plane) go through the
 * support functions of the two convex primitives: GJK (Gilbert, Johnson and
 * Keerthi 1988, with Ericson's Voronoi-region closest points on the simplex)
 * gives the distance of separated shapes; for overlapping shapes EPA (van den
 * Bergen 2001) expands the final GJK simplex into a polytope of the Minkowski
 * difference until its face nearest the origin is on the boundary. One
 * contact per pair; the normal points from geom1 to geom2. */
typedef struct { int type; const double *x, *R, *s; } cvx_t;
typedef struct { double w[3], a[3], b[3]; } svert_t;

static void cvx_support(const cvx_t* g, const double d[3], double out[3]) {
  double dl[3], pl[3] = {0, 0, 0};
  mulmatTvec3(g->R, d, dl);
  double n = sqrt(dot3(dl, dl));
  if (n < 1e-300) { dl[0] = 1; dl[1] = dl[2] = 0; n = 1; }
  const double* s = g->s;
  if (g->type == GSPHERE || g->type == GCAPSULE) {
    for (int i = 0; i < 3; i++) pl[i] = s[0] * dl[i] / n;
    if (g->type == GCAPSULE) pl[2] += dl[2] >= 0 ? s[1] : -s[1];
  } else if (g->type == GELLIPSOID) {
    double q[3], den = 0;
    for (int i = 0; i < 3; i++) { q[i] = s[i] * s[i] * dl[i]; den += s[i] * s[i] * dl[i] * dl[i]; }
    den = sqrt(den);
    for (int i = 0; i < 3; i++) pl[i] = q[i] / den;
  } else if (g->type == GCYLINDER) {
    double r = sqrt(dl[0] * dl[0] + dl[1] * dl[1]);
    if (r > 1e-300) { pl[0] = s[0] * dl[0] / r; pl[1] = s[0] * dl[1] / r; }
    pl[2] = dl[2] >= 0 ? s[1] : -s[1];
  } else {  /* box */
    for (int i = 0; i < 3; i++) pl[i] = dl[i] >= 0 ? s[i] : -s[i];
  }
  mulmatvec3(g->R, pl, out);
  for (int i = 0; i < 3; i++) out[i] += g->x[i];
}

/* support of the Minkowski difference A - B in direction d */
static void mink_support(const cvx_t* A, const cvx_t* B, const double d[3], svert_t* v) {
  double nd[3] = {-d[0], -d[1], -d[2]};
  cvx_support(A, d, v->a);
  cvx_support(B, nd, v->b);
  for (int i = 0; i < 3; i++) v->w[i] = v->a[i] - v->b[i];
}

/* barycentric coordinates of the point of triangle (a, b, c) nearest the origin */
static void tri_nearest(const double a[3], const double b[3], const double c[3], double l[3]) {
  double ab[3], ac[3], ap[3] = {-a[0], -a[1], -a[2]}, bp[3] = {-b[0], -b[1], -b[2]}, cp[3] = {-c[0], -c[1], -c[2]};
  for (int i = 0; i < 3; i++) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; }
  double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0 && d2 <= 0) { l[0] = 1; l[1] = l[2] = 0; return; }
  double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0 && d4 <= d3) { l[1] = 1; l[0] = l[2] = 0; return; }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0 && d1 >= 0 && d3 <= 0) { double v = d1 / (d1 - d3); l[0] = 1 - v; l[1] = v; l[2] = 0; return; }
  double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0 && d5 <= d6) { l[2] = 1; l[0] = l[1] = 0; return; }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0 && d2 >= 0 && d6 <= 0) { double w = d2 / (d2 - d6); l[0] = 1 - w; l[1] = 0; l[2] = w; return; }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
    double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    l[0] = 0; l[1] = 1 - w; l[2] = w; return;
  }
  double den = 1 / (va + vb + vc), v = vb * den, w = vc * den;
  l[0] = 1 - v - w; l[1] = v; l[2] = w;
}

/* nearest point of the simplex to the origin; the simplex shrinks to the
 * vertices with non-zero weight. Returns 1 when a tetrahedron holds the origin. */
static int simplex_nearest(svert_t* S, int* n, double lam[4], double v[3]) {
  double l[4] = {0, 0, 0, 0};
  if (*n == 1) { l[0] = 1; }
  else if (*n == 2) {
    double e[3], ee = 0, t = 0;
    for (int i = 0; i < 3; i++) { e[i] = S[1].w[i] - S[0].w[i]; ee += e[i] * e[i]; t -= S[0].w[i] * e[i]; }
    t = ee > 0 ? t / ee : 0;
    if (t <= 0) { l[0] = 1; } else if (t >= 1) { l[1] = 1; } else { l[0] = 1 - t; l[1] = t; }
  } else if (*n == 3) {
    tri_nearest(S[0].w, S[1].w, S[2].w, l);
  } else {
    static const int F[4][4] = {{0, 1, 2, 3}, {0, 1, 3, 2}, {0, 2, 3, 1}, {1, 2, 3, 0}};
    double best = 1e300;
    int outside = 0;
    for (int f = 0; f < 4; f++) {
      const double *a = S[F[f][0]].w, *b = S[F[f][1]].w, *c = S[F[f][2]].w, *o = S[F[f][3]].w;
      double ab[3], ac[3], nrm[3], ao[3];
      for (int i = 0; i < 3; i++) { ab[i] = b[i] - a[i]; ac[i] = c[i] - a[i]; ao[i] = o[i] - a[i]; }
      cross3(ab, ac, nrm);
      double so = -dot3(nrm, a), sl = dot3(nrm, ao);
      if (so * sl >= 0) continue;  /* origin on the same side as the opposite vertex */
      outside = 1;
      double tl[3], p[3];
      tri_nearest(a, b, c, tl);
      for (int i = 0; i < 3; i++) p[i] = tl[0] * a[i] + tl[1] * b[i] + tl[2] * c[i];
      double d = dot3(p, p);
      if (d < best) {
        best = d;
        l[0] = l[1] = l[2] = l[3] = 0;
        l[F[f][0]] = tl[0]; l[F[f][1]] = tl[1]; l[F[f][2]] = tl[2];
      }
    }
    if (!outside) {
      v[0] = v[1] = v[2] = 0;
      return 1;
    }
  }
  v[0] = v[1] = v[2] = 0;
  int k = 0;
  for (int i = 0; i < *n; i++) {
    if (l[i] <= 0) continue;
    for (int q = 0; q < 3; q++) v[q] += l[i] * S[i].w[q];
    S[k] = S[i];
    lam[k++] = l[i];
  }
  *n = k;
  return 0;
}

enum { EPA_MAXV = 64, EPA_MAXF = 128 };

/* GJK distance; on overlap, EPA depth. Returns the number of contacts (0/1). */
static int gjk_epa(const cvx_t* A, const cvx_t* B, double margin, rawc_t* out) {
  svert_t S[4];
  int n = 1;
  double d0[3] = {B->x[0] - A->x[0], B->x[1] - A->x[1], B->x[2] - A->x[2]};
  if (dot3(d0, d0) < 1e-24) { d0[0] = 1; d0[1] = d0[2] = 0; }
  mink_support(A, B, d0, &S[0]);
  double v[3] = {S[0].w[0], S[0].w[1], S[0].w[2]}, lam[4] = {1, 0, 0, 0};
  int overlap = 0;
  for (int it = 0; it < 100; it++) {
    double vv = dot3(v, v);
    if (vv < 1e-24) { overlap = 1; break; }
    svert_t w;
    double nv_[3] = {-v[0], -v[1], -v[2]};
    mink_support(A, B, nv_, &w);
    if (vv - dot3(v, w.w) <= 1e-12 * vv) break;  /* no progress: v is the nearest point */
    int dup = 0;
    for (int i = 0; i < n; i++)
      if (fabs(S[i].w[0] - w.w[0]) + fabs(S[i].w[1] - w.w[1]) + fabs(S[i].w[2] - w.w[2]) < 1e-14) dup = 1;
    if (dup) break;
    S[n++] = w;
    if (simplex_nearest(S, &n, lam, v)) { overlap = 1; break; }
  }
  if (!overlap) {
    double pa[3] = {0, 0, 0}, pb[3] = {0, 0, 0};
    for (int i = 0; i < n; i++)
      for (int q = 0; q < 3; q++) { pa[q] += lam[i] * S[i].a[q]; pb[q] += lam[i] * S[i].b[q]; }
    double dv[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
    double dist = sqrt(dot3(dv, dv));
    if (dist > margin) return 0;
    out->dist = dist;
    for (int q = 0; q < 3; q++) { out->n[q] = dv[q] / dist; out->pos[q] = 0.5 * (pa[q] + pb[q]); }
    return 1;
  }
  /* EPA: a tetrahedron around the origin from the final simplex */
  svert_t V[EPA_MAXV];
  int nvtx = n;
  for (int i = 0; i < n; i++) V[i] = S[i];
  static const double ax[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int k = 0; nvtx < 4 && k < 6; k++) {
    double dir[3] = {ax[k][0], ax[k][1], ax[k][2]};
    if (nvtx == 2) {
      double e[3] = {V[1].w[0] - V[0].w[0], V[1].w[1] - V[0].w[1], V[1].w[2] - V[0].w[2]};
      cross3(e, ax[k], dir);
    } else if (nvtx == 3) {
      double e1[3], e2[3];
      for (int q = 0; q < 3; q++) { e1[q] = V[1].w[q] - V[0].w[q]; e2[q] = V[2].w[q] - V[0].w[q]; }
      cross3(e1, e2, dir);
      if (k & 1) for (int q = 0; q < 3; q++) dir[q] = -dir[q];
    }
    if (dot3(dir, dir) < 1e-24) continue;
    svert_t w;
    mink_support(A, B, dir, &w);
    /* accept only a point that adds a dimension */
    double ok = 0;
    if (nvtx == 1) { double e[3] = {w.w[0] - V[0].w[0], w.w[1] - V[0].w[1], w.w[2] - V[0].w[2]}; ok = dot3(e, e); }
    else if (nvtx == 2) {
      double e1[3], e2[3], c[3];
      for (int q = 0; q < 3; q++) { e1[q] = V[1].w[q] - V[0].w[q]; e2[q] = w.w[q] - V[0].w[q]; }
      cross3(e1, e2, c);
      ok = dot3(c, c);
    } else {
      double e1[3], e2[3], e3[3], c[3];
      for (int q = 0; q < 3; q++) { e1[q] = V[1].w[q] - V[0].w[q]; e2[q] = V[2].w[q] - V[0].w[q]; e3[q] = w.w[q] - V[0].w[q]; }
      cross3(e1, e2, c);
      ok = fabs(dot3(c, e3));
    }
    if (ok > 1e-18) V[nvtx++] = w;
  }
  if (nvtx < 4) {  /* degenerate: touching */
    out->dist = 0;
    double dn = sqrt(dot3(d0, d0));
    for (int q = 0; q < 3; q++) { out->n[q] = d0[q] / dn; out->pos[q] = 0.5 * (V[0].a[q] + V[0].b[q]); }
    return margin >= 0 ? 1 : 0;
  }
  int F[EPA_MAXF][3], nf = 0;
  double Fn[EPA_MAXF][3], Fd[EPA_MAXF];
  /* faces keep their winding outward: the initial ones are oriented away from
   * the opposite vertex, a new face takes the winding of the removed face
   * whose horizon edge it closes */
#define EPA_FACE(i0, i1, i2)                                                              \
  do {                                                                                    \
    int a_ = (i0), b_ = (i1), c_ = (i2);                                                  \
    double e1_[3], e2_[3], n_[3];                                                         \
    for (int q = 0; q < 3; q++) { e1_[q] = V[b_].w[q] - V[a_].w[q]; e2_[q] = V[c_].w[q] - V[a_].w[q]; } \
    cross3(e1_, e2_, n_);                                                                 \
    double l_ = sqrt(dot3(n_, n_));                                                       \
    if (l_ > 1e-300) for (int q = 0; q < 3; q++) n_[q] /= l_;                            \
    F[nf][0] = a_; F[nf][1] = b_; F[nf][2] = c_;                                          \
    for (int q = 0; q < 3; q++) Fn[nf][q] = n_[q];                                        \
    Fd[nf] = dot3(n_, V[a_].w); nf++;                                                     \
  } while (0)
  {
    static const int T[4][4] = {{0, 1, 2, 3}, {0, 3, 1, 2}, {0, 2, 3, 1}, {1, 3, 2, 0}};
    for (int f = 0; f < 4; f++) {
      int a = T[f][0], b = T[f][1], c = T[f][2], o = T[f][3];
      double e1[3], e2[3], nn[3], eo[3];
      for (int q = 0; q < 3; q++) { e1[q] = V[b].w[q] - V[a].w[q]; e2[q] = V[c].w[q] - V[a].w[q]; eo[q] = V[o].w[q] - V[a].w[q]; }
      cross3(e1, e2, nn);
      if (dot3(nn, eo) > 0) EPA_FACE(a, c, b); else EPA_FACE(a, b, c);
    }
  }
  int best = 0;
  for (int it = 0; it < 64; it++) {
    best = 0;
    for (int f = 1; f < nf; f++) if (Fd[f] < Fd[best]) best = f;
    svert_t w;
    mink_support(A, B, Fn[best], &w);
    double gap = dot3(Fn[best], w.w) - Fd[best];
    if (gap <= 1e-10 * (1 + Fd[best]) || nvtx >= EPA_MAXV) break;
    /* remove the faces w sees, keep their unshared edges (the horizon) */
    int E[3 * EPA_MAXF][2], ne = 0;
    int keep = 0;
    for (int f = 0; f < nf; f++) {
      double rel[3] = {w.w[0] - V[F[f][0]].w[0], w.w[1] - V[F[f][0]].w[1], w.w[2] - V[F[f][0]].w[2]};
      if (dot3(Fn[f], rel) > 1e-14) {
        for (int e = 0; e < 3; e++) {
          int a = F[f][e], b = F[f][(e + 1) % 3], found = -1;
          for (int k = 0; k < ne; k++) if (E[k][0] == b && E[k][1] == a) { found = k; break; }
          if (found >= 0) { E[found][0] = E[ne - 1][0]; E[found][1] = E[ne - 1][1]; ne--; }
          else { E[ne][0] = a; E[ne][1] = b; ne++; }
        }
      } else {
        if (keep != f) {
          memcpy(F[keep], F[f], sizeof(F[f])); memcpy(Fn[keep], Fn[f], sizeof(Fn[f])); Fd[keep] = Fd[f];
        }
        keep++;
      }
    }
    nf = keep;
    if (nf + ne > EPA_MAXF) break;
    V[nvtx] = w;
    for (int k = 0; k < ne; k++) EPA_FACE(E[k][0], E[k][1], nvtx);
    nvtx++;
  }
#undef EPA_FACE
  for (int f = 1; f < nf; f++) if (Fd[f] < Fd[best]) best = f;
  double depth = Fd[best], p[3] = {Fn[best][0] * depth, Fn[best][1] * depth, Fn[best][2] * depth};
  /* barycentric coordinates of p on the face */
  const double *a = V[F[best][0]].w, *b = V[F[best][1]].w, *c = V[F[best][2]].w;
  double v0[3], v1[3], v2[3];
  for (int q = 0; q < 3; q++) { v0[q] = b[q] - a[q]; v1[q] = c[q] - a[q]; v2[q] = p[q] - a[q]; }
  double d00 = dot3(v0, v0), d01 = dot3(v0, v1), d11 = dot3(v1, v1), d20 = dot3(v2, v0), d21 = dot3(v2, v1);
  double den = d00 * d11 - d01 * d01;
  double lv = den != 0 ? (d11 * d20 - d01 * d21) / den : 0, lw = den != 0 ? (d00 * d21 - d01 * d20) / den : 0, lu = 1 - lv - lw;
  double pa[3], pb[3];
  for (int q = 0; q < 3; q++) {
    pa[q] = lu * V[F[best][0]].a[q] + lv * V[F[best][1]].a[q] + lw * V[F[best][2]].a[q];
    pb[q] = lu * V[F[best][0]].b[q] + lv * V[F[best][1]].b[q] + lw * V[F[best][2]].b[q];
  }
  /* the face distance bounds the depth from below with a first-order error in
   * the normal; the supports along the normal give the second-order value
   * h_A(n) + h_B(-n), and the points move onto the two supporting planes */
  svert_t sw;
  mink_support(A, B, Fn[best], &sw);
  double ha = dot3(Fn[best], sw.a), hb = dot3(Fn[best], sw.b);
  double ea = ha - dot3(Fn[best], pa), eb = hb - dot3(Fn[best], pb);
  for (int q = 0; q < 3; q++) { pa[q] += ea * Fn[best][q]; pb[q] += eb * Fn[best][q]; }
  depth = ha - hb;
  out->dist = -depth;
  for (int q = 0; q < 3; q++) { out->n[q] = Fn[best][q]; out->pos[q] = 0.5 * (pa[q] + pb[q]); }
  return 1;
}

/* narrow phase for one pair; geom1 has the lower type. Returns the number of
 * raw contacts written (normal from geom1 to geom2). */
static int narrowphase(int t1, int t2, const double* s1, const double* s2, const double* x1, const double* R1,
                       const double* x2, const double* R2, double margin, rawc_t* out, int maxc) {
  (void)maxc;
  if (t1 == GPLANE) {
    double pn[3] = {R1[2], R1[5], R1[8]};
    if (t2 == GSPHERE) return plane_point(x1, pn, x2, s2[0], margin, out);
    if (t2 == GELLIPSOID || t2 == GCYLINDER) {  /* the deepest support point */
      cvx_t g = {t2, x2, R2, s2};
      double dn[3] = {-pn[0], -pn[1], -pn[2]}, sp[3];
      cvx_support(&g, dn, sp);
      return plane_point(x1, pn, sp, 0.0, margin, out);
    }
    if (t2 == GCAPSULE) {
      double p0[3], dd[3];
      capsule_ends(x2, R2, s2[1], p0, dd);
      double pA[3] = {p0[0] + dd[0], p0[1] + dd[1], p0[2] + dd[2]};
      int n = 0;
      n += plane_point(x1, pn, pA, s2[0], margin, out + n);
      n += plane_point(x1, pn, p0, s2[0], margin, out + n);
      return n;
    }
    if (t2 == GBOX) {
      int n = 0;
      for (int k = 0; k < 8 && n < 4; k++) {
        double cl[3] = {(k & 1) ? s2[0] : -s2[0], (k & 2) ? s2[1] : -s2[1], (k & 4) ? s2[2] : -s2[2]}, cw[3];
        mulmatvec3(R2, cl, cw);
        for (int i = 0; i < 3; i++) cw[i] += x2[i];
        n += plane_point(x1, pn, cw, 0.0, margin, out + n);
      }
      return n;
    }
    return -1;
  }
  if (t1 == GELLIPSOID || t1 == GCYLINDER || t2 == GELLIPSOID || t2 == GCYLINDER) {
    cvx_t A = {t1, x1, R1, s1}, B = {t2, x2, R2, s2};
    return gjk_epa(&A, &B, margin, out);
  }
  if (t1 == GSPHERE) {
    if (t2 == GSPHERE) return sphere_sphere(x1, s1[0], x2, s2[0], margin, out);
    if (t2 == GCAPSULE) {
      double p0[3], dd[3];
      capsule_ends(x2, R2, s2[1], p0, dd);
      double rel[3] = {x1[0] - p0[0], x1[1] - p0[1], x1[2] - p0[2]};
      double a = dot3(dd, dd);
      double t = a > MINVAL ? clampd(dot3(rel, dd) / a, 0, 1) : 0;
      double c2[3] = {p0[0] + t * dd[0], p0[1] + t * dd[1], p0[2] + t * dd[2]};
      return sphere_sphere(x1, s1[0], c2, s2[0], margin, out);
    }
    if (t2 == GBOX) return sphere_box(x1, s1[0], x2, R2, s2, margin, out);
    return -1;
  }
  if (t1 == GCAPSULE && t2 == GBOX) return capsule_box(x1, R1, s1, x2, R2, s2, margin, out);
  if (t1 == GBOX && t2 == GBOX) return box_box(x1, R1, s1, x2, R2, s2, margin, out);
  if (t1 == GCAPSULE && t2 == GCAPSULE) {
    double a0[3], ad[3], b0[3], bd[3], s, t;
    capsule_ends(x1, R1, s1[1], a0, ad);
    capsule_ends(x2, R2, s2[1], b0, bd);
    seg_seg(a0, ad, b0, bd, &s, &t);
    double c1[3] = {a0[0] + s * ad[0], a0[1] + s * ad[1], a0[2] + s * ad[2]};
    double c2[3] = {b0[0] + t * bd[0], b0[1] + t * bd[1], b0[2] + t * bd[2]};
    return sphere_sphere(c1, s1[0], c2, s2[0], margin, out);
  }
  return -1;
}

/* Replace the narrow phase of the following forwards by a given contact list
 * (pair index, distance, position, normal from geom1 to geom2 per contact; the
 * frame and the pair's friction, solref, solimp, condim and margin as the
 * narrow phase would set them); n < 0 restores the narrow phase. Used to
 * solve another implementation's contact set in FP64. */
void orc_inject_contacts(const orc_prep_t* p, orc_data_t* d, int n, const int* pair, const double* dist,
                         const double* pos, const double* normal) {
  free(d->inject);
  d->inject = NULL;
  d->ninject = n < 0 ? -1 : n;
  if (n <= 0) return;
  d->inject = (orc_contact_t*)calloc((size_t)n, sizeof(orc_contact_t));
  for (int k = 0; k < n; k++) {
    const orc_pair_t* pr = &p->pairs[pair[k]];
    orc_contact_t* c = &d->inject[k];
    c->dist = dist[k];
    memcpy(c->pos, pos + 3 * k, 3 * sizeof(double));
    make_frame(normal + 3 * k, c->frame);
    c->geom1 = pr->g1; c->geom2 = pr->g2; c->pair = pair[k]; c->dim = pr->condim;
    memcpy(c->friction, pr->friction, 5 * sizeof(double));
    memcpy(c->solref, pr->solref, 2 * sizeof(double));
    memcpy(c->solimp, pr->solimp, 5 * sizeof(double));
    c->includemargin = pr->margin - pr->gap;
  }
}

int orc_collide_pair(const orc_prep_t* p, int pi, const double* xpos1, const double* xmat1, const double* xpos2,
                     const double* xmat2, orc_contact_t* out, int maxc) {
  const orc_model_t* m = &p->m;
  const orc_pair_t* pr = &p->pairs[pi];
  rawc_t raw[8];
  int n = narrowphase(m->geom_type[pr->g1], m->geom_type[pr->g2], m->geom_size + 3 * pr->g1, m->geom_size + 3 * pr->g2,
                      xpos1, xmat1, xpos2, xmat2, pr->margin, raw, 8);
  if (n < 0) return n;
  if (n > maxc) n = maxc;
  for (int k = 0; k < n; k++) {
    orc_contact_t* c = &out[k];
    memset(c, 0, sizeof(*c));
    c->dist = raw[k].dist;
    memcpy(c->pos, raw[k].pos, 3 * sizeof(double));
    make_frame(raw[k].n, c->frame);
    c->geom1 = pr->g1; c->geom2 = pr->g2; c->pair = pi; c->dim = pr->condim;
    memcpy(c->friction, pr->friction, 5 * sizeof(double));
    memcpy(c->solref, pr->solref, 2 * sizeof(double));
    memcpy(c->solimp, pr->solimp, 5 * sizeof(double));
    c->includemargin = pr->margin - pr->gap;
  }
  return n;
}

/* mj_collision over the static pair table with the bounding-sphere cull of
 * reference collision.hpp:301-305 (planes are never culled). Contacts are
 * ordered by pair index, then by the narrow phase's own order. */
static int collision(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  d->ncon = 0;
  int status = 0;
  if (m->disableflags & (DS_CONTACT | DS_CONSTRAINT)) return 0;
  if (d->ninject >= 0) {
    for (int k = 0; k < d->ninject && k < ORC_MAXCON; k++) {
      d->con[d->ncon] = d->inject[k];
      if (d->has_rand)
        for (int i = 0; i < 5; i++) d->con[d->ncon].friction[i] *= d->rand_fric;
      d->ncon++;
    }
    return d->ninject > ORC_MAXCON ? 1 : 0;
  }
  for (int pi = 0; pi < p->npair; pi++) {
    const orc_pair_t* pr = &p->pairs[pi];
    const double *x1 = d->geom_xpos + 3 * pr->g1, *x2 = d->geom_xpos + 3 * pr->g2;
    double rb1 = m->geom_rbound[pr->g1], rb2 = m->geom_rbound[pr->g2];
    if (rb1 > 0 && rb2 > 0) {
      double dv[3] = {x1[0] - x2[0], x1[1] - x2[1], x1[2] - x2[2]};
      if (sqrt(dot3(dv, dv)) > rb1 + rb2 + pr->margin) continue;
    }
    orc_contact_t tmp[8];
    int n = orc_collide_pair(p, pi, x1, d->geom_xmat + 9 * pr->g1, x2, d->geom_xmat + 9 * pr->g2, tmp, 8);
    if (n < 0) { status |= 4; continue; }
    for (int k = 0; k < n; k++) {
      if (d->ncon >= (p->con_cap > 0 && p->con_cap < ORC_MAXCON ? p->con_cap : ORC_MAXCON)) { status |= 1; break; }
      if (d->has_rand) /* reference solver.hpp:435-436,530: every friction coefficient scaled */
        for (int i = 0; i < 5; i++) tmp[k].friction[i] *= d->rand_fric;
      d->con[d->ncon++] = tmp[k];
    }
  }
  return status;
}

/* ------------------------------------------------------------------ constraints */
/* translational / rotational Jacobian of a world point moving with body b
 * (mj_jac from cdof about the root subtree COM; reference solver.hpp:474-508) */
static void jac_point(const orc_prep_t* p, const orc_data_t* d, int b, const double pt[3], double* jp, double* jr, double sgn) {
  const orc_model_t* m = &p->m;
  int nv = m->nv;
  int bb = b;
  while (bb > 0 && m->body_dofnum[bb] == 0) bb = m->body_parentid[bb];
  if (bb == 0) return;
  const double* c = d->subtree_com + 3 * m->body_rootid[b];
  double off[3] = {pt[0] - c[0], pt[1] - c[1], pt[2] - c[2]};
  for (int k = m->body_dofadr[bb] + m->body_dofnum[bb] - 1; k >= 0; k = m->dof_parentid[k]) {
    const double* cd = d->cdof + 6 * k;
    double t[3];
    cross3(cd, off, t);
    for (int i = 0; i < 3; i++) {
      jp[i * nv + k] += sgn * (cd[3 + i] + t[i]);
      jr[i * nv + k] += sgn * cd[i];
    }
  }
}

/* MuJoCo impedance: solimp sigmoid of |pos - margin| / width, clamped to
 * [dmin, dmax] (reference solver.hpp:120-139) */
static double impedance(const double si[5], double x) {
  double dmin = clampd(si[0], MINIMP, MAXIMP), dmax = clampd(si[1], MINIMP, MAXIMP);
  double width = si[2], mid = clampd(si[3], MINIMP, MAXIMP), power = si[4] < 1 ? 1 : si[4];
  if (dmin == dmax || width <= MINVAL) return dmax;
  double xx = fabs(x) / width;
  if (xx >= 1) return dmax;
  double y;
  if (power == 1) y = xx;
  else if (xx <= mid) y = pow(xx, power) / pow(mid, power - 1);
  else y = 1 - pow(1 - xx, power) / pow(1 - mid, power - 1);
  return dmin + y * (dmax - dmin);
}

static int add_row(orc_data_t* d, int nv, int type, int id, double pos, double margin, double dA, double floss) {
  if (d->nefc >= ORC_MAXEFC) return -1;
  int r = d->nefc++;
  memset(d->efc_J + (size_t)r * nv, 0, sizeof(double) * nv);
  d->efc_type[r] = type; d->efc_id[r] = id; d->efc_pos[r] = pos; d->efc_margin[r] = margin;
  d->efc_diagApprox[r] = dA; d->efc_floss[r] = floss;
  return r;
}

/* mj_makeConstraint + mj_makeImpedance + mj_referenceConstraint for dof
 * friction, joint limits and contacts (pyramidal / frictionless / elliptic
 * rows). Row order: friction, limits, contacts. Per-row k, b, imp, R and aref
 * follow reference solver.hpp:120-139,346-371,510-587; pyramidal rows share
 * R = 2 mu^2 / impratio * R_first with R_first from diagApprox = tran*(1+mu^2)
 * (reference solver.hpp:519-525). */
static int make_constraint(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  int nv = m->nv;
  int status = 0;
  d->nefc = 0;
  if (m->disableflags & DS_CONSTRAINT) return 0;
  /* solref / solimp per row, kept in local arrays */
  static __thread double sref[ORC_MAXEFC][2], simp[ORC_MAXEFC][5];
  if (!(m->disableflags & DS_FRICTIONLOSS)) {
    for (int k = 0; k < nv; k++) {
      if (m->dof_frictionloss[k] <= 0) continue;
      if (p->efc_cap > 0 && d->nefc >= p->efc_cap - 4) { status |= 2; break; }
      int r = add_row(d, nv, EFC_FRICTION_DOF, k, 0, 0, p->dof_invweight0[k], m->dof_frictionloss[k]);
      if (r < 0) { status |= 2; break; }
      d->efc_J[(size_t)r * nv + k] = 1;
      memcpy(sref[r], m->dof_solref + 2 * k, 2 * sizeof(double));
      memcpy(simp[r], m->dof_solimp + 5 * k, 5 * sizeof(double));
    }
  }
  if (!(m->disableflags & DS_LIMIT)) {
    for (int j = 0; j < m->njnt; j++) {
      int t = m->jnt_type[j];
      if (!m->jnt_limited[j] || (t != JHINGE && t != JSLIDE)) continue;
      int qa = m->jnt_qposadr[j], da = m->jnt_dofadr[j];
      double v = d->qpos[qa];
      for (int side = 0; side < 2; side++) {
        double dist = side == 0 ? v - m->jnt_range[2 * j] : m->jnt_range[2 * j + 1] - v;
        if (dist < m->jnt_margin[j]) {
          if (p->efc_cap > 0 && d->nefc >= p->efc_cap - 4) { status |= 2; break; }
          int r = add_row(d, nv, EFC_LIMIT_JOINT, j, dist, m->jnt_margin[j], p->dof_invweight0[da], 0);
          if (r < 0) { status |= 2; break; }
          d->efc_J[(size_t)r * nv + da] = side == 0 ? 1.0 : -1.0;
          memcpy(sref[r], m->jnt_solref + 2 * j, 2 * sizeof(double));
          memcpy(simp[r], m->jnt_solimp + 5 * j, 5 * sizeof(double));
        }
      }
    }
  }
  double* jac = (double*)calloc((size_t)6 * nv, sizeof(double));
  if (!(m->disableflags & DS_CONTACT)) {
    for (int ci = 0; ci < d->ncon; ci++) {
      orc_contact_t* c = &d->con[ci];
      if (!(c->dist < c->includemargin)) continue;
      if (p->efc_cap > 0 && d->nefc >= p->efc_cap - 10) { status |= 2; break; }
      const orc_pair_t* pr = &p->pairs[c->pair];
      memset(jac, 0, sizeof(double) * 6 * nv);
      jac_point(p, d, pr->b2, c->pos, jac, jac + 3 * nv, 1.0);
      jac_point(p, d, pr->b1, c->pos, jac, jac + 3 * nv, -1.0);
      /* rows of the contact frame: translational (n, t1, t2) and rotational (n, t1, t2) */
      double Jf[6][64 * 4];
      double* Jrow[6];
      double* big = NULL;
      if (nv > 64 * 4) { big = (double*)calloc((size_t)6 * nv, sizeof(double)); for (int i = 0; i < 6; i++) Jrow[i] = big + (size_t)i * nv; }
      else for (int i = 0; i < 6; i++) Jrow[i] = Jf[i];
      for (int a = 0; a < 3; a++) {
        const double* ax = c->frame + 3 * a;
        for (int k = 0; k < nv; k++) {
          Jrow[a][k] = ax[0] * jac[0 * nv + k] + ax[1] * jac[1 * nv + k] + ax[2] * jac[2 * nv + k];
          Jrow[3 + a][k] = ax[0] * jac[3 * nv + k] + ax[1] * jac[4 * nv + k] + ax[2] * jac[5 * nv + k];
        }
      }
      double tran = p->body_invweight0[2 * pr->b1] + p->body_invweight0[2 * pr->b2];
      int dim = c->dim;
      if (dim == 1) {
        int r = add_row(d, nv, EFC_CONTACT_FRICTIONLESS, ci, c->dist, c->includemargin, tran, 0);
        if (r < 0) { status |= 2; free(big); break; }
        memcpy(d->efc_J + (size_t)r * nv, Jrow[0], sizeof(double) * nv);
        memcpy(sref[r], c->solref, 2 * sizeof(double)); memcpy(simp[r], c->solimp, 5 * sizeof(double));
      } else if (m->cone == 0) {
        int nrow = 2 * (dim - 1);
        if (d->nefc + nrow > ORC_MAXEFC) { status |= 2; free(big); break; }
        double dA = tran * (1 + c->friction[0] * c->friction[0]);
        for (int k = 0; k < dim - 1; k++) {
          /* direction k: 0,1 tangents (translational); 2 spin (rotational normal); 3,4 rolling */
          const double* Jk = k < 2 ? Jrow[1 + k] : (k == 2 ? Jrow[3] : Jrow[4 + (k - 3)]);
          for (int sg = 0; sg < 2; sg++) {
            int r = add_row(d, nv, EFC_CONTACT_PYRAMIDAL, ci, c->dist, c->includemargin, dA, 0);
            double mu = c->friction[k] * (sg == 0 ? 1.0 : -1.0);
            for (int q = 0; q < nv; q++) d->efc_J[(size_t)r * nv + q] = Jrow[0][q] + mu * Jk[q];
            memcpy(sref[r], c->solref, 2 * sizeof(double)); memcpy(simp[r], c->solimp, 5 * sizeof(double));
          }
        }
      } else {
        /* elliptic: normal row, then t1, t2, spin, roll1, roll2 (dim - 1 of them);
         * friction rows carry no position term */
        if (d->nefc + dim > ORC_MAXEFC) { status |= 2; free(big); break; }
        const double* rows[6] = {Jrow[0], Jrow[1], Jrow[2], Jrow[3], Jrow[4], Jrow[5]};
        for (int k = 0; k < dim; k++) {
          int r = add_row(d, nv, EFC_CONTACT_ELLIPTIC, ci, k == 0 ? c->dist : 0.0, k == 0 ? c->includemargin : 0.0, tran, 0);
          memcpy(d->efc_J + (size_t)r * nv, rows[k], sizeof(double) * nv);
          memcpy(sref[r], c->solref, 2 * sizeof(double)); memcpy(simp[r], c->solimp, 5 * sizeof(double));
        }
      }
      free(big);
    }
  }
  free(jac);
  /* impedance, reference acceleration, regulariser */
  for (int r = 0; r < d->nefc; r++) {
    double v = 0;
    for (int k = 0; k < nv; k++) v += d->efc_J[(size_t)r * nv + k] * d->qvel[k];
    d->efc_vel[r] = v;
    double tc = sref[r][0], dr = sref[r][1];
    double dmax = clampd(simp[r][1], MINIMP, MAXIMP);
    double kk, bb;
    if (tc > 0) {
      if (!(m->disableflags & DS_REFSAFE)) tc = fmax(tc, 2 * m->timestep);
      kk = 1.0 / (dmax * dmax * tc * tc * dr * dr);
      bb = 2.0 / (dmax * tc);
    } else {
      kk = -tc / (dmax * dmax);
      bb = -dr / dmax;
    }
    double x = d->efc_pos[r] - d->efc_margin[r];
    double imp = impedance(simp[r], x);
    d->efc_imp[r] = imp;
    d->efc_aref[r] = -bb * v - kk * imp * x;
    d->efc_R[r] = fmax(d->efc_diagApprox[r] * (1 - imp) / imp, MINVAL);
  }
  /* pyramidal contacts: one R for all edges of a contact */
  for (int r = 0; r < d->nefc;) {
    if (d->efc_type[r] != EFC_CONTACT_PYRAMIDAL) { r++; continue; }
    const orc_contact_t* c = &d->con[d->efc_id[r]];
    int nrow = 2 * (c->dim - 1);
    double mu0 = c->friction[0];
    double Rpy = fmax(2 * mu0 * mu0 / m->impratio * d->efc_R[r], MINVAL);
    for (int k = 0; k < nrow; k++) d->efc_R[r + k] = Rpy;
    r += nrow;
  }
  /* elliptic contacts (MuJoCo mj_makeImpedance): R_1 = R_0 / impratio for the
   * first friction row, R_k = R_1 mu_0^2 / mu_(k-1)^2 for the others */
  for (int r = 0; r < d->nefc;) {
    if (d->efc_type[r] != EFC_CONTACT_ELLIPTIC) { r++; continue; }
    const orc_contact_t* c = &d->con[d->efc_id[r]];
    if (c->dim > 1) {
      double R1 = d->efc_R[r] / fmax(MINVAL, m->impratio);
      for (int k = 1; k < c->dim; k++)
        d->efc_R[r + k] = fmax(k == 1 ? R1 : R1 * c->friction[0] * c->friction[0] / (c->friction[k - 1] * c->friction[k - 1]), MINVAL);
    }
    r += c->dim;
  }
  for (int r = 0; r < d->nefc; r++) d->efc_D[r] = 1.0 / d->efc_R[r];
  return status;
}

/* Elliptic cone of one contact (rows r .. r+dim-1, row values x): with
 * U_0 = mu x_0, U_j = f_j x_j (f the contact's friction, mu = f_0),
 * N = U_0 and T = |U_1..|:
 *   top (N >= mu T, or T = 0 and N >= 0): no force, no cost;
 *   bottom (mu N + T <= 0, or T = 0 and N < 0): every row quadratic;
 *   middle: cost 1/2 Dm (N - mu T)^2, Dm = D_0 / (mu^2 (1 + mu^2))
 * (MuJoCo's primal elliptic cone). Returns the zone (0 top, 1 bottom, 2
 * middle) and writes the cost and, if asked, the forces and the dim x dim
 * Hessian of the cost in the row values. */
static int ellip_state(const orc_data_t* d, int r, int dim, const double* x, double* cost, double* force, double* Hc) {
  const orc_contact_t* c = &d->con[d->efc_id[r]];
  double mu = c->friction[0], U[6], T = 0;
  U[0] = mu * x[0];
  for (int j = 1; j < dim; j++) { U[j] = c->friction[j - 1] * x[j]; T += U[j] * U[j]; }
  T = sqrt(T);
  double N = U[0];
  if (Hc) memset(Hc, 0, sizeof(double) * dim * dim);
  if (N >= mu * T || (T <= 0 && N >= 0)) {
    *cost = 0;
    if (force) for (int j = 0; j < dim; j++) force[j] = 0;
    return 0;
  }
  if (mu * N + T <= 0 || (T <= 0 && N < 0)) {
    *cost = 0;
    for (int j = 0; j < dim; j++) {
      double D = d->efc_D[r + j];
      *cost += 0.5 * D * x[j] * x[j];
      if (force) force[j] = -D * x[j];
      if (Hc) Hc[j * dim + j] = D;
    }
    return 1;
  }
  double Dm = d->efc_D[r] / (mu * mu * (1 + mu * mu)), g = N - mu * T;
  *cost = 0.5 * Dm * g * g;
  double a[6];
  a[0] = mu;
  for (int j = 1; j < dim; j++) a[j] = -mu * c->friction[j - 1] * U[j] / T;
  if (force) for (int j = 0; j < dim; j++) force[j] = -Dm * g * a[j];
  if (Hc) {
    for (int i = 0; i < dim; i++)
      for (int j = 0; j < dim; j++) Hc[i * dim + j] = Dm * a[i] * a[j];
    for (int i = 1; i < dim; i++)
      for (int j = 1; j < dim; j++) {
        double B = -mu * c->friction[i - 1] * c->friction[j - 1] * ((i == j ? 1.0 : 0.0) - U[i] * U[j] / (T * T)) / T;
        Hc[i * dim + j] += Dm * g * B;
      }
  }
  return 2;
}

/* value, slope and curvature of one elliptic contact's cost along x + alpha v */
static int ellip_line(const orc_data_t* d, int r, int dim, const double* x0, const double* v, double alpha, double* f,
                      double* d1, double* d2) {
  double x[6] = {0, 0, 0, 0, 0, 0};
  for (int j = 0; j < dim; j++) x[j] = x0[j] + alpha * v[j];
  double cst;
  int z = ellip_state(d, r, dim, x, &cst, NULL, NULL);
  *f = cst; *d1 = 0; *d2 = 0;
  if (z == 1) {
    for (int j = 0; j < dim; j++) { double D = d->efc_D[r + j]; *d1 += D * x[j] * v[j]; *d2 += D * v[j] * v[j]; }
  } else if (z == 2) {
    const orc_contact_t* c = &d->con[d->efc_id[r]];
    double mu = c->friction[0], T = 0, TV = 0, VV = 0;
    for (int j = 1; j < dim; j++) {
      double U = c->friction[j - 1] * x[j], V = c->friction[j - 1] * v[j];
      T += U * U; TV += U * V; VV += V * V;
    }
    T = sqrt(T);
    double Tp = TV / T, Tpp = (VV - Tp * Tp) / T;
    double Dm = d->efc_D[r] / (mu * mu * (1 + mu * mu));
    double g = mu * x[0] - mu * T, gp = mu * v[0] - mu * Tp, gpp = -mu * Tpp;
    *d1 = Dm * g * gp;
    *d2 = Dm * (gp * gp + g * gpp);
  }
  return z;
}

/* ------------------------------------------------------------------ solver */
/* Constraint state at Jaref = J*qacc - aref: force, active flag and cost.
 * One-sided rows are active for x < 0; friction rows are quadratic inside
 * |D x| < floss and linear outside. */
static double update_constraint(orc_data_t* d) {
  double cost = 0;
  for (int r = 0; r < d->nefc; r++) {
    double x = d->efc_Jaref[r], D = d->efc_D[r];
    int t = d->efc_type[r];
    if (t == EFC_CONTACT_ELLIPTIC) {
      int dim = d->con[d->efc_id[r]].dim;
      double cst;
      int z = ellip_state(d, r, dim, d->efc_Jaref + r, &cst, d->efc_force + r, NULL);
      for (int j = 0; j < dim; j++) d->efc_active[r + j] = z;
      cost += cst;
      r += dim - 1;
    } else if (t == EFC_EQUALITY) {
      d->efc_active[r] = 1; d->efc_force[r] = -D * x; cost += 0.5 * D * x * x;
    } else if (t == EFC_FRICTION_DOF) {
      double f = d->efc_floss[r], R = d->efc_R[r];
      if (x <= -R * f) { d->efc_active[r] = 0; d->efc_force[r] = f; cost += -f * x - 0.5 * R * f * f; }
      else if (x >= R * f) { d->efc_active[r] = 0; d->efc_force[r] = -f; cost += f * x - 0.5 * R * f * f; }
      else { d->efc_active[r] = 1; d->efc_force[r] = -D * x; cost += 0.5 * D * x * x; }
    } else {
      if (x < 0) { d->efc_active[r] = 1; d->efc_force[r] = -D * x; cost += 0.5 * D * x * x; }
      else { d->efc_active[r] = 0; d->efc_force[r] = 0; }
    }
  }
  return cost;
}

/* total cost and gradient at the current qacc (Ma, Jaref already set) */
static void cost_grad(const orc_prep_t* p, orc_data_t* d) {
  int nv = p->m.nv;
  double cc = update_constraint(d);
  double g = 0;
  for (int k = 0; k < nv; k++) g += 0.5 * (d->Ma[k] - d->qfrc_smooth[k]) * (d->qacc[k] - d->qacc_smooth[k]);
  d->cost = g + cc;
  for (int k = 0; k < nv; k++) d->grad[k] = d->Ma[k] - d->qfrc_smooth[k];
  for (int r = 0; r < d->nefc; r++) {
    double f = d->efc_force[r];
    if (f == 0) continue;
    const double* J = d->efc_J + (size_t)r * nv;
    for (int k = 0; k < nv; k++) d->grad[k] -= J[k] * f;
  }
}

static void mul_m(const orc_data_t* d, int nv, const double* x, double* y) {
  for (int i = 0; i < nv; i++) {
    double s = 0;
    for (int j = 0; j < nv; j++) s += d->qM[i * nv + j] * x[j];
    y[i] = s;
  }
}

static void set_qacc(const orc_prep_t* p, orc_data_t* d, const double* qacc) {
  int nv = p->m.nv;
  memcpy(d->qacc, qacc, sizeof(double) * nv);
  mul_m(d, nv, d->qacc, d->Ma);
  for (int r = 0; r < d->nefc; r++) {
    const double* J = d->efc_J + (size_t)r * nv;
    double s = 0;
    for (int k = 0; k < nv; k++) s += J[k] * qacc[k];
    d->efc_Jaref[r] = s - d->efc_aref[r];
  }
}

/* 1-D cost along the search direction: value and first/second derivative at
 * step alpha (piecewise quadratic; rows switch with the sign / zone of x).
 * nchange counts rows whose zone at alpha differs from their zone at ref. */
typedef struct { double alpha, f, d1, d2; int nchange; } lspt_t;
static lspt_t ls_eval(const orc_data_t* d, double alpha, double ref, const double gq[3]) {
  lspt_t pt;
  pt.alpha = alpha;
  pt.f = gq[0] + alpha * gq[1] + 0.5 * alpha * alpha * gq[2];
  pt.d1 = gq[1] + alpha * gq[2];
  pt.d2 = gq[2];
  pt.nchange = 0;
  for (int r = 0; r < d->nefc; r++) {
    double jv = d->efc_Jv[r];
    double x = d->efc_Jaref[r] + alpha * jv, xr = d->efc_Jaref[r] + ref * jv, D = d->efc_D[r];
    int t = d->efc_type[r];
    if (t == EFC_CONTACT_ELLIPTIC) {
      /* the middle zone is not quadratic: it never counts as one exact piece */
      int dim = d->con[d->efc_id[r]].dim;
      double f, g1, g2, fr, h1, h2;
      int z = ellip_line(d, r, dim, d->efc_Jaref + r, d->efc_Jv + r, alpha, &f, &g1, &g2);
      int zr = ellip_line(d, r, dim, d->efc_Jaref + r, d->efc_Jv + r, ref, &fr, &h1, &h2);
      pt.nchange += (z != zr) || z == 2;
      pt.f += f; pt.d1 += g1; pt.d2 += g2;
      r += dim - 1;
    } else if (t == EFC_FRICTION_DOF) {
      double f = d->efc_floss[r], R = d->efc_R[r];
      int z = x <= -R * f ? 0 : (x >= R * f ? 2 : 1), zr = xr <= -R * f ? 0 : (xr >= R * f ? 2 : 1);
      pt.nchange += z != zr;
      if (z == 0) { pt.f += -f * x - 0.5 * R * f * f; pt.d1 += -f * jv; }
      else if (z == 2) { pt.f += f * x - 0.5 * R * f * f; pt.d1 += f * jv; }
      else { pt.f += 0.5 * D * x * x; pt.d1 += D * x * jv; pt.d2 += D * jv * jv; }
    } else if (t == EFC_EQUALITY) {
      pt.f += 0.5 * D * x * x; pt.d1 += D * x * jv; pt.d2 += D * jv * jv;
    } else {
      pt.nchange += (x < 0) != (xr < 0);
      if (x < 0) { pt.f += 0.5 * D * x * x; pt.d1 += D * x * jv; pt.d2 += D * jv * jv; }
    }
  }
  return pt;
}

/* Exact line search on the convex piecewise-quadratic cost: Newton steps
 * until the derivative changes sign, then Newton-from-the-left / secant on
 * the bracket. Stops when |f'| <= gtol (MuJoCo's scaled ls_tolerance) or when
 * a Newton step ends in the same quadratic piece it started from (then it is
 * the exact minimiser). Returns 0 when search is not a descent direction. */
static double linesearch(const orc_prep_t* p, orc_data_t* d, int* exact) {
  const orc_model_t* m = &p->m;
  int nv = m->nv;
  double snorm = 0;
  for (int k = 0; k < nv; k++) snorm += d->search[k] * d->search[k];
  snorm = sqrt(snorm);
  if (snorm < MINVAL) return 0;
  double scale = 1.0 / (p->meaninertia * (nv > 1 ? nv : 1));
  double gtol = m->tolerance * m->ls_tolerance * snorm / scale;
  mul_m(d, nv, d->search, d->Mv);
  for (int r = 0; r < d->nefc; r++) {
    const double* J = d->efc_J + (size_t)r * nv;
    double s = 0;
    for (int k = 0; k < nv; k++) s += J[k] * d->search[k];
    d->efc_Jv[r] = s;
  }
  double gq[3] = {0, 0, 0};
  for (int k = 0; k < nv; k++) {
    gq[0] += 0.5 * (d->Ma[k] - d->qfrc_smooth[k]) * (d->qacc[k] - d->qacc_smooth[k]);
    gq[1] += (d->Ma[k] - d->qfrc_smooth[k]) * d->search[k];
    gq[2] += d->search[k] * d->Mv[k];
  }
  *exact = 0;
  lspt_t p0 = ls_eval(d, 0, 0, gq);
  if (!(p0.d1 < 0)) return 0;
  int it = 1, maxit = m->ls_iterations;
  lspt_t lo = p0;
  lspt_t pt = ls_eval(d, -p0.d1 / p0.d2, 0, gq);
  if (pt.nchange == 0) { *exact = 1; return pt.alpha; }
  if (fabs(pt.d1) <= gtol) return pt.alpha;
  while (pt.d1 < 0 && it < maxit) {
    lo = pt;
    pt = ls_eval(d, pt.alpha - pt.d1 / pt.d2, lo.alpha, gq);
    it++;
    if (pt.nchange == 0 || fabs(pt.d1) <= gtol) return pt.alpha;
  }
  if (pt.d1 < 0) return pt.alpha;
  lspt_t hi = pt;
  while (it < maxit) {
    double a = lo.alpha - lo.d1 / lo.d2;
    int newton = a > lo.alpha && a < hi.alpha;
    if (!newton) a = lo.alpha - lo.d1 * (hi.alpha - lo.alpha) / (hi.d1 - lo.d1);
    if (!(a > lo.alpha && a < hi.alpha)) a = 0.5 * (lo.alpha + hi.alpha);
    pt = ls_eval(d, a, lo.alpha, gq);
    it++;
    if ((newton && pt.nchange == 0) || fabs(pt.d1) <= gtol) return pt.alpha;
    if (pt.d1 < 0) lo = pt; else hi = pt;
    if (!(hi.alpha > lo.alpha)) break;
  }
  return lo.f <= hi.f ? lo.alpha : hi.alpha;
}

/* dense Cholesky A = L L^T in place (lower triangle), n x n row-major */
static int chol_dense(double* A, int n) {
  for (int k = 0; k < n; k++) {
    double s = A[k * n + k];
    for (int q = 0; q < k; q++) s -= A[k * n + q] * A[k * n + q];
    if (!(s > 0)) return -1;
    double l = sqrt(s);
    A[k * n + k] = l;
    for (int i = k + 1; i < n; i++) {
      double t = A[i * n + k];
      for (int q = 0; q < k; q++) t -= A[i * n + q] * A[k * n + q];
      A[i * n + k] = t / l;
    }
    for (int j = k + 1; j < n; j++) A[k * n + j] = 0;
  }
  return 0;
}
/* x = (L L^T)^-1 b with L from chol_dense */
static void chol_solve(const double* L, int n, const double* b, double* x) {
  for (int i = 0; i < n; i++) {
    double s = b[i];
    for (int q = 0; q < i; q++) s -= L[i * n + q] * x[q];
    x[i] = s / L[i * n + i];
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = x[i];
    for (int q = i + 1; q < n; q++) s -= L[q * n + i] * x[q];
    x[i] = s / L[i * n + i];
  }
}

/* H = M + J^T diag(D * active) J, dense Cholesky in place; returns 0 on success */
static int hessian_factor(const orc_prep_t* p, orc_data_t* d) {
  int nv = p->m.nv;
  double* H = d->H;
  memcpy(H, d->qM, sizeof(double) * nv * nv);
  for (int r = 0; r < d->nefc; r++) {
    if (d->efc_type[r] == EFC_CONTACT_ELLIPTIC) {
      /* H += J_c^T Hc J_c over the contact's rows */
      int dim = d->con[d->efc_id[r]].dim;
      double Hc[36], cst;
      ellip_state(d, r, dim, d->efc_Jaref + r, &cst, NULL, Hc);
      for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) {
          double h = Hc[a * dim + b];
          if (h == 0) continue;
          const double *Ja = d->efc_J + (size_t)(r + a) * nv, *Jb = d->efc_J + (size_t)(r + b) * nv;
          for (int i = 0; i < nv; i++)
            for (int j = 0; j <= i; j++) H[i * nv + j] += Ja[i] * h * Jb[j];
        }
      r += dim - 1;
      continue;
    }
    if (!d->efc_active[r]) continue;
    const double* J = d->efc_J + (size_t)r * nv;
    double D = d->efc_D[r];
    for (int i = 0; i < nv; i++) {
      if (J[i] == 0) continue;
      double a = D * J[i];
      for (int j = 0; j <= i; j++) H[i * nv + j] += a * J[j];
    }
  }
  return chol_dense(H, nv);
}
static void hessian_solve(const orc_prep_t* p, const orc_data_t* d, const double* b, double* x) {
  chol_solve(d->H, p->m.nv, b, x);
}

/* mj_fwdConstraint: warm start (previous qacc vs qacc_smooth by total cost),
 * then MuJoCo's primal Newton (H^-1 direction) or nonlinear CG (M^-1
 * preconditioned Polak-Ribiere) with exact line search; stop when
 * scale*(cost decrease) < tol or scale*|grad| < tol, scale = 1/(meaninertia*nv). */
static int solve(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  int nv = m->nv;
  d->niter = 0;
  if (d->nefc == 0) {
    memcpy(d->qacc, d->qacc_smooth, sizeof(double) * nv);
    memset(d->qfrc_constraint, 0, sizeof(double) * nv);
    return 0;
  }
  set_qacc(p, d, d->qacc_smooth);
  cost_grad(p, d);
  if (!(m->disableflags & DS_WARMSTART)) {
    double c_smooth = d->cost;
    double* keep = d->tmp2;
    memcpy(keep, d->qacc_smooth, sizeof(double) * nv);
    set_qacc(p, d, d->qacc_ws);
    cost_grad(p, d);
    if (d->cost > c_smooth) { set_qacc(p, d, keep); cost_grad(p, d); }
  }
  double scale = 1.0 / (p->meaninertia * (nv > 1 ? nv : 1));
  int newton = m->solver == 2;
  int status = 0;
  if (newton) {
    if (hessian_factor(p, d) != 0) return 16;
    hessian_solve(p, d, d->grad, d->Mgrad);
  } else {
    memcpy(d->Mgrad, d->grad, sizeof(double) * nv);
    solve_m(p, d, d->Mgrad);
  }
  for (int k = 0; k < nv; k++) d->search[k] = -d->Mgrad[k];
  double* grad_old = (double*)calloc(nv, sizeof(double));
  double* mgrad_old = (double*)calloc(nv, sizeof(double));
  for (int it = 0; it < m->iterations; it++) {
    int exact = 0;
    double alpha = linesearch(p, d, &exact);
    d->niter = it + 1;
    if (alpha == 0) break;
    for (int k = 0; k < nv; k++) {
      d->qacc[k] += alpha * d->search[k];
      d->Ma[k] += alpha * d->Mv[k];
    }
    for (int r = 0; r < d->nefc; r++) d->efc_Jaref[r] += alpha * d->efc_Jv[r];
    double oldcost = d->cost;
    memcpy(grad_old, d->grad, sizeof(double) * nv);
    memcpy(mgrad_old, d->Mgrad, sizeof(double) * nv);
    cost_grad(p, d);
    double gn = 0;
    for (int k = 0; k < nv; k++) gn += d->grad[k] * d->grad[k];
    double improvement = scale * (oldcost - d->cost), gradient = scale * sqrt(gn);
    if (improvement < m->tolerance || gradient < m->tolerance) break;
    /* a Newton step that stayed inside one quadratic piece landed on that
     * piece's stationary point, which is the exact minimiser */
    if (newton && exact) break;
    if (newton) {
      if (hessian_factor(p, d) != 0) { status = 16; break; }
      hessian_solve(p, d, d->grad, d->Mgrad);
    } else {
      memcpy(d->Mgrad, d->grad, sizeof(double) * nv);
      solve_m(p, d, d->Mgrad);
    }
    if (newton) {
      for (int k = 0; k < nv; k++) d->search[k] = -d->Mgrad[k];
    } else {
      double num = 0, den = 0;
      for (int k = 0; k < nv; k++) { num += d->grad[k] * (d->Mgrad[k] - mgrad_old[k]); den += grad_old[k] * mgrad_old[k]; }
      double beta = num / fmax(MINVAL, den);
      if (beta < 0) beta = 0;
      for (int k = 0; k < nv; k++) d->search[k] = -d->Mgrad[k] + beta * d->search[k];
    }
  }
  free(grad_old); free(mgrad_old);
  for (int k = 0; k < nv; k++) {
    double s = 0;
    for (int r = 0; r < d->nefc; r++) s += d->efc_J[(size_t)r * nv + k] * d->efc_force[r];
    d->qfrc_constraint[k] = s;
  }
  return status;
}

/* ------------------------------------------------------------------ pipeline */
int orc_forward(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  int nv = m->nv;
  int status = 0;
  kinematics(p, d);
  com_pos(p, d);
  crb_mass(p, d);
  if (nv > 0 && factor_m(p, d) != 0) return 32;
  status |= collision(p, d);
  status |= make_constraint(p, d);
  com_vel(p, d);
  passive(p, d);
  rne_bias(p, d);
  actuation(p, d);
  for (int k = 0; k < nv; k++) d->qfrc_smooth[k] = d->qfrc_passive[k] - d->qfrc_bias[k] + d->qfrc_actuator[k];
  if (d->has_rand) free_drag(p, d);
  memcpy(d->qacc_smooth, d->qfrc_smooth, sizeof(double) * nv);
  if (nv > 0) solve_m(p, d, d->qacc_smooth);
  /* aref needs efc_vel which needs qvel only: already computed in make_constraint */
  status |= solve(p, d);
  return status;
}

/* ------------------------------------------------------------------ rays */
/* mj_ray against one primitive in its own frame: smallest x >= 0 with p + x v
 * on the surface, else -1 (x in units of |v|). Planes are hit from +z only,
 * inside nonzero half extents; from inside a closed primitive the exit point
 * is hit. Reference src/batched.cpp:1739-1806. */
static double ray_quadric(double a, double b, double c) {
  double disc = b * b - a * c;
  if (disc < 1e-12 || fabs(a) < 1e-12) return -1;
  double s = sqrt(disc), x0 = (-b - s) / a, x1 = (-b + s) / a;
  return x0 >= 0 ? x0 : (x1 >= 0 ? x1 : -1);
}
static double ray_closer(double best, double x) { return x >= 0 && (best < 0 || x < best) ? x : best; }
double orc_ray_geom(int type, const double* sz, const double* p, const double* v) {
  if (type == GPLANE) {
    if (v[2] > -1e-9) return -1;
    double x = -p[2] / v[2];
    if (x < 0) return -1;
    double hx = p[0] + x * v[0], hy = p[1] + x * v[1];
    if ((sz[0] > 0 && fabs(hx) > sz[0]) || (sz[1] > 0 && fabs(hy) > sz[1])) return -1;
    return x;
  }
  if (type == GSPHERE) return ray_quadric(dot3(v, v), dot3(p, v), dot3(p, p) - sz[0] * sz[0]);
  if (type == GELLIPSOID) {
    double ps[3], vs[3];
    for (int i = 0; i < 3; i++) { ps[i] = p[i] / sz[i]; vs[i] = v[i] / sz[i]; }
    return ray_quadric(dot3(vs, vs), dot3(ps, vs), dot3(ps, ps) - 1);
  }
  if (type == GCAPSULE || type == GCYLINDER) {
    double r = sz[0], h = sz[1], best = -1;
    double a = v[0] * v[0] + v[1] * v[1], b = p[0] * v[0] + p[1] * v[1], c = p[0] * p[0] + p[1] * p[1] - r * r;
    double disc = b * b - a * c;
    if (a > 1e-12 && disc >= 0) {
      for (int k = -1; k <= 1; k += 2) {
        double x = (-b + k * sqrt(disc)) / a;
        if (fabs(p[2] + x * v[2]) <= h) best = ray_closer(best, x);
      }
    }
    for (int e = -1; e <= 1; e += 2) {
      if (type == GCYLINDER) {
        if (fabs(v[2]) <= 1e-12) continue;
        double x = (e * h - p[2]) / v[2], qx = p[0] + x * v[0], qy = p[1] + x * v[1];
        if (x >= 0 && qx * qx + qy * qy <= r * r) best = ray_closer(best, x);
      } else {
        double pc[3] = {p[0], p[1], p[2] - e * h};
        double a2 = dot3(v, v), b2 = dot3(pc, v), c2 = dot3(pc, pc) - r * r, d2 = b2 * b2 - a2 * c2;
        if (d2 < 0) continue;
        for (int k = -1; k <= 1; k += 2) {
          double x = (-b2 + k * sqrt(d2)) / a2;
          if (e * (p[2] + x * v[2] - e * h) >= 0) best = ray_closer(best, x);
        }
      }
    }
    return best;
  }
  if (type == GBOX) {
    double tn = -1e300, tf = 1e300;
    for (int i = 0; i < 3; i++) {
      if (fabs(v[i]) < 1e-12) {
        if (fabs(p[i]) > sz[i]) return -1;
        continue;
      }
      double t0 = (-sz[i] - p[i]) / v[i], t1 = (sz[i] - p[i]) / v[i];
      tn = fmax(tn, fmin(t0, t1));
      tf = fmin(tf, fmax(t0, t1));
    }
    if (tn > tf || tf < 0) return -1;
    return tn >= 0 ? tn : tf;
  }
  return -1;
}

/* nearest geom allowed by geom_ok along pnt + x vec at d's geom poses (call
 * orc_kinematics first); returns x or -1 and the geom id or -1 */
double orc_ray(const orc_prep_t* p, const orc_data_t* d, const double* pnt, const double* vec, const int* geom_ok,
               int* geomid) {
  const orc_model_t* m = &p->m;
  double best = -1;
  *geomid = -1;
  for (int g = 0; g < m->ngeom; g++) {
    if (!geom_ok[g]) continue;
    const double *x = d->geom_xpos + 3 * g, *R = d->geom_xmat + 9 * g;
    double dd[3] = {pnt[0] - x[0], pnt[1] - x[1], pnt[2] - x[2]}, lp[3], lv[3];
    for (int k = 0; k < 3; k++) {
      lp[k] = R[k] * dd[0] + R[3 + k] * dd[1] + R[6 + k] * dd[2];
      lv[k] = R[k] * vec[0] + R[3 + k] * vec[1] + R[6 + k] * vec[2];
    }
    double t = orc_ray_geom(m->geom_type[g], m->geom_size + 3 * g, lp, lv);
    if (t >= 0 && (best < 0 || t < best)) { best = t; *geomid = g; }
  }
  return best;
}

void orc_kinematics(const orc_prep_t* p, orc_data_t* d) {
  kinematics(p, d);
  com_pos(p, d);
  crb_mass(p, d);
  if (p->m.nv > 0) factor_m(p, d);
  com_vel(p, d);
}

/* q <- q * exp(h*omega/2), omega in the local frame, renormalised */
static void quat_integrate(double* q, const double* omega, double h) {
  double w[3] = {omega[0], omega[1], omega[2]};
  double ang = h * normalize3(w);
  double qr[4];
  if (ang == 0) { qr[0] = 1; qr[1] = qr[2] = qr[3] = 0; }
  else axisangle2quat(w, ang, qr);
  normalize4(q);
  mulquat(q, qr, q);
  normalize4(q);
}

/* mj_integratePos: hinge/slide add h*qvel; ball/free rotate the quaternion by
 * exp(h*omega/2) with omega in the local frame (reference euler.hpp:55-104). */
static void integrate_pos(const orc_model_t* m, double* qpos, const double* qvel, double h) {
  for (int j = 0; j < m->njnt; j++) {
    int qa = m->jnt_qposadr[j], da = m->jnt_dofadr[j];
    int t = m->jnt_type[j];
    if (t == JFREE) {
      for (int i = 0; i < 3; i++) qpos[qa + i] += h * qvel[da + i];
      qa += 3; da += 3;
    }
    if (t == JFREE || t == JBALL) {
      quat_integrate(qpos + qa, qvel + da, h);
    } else {
      qpos[qa] += h * qvel[da];
    }
  }
}

/* Euler (implicit joint damping when present) and implicitfast (actuator and
 * passive velocity derivatives, Coriolis dropped); then qvel += h*qacc,
 * qpos integrated with the new qvel, qacc_warmstart = solver qacc. */
static int integrate(const orc_prep_t* p, orc_data_t* d) {
  const orc_model_t* m = &p->m;
  int nv = m->nv;
  double h = m->timestep;
  double* acc = d->tmp1;
  memcpy(acc, d->qacc, sizeof(double) * nv);
  int damp = 0;
  for (int k = 0; k < nv; k++) if (m->dof_damping[k] > 0) damp = 1;
  int implicit = m->integrator == 3 || m->integrator == 2;
  if (implicit || (damp && !(m->disableflags & DS_EULERDAMP))) {
    double* A = (double*)calloc((size_t)nv * nv, sizeof(double));
    memcpy(A, d->qM, sizeof(double) * nv * nv);
    for (int k = 0; k < nv; k++) A[k * nv + k] += h * m->dof_damping[k];
    if (implicit) {
      for (int u = 0; u < m->nu; u++) {
        if (m->disableflags & DS_ACTUATION) break;
        int j = m->actuator_trnid[u], da = m->jnt_dofadr[j];
        double gear = m->actuator_gear[u];
        double f = d->actuator_force[u];
        if (m->actuator_forcelimited[u] && (f <= m->actuator_forcerange[2 * u] || f >= m->actuator_forcerange[2 * u + 1])) continue;
        A[da * nv + da] -= h * gear * gear * m->actuator_biasprm[3 * u + 2];
      }
    }
    double* rhs = d->tmp2;
    for (int k = 0; k < nv; k++) rhs[k] = d->qfrc_smooth[k] + d->qfrc_constraint[k];
    if (factor_m_into(m, A, A) != 0) { free(A); return 32; }
    memcpy(acc, rhs, sizeof(double) * nv);
    solve_ld(m, A, acc);
    free(A);
  }
  for (int k = 0; k < nv; k++) d->qvel[k] += h * acc[k];
  integrate_pos(m, d->qpos, d->qvel, h);
  memcpy(d->qacc_ws, d->qacc, sizeof(double) * nv);
  return 0;
}

int orc_step(const orc_prep_t* p, orc_data_t* d) {
  int st = orc_forward(p, d);
  st |= integrate(p, d);
  d->status = st;
  return st;
}

int orc_ncon(const orc_data_t* d) { return d->ncon; }
const orc_contact_t* orc_contacts(const orc_data_t* d) { return d->con; }
int orc_nefc(const orc_data_t* d) { return d->nefc; }
/* Total cost, gradient and full Hessian (nv x nv) of the constraint problem
 * built by the last orc_forward, at qacc. The constraint state (forces,
 * zones in efc_active) is left at qacc. */
int orc_cost(const orc_prep_t* p, orc_data_t* d, const double* qacc, double* cost, double* grad, double* Hout) {
  int nv = p->m.nv;
  set_qacc(p, d, qacc);
  cost_grad(p, d);
  if (cost) *cost = d->cost;
  if (grad) memcpy(grad, d->grad, sizeof(double) * nv);
  if (Hout) {
    double* keep = (double*)malloc(sizeof(double) * nv * nv);
    memcpy(keep, d->H, sizeof(double) * nv * nv);
    double* H = d->H;
    memcpy(H, d->qM, sizeof(double) * nv * nv);
    for (int r = 0; r < d->nefc; r++) {
      if (d->efc_type[r] == EFC_CONTACT_ELLIPTIC) {
        int dim = d->con[d->efc_id[r]].dim;
        double Hc[36], cst;
        ellip_state(d, r, dim, d->efc_Jaref + r, &cst, NULL, Hc);
        for (int a = 0; a < dim; a++)
          for (int b = 0; b < dim; b++)
            for (int i = 0; i < nv; i++)
              for (int j = 0; j < nv; j++)
                H[i * nv + j] += d->efc_J[(size_t)(r + a) * nv + i] * Hc[a * dim + b] * d->efc_J[(size_t)(r + b) * nv + j];
        r += dim - 1;
        continue;
      }
      if (!d->efc_active[r]) continue;
      for (int i = 0; i < nv; i++)
        for (int j = 0; j < nv; j++)
          H[i * nv + j] += d->efc_D[r] * d->efc_J[(size_t)r * nv + i] * d->efc_J[(size_t)r * nv + j];
    }
    memcpy(Hout, H, sizeof(double) * nv * nv);
    memcpy(d->H, keep, sizeof(double) * nv * nv);
    free(keep);
  }
  return 0;
}

int orc_solver_niter(const orc_data_t* d) { return d->niter; }

int orc_get(const orc_data_t* d, const char* name, double* out, int cap) {
  struct { const char* n; const double* p; int len; } f[] = {
      {"qpos", d->qpos, d->nq}, {"qvel", d->qvel, d->nv}, {"qacc", d->qacc, d->nv},
      {"qacc_smooth", d->qacc_smooth, d->nv}, {"qacc_warmstart", d->qacc_ws, d->nv},
      {"xpos", d->xpos, 3 * d->nbody}, {"xquat", d->xquat, 4 * d->nbody}, {"xmat", d->xmat, 9 * d->nbody},
      {"xipos", d->xipos, 3 * d->nbody}, {"ximat", d->ximat, 9 * d->nbody},
      {"xanchor", d->xanchor, 3 * d->njnt}, {"xaxis", d->xaxis, 3 * d->njnt},
      {"geom_xpos", d->geom_xpos, 3 * d->ngeom}, {"geom_xmat", d->geom_xmat, 9 * d->ngeom},
      {"subtree_com", d->subtree_com, 3 * d->nbody}, {"cinert", d->cinert, 10 * d->nbody},
      {"cdof", d->cdof, 6 * d->nv}, {"cvel", d->cvel, 6 * d->nbody}, {"cdof_dot", d->cdof_dot, 6 * d->nv},
      {"crb", d->crb, 10 * d->nbody}, {"qM", d->qM, d->nv * d->nv}, {"qLD", d->qLD, d->nv * d->nv},
      {"qfrc_bias", d->qfrc_bias, d->nv}, {"qfrc_passive", d->qfrc_passive, d->nv},
      {"qfrc_actuator", d->qfrc_actuator, d->nv}, {"qfrc_smooth", d->qfrc_smooth, d->nv},
      {"qfrc_constraint", d->qfrc_constraint, d->nv}, {"actuator_force", d->actuator_force, d->nu},
      {"efc_J", d->efc_J, d->nefc * d->nv}, {"efc_D", d->efc_D, d->nefc}, {"efc_R", d->efc_R, d->nefc},
      {"efc_aref", d->efc_aref, d->nefc}, {"efc_force", d->efc_force, d->nefc}, {"efc_pos", d->efc_pos, d->nefc},
      {"efc_vel", d->efc_vel, d->nefc}, {"efc_imp", d->efc_imp, d->nefc}, {"efc_diagApprox", d->efc_diagApprox, d->nefc},
  };
  for (size_t i = 0; i < sizeof(f) / sizeof(f[0]); i++) {
    if (strcmp(f[i].n, name) == 0) {
      int n = f[i].len < cap ? f[i].len : cap;
      if (out && n > 0) memcpy(out, f[i].p, sizeof(double) * n);
      return f[i].len;
    }
  }
  if (strcmp(name, "efc_active") == 0) {
    for (int i = 0; i < d->nefc && i < cap; i++) out[i] = d->efc_active[i];
    return d->nefc;
  }
  if (strcmp(name, "efc_type") == 0 || strcmp(name, "efc_id") == 0) {
    const int* src = name[4] == 't' ? d->efc_type : d->efc_id;
    int n = d->nefc < cap ? d->nefc : cap;
    for (int i = 0; i < n; i++) out[i] = src[i];
    return d->nefc;
  }
  return -1;
}

/* ------------------------------------------------------------------ batch */
typedef struct {
  const orc_prep_t* p;
  int nenv, nsub, tid, nthreads;
  double *qpos, *qvel, *qacc_ws;
  const double* ctrl;
  int* niter;
  int bad;
} job_t;

static void* batch_worker(void* arg) {
  job_t* j = (job_t*)arg;
  const orc_model_t* m = &j->p->m;
  orc_data_t* d = orc_data_new(j->p);
  for (int e = j->tid; e < j->nenv; e += j->nthreads) {
    orc_set_state(d, j->qpos + (size_t)e * m->nq, j->qvel + (size_t)e * m->nv, j->qacc_ws + (size_t)e * m->nv,
                  j->ctrl ? j->ctrl + (size_t)e * m->nu : NULL);
    int st = 0, it = 0;
    for (int s = 0; s < j->nsub; s++) { st |= orc_step(j->p, d); it += d->niter; }
    if (st) j->bad++;
    if (j->niter) j->niter[e] = it;
    orc_get_state(d, j->qpos + (size_t)e * m->nq, j->qvel + (size_t)e * m->nv, j->qacc_ws + (size_t)e * m->nv);
  }
  orc_data_free(d);
  return NULL;
}

int orc_batch_step(const orc_prep_t* p, int nenv, double* qpos, double* qvel, double* qacc_ws, const double* ctrl,
                   int nsub, int nthreads, int* niter_out) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > nenv) nthreads = nenv > 0 ? nenv : 1;
  pthread_t th[256];
  job_t jobs[256];
  if (nthreads > 256) nthreads = 256;
  for (int t = 0; t < nthreads; t++) {
    jobs[t].p = p; jobs[t].nenv = nenv; jobs[t].nsub = nsub; jobs[t].tid = t; jobs[t].nthreads = nthreads;
    jobs[t].qpos = qpos; jobs[t].qvel = qvel; jobs[t].qacc_ws = qacc_ws; jobs[t].ctrl = ctrl;
    jobs[t].niter = niter_out; jobs[t].bad = 0;
  }
  if (nthreads == 1) batch_worker(&jobs[0]);
  else {
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
  }
  int bad = 0;
  for (int t = 0; t < nthreads; t++) bad += jobs[t].bad;
  return bad;
}

/* ------------------------------------------------------------------ math KAT entry */
/* Exposes the oracle's own math routines for known-answer tests. */
int orc_math(const char* op, const double* in, double* out) {
  if (strcmp(op, "quat_mul") == 0) { mulquat(in, in + 4, out); return 4; }
  if (strcmp(op, "quat2mat") == 0) { quat2mat(in, out); return 9; }
  if (strcmp(op, "quat_rotate") == 0) { double R[9]; quat2mat(in, R); mulmatvec3(R, in + 4, out); return 3; }
  if (strcmp(op, "axisangle") == 0) { axisangle2quat(in, in[3], out); return 4; }
  if (strcmp(op, "quat_integrate") == 0) { memcpy(out, in, 4 * sizeof(double)); quat_integrate(out, in + 4, in[7]); return 4; }
  if (strcmp(op, "cross") == 0) { cross3(in, in + 3, out); return 3; }
  if (strcmp(op, "mul_inert") == 0) { mul_inert_vec(in, in + 10, out); return 6; }
  if (strcmp(op, "cross_motion") == 0) { cross_motion(in, in + 6, out); return 6; }
  if (strcmp(op, "cross_force") == 0) { cross_force(in, in + 6, out); return 6; }
  if (strcmp(op, "seg_seg") == 0) { seg_seg(in, in + 3, in + 6, in + 9, out, out + 1); return 2; }
  if (strcmp(op, "make_frame") == 0) { make_frame(in, out); return 9; }
  if (strcmp(op, "impedance") == 0) { out[0] = impedance(in, in[5]); return 1; }
  if (strcmp(op, "cholesky") == 0) {
    int n = (int)in[0];
    memcpy(out, in + 1, sizeof(double) * n * n);
    return chol_dense(out, n) == 0 ? n * n : -1;
  }
  if (strcmp(op, "chol_solve") == 0) {
    int n = (int)in[0];
    double* L = (double*)malloc(sizeof(double) * n * n);
    memcpy(L, in + 1, sizeof(double) * n * n);
    int ok = chol_dense(L, n);
    if (ok == 0) chol_solve(L, n, in + 1 + n * n, out);
    free(L);
    return ok == 0 ? n : -1;
  }
  return -1;
}
