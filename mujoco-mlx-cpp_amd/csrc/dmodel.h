// Flat FP32 model the kernels read, and the per-env LDS workspace layout.
//
// The reference bakes the same model data into its GLSL kernels as constant
// arrays (reference src/compat/mkx_kernels/forward.hpp:108-141,
// collision.hpp:506-517) and passes the pair/limit tables as float buffers
// (reference src/batched.cpp:317-561). Here the tables live in one device
// buffer; integer tables stay integers.
#pragma once
#include <stdint.h>

namespace mjk {

enum { JFREE = 0, JBALL = 1, JSLIDE = 2, JHINGE = 3 };
enum { GPLANE = 0, GSPHERE = 2, GCAPSULE = 3, GELLIPSOID = 4, GCYLINDER = 5, GBOX = 6 };
enum { EFC_EQUALITY = 0, EFC_FRICTION_DOF = 1, EFC_LIMIT_JOINT = 3, EFC_CONTACT_FRICTIONLESS = 5,
       EFC_CONTACT_PYRAMIDAL = 6, EFC_CONTACT_ELLIPTIC = 7 };
enum {
  DS_CONSTRAINT = 1 << 0, DS_FRICTIONLOSS = 1 << 2, DS_LIMIT = 1 << 3, DS_CONTACT = 1 << 4,
  DS_PASSIVE = 1 << 5, DS_GRAVITY = 1 << 6, DS_CLAMPCTRL = 1 << 7, DS_WARMSTART = 1 << 8,
  DS_ACTUATION = 1 << 10, DS_REFSAFE = 1 << 11, DS_EULERDAMP = 1 << 15,
};
// per-env status bits
enum { ST_CON_OVERFLOW = 1, ST_EFC_OVERFLOW = 2, ST_BAD_PAIR = 4, ST_NOT_PD = 16, ST_M_NOT_PD = 32,
       ST_NONFINITE = 64 };

// contact record in the workspace (floats)
enum { CON_POS = 0, CON_FRAME = 3, CON_DIST = 12, CON_PAIR = 13, CON_NROW = 14, CON_ROW0 = 15, CON_STRIDE = 16 };
// learned warm-start features: the first FEAT_NCON contacts, each as
// (frame normal 3, penetration depth, limb one-hot 4)
enum { FEAT_NCON = 16, FEAT_PER_CON = 8, FEAT_NCLASS = 4 };
// max contacts one narrow-phase call emits
enum { MAXC_PAIR = 4 };
enum { MAX_DOF_DEPTH = 64 };
// beyond this many dofs the constraint rows are stored compressed: each row
// keeps only the dofs of its source (a friction / limit row its dof, a
// contact row the union of its two bodies' dof chains), jw values and dof
// indices per row. Up to it the rows stay dense for the Hessian's matrix-core
// tiles (32 x 32).
enum { SPARSE_ROWS_NV = 32 };
// beyond this many dofs qM and its LDL^T factor are stored as compact rows
// (MuJoCo's sparse qM): row k holds its diagonal, then its dof ancestors
// nearest first, at dof_madr[k]; entry (k, j) of an ancestor j sits at
// dof_madr[k] + dof_depth[k] - dof_depth[j]. Up to it both are dense nv x nv.
enum { COMPACT_M_NV = 64 };
// compact qM whose dof trees are all paths (each dof's parent the dof
// before it) of at most this many dofs: one lane solves and multiplies a
// whole tree in registers (free bodies: 6 dofs each)
enum { PATH_TREE_MAX = 8 };

// Device passes see the model's arrays as global memory (address space 1), so
// loads through them stay global loads wherever the struct itself is read
// from (a kernel argument or a device-memory copy); host code and the CPU
// interpreter see plain pointers of the same size.
#if defined(__HIP_DEVICE_COMPILE__)
#define MJK_G __attribute__((address_space(1)))
#else
#define MJK_G
#endif

struct DModel {
  int nq, nv, nu, nbody, njnt, ngeom, npair;
  int nlevel, ndoflevel, maxdepth, nfloss, nlim;
  int nflev;  // levels of the level-parallel LDL of qM (deepest dof level first)
  int iterations, ls_iterations, solver, cone, integrator, disableflags;
  int has_damping;
  float timestep, gravity[3], tolerance, ls_tolerance, impratio, meaninertia;
  int nfeat;  // learned warm-start input width: nq + nv + nu + FEAT_NCON * FEAT_PER_CON
  int nsensor, nsensordata, nsens_acc;  // sensors, sensordata width, accelerometers
  int jw;  // entries per compressed constraint row (0: dense rows of nv)
  int nM;  // stored entries of qM (and of its factor): nv * nv, or the compact rows' total
  int ntree;  // dof trees when every tree is a path of at most PATH_TREE_MAX dofs, else 0
  float magnetic[3];
  // pointers only from here on: Blob::bind relocates every word after body_parentid
  // bodies
  const MJK_G int *body_parentid, *body_rootid, *body_jntnum, *body_jntadr, *body_dofnum, *body_dofadr,
      *body_subtree_end, *level_adr, *level_body, *body_lastdof;
  const MJK_G float *body_pos, *body_quat, *body_ipos, *body_iquat, *body_mass, *body_inertia,
      *body_subtreemass, *body_invweight0;
  // joints
  const MJK_G int *jnt_type, *jnt_qposadr, *jnt_dofadr, *jnt_bodyid;
  const MJK_G float *jnt_pos, *jnt_axis, *jnt_range, *jnt_margin, *jnt_solref, *jnt_solimp, *jnt_stiffness;
  const MJK_G float *qpos0, *qpos_spring;
  // dofs
  const MJK_G int *dof_bodyid, *dof_jntid, *dof_parentid, *dof_depth, *dof_subtree_end, *doflevel_adr,
      *doflevel_dof, *dof_chain, *dof_nchain, *dof_act_adr, *dof_act_num, *dof_act_list, *dof_madr,
      *tree_adr;  // first dof of each path tree [ntree + 1]
  // level-parallel LDL of qM: level l updates entries [flev_adr[l], flev_adr[l+1]),
  // entry e is (flev_i[e], flev_j[e]) and sums over dofs flev_k[flev_kadr[e] .. + flev_kcnt[e])
  const MJK_G int *flev_adr, *flev_i, *flev_j, *flev_kadr, *flev_kcnt, *flev_k;
  const MJK_G float *dof_armature, *dof_damping, *dof_frictionloss, *dof_invweight0, *dof_solref, *dof_solimp;
  // geoms
  const MJK_G int *geom_type, *geom_bodyid, *geom_fclass;
  const MJK_G float *geom_size, *geom_pos, *geom_quat, *geom_rbound;
  // static pair table (geom1 has the lower type)
  const MJK_G int *pair_g1, *pair_g2, *pair_b1, *pair_b2, *pair_condim;
  const MJK_G float *pair_friction, *pair_solref, *pair_solimp, *pair_margin, *pair_gap;
  // actuators
  const MJK_G int *act_dof, *act_qadr, *act_ctrllimited, *act_forcelimited;
  const MJK_G float *act_gear, *act_gain, *act_bias, *act_ctrlrange, *act_forcerange;
  // constraint source tables
  const MJK_G int *floss_dof, *lim_jnt;
  // sensors (MuJoCo mjtSensor ids): sensordata address and width, frame or
  // subtree body (-1 for joint sensors) with the frame's pose in that body,
  // qpos / dof address and gear of joint and actuator sensors, accelerometer
  // slot (-1 for other sensors)
  const MJK_G int *sens_type, *sens_adr, *sens_dim, *sens_body, *sens_qadr, *sens_dadr, *sens_acc;
  const MJK_G float *sens_pos, *sens_quat, *sens_gear;
};
// sensor kinds (MuJoCo mjtSensor)
enum { SENS_ACCELEROMETER = 1, SENS_VELOCIMETER = 2, SENS_GYRO = 3, SENS_MAGNETOMETER = 6, SENS_JOINTPOS = 9,
       SENS_JOINTVEL = 10, SENS_ACTUATORPOS = 13, SENS_ACTUATORVEL = 14, SENS_JOINTACTFRC = 16, SENS_BALLQUAT = 18,
       SENS_BALLANGVEL = 19, SENS_FRAMEPOS = 26, SENS_FRAMEQUAT = 27, SENS_FRAMEXAXIS = 28, SENS_FRAMEYAXIS = 29,
       SENS_FRAMEZAXIS = 30, SENS_FRAMELINVEL = 31, SENS_FRAMEANGVEL = 32, SENS_SUBTREECOM = 35 };

// Per-env workspace layout in floats. Region A (kinematics / dynamics
// intermediates) is dead once the constraint Jacobian is built, so the FP64
// Newton Hessian overlays it.
struct Layout {
  // maxcon / maxefc: buffer sizes; efc_cap: the row budget that admits rows
  // (limit rows while fewer than efc_cap - 4 rows exist, a contact while fewer
  // than efc_cap - 10), maxefc >= min(rows the model can produce, efc_cap);
  // con_cap: contacts counted per env (maxcon <= con_cap of them are stored:
  // at least every contact the row budget can admit)
  int maxcon, maxefc, efc_cap, con_cap, nv, total;
  int nM;    // floats of qM and of qLD (DModel::nM)
  int rand;  // per-env physical parameters (rand_size floats), or -1 without domain randomisation
  int qpos, qvel, ctrl, qacc_ws, qacc, qM, qLD, qfrc_smooth, qacc_smooth, qfrc_constraint, act_force, misc;
  int xpos, xquat, xmat, xipos, xanchor, xaxis, gxpos, gxmat, scom, cinert, cdof, cvel, cdofdot, crb, cacc, cfrc, buf6;
  int H;
  int con, efc_J, efc_D, efc_aref, efc_Jaref, efc_Jv, efc_force, efc_pos, efc_margin,
      efc_type, efc_act, efc_id, efc_Dact;
  int jw;        // compressed rows: entries per row (0: dense rows of nv entries)
  int efc_Jidx;  // compressed rows: the dof of each entry (-1 past the row's end), ints
  int wrench;    // compressed rows: per-body contact wrench scratch of J^T f [nbody][6]
  int cwrench;   // compressed rows: per-contact wrenches on body2 and body1 [maxcon][12]
  int bcon_adr;  // compressed rows: per-body contact lists, CSR offsets [nbody + 1] (ints)
  int bcon;      // compressed rows: per-body contact lists, entries 2 c + side [2 maxcon] (ints)
  int Ma, grad, Mgrad, search, Mv, grad_old, Mgrad_old, tmpv, dvec;
};

// misc ints (stored as int bits in the float workspace)
enum { MISC_NCON = 0, MISC_NEFC = 1, MISC_STATUS = 2, MISC_NITER = 3, MISC_NFLOSSROWS = 4, MISC_NLIMROWS = 5,
       MISC_CONSUB = 6, MISC_NCON_ALL = 7, MISC_SIZE = 8 };

#if defined(__HIPCC__)
#define MJK_HD __host__ __device__
#else
#define MJK_HD
#endif

MJK_HD inline int al2(int x) { return (x + 1) & ~1; }

// Per-env physical parameters of domain randomisation, one block per env:
// mass scale per body [nbody], body-frame COM shift per body [nbody][3],
// friction scale [1], actuator force scale per actuator [nu], quadratic drag
// coefficient per free joint [nfree] (in joint order).
MJK_HD inline int rand_com(int nbody) { return nbody; }
MJK_HD inline int rand_fric(int nbody) { return 4 * nbody; }
MJK_HD inline int rand_gain(int nbody) { return 4 * nbody + 1; }
MJK_HD inline int rand_drag(int nbody, int nu) { return 4 * nbody + 1 + nu; }
MJK_HD inline int rand_size(int nbody, int nu, int nfree) { return 4 * nbody + 1 + nu + nfree; }

// qld_shared: the solver never reads the LDL factor of qM (Newton), so the
// factor lives in the efc Jacobian's space: it is dead once qacc_smooth is
// solved, before the Jacobian is built, and the integrator reuses the space
// as scratch after the solve.
// jw > 0: compressed constraint rows of jw entries (models beyond
// SPARSE_ROWS_NV dofs); subtree com and cdof then stay live through the solve
// (J^T f is accumulated from per-body contact wrenches).
MJK_HD inline Layout make_layout(int nq, int nv, int nu, int nbody, int njnt, int ngeom, int maxcon, int maxefc,
                                 bool qld_shared = false, int nrand = 0, int jw = 0, int nM = 0) {
  Layout L;
  L.maxcon = maxcon; L.maxefc = maxefc; L.efc_cap = maxefc; L.con_cap = maxcon; L.nv = nv;
  L.jw = jw;
  int nvp = nv > 0 ? nv : 1, nup = nu > 0 ? nu : 1, njp = njnt > 0 ? njnt : 1, ngp = ngeom > 0 ? ngeom : 1;
  if (nM <= 0) nM = nvp * nvp;
  L.nM = nM;
  int ncon = maxcon > 0 ? maxcon : 1;
  int o = 0;
  // persistent across the substep
  L.qpos = o; o += al2(nq > 0 ? nq : 1);
  L.qvel = o; o += al2(nvp);
  L.ctrl = o; o += al2(nup);
  L.qacc_ws = o; o += al2(nvp);
  L.qacc = o; o += al2(nvp);
  L.qM = o; o += al2(nM);
  if (!qld_shared) { L.qLD = o; o += al2(nM); }
  L.qfrc_smooth = o; o += al2(nvp);
  L.qacc_smooth = o; o += al2(nvp);
  L.qfrc_constraint = o; o += al2(nvp);
  L.act_force = o; o += al2(nup);
  L.misc = o; o += MISC_SIZE;
  L.rand = nrand > 0 ? o : -1;
  o += nrand;
  o = al2(o);
  // The rest is packed by lifetime within a substep:
  //   S  scom, cdof          com positions -> constraint rows
  //   G  geom poses          kinematics -> collision
  //   KC the other kinematics and dynamics intermediates, kinematics -> smooth
  //   con, rows              collision / constraint rows -> end of the substep
  //   V  solver vectors      solve (tmpv also integration)
  //   H  FP64 Hessian        solve (Newton only, lower triangle packed); the
  //                          LDL factor of qM (factor -> smooth, and the
  //                          integrator's factor) uses its space
  // con starts where KC starts (KC is dead once collision runs); the rows
  // follow con; V takes S and G's space when it fits (both dead in the solve).
  int r0 = o;
  L.scom = o; o += 3 * nbody;
  L.cdof = o; o += 6 * nvp;
  L.gxpos = o; o += 3 * ngp;
  L.gxmat = o; o += 9 * ngp;
  o = al2(o);
  int kc0 = o;
  L.xpos = o; o += 3 * nbody;
  L.xquat = o; o += 4 * nbody;
  L.xmat = o; o += 9 * nbody;
  L.xipos = o; o += 3 * nbody;
  L.xanchor = o; o += 3 * njp;
  L.xaxis = o; o += 3 * njp;
  L.cinert = o; o += 10 * nbody;
  L.cvel = o; o += 6 * nbody;
  L.cdofdot = o; o += 6 * nvp;
  L.crb = o; o += 10 * nbody;
  L.cacc = o; o += 6 * nbody;
  L.cfrc = o; o += 6 * nbody;
  L.buf6 = o; o += 6 * nvp;
  int kc_end = al2(o);
  o = kc0;
  L.con = o; o += CON_STRIDE * ncon;
  L.efc_J = o; o += maxefc * (jw > 0 ? jw : nvp);
  L.efc_Jidx = -1;
  if (jw > 0) { L.efc_Jidx = o; o += maxefc * jw; }
  L.efc_D = o; o += maxefc;
  L.efc_aref = o; o += maxefc;
  L.efc_Jaref = o; o += maxefc;
  L.efc_Jv = o; o += maxefc;
  L.efc_force = o; o += maxefc;
  L.efc_type = o; o += maxefc;
  L.efc_act = o; o += maxefc;
  L.efc_id = o; o += maxefc;
  L.efc_Dact = o; o += maxefc;
  // contact distance and margin are read only while the rows are built; the
  // solver-only vectors Jv and Dact take their place afterwards
  L.efc_pos = L.efc_Jv;
  L.efc_margin = L.efc_Dact;
  int rows_end = al2(o);
  o = al2(kc_end > rows_end ? kc_end : rows_end);
  // the FP64 Newton Hessian, its lower triangle packed row by row (Newton
  // only; its nv (nv + 1) floats also hold the LDL factor of qM, nM <= nv (nv + 1))
  L.H = o;
  if (qld_shared) { o += al2(nvp * (nvp + 1)); L.qLD = L.H; }
  int nvec = al2(8 * nvp) + 4 * nvp;  // eight FP32 vectors, then two FP64 ones
  int vbase = jw == 0 && nvec <= kc0 - r0 ? r0 : o;
  if (vbase == o) o += nvec;
  int v = vbase;
  L.Ma = v; v += nvp;
  L.grad = v; v += nvp;
  L.Mgrad = v; v += nvp;
  L.search = v; v += nvp;
  L.Mv = v; v += nvp;
  L.grad_old = v; v += nvp;
  L.Mgrad_old = v; v += nvp;
  L.tmpv = v; v += nvp;
  v = al2(v);
  L.dvec = v;  // two FP64 vectors: inverse pivots and a solve vector
  L.wrench = L.cwrench = L.bcon_adr = L.bcon = -1;
  if (jw > 0) {
    L.wrench = o; o += 6 * nbody;
    L.cwrench = o; o += 12 * ncon;
    L.bcon_adr = o; o += nbody + 1;
    L.bcon = o; o += 2 * ncon;
  }
  L.total = al2(o);
  return L;
}

}  // namespace mjk
