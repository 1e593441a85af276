/* ref64: FP64 CPU oracle for the batched MuJoCo step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The engine never links, loads or calls it.
 *
 * It restates, in double precision, the MuJoCo C 3.5.0 pipeline that the
 * reference uses as its oracle (reference src/mjb.cpp:807-839 calls mj_step per
 * env; tests/test_gpu_vs_cpu.cpp compares against it), because MuJoCo C itself
 * is absent from this environment. Each function cites the reference file:line
 * whose semantics it follows, or states the MuJoCo behaviour it restates.
 *
 * Pinned by checks under tests/: kinematics, mass matrix, bias forces and
 * integration (closed forms, brute force, finite differences of the
 * Lagrangian); the reference's math and Cholesky known-answer tests;
 * primitive, box and convex collision (closed forms, sampling, the
 * support-function duality); impedance and constraint rows (direct
 * evaluation); the solvers (KKT optimality, Newton == CG, finite differences
 * of the elliptic cone cost); the reference's capacity rules (row counts).
 *
 * Parity unpinned (MuJoCo C would be needed to check them): MuJoCo's own
 * contact points, counts and frames for box-box, capsule-box and convex
 * (ccd) pairs beyond the closed forms; its line-search details and tolerance
 * scaling; warm-start selection; the elliptic cone's mu regularisation with
 * impratio and the friction rows' R; the implicitfast derivative terms;
 * mj_setConst's invweight0 / meaninertia (checked only against this file's
 * own FP64 setConst).
 */
#ifndef REF64_H
#define REF64_H

#ifdef __cplusplus
extern "C" {
#endif

/* Raw compiled model (mjModel field names). Pointers are borrowed. */
typedef struct {
  int nq, nv, nu, nbody, njnt, ngeom;
  double timestep, gravity[3], tolerance, ls_tolerance, impratio;
  int iterations, ls_iterations, solver, cone, integrator, disableflags;
  const int *body_parentid, *body_rootid, *body_weldid, *body_jntnum, *body_jntadr, *body_dofnum, *body_dofadr;
  const double *body_pos, *body_quat, *body_ipos, *body_iquat, *body_mass, *body_inertia;
  const int *jnt_type, *jnt_qposadr, *jnt_dofadr, *jnt_bodyid, *jnt_limited;
  const double *jnt_pos, *jnt_axis, *jnt_range, *jnt_margin, *jnt_solref, *jnt_solimp, *jnt_stiffness;
  const int *dof_bodyid, *dof_jntid, *dof_parentid;
  const double *dof_armature, *dof_damping, *dof_frictionloss, *dof_solref, *dof_solimp;
  const int *geom_type, *geom_contype, *geom_conaffinity, *geom_condim, *geom_bodyid, *geom_priority;
  const double *geom_size, *geom_pos, *geom_quat, *geom_friction, *geom_solref, *geom_solimp,
      *geom_margin, *geom_gap, *geom_solmix, *geom_rbound;
  const int *actuator_trnid, *actuator_ctrllimited, *actuator_forcelimited;
  const double *actuator_gear, *actuator_gainprm, *actuator_biasprm, *actuator_ctrlrange, *actuator_forcerange;
  int nexclude;
  const int* exclude_signature;
  int npair;
  const int *pair_geom1, *pair_geom2, *pair_dim;
  const double *pair_friction, *pair_solref, *pair_solimp, *pair_margin, *pair_gap;
  const double *qpos0, *qpos_spring;
} orc_model_t;

/* Limits of the oracle's per-env work arrays. */
#define ORC_MAXCON 512
#define ORC_MAXEFC 2048

typedef struct {
  int g1, g2, b1, b2, condim;
  double friction[5], solref[2], solimp[5], margin, gap;
} orc_pair_t;

typedef struct {
  double dist, pos[3], frame[9];
  int geom1, geom2, pair, dim;
  double friction[5], solref[2], solimp[5], includemargin;
} orc_contact_t;

/* Prepared model: raw fields + derived tables (pair table, invweight0,
 * meaninertia) computed by the oracle itself from the raw fields. */
typedef struct orc_prep orc_prep_t;

orc_prep_t* orc_prepare(const orc_model_t* m);
void orc_free(orc_prep_t* p);
void orc_set_caps(orc_prep_t* p, int con_cap, int efc_cap);
int orc_npair(const orc_prep_t* p);
const orc_pair_t* orc_pairs(const orc_prep_t* p);
const double* orc_body_invweight0(const orc_prep_t* p);
const double* orc_dof_invweight0(const orc_prep_t* p);
double orc_meaninertia(const orc_prep_t* p);

/* Per-env working data of one step (everything the pipeline computes). */
typedef struct orc_data orc_data_t;
orc_data_t* orc_data_new(const orc_prep_t* p);
void orc_data_free(orc_data_t* d);

/* state in / out */
void orc_set_state(orc_data_t* d, const double* qpos, const double* qvel, const double* qacc_ws, const double* ctrl);
void orc_get_state(const orc_data_t* d, double* qpos, double* qvel, double* qacc_ws);
/* domain randomisation of this data (reference src/batched.cpp:1508-1538):
 * mass_scale [nbody] (mass and inertia), com_delta [nbody][3] (body-frame COM
 * shift), friction_scale [1] (all contact friction coefficients),
 * gain_scale [nu] (actuator force), drag [nfree] (F = -k |v| v on each free
 * joint's translation). NULL leaves a parameter at its current value
 * (initially the identity). */
void orc_set_env_params(orc_data_t* d, const double* mass_scale, const double* com_delta, const double* friction_scale,
                        const double* gain_scale, const double* drag);
int orc_nfree(const orc_data_t* d);
/* mj_ray: one primitive in its own frame (x along p + x v, or -1), and the
 * nearest geom with geom_ok[g] != 0 at d's geom poses (after orc_kinematics) */
double orc_ray_geom(int type, const double* size, const double* p, const double* v);
double orc_ray(const orc_prep_t* p, const orc_data_t* d, const double* pnt, const double* vec, const int* geom_ok,
               int* geomid);

/* One physics step (mj_step): forward pipeline, solver, integrator.
 * Returns 0, or a negative status (>0 capacity overflow bits). */
int orc_step(const orc_prep_t* p, orc_data_t* d);
void orc_inject_contacts(const orc_prep_t* p, orc_data_t* d, int n, const int* pair, const double* dist,
                         const double* pos, const double* normal);
/* Forward pipeline only (no integration), leaving every field readable. */
int orc_forward(const orc_prep_t* p, orc_data_t* d);
/* Kinematics-only pieces (for unit checks). */
void orc_kinematics(const orc_prep_t* p, orc_data_t* d);

/* Field access: copies a named field into out (capacity cap doubles);
 * returns the number of doubles, or -1 for an unknown name. */
int orc_get(const orc_data_t* d, const char* name, double* out, int cap);
int orc_ncon(const orc_data_t* d);
const orc_contact_t* orc_contacts(const orc_data_t* d);
int orc_nefc(const orc_data_t* d);
int orc_solver_niter(const orc_data_t* d);
/* cost, gradient and Hessian of the last forward's constraint problem at qacc */
int orc_cost(const orc_prep_t* p, orc_data_t* d, const double* qacc, double* cost, double* grad, double* H);

/* Narrow-phase entry for unit tests: collide one prepared pair given the two
 * geoms' world poses; writes up to maxc contacts, returns the count. */
int orc_collide_pair(const orc_prep_t* p, int pair, const double* xpos1, const double* xmat1,
                     const double* xpos2, const double* xmat2, orc_contact_t* out, int maxc);

/* Batched stepping: nsub substeps per env with the same ctrl, over nthreads
 * POSIX threads (env i goes to thread i % nthreads, as reference
 * src/mjb.cpp:826-838 assigns envs). Arrays are env-major. Returns the number
 * of envs whose step reported an overflow. */
int orc_batch_step(const orc_prep_t* p, int nenv, double* qpos, double* qvel, double* qacc_ws,
                   const double* ctrl, int nsub, int nthreads, int* niter_out);

/* The oracle's own math routines, for known-answer tests: quat_mul,
 * quat2mat, quat_rotate, axisangle, quat_integrate, cross, mul_inert,
 * cross_motion, cross_force, seg_seg, make_frame, impedance, cholesky,
 * chol_solve. Returns the output length, -1 on failure. */
int orc_math(const char* op, const double* in, double* out);

#ifdef __cplusplus
}
#endif
#endif
