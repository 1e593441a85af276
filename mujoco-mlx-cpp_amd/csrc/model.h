// Host-side model of the HIP engine: the MJCF subset compiled into flat FP64
// arrays with MuJoCo field names, plus the static tables the kernels read.
//
// Field meaning follows mjModel (MuJoCo 3.5.0, the version the reference pins
// at pyproject.toml:2,12); the reference converts the same fields to FP32 in
// convert_model (src/io.cpp:255-499) and derives its collision pair table in
// Model::init_cache (src/io.cpp:914-1070).
#pragma once

#include <map>
#include <string>
#include <vector>

namespace mjh {

enum GeomType { GEOM_PLANE = 0, GEOM_HFIELD = 1, GEOM_SPHERE = 2, GEOM_CAPSULE = 3,
                GEOM_ELLIPSOID = 4, GEOM_CYLINDER = 5, GEOM_BOX = 6, GEOM_MESH = 7 };
enum JointType { JNT_FREE = 0, JNT_BALL = 1, JNT_SLIDE = 2, JNT_HINGE = 3 };
enum SolverType { SOL_PGS = 0, SOL_CG = 1, SOL_NEWTON = 2 };
enum ConeType { CONE_PYRAMIDAL = 0, CONE_ELLIPTIC = 1 };
enum IntegratorType { INT_EULER = 0, INT_RK4 = 1, INT_IMPLICIT = 2, INT_IMPLICITFAST = 3 };
enum DisableBit {
  DSBL_CONSTRAINT = 1 << 0, DSBL_EQUALITY = 1 << 1, DSBL_FRICTIONLOSS = 1 << 2,
  DSBL_LIMIT = 1 << 3, DSBL_CONTACT = 1 << 4, DSBL_PASSIVE = 1 << 5,
  DSBL_GRAVITY = 1 << 6, DSBL_CLAMPCTRL = 1 << 7, DSBL_WARMSTART = 1 << 8,
  DSBL_FILTERPARENT = 1 << 9, DSBL_ACTUATION = 1 << 10, DSBL_REFSAFE = 1 << 11,
  DSBL_EULERDAMP = 1 << 15,
};

struct Model {
  std::string name;
  // <option>
  double timestep = 0.002;
  double gravity[3] = {0, 0, -9.81};
  double magnetic[3] = {0, -0.5, 0};
  double tolerance = 1e-8, ls_tolerance = 0.01, impratio = 1.0;
  int iterations = 100, ls_iterations = 50;
  int solver = SOL_NEWTON, cone = CONE_PYRAMIDAL, integrator = INT_EULER;
  int disableflags = 0;

  int nq = 0, nv = 0, nu = 0, nbody = 0, njnt = 0, ngeom = 0, nsite = 0;

  // bodies (world is body 0)
  std::vector<int> body_parentid, body_rootid, body_weldid, body_jntnum, body_jntadr,
      body_dofnum, body_dofadr, body_geomnum, body_geomadr;
  std::vector<double> body_pos, body_quat, body_ipos, body_iquat, body_mass,
      body_inertia, body_invweight0, body_subtreemass;
  std::vector<std::string> body_name;

  // joints
  std::vector<int> jnt_type, jnt_qposadr, jnt_dofadr, jnt_bodyid, jnt_limited;
  std::vector<double> jnt_pos, jnt_axis, jnt_range, jnt_margin, jnt_solref, jnt_solimp,
      jnt_stiffness;
  std::vector<std::string> jnt_name;

  // dofs
  std::vector<int> dof_bodyid, dof_jntid, dof_parentid;
  std::vector<double> dof_armature, dof_damping, dof_frictionloss, dof_invweight0,
      dof_solref, dof_solimp;

  // geoms
  std::vector<int> geom_type, geom_contype, geom_conaffinity, geom_condim, geom_bodyid,
      geom_priority, geom_group;
  std::vector<double> geom_size, geom_pos, geom_quat, geom_friction, geom_solref,
      geom_solimp, geom_margin, geom_gap, geom_solmix, geom_rbound;
  std::vector<std::string> geom_name;

  // actuators (joint transmission only)
  std::vector<int> actuator_trnid, actuator_ctrllimited, actuator_forcelimited;
  std::vector<double> actuator_gear, actuator_gainprm, actuator_biasprm,
      actuator_ctrlrange, actuator_forcerange;
  std::vector<std::string> actuator_name;

  // sites (frames fixed to a body)
  std::vector<int> site_bodyid;
  std::vector<double> site_pos, site_quat;
  std::vector<std::string> site_name;

  // sensors: MuJoCo mjtSensor / mjtObj ids, sensordata address and width
  int nsensor = 0, nsensordata = 0;
  std::vector<int> sensor_type, sensor_objtype, sensor_objid, sensor_adr, sensor_dim;
  std::vector<std::string> sensor_name;

  // explicit <pair> and <exclude>
  std::vector<int> pair_geom1, pair_geom2, pair_dim;
  std::vector<double> pair_friction, pair_solref, pair_solimp, pair_margin, pair_gap;
  std::vector<int> exclude_signature;  // (min(b1,b2) << 16) | max(b1,b2)

  std::vector<double> qpos0, qpos_spring;
  double meaninertia = 1.0;

  // MuJoCo mjtObj ids: 1 body (2 xbody), 3 joint, 5 geom, 6 site, 19 actuator, 20 sensor
  int name2id(int objtype, const std::string& nm) const;
  // the name of object id, or nullptr (unnamed, out of range, other types)
  const char* id2name(int objtype, int id) const;
};

// One candidate geom pair of the static broadphase table, with the mixed
// contact parameters (MuJoCo mj_contactParam semantics; the reference restates
// them in src/io.cpp:922-959 and bakes the pair table at src/io.cpp:961-1070).
struct CollisionPair {
  int g1, g2;        // g1 has the lower geom type (normal points g1 -> g2)
  int b1, b2;        // body ids
  int condim;
  double friction[5];
  double solref[2], solimp[5];
  double margin, gap;
};

// Compile an MJCF string (subset: compiler/option/default/worldbody/body/
// joint/freejoint/geom/inertial/site/actuator{position,motor,velocity,general}/
// contact{pair,exclude}). Throws std::runtime_error with a message on error.
struct LoadOptions { int self_collision = 1; int max_condim = 0; };
Model compile_mjcf(const std::string& xml, const LoadOptions& opt);

// Static tables derived from a compiled model.
std::vector<CollisionPair> build_pairs(const Model& m);

// FP64 host kinematics/CRB at qpos0 used by the compiler for body/dof
// invweight0 and meaninertia (MuJoCo mj_setConst).
void set_const(Model& m);

// small math (wxyz quaternions, row-major 3x3)
void quat2mat(const double q[4], double R[9]);
void mulquat(const double a[4], const double b[4], double out[4]);
void rotvec(const double q[4], const double v[3], double out[3]);
void mat2quat(const double R[9], double q[4]);
void normquat(double q[4]);

}  // namespace mjh
