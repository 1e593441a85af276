// MJCF-subset parser and compiler for the HIP engine.
//
// The reference delegates parsing and compilation to MuJoCo C (mj_loadXML /
// mj_parseXML + mj_compile, reference src/io.cpp:563-601) and converts the
// result to FP32 (src/io.cpp:255-499). MuJoCo C is not part of this engine, so
// the subset of MJCF used by the benchmark scenes is compiled here with MuJoCo's
// documented semantics: default classes, body/joint/geom frames, inertia from
// geoms (density * volume), qpos0, actuator gain/bias parameters and the
// load-time options of mjmlx_load_model_ex (reference src/io.cpp:567-594).
#include "model.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace mjh {

// ---------------------------------------------------------------- small math
void quat2mat(const double q[4], double R[9]) {
  double w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = w * w + x * x - y * y - z * z; R[1] = 2 * (x * y - w * z);         R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z);         R[4] = w * w - x * x + y * y - z * z; R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y);         R[7] = 2 * (y * z + w * x);         R[8] = w * w - x * x - y * y + z * z;
}
void mulquat(const double a[4], const double b[4], double o[4]) {
  double r0 = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  double r1 = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  double r2 = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  double r3 = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
  o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3;
}
void rotvec(const double q[4], const double v[3], double o[3]) {
  double R[9]; quat2mat(q, R);
  double t0 = R[0] * v[0] + R[1] * v[1] + R[2] * v[2];
  double t1 = R[3] * v[0] + R[4] * v[1] + R[5] * v[2];
  double t2 = R[6] * v[0] + R[7] * v[1] + R[8] * v[2];
  o[0] = t0; o[1] = t1; o[2] = t2;
}
void normquat(double q[4]) {
  double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (n < 1e-15) { q[0] = 1; q[1] = q[2] = q[3] = 0; return; }
  for (int i = 0; i < 4; i++) q[i] /= n;
}
void mat2quat(const double R[9], double q[4]) {
  double tr = R[0] + R[4] + R[8];
  if (tr > 0) {
    double s = std::sqrt(tr + 1.0) * 2;
    q[0] = 0.25 * s; q[1] = (R[7] - R[5]) / s; q[2] = (R[2] - R[6]) / s; q[3] = (R[3] - R[1]) / s;
  } else if (R[0] > R[4] && R[0] > R[8]) {
    double s = std::sqrt(1.0 + R[0] - R[4] - R[8]) * 2;
    q[0] = (R[7] - R[5]) / s; q[1] = 0.25 * s; q[2] = (R[1] + R[3]) / s; q[3] = (R[2] + R[6]) / s;
  } else if (R[4] > R[8]) {
    double s = std::sqrt(1.0 + R[4] - R[0] - R[8]) * 2;
    q[0] = (R[2] - R[6]) / s; q[1] = (R[1] + R[3]) / s; q[2] = 0.25 * s; q[3] = (R[5] + R[7]) / s;
  } else {
    double s = std::sqrt(1.0 + R[8] - R[0] - R[4]) * 2;
    q[0] = (R[3] - R[1]) / s; q[1] = (R[2] + R[6]) / s; q[2] = (R[5] + R[7]) / s; q[3] = 0.25 * s;
  }
  if (q[0] < 0) for (int i = 0; i < 4; i++) q[i] = -q[i];
  normquat(q);
}
static void axisangle2quat(const double ax[3], double a, double q[4]) {
  double n = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
  if (n < 1e-15) { q[0] = 1; q[1] = q[2] = q[3] = 0; return; }
  double s = std::sin(0.5 * a) / n;
  q[0] = std::cos(0.5 * a); q[1] = ax[0] * s; q[2] = ax[1] * s; q[3] = ax[2] * s;
}
// minimal rotation taking +z to vec
static void quatz2vec(const double v[3], double q[4]) {
  double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  q[0] = 1; q[1] = q[2] = q[3] = 0;
  if (n < 1e-15) return;
  double d[3] = {v[0] / n, v[1] / n, v[2] / n};
  double c = d[2];  // dot(z, d)
  double ax[3] = {-d[1], d[0], 0};  // z x d
  double s = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1]);
  if (s < 1e-10) {
    if (c < 0) { q[0] = 0; q[1] = 1; q[2] = 0; q[3] = 0; }
    return;
  }
  axisangle2quat(ax, std::atan2(s, c), q);
}

// symmetric 3x3 eigen-decomposition (cyclic Jacobi); columns of V are eigenvectors
static void eig3(const double A_in[9], double w[3], double V[9]) {
  double A[9]; std::memcpy(A, A_in, sizeof(A));
  for (int i = 0; i < 9; i++) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 50; sweep++) {
    double off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
    if (off < 1e-30) break;
    for (int p = 0; p < 2; p++) for (int qi = p + 1; qi < 3; qi++) {
      double apq = A[3 * p + qi];
      if (std::fabs(apq) < 1e-300) continue;
      double app = A[3 * p + p], aqq = A[3 * qi + qi];
      double theta = 0.5 * (aqq - app) / apq;
      double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
      double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
      for (int k = 0; k < 3; k++) {  // A = J^T A J
        double akp = A[3 * k + p], akq = A[3 * k + qi];
        A[3 * k + p] = c * akp - s * akq; A[3 * k + qi] = s * akp + c * akq;
      }
      for (int k = 0; k < 3; k++) {
        double apk = A[3 * p + k], aqk = A[3 * qi + k];
        A[3 * p + k] = c * apk - s * aqk; A[3 * qi + k] = s * apk + c * aqk;
      }
      for (int k = 0; k < 3; k++) {
        double vkp = V[3 * k + p], vkq = V[3 * k + qi];
        V[3 * k + p] = c * vkp - s * vkq; V[3 * k + qi] = s * vkp + c * vkq;
      }
    }
  }
  w[0] = A[0]; w[1] = A[4]; w[2] = A[8];
  // right-handed
  double det = V[0] * (V[4] * V[8] - V[5] * V[7]) - V[1] * (V[3] * V[8] - V[5] * V[6]) + V[2] * (V[3] * V[7] - V[4] * V[6]);
  if (det < 0) for (int k = 0; k < 3; k++) V[3 * k + 2] = -V[3 * k + 2];
}

// ---------------------------------------------------------------- XML
struct XNode {
  std::string tag;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<std::unique_ptr<XNode>> kids;
  const std::string* get(const std::string& k) const {
    for (auto& a : attrs) if (a.first == k) return &a.second;
    return nullptr;
  }
};

class XmlParser {
 public:
  explicit XmlParser(const std::string& s) : s_(s) {}
  std::unique_ptr<XNode> parse() {
    skip_misc();
    auto root = element();
    return root;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;
  [[noreturn]] void fail(const std::string& msg) {
    throw std::runtime_error("MJCF parse error at offset " + std::to_string(p_) + ": " + msg);
  }
  void ws() { while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) p_++; }
  bool starts(const char* t) { return s_.compare(p_, std::strlen(t), t) == 0; }
  void skip_misc() {
    for (;;) {
      ws();
      if (starts("<?")) { size_t e = s_.find("?>", p_); if (e == std::string::npos) fail("unterminated <?"); p_ = e + 2; }
      else if (starts("<!--")) { size_t e = s_.find("-->", p_); if (e == std::string::npos) fail("unterminated comment"); p_ = e + 3; }
      else if (starts("<!")) { size_t e = s_.find('>', p_); if (e == std::string::npos) fail("unterminated <!"); p_ = e + 1; }
      else break;
    }
  }
  std::string name() {
    size_t b = p_;
    while (p_ < s_.size() && (std::isalnum((unsigned char)s_[p_]) || s_[p_] == '_' || s_[p_] == '-' || s_[p_] == ':' || s_[p_] == '.')) p_++;
    if (b == p_) fail("expected a name");
    return s_.substr(b, p_ - b);
  }
  std::unique_ptr<XNode> element() {
    if (p_ >= s_.size() || s_[p_] != '<') fail("expected '<'");
    p_++;
    auto n = std::make_unique<XNode>();
    n->tag = name();
    for (;;) {
      ws();
      if (p_ >= s_.size()) fail("unexpected end in tag");
      if (starts("/>")) { p_ += 2; return n; }
      if (s_[p_] == '>') { p_++; break; }
      std::string k = name();
      ws();
      if (p_ >= s_.size() || s_[p_] != '=') fail("expected '=' after attribute " + k);
      p_++; ws();
      char qc = s_[p_];
      if (qc != '"' && qc != '\'') fail("expected quote");
      size_t e = s_.find(qc, p_ + 1);
      if (e == std::string::npos) fail("unterminated attribute");
      n->attrs.emplace_back(k, s_.substr(p_ + 1, e - p_ - 1));
      p_ = e + 1;
    }
    for (;;) {
      // text content is ignored
      while (p_ < s_.size() && s_[p_] != '<') p_++;
      if (p_ >= s_.size()) fail("unterminated element " + n->tag);
      if (starts("<!--")) { size_t e = s_.find("-->", p_); if (e == std::string::npos) fail("unterminated comment"); p_ = e + 3; continue; }
      if (starts("</")) {
        p_ += 2; std::string t = name(); ws();
        if (t != n->tag) fail("mismatched </" + t + "> for <" + n->tag + ">");
        if (p_ >= s_.size() || s_[p_] != '>') fail("expected '>'");
        p_++;
        return n;
      }
      n->kids.push_back(element());
    }
  }
};

// ---------------------------------------------------------------- attribute helpers
static std::vector<double> parse_nums(const std::string& s) {
  std::vector<double> v;
  const char* c = s.c_str();
  char* end;
  for (;;) {
    while (*c && (std::isspace((unsigned char)*c) || *c == ',')) c++;
    if (!*c) break;
    double x = std::strtod(c, &end);
    if (end == c) throw std::runtime_error("MJCF: bad number list '" + s + "'");
    v.push_back(x);
    c = end;
  }
  return v;
}

using AttrMap = std::map<std::string, std::string>;

struct DefaultClass {
  std::string parent;
  std::map<std::string, AttrMap> el;  // tag -> attrs
};

struct Ctx {
  std::map<std::string, DefaultClass> classes;
  bool radian = false;
  bool autolimits = true;
  std::string eulerseq = "xyz";
};

static void read_defaults(Ctx& cx, const XNode& n, const std::string& cls, const std::string& parent) {
  DefaultClass& dc = cx.classes[cls];
  dc.parent = parent;
  if (!parent.empty()) dc.el = cx.classes[parent].el;  // inherit
  for (auto& k : n.kids) {
    if (k->tag == "default") {
      const std::string* c = k->get("class");
      if (!c) throw std::runtime_error("MJCF: nested <default> without class");
      read_defaults(cx, *k, *c, cls);
    } else {
      AttrMap& am = dc.el[k->tag];
      for (auto& a : k->attrs) am[a.first] = a.second;
    }
  }
}

// merged attributes of element `n` given the active class
static AttrMap resolve(const Ctx& cx, const XNode& n, const std::string& active_cls) {
  std::string cls = active_cls;
  if (const std::string* c = n.get("class")) cls = *c;
  AttrMap am;
  auto it = cx.classes.find(cls);
  if (it == cx.classes.end()) {
    if (cls != "main") throw std::runtime_error("MJCF: unknown default class '" + cls + "'");
  } else {
    // actuator shortcuts share the <general> defaults of their class
    std::vector<std::string> tags = {n.tag};
    if (n.tag == "position" || n.tag == "motor" || n.tag == "velocity") tags.insert(tags.begin(), "general");
    for (auto& t : tags) {
      auto e = it->second.el.find(t);
      if (e != it->second.el.end()) for (auto& kv : e->second) am[kv.first] = kv.second;
    }
  }
  for (auto& a : n.attrs) am[a.first] = a.second;
  return am;
}

static bool has(const AttrMap& a, const char* k) { return a.count(k) > 0; }
static std::vector<double> nums(const AttrMap& a, const char* k) { return parse_nums(a.at(k)); }
static double num(const AttrMap& a, const char* k, double def) {
  auto it = a.find(k);
  if (it == a.end()) return def;
  auto v = parse_nums(it->second);
  if (v.empty()) throw std::runtime_error(std::string("MJCF: empty attribute ") + k);
  return v[0];
}
static void vec(const AttrMap& a, const char* k, double* out, int n, const double* def) {
  auto it = a.find(k);
  if (it == a.end()) { for (int i = 0; i < n; i++) out[i] = def[i]; return; }
  auto v = parse_nums(it->second);
  if ((int)v.size() < n) {
    // MuJoCo fills missing trailing values with defaults
    for (int i = 0; i < n; i++) out[i] = i < (int)v.size() ? v[i] : def[i];
    return;
  }
  for (int i = 0; i < n; i++) out[i] = v[i];
}
static std::string str(const AttrMap& a, const char* k, const std::string& def) {
  auto it = a.find(k);
  return it == a.end() ? def : it->second;
}
static bool boolean(const AttrMap& a, const char* k, bool def) {
  auto it = a.find(k);
  if (it == a.end()) return def;
  return it->second == "true" || it->second == "1";
}

// frame orientation from quat / axisangle / euler / zaxis attributes
static void orientation(const Ctx& cx, const AttrMap& a, double q[4]) {
  q[0] = 1; q[1] = q[2] = q[3] = 0;
  double angs = cx.radian ? 1.0 : M_PI / 180.0;
  if (has(a, "quat")) {
    auto v = nums(a, "quat");
    if (v.size() != 4) throw std::runtime_error("MJCF: quat needs 4 numbers");
    for (int i = 0; i < 4; i++) q[i] = v[i];
    normquat(q);
  } else if (has(a, "axisangle")) {
    auto v = nums(a, "axisangle");
    if (v.size() != 4) throw std::runtime_error("MJCF: axisangle needs 4 numbers");
    axisangle2quat(v.data(), v[3] * angs, q);
  } else if (has(a, "euler")) {
    auto v = nums(a, "euler");
    if (v.size() != 3) throw std::runtime_error("MJCF: euler needs 3 numbers");
    for (int i = 0; i < 3; i++) {
      char c = cx.eulerseq[i];
      double ax[3] = {0, 0, 0};
      ax[std::tolower(c) - 'x'] = 1;
      double qi[4]; axisangle2quat(ax, v[i] * angs, qi);
      if (std::islower((unsigned char)c)) mulquat(q, qi, q);  // intrinsic: rotate with the frame
      else mulquat(qi, q, q);                                 // extrinsic: fixed axes
    }
  } else if (has(a, "zaxis")) {
    auto v = nums(a, "zaxis");
    quatz2vec(v.data(), q);
  }
}

// ---------------------------------------------------------------- compiler
struct GeomInertia { double mass; double pos[3]; double R[9]; double I[3]; };

static double geom_volume(int type, const double* s) {
  switch (type) {
    case GEOM_SPHERE: return 4.0 / 3.0 * M_PI * s[0] * s[0] * s[0];
    case GEOM_CAPSULE: return M_PI * s[0] * s[0] * 2 * s[1] + 4.0 / 3.0 * M_PI * s[0] * s[0] * s[0];
    case GEOM_CYLINDER: return M_PI * s[0] * s[0] * 2 * s[1];
    case GEOM_ELLIPSOID: return 4.0 / 3.0 * M_PI * s[0] * s[1] * s[2];
    case GEOM_BOX: return 8.0 * s[0] * s[1] * s[2];
    default: return 0;
  }
}
// principal inertia about the geom centre, in the geom frame
static void geom_inertia(int type, const double* s, double mass, double density, double I[3]) {
  switch (type) {
    case GEOM_SPHERE: I[0] = I[1] = I[2] = 0.4 * mass * s[0] * s[0]; break;
    case GEOM_BOX:
      I[0] = mass / 3.0 * (s[1] * s[1] + s[2] * s[2]);
      I[1] = mass / 3.0 * (s[0] * s[0] + s[2] * s[2]);
      I[2] = mass / 3.0 * (s[0] * s[0] + s[1] * s[1]);
      break;
    case GEOM_CYLINDER:
      I[0] = I[1] = mass * (s[0] * s[0] / 4.0 + s[1] * s[1] / 3.0);
      I[2] = mass * s[0] * s[0] / 2.0;
      break;
    case GEOM_ELLIPSOID:
      I[0] = mass / 5.0 * (s[1] * s[1] + s[2] * s[2]);
      I[1] = mass / 5.0 * (s[0] * s[0] + s[2] * s[2]);
      I[2] = mass / 5.0 * (s[0] * s[0] + s[1] * s[1]);
      break;
    case GEOM_CAPSULE: {
      // cylinder of half-length h plus two hemispherical caps; each cap's
      // centroid sits 3r/8 beyond the cylinder end (parallel-axis theorem)
      double r = s[0], h = s[1];
      double mc = density * M_PI * r * r * 2 * h, ms = density * 4.0 / 3.0 * M_PI * r * r * r;
      double tot = mc + ms;
      double sc = tot > 0 ? mass / tot : 0;  // honours an explicit geom mass
      mc *= sc; ms *= sc;
      I[0] = I[1] = mc * (r * r / 4.0 + h * h / 3.0) + ms * (0.4 * r * r + h * h + 0.75 * h * r);
      I[2] = mc * r * r / 2.0 + ms * 0.4 * r * r;
      break;
    }
    default: I[0] = I[1] = I[2] = 0;
  }
}

struct BodyBuild {
  std::vector<GeomInertia> gi;
  bool explicit_inertial = false;
};

class Compiler {
 public:
  Compiler(const std::string& xml, const LoadOptions& o) : xml_(xml), opt_(o) {}
  Model run();

 private:
  const std::string& xml_;
  LoadOptions opt_;
  Ctx cx_;
  Model m_;
  std::vector<BodyBuild> bb_;
  std::vector<bool> body_robot_;  // has a joint in itself or an ancestor

  void option(const XNode& n);
  void body(const XNode& n, int parent, const std::string& childclass, bool robot);
  void add_geom(const XNode& n, int b, const std::string& cls, bool robot);
  void add_joint(const AttrMap& a, int b, bool is_free);
  void actuators(const XNode& n);
  void add_site(const XNode& n, int b, const std::string& cls);
  void sensors(const XNode& n);
  void contact(const XNode& n);
  void finish_inertia();
};

void Compiler::option(const XNode& n) {
  AttrMap a; for (auto& kv : n.attrs) a[kv.first] = kv.second;
  m_.timestep = num(a, "timestep", m_.timestep);
  if (has(a, "gravity")) { auto g = nums(a, "gravity"); for (int i = 0; i < 3 && i < (int)g.size(); i++) m_.gravity[i] = g[i]; }
  if (has(a, "magnetic")) { auto g = nums(a, "magnetic"); for (int i = 0; i < 3 && i < (int)g.size(); i++) m_.magnetic[i] = g[i]; }
  m_.iterations = (int)num(a, "iterations", m_.iterations);
  m_.ls_iterations = (int)num(a, "ls_iterations", m_.ls_iterations);
  m_.tolerance = num(a, "tolerance", m_.tolerance);
  m_.ls_tolerance = num(a, "ls_tolerance", m_.ls_tolerance);
  m_.impratio = num(a, "impratio", m_.impratio);
  std::string s = str(a, "solver", "Newton");
  m_.solver = s == "PGS" ? SOL_PGS : s == "CG" ? SOL_CG : SOL_NEWTON;
  s = str(a, "cone", "pyramidal");
  m_.cone = s == "elliptic" ? CONE_ELLIPTIC : CONE_PYRAMIDAL;
  s = str(a, "integrator", "Euler");
  m_.integrator = s == "RK4" ? INT_RK4 : s == "implicit" ? INT_IMPLICIT : s == "implicitfast" ? INT_IMPLICITFAST : INT_EULER;
  for (auto& k : n.kids) {
    if (k->tag != "flag") continue;
    static const std::pair<const char*, int> flags[] = {
        {"constraint", DSBL_CONSTRAINT}, {"equality", DSBL_EQUALITY}, {"frictionloss", DSBL_FRICTIONLOSS},
        {"limit", DSBL_LIMIT}, {"contact", DSBL_CONTACT}, {"passive", DSBL_PASSIVE}, {"gravity", DSBL_GRAVITY},
        {"clampctrl", DSBL_CLAMPCTRL}, {"warmstart", DSBL_WARMSTART}, {"filterparent", DSBL_FILTERPARENT},
        {"actuation", DSBL_ACTUATION}, {"refsafe", DSBL_REFSAFE}, {"eulerdamp", DSBL_EULERDAMP}};
    for (auto& f : flags) {
      const std::string* v = k->get(f.first);
      if (v && *v == "disable") m_.disableflags |= f.second;
      if (v && *v == "enable") m_.disableflags &= ~f.second;
    }
  }
}

void Compiler::add_joint(const AttrMap& a, int b, bool is_free) {
  int id = m_.njnt++;
  std::string ts = is_free ? "free" : str(a, "type", "hinge");
  int type = ts == "free" ? JNT_FREE : ts == "ball" ? JNT_BALL : ts == "slide" ? JNT_SLIDE : ts == "hinge" ? JNT_HINGE : -1;
  if (type < 0) throw std::runtime_error("MJCF: unsupported joint type " + ts);
  if (type == JNT_FREE && m_.body_parentid[b] != 0) throw std::runtime_error("MJCF: free joint must belong to a child of the world body");
  m_.jnt_type.push_back(type);
  m_.jnt_bodyid.push_back(b);
  m_.jnt_name.push_back(str(a, "name", ""));
  m_.jnt_qposadr.push_back(m_.nq);
  m_.jnt_dofadr.push_back(m_.nv);
  double z3[3] = {0, 0, 0}, zax[3] = {0, 0, 1};
  double pos[3], axis[3];
  vec(a, "pos", pos, 3, z3);
  vec(a, "axis", axis, 3, zax);
  double an = std::sqrt(axis[0] * axis[0] + axis[1] * axis[1] + axis[2] * axis[2]);
  if (an < 1e-12) throw std::runtime_error("MJCF: zero joint axis");
  for (int i = 0; i < 3; i++) { m_.jnt_pos.push_back(type == JNT_FREE ? 0.0 : pos[i]); m_.jnt_axis.push_back(axis[i] / an); }
  double range[2] = {0, 0};
  bool limited = false;
  if (type == JNT_HINGE || type == JNT_SLIDE) {
    double zr[2] = {0, 0};
    vec(a, "range", range, 2, zr);
    if (type == JNT_HINGE && !cx_.radian) { range[0] *= M_PI / 180.0; range[1] *= M_PI / 180.0; }
    std::string lim = str(a, "limited", "auto");
    limited = lim == "true" || (lim == "auto" && cx_.autolimits && has(a, "range"));
    if (limited && range[0] >= range[1]) throw std::runtime_error("MJCF: joint range lower must be < upper");
  }
  m_.jnt_range.push_back(range[0]); m_.jnt_range.push_back(range[1]);
  m_.jnt_limited.push_back(limited ? 1 : 0);
  m_.jnt_margin.push_back(num(a, "margin", 0.0));
  double dsr[2] = {0.02, 1.0}, dsi[5] = {0.9, 0.95, 0.001, 0.5, 2.0};
  double sr[2], si[5];
  vec(a, "solreflimit", sr, 2, dsr); vec(a, "solimplimit", si, 5, dsi);
  for (int i = 0; i < 2; i++) m_.jnt_solref.push_back(sr[i]);
  for (int i = 0; i < 5; i++) m_.jnt_solimp.push_back(si[i]);
  m_.jnt_stiffness.push_back(num(a, "stiffness", 0.0));
  int nqj = type == JNT_FREE ? 7 : type == JNT_BALL ? 4 : 1;
  int nvj = type == JNT_FREE ? 6 : type == JNT_BALL ? 3 : 1;
  double ref = num(a, "ref", 0.0), springref = num(a, "springref", 0.0);
  if (type == JNT_HINGE && !cx_.radian) { ref *= M_PI / 180.0; springref *= M_PI / 180.0; }
  if (type == JNT_FREE) {
    // qpos0 of a free joint is the body's initial pose (its parent is the world)
    for (int i = 0; i < 3; i++) { m_.qpos0.push_back(m_.body_pos[3 * b + i]); }
    for (int i = 0; i < 4; i++) { m_.qpos0.push_back(m_.body_quat[4 * b + i]); }
  } else if (type == JNT_BALL) {
    double q1[4] = {1, 0, 0, 0};
    for (int i = 0; i < 4; i++) m_.qpos0.push_back(q1[i]);
  } else {
    m_.qpos0.push_back(ref);
  }
  for (int i = 0; i < nqj; i++) m_.qpos_spring.push_back(nqj == 1 ? springref : m_.qpos0[m_.nq + i]);
  double srf[2], sif[5];
  vec(a, "solreffriction", srf, 2, dsr); vec(a, "solimpfriction", sif, 5, dsi);
  int first_dof = m_.nv;
  for (int i = 0; i < nvj; i++) {
    int d = m_.nv + i;
    m_.dof_bodyid.push_back(b);
    m_.dof_jntid.push_back(id);
    // parent dof: previous dof of this body, else last dof of the nearest ancestor with dofs
    int par = -1;
    if (d > first_dof || m_.body_dofnum[b] > 0) par = d - 1;
    else {
      int pb = m_.body_parentid[b];
      while (pb > 0 && m_.body_dofnum[pb] == 0) pb = m_.body_parentid[pb];
      if (pb > 0) par = m_.body_dofadr[pb] + m_.body_dofnum[pb] - 1;
    }
    m_.dof_parentid.push_back(par);
    m_.dof_armature.push_back(num(a, "armature", 0.0));
    m_.dof_damping.push_back(num(a, "damping", 0.0));
    m_.dof_frictionloss.push_back(num(a, "frictionloss", 0.0));
    for (int k = 0; k < 2; k++) m_.dof_solref.push_back(srf[k]);
    for (int k = 0; k < 5; k++) m_.dof_solimp.push_back(sif[k]);
  }
  if (m_.body_jntnum[b] == 0) { m_.body_jntadr[b] = id; m_.body_dofadr[b] = m_.nv; }
  m_.body_jntnum[b]++;
  m_.body_dofnum[b] += nvj;
  m_.nq += nqj;
  m_.nv += nvj;
}

void Compiler::add_geom(const XNode& n, int b, const std::string& cls, bool robot) {
  AttrMap a = resolve(cx_, n, cls);
  std::string ts = str(a, "type", "sphere");
  int type = ts == "plane" ? GEOM_PLANE : ts == "sphere" ? GEOM_SPHERE : ts == "capsule" ? GEOM_CAPSULE
           : ts == "ellipsoid" ? GEOM_ELLIPSOID : ts == "cylinder" ? GEOM_CYLINDER : ts == "box" ? GEOM_BOX : -1;
  if (type < 0) throw std::runtime_error("MJCF: unsupported geom type " + ts);
  double size[3] = {0, 0, 0};
  if (has(a, "size")) { auto v = nums(a, "size"); for (int i = 0; i < 3 && i < (int)v.size(); i++) size[i] = v[i]; }
  double pos[3] = {0, 0, 0}, q[4];
  if (has(a, "pos")) { auto v = nums(a, "pos"); for (int i = 0; i < 3; i++) pos[i] = v.at(i); }
  orientation(cx_, a, q);
  if (has(a, "fromto")) {
    if (type != GEOM_CAPSULE && type != GEOM_CYLINDER && type != GEOM_BOX && type != GEOM_ELLIPSOID)
      throw std::runtime_error("MJCF: fromto on unsupported geom type");
    auto v = nums(a, "fromto");
    if (v.size() != 6) throw std::runtime_error("MJCF: fromto needs 6 numbers");
    double d[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int i = 0; i < 3; i++) pos[i] = 0.5 * (v[i] + v[3 + i]);
    quatz2vec(d, q);
    if (type == GEOM_CAPSULE || type == GEOM_CYLINDER) size[1] = 0.5 * len;
    else { size[1] = size[0]; size[2] = 0.5 * len; }
  }
  for (int i = 0; i < 3; i++) if (type != GEOM_PLANE && i < (type == GEOM_SPHERE ? 1 : type == GEOM_CAPSULE || type == GEOM_CYLINDER ? 2 : 3) && size[i] <= 0)
    throw std::runtime_error("MJCF: geom size must be positive");
  if (type == GEOM_PLANE && b != 0 && body_robot_[b]) throw std::runtime_error("MJCF: planes must be static");
  int id = m_.ngeom++;
  m_.geom_type.push_back(type);
  m_.geom_bodyid.push_back(b);
  m_.geom_name.push_back(str(a, "name", ""));
  for (int i = 0; i < 3; i++) { m_.geom_size.push_back(size[i]); m_.geom_pos.push_back(pos[i]); }
  for (int i = 0; i < 4; i++) m_.geom_quat.push_back(q[i]);
  int contype = (int)num(a, "contype", 1), conaff = (int)num(a, "conaffinity", 1), condim = (int)num(a, "condim", 3);
  // load-time options (reference src/io.cpp:567-594): only geoms of bodies that
  // move (a joint on the body or an ancestor) are rewritten
  if (robot && !opt_.self_collision) conaff = 0;
  if (robot && opt_.max_condim > 0) condim = std::min(condim, opt_.max_condim);
  if (condim != 1 && condim != 3 && condim != 4 && condim != 6) throw std::runtime_error("MJCF: condim must be 1, 3, 4 or 6");
  m_.geom_contype.push_back(contype);
  m_.geom_conaffinity.push_back(conaff);
  m_.geom_condim.push_back(condim);
  m_.geom_priority.push_back((int)num(a, "priority", 0));
  m_.geom_group.push_back((int)num(a, "group", 0));
  double df[3] = {1, 0.005, 0.0001}, dsr[2] = {0.02, 1}, dsi[5] = {0.9, 0.95, 0.001, 0.5, 2};
  double f[3], sr[2], si[5];
  vec(a, "friction", f, 3, df); vec(a, "solref", sr, 2, dsr); vec(a, "solimp", si, 5, dsi);
  for (int i = 0; i < 3; i++) m_.geom_friction.push_back(f[i]);
  for (int i = 0; i < 2; i++) m_.geom_solref.push_back(sr[i]);
  for (int i = 0; i < 5; i++) m_.geom_solimp.push_back(si[i]);
  m_.geom_margin.push_back(num(a, "margin", 0));
  m_.geom_gap.push_back(num(a, "gap", 0));
  m_.geom_solmix.push_back(num(a, "solmix", 1));
  double rb = 0;
  switch (type) {
    case GEOM_SPHERE: rb = size[0]; break;
    case GEOM_CAPSULE: rb = size[0] + size[1]; break;
    case GEOM_CYLINDER: rb = std::sqrt(size[0] * size[0] + size[1] * size[1]); break;
    case GEOM_ELLIPSOID: rb = std::max(size[0], std::max(size[1], size[2])); break;
    case GEOM_BOX: rb = std::sqrt(size[0] * size[0] + size[1] * size[1] + size[2] * size[2]); break;
    default: rb = 0;
  }
  m_.geom_rbound.push_back(rb);
  if (m_.body_geomnum[b] == 0) m_.body_geomadr[b] = id;
  m_.body_geomnum[b]++;
  // mass properties
  if (type != GEOM_PLANE && b != 0) {
    double vol = geom_volume(type, size);
    double density = num(a, "density", 1000.0);
    double mass = has(a, "mass") ? num(a, "mass", 0) : density * vol;
    if (has(a, "mass") && vol > 0) density = mass / vol;
    GeomInertia g; g.mass = mass;
    for (int i = 0; i < 3; i++) g.pos[i] = pos[i];
    quat2mat(q, g.R);
    geom_inertia(type, size, mass, density, g.I);
    bb_[b].gi.push_back(g);
  }
}

void Compiler::body(const XNode& n, int parent, const std::string& childclass, bool robot_parent) {
  int b = m_.nbody++;
  AttrMap a = resolve(cx_, n, childclass);  // only frame attributes are used
  std::string cls = childclass;
  if (const std::string* c = n.get("childclass")) cls = *c;
  m_.body_parentid.push_back(parent);
  m_.body_name.push_back(str(a, "name", ""));
  double pos[3] = {0, 0, 0}, q[4];
  if (has(a, "pos")) { auto v = nums(a, "pos"); for (int i = 0; i < 3; i++) pos[i] = v.at(i); }
  orientation(cx_, a, q);
  for (int i = 0; i < 3; i++) m_.body_pos.push_back(pos[i]);
  for (int i = 0; i < 4; i++) m_.body_quat.push_back(q[i]);
  m_.body_jntnum.push_back(0); m_.body_jntadr.push_back(-1);
  m_.body_dofnum.push_back(0); m_.body_dofadr.push_back(-1);
  m_.body_geomnum.push_back(0); m_.body_geomadr.push_back(-1);
  bb_.emplace_back();
  bool has_joint = false;
  for (auto& k : n.kids) if (k->tag == "joint" || k->tag == "freejoint") has_joint = true;
  body_robot_.push_back(robot_parent || has_joint);
  bool robot = robot_parent || has_joint;
  // inertial frame defaults (overwritten by finish_inertia)
  m_.body_ipos.insert(m_.body_ipos.end(), {0, 0, 0});
  m_.body_iquat.insert(m_.body_iquat.end(), {1, 0, 0, 0});
  m_.body_mass.push_back(0);
  m_.body_inertia.insert(m_.body_inertia.end(), {0, 0, 0});
  for (auto& k : n.kids) {
    if (k->tag == "joint") add_joint(resolve(cx_, *k, cls), b, false);
    else if (k->tag == "freejoint") { AttrMap fa; for (auto& kv : k->attrs) fa[kv.first] = kv.second; add_joint(fa, b, true); }
  }
  for (auto& k : n.kids) {
    if (k->tag == "geom") add_geom(*k, b, cls, robot);
    else if (k->tag == "inertial") {
      AttrMap ia = resolve(cx_, *k, cls);
      bb_[b].explicit_inertial = true;
      double p[3] = {0, 0, 0}, iq[4];
      if (has(ia, "pos")) { auto v = nums(ia, "pos"); for (int i = 0; i < 3; i++) p[i] = v.at(i); }
      orientation(cx_, ia, iq);
      double mass = num(ia, "mass", 0);
      double I[3] = {0, 0, 0};
      if (has(ia, "diaginertia")) {
        auto v = nums(ia, "diaginertia"); for (int i = 0; i < 3; i++) I[i] = v.at(i);
        for (int i = 0; i < 3; i++) m_.body_ipos[3 * b + i] = p[i];
        for (int i = 0; i < 4; i++) m_.body_iquat[4 * b + i] = iq[i];
        for (int i = 0; i < 3; i++) m_.body_inertia[3 * b + i] = I[i];
      } else if (has(ia, "fullinertia")) {
        auto v = nums(ia, "fullinertia");  // xx yy zz xy xz yz
        double A[9] = {v.at(0), v.at(3), v.at(4), v.at(3), v.at(1), v.at(5), v.at(4), v.at(5), v.at(2)};
        double w[3], V[9]; eig3(A, w, V);
        double R[9]; quat2mat(iq, R);
        double RV[9];
        for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) RV[3 * r + c] = R[3 * r] * V[c] + R[3 * r + 1] * V[3 + c] + R[3 * r + 2] * V[6 + c];
        double qq[4]; mat2quat(RV, qq);
        for (int i = 0; i < 3; i++) m_.body_ipos[3 * b + i] = p[i];
        for (int i = 0; i < 4; i++) m_.body_iquat[4 * b + i] = qq[i];
        for (int i = 0; i < 3; i++) m_.body_inertia[3 * b + i] = w[i];
      } else {
        throw std::runtime_error("MJCF: <inertial> needs diaginertia or fullinertia");
      }
      m_.body_mass[b] = mass;
    } else if (k->tag == "site") {
      add_site(*k, b, cls);
    } else if (k->tag == "camera" || k->tag == "light") {
      // ignored: rendering only
    } else if (k->tag == "body" || k->tag == "joint" || k->tag == "freejoint") {
      // handled elsewhere
    } else {
      throw std::runtime_error("MJCF: unsupported element <" + k->tag + "> in body");
    }
  }
  for (auto& k : n.kids) if (k->tag == "body") body(*k, b, cls, robot);
}

void Compiler::finish_inertia() {
  for (int b = 1; b < m_.nbody; b++) {
    if (bb_[b].explicit_inertial) continue;
    auto& gs = bb_[b].gi;
    double M = 0, c[3] = {0, 0, 0};
    for (auto& g : gs) { M += g.mass; for (int i = 0; i < 3; i++) c[i] += g.mass * g.pos[i]; }
    if (M <= 0) continue;
    for (int i = 0; i < 3; i++) c[i] /= M;
    double A[9] = {0};
    for (auto& g : gs) {
      // R diag(I) R^T + m (|d|^2 E - d d^T)
      for (int r = 0; r < 3; r++) for (int cc = 0; cc < 3; cc++) {
        double s = 0;
        for (int k = 0; k < 3; k++) s += g.R[3 * r + k] * g.I[k] * g.R[3 * cc + k];
        A[3 * r + cc] += s;
      }
      double d[3] = {g.pos[0] - c[0], g.pos[1] - c[1], g.pos[2] - c[2]};
      double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      for (int r = 0; r < 3; r++) for (int cc = 0; cc < 3; cc++) A[3 * r + cc] += g.mass * ((r == cc ? dd : 0) - d[r] * d[cc]);
    }
    double w[3], V[9];
    eig3(A, w, V);
    double q[4]; mat2quat(V, q);
    for (int i = 0; i < 3; i++) { m_.body_ipos[3 * b + i] = c[i]; m_.body_inertia[3 * b + i] = w[i]; }
    for (int i = 0; i < 4; i++) m_.body_iquat[4 * b + i] = q[i];
    m_.body_mass[b] = M;
  }
  for (int b = 1; b < m_.nbody; b++) {
    if (m_.body_dofnum[b] > 0 || true) {
      // a moving body needs mass unless a descendant supplies it; MuJoCo errors
      // on massless leaves, which the dof check below catches through M > 0
    }
  }
}

void Compiler::actuators(const XNode& n) {
  for (auto& k : n.kids) {
    AttrMap a = resolve(cx_, *k, "main");
    if (!has(a, "joint")) throw std::runtime_error("MJCF: only joint transmissions are supported (actuator " + str(a, "name", "") + ")");
    int j = m_.name2id(3, a.at("joint"));
    if (j < 0) throw std::runtime_error("MJCF: actuator joint '" + a.at("joint") + "' not found");
    if (m_.jnt_type[j] == JNT_FREE || m_.jnt_type[j] == JNT_BALL) throw std::runtime_error("MJCF: actuators on free/ball joints are not supported");
    double gain[3] = {0, 0, 0}, bias[3] = {0, 0, 0};
    if (k->tag == "motor") { gain[0] = 1; }
    else if (k->tag == "position") { double kp = num(a, "kp", 1), kv = num(a, "kv", 0); gain[0] = kp; bias[1] = -kp; bias[2] = -kv; }
    else if (k->tag == "velocity") { double kv = num(a, "kv", 1); gain[0] = kv; bias[2] = -kv; }
    else if (k->tag == "general") {
      double dg[3] = {1, 0, 0}, db[3] = {0, 0, 0};
      vec(a, "gainprm", gain, 3, dg); vec(a, "biasprm", bias, 3, db);
      if (str(a, "gaintype", "fixed") != "fixed" || (str(a, "biastype", "none") != "none" && str(a, "biastype", "none") != "affine"))
        throw std::runtime_error("MJCF: unsupported general actuator gain/bias type");
      if (str(a, "biastype", "none") == "none") bias[0] = bias[1] = bias[2] = 0;
    } else {
      throw std::runtime_error("MJCF: unsupported actuator <" + k->tag + ">");
    }
    if (has(a, "dyntype") && a.at("dyntype") != "none") throw std::runtime_error("MJCF: actuator dynamics are not supported");
    m_.actuator_name.push_back(str(a, "name", ""));
    m_.actuator_trnid.push_back(j);
    double g6[6] = {1, 0, 0, 0, 0, 0};
    double gear[6]; vec(a, "gear", gear, 6, g6);
    m_.actuator_gear.push_back(gear[0]);
    for (int i = 0; i < 3; i++) { m_.actuator_gainprm.push_back(gain[i]); m_.actuator_biasprm.push_back(bias[i]); }
    double z2[2] = {0, 0}, cr[2], fr[2];
    vec(a, "ctrlrange", cr, 2, z2); vec(a, "forcerange", fr, 2, z2);
    std::string cl = str(a, "ctrllimited", "auto"), fl = str(a, "forcelimited", "auto");
    bool ctrllim = cl == "true" || (cl == "auto" && cx_.autolimits && has(a, "ctrlrange"));
    bool forcelim = fl == "true" || (fl == "auto" && cx_.autolimits && has(a, "forcerange"));
    if (ctrllim && cr[0] >= cr[1]) throw std::runtime_error("MJCF: invalid ctrlrange");
    m_.actuator_ctrllimited.push_back(ctrllim);
    m_.actuator_forcelimited.push_back(forcelim);
    m_.actuator_ctrlrange.push_back(cr[0]); m_.actuator_ctrlrange.push_back(cr[1]);
    m_.actuator_forcerange.push_back(fr[0]); m_.actuator_forcerange.push_back(fr[1]);
    m_.nu++;
  }
}

void Compiler::add_site(const XNode& n, int b, const std::string& cls) {
  AttrMap a = resolve(cx_, n, cls);
  double pos[3] = {0, 0, 0}, q[4];
  if (has(a, "pos")) { auto v = nums(a, "pos"); for (int i = 0; i < 3; i++) pos[i] = v.at(i); }
  orientation(cx_, a, q);
  if (has(a, "fromto")) throw std::runtime_error("MJCF: site fromto is not supported");
  m_.site_bodyid.push_back(b);
  m_.site_name.push_back(str(a, "name", ""));
  for (int i = 0; i < 3; i++) m_.site_pos.push_back(pos[i]);
  for (int i = 0; i < 4; i++) m_.site_quat.push_back(q[i]);
  m_.nsite++;
}

// <sensor>: the kinds the reference computes on its batched path (reference
// src/sensor.cpp:126-183), with MuJoCo's mjtSensor ids and sensordata widths
void Compiler::sensors(const XNode& n) {
  struct Kind { const char* tag; int type, dim; const char* attr; int objtype; };
  static const Kind kinds[] = {
      {"accelerometer", 1, 3, "site", 6}, {"velocimeter", 2, 3, "site", 6}, {"gyro", 3, 3, "site", 6},
      {"magnetometer", 6, 3, "site", 6}, {"jointpos", 9, 1, "joint", 3}, {"jointvel", 10, 1, "joint", 3},
      {"actuatorpos", 13, 1, "actuator", 19}, {"actuatorvel", 14, 1, "actuator", 19},
      {"jointactuatorfrc", 16, 1, "joint", 3}, {"ballquat", 18, 4, "joint", 3}, {"ballangvel", 19, 3, "joint", 3},
      {"framepos", 26, 3, nullptr, 0}, {"framequat", 27, 4, nullptr, 0}, {"framexaxis", 28, 3, nullptr, 0},
      {"frameyaxis", 29, 3, nullptr, 0}, {"framezaxis", 30, 3, nullptr, 0}, {"framelinvel", 31, 3, nullptr, 0},
      {"frameangvel", 32, 3, nullptr, 0}, {"subtreecom", 35, 3, "body", 1},
  };
  for (auto& k : n.kids) {
    AttrMap a = resolve(cx_, *k, "main");
    const Kind* kd = nullptr;
    for (auto& c : kinds) if (k->tag == c.tag) kd = &c;
    if (!kd) throw std::runtime_error("MJCF: unsupported sensor <" + k->tag + ">");
    int objtype = kd->objtype, objid = -1;
    if (kd->attr) {
      objid = m_.name2id(objtype, str(a, kd->attr, ""));
    } else {  // frame sensors name their object by objtype / objname
      std::string ot = str(a, "objtype", "");
      objtype = ot == "body" ? 1 : ot == "xbody" ? 2 : ot == "geom" ? 5 : ot == "site" ? 6 : -1;
      if (objtype < 0) throw std::runtime_error("MJCF: frame sensor objtype must be body, xbody, geom or site");
      if (has(a, "reftype") || has(a, "refname")) throw std::runtime_error("MJCF: frame sensors relative to a reference frame are not supported");
      objid = m_.name2id(objtype, str(a, "objname", ""));
    }
    if (objid < 0) throw std::runtime_error("MJCF: sensor <" + k->tag + "> names an unknown object");
    if (kd->objtype == 3) {
      int jt = m_.jnt_type[objid];
      bool ball = kd->type == 18 || kd->type == 19;
      if (ball ? jt != JNT_BALL : (jt != JNT_HINGE && jt != JNT_SLIDE))
        throw std::runtime_error("MJCF: sensor <" + k->tag + "> on the wrong joint type");
    }
    m_.sensor_type.push_back(kd->type);
    m_.sensor_objtype.push_back(objtype);
    m_.sensor_objid.push_back(objid);
    m_.sensor_adr.push_back(m_.nsensordata);
    m_.sensor_dim.push_back(kd->dim);
    m_.sensor_name.push_back(str(a, "name", ""));
    m_.nsensordata += kd->dim;
    m_.nsensor++;
  }
}

void Compiler::contact(const XNode& n) {
  for (auto& k : n.kids) {
    AttrMap a = resolve(cx_, *k, "main");
    if (k->tag == "exclude") {
      int b1 = m_.name2id(1, str(a, "body1", "")), b2 = m_.name2id(1, str(a, "body2", ""));
      if (b1 < 0 || b2 < 0) throw std::runtime_error("MJCF: exclude body not found");
      m_.exclude_signature.push_back((std::min(b1, b2) << 16) | std::max(b1, b2));
    } else if (k->tag == "pair") {
      int g1 = m_.name2id(5, str(a, "geom1", "")), g2 = m_.name2id(5, str(a, "geom2", ""));
      if (g1 < 0 || g2 < 0) throw std::runtime_error("MJCF: pair geom not found");
      m_.pair_geom1.push_back(g1); m_.pair_geom2.push_back(g2);
      // unspecified pair parameters follow the mixed geom parameters
      double f5[5] = {std::max(m_.geom_friction[3 * g1], m_.geom_friction[3 * g2]),
                      std::max(m_.geom_friction[3 * g1], m_.geom_friction[3 * g2]),
                      std::max(m_.geom_friction[3 * g1 + 1], m_.geom_friction[3 * g2 + 1]),
                      std::max(m_.geom_friction[3 * g1 + 2], m_.geom_friction[3 * g2 + 2]),
                      std::max(m_.geom_friction[3 * g1 + 2], m_.geom_friction[3 * g2 + 2])};
      double f[5], sr[2], si[5];
      double dsr[2] = {0.02, 1}, dsi[5] = {0.9, 0.95, 0.001, 0.5, 2};
      vec(a, "friction", f, 5, f5); vec(a, "solref", sr, 2, dsr); vec(a, "solimp", si, 5, dsi);
      for (int i = 0; i < 5; i++) { m_.pair_friction.push_back(f[i]); m_.pair_solimp.push_back(si[i]); }
      for (int i = 0; i < 2; i++) m_.pair_solref.push_back(sr[i]);
      m_.pair_dim.push_back((int)num(a, "condim", std::max(m_.geom_condim[g1], m_.geom_condim[g2])));
      m_.pair_margin.push_back(num(a, "margin", std::max(m_.geom_margin[g1], m_.geom_margin[g2])));
      m_.pair_gap.push_back(num(a, "gap", std::max(m_.geom_gap[g1], m_.geom_gap[g2])));
    } else {
      throw std::runtime_error("MJCF: unsupported <contact> child <" + k->tag + ">");
    }
  }
}

Model Compiler::run() {
  XmlParser px(xml_);
  auto root = px.parse();
  if (root->tag != "mujoco") throw std::runtime_error("MJCF: root element must be <mujoco>");
  m_.name = root->get("model") ? *root->get("model") : "";
  cx_.classes["main"];
  for (auto& k : root->kids) {
    if (k->tag == "compiler") {
      if (auto v = k->get("angle")) cx_.radian = (*v == "radian");
      if (auto v = k->get("autolimits")) cx_.autolimits = (*v == "true");
      if (auto v = k->get("eulerseq")) cx_.eulerseq = *v;
      if (auto v = k->get("inertiafromgeom")) if (*v == "false") throw std::runtime_error("MJCF: inertiafromgeom=false is not supported");
    } else if (k->tag == "default") {
      read_defaults(cx_, *k, "main", "");
    }
  }
  for (auto& k : root->kids) if (k->tag == "option") option(*k);
  // world body
  m_.nbody = 1;
  m_.body_parentid.push_back(0); m_.body_name.push_back("world");
  m_.body_pos.insert(m_.body_pos.end(), {0, 0, 0}); m_.body_quat.insert(m_.body_quat.end(), {1, 0, 0, 0});
  m_.body_ipos.insert(m_.body_ipos.end(), {0, 0, 0}); m_.body_iquat.insert(m_.body_iquat.end(), {1, 0, 0, 0});
  m_.body_mass.push_back(0); m_.body_inertia.insert(m_.body_inertia.end(), {0, 0, 0});
  m_.body_jntnum.push_back(0); m_.body_jntadr.push_back(-1); m_.body_dofnum.push_back(0); m_.body_dofadr.push_back(-1);
  m_.body_geomnum.push_back(0); m_.body_geomadr.push_back(-1);
  bb_.emplace_back(); body_robot_.push_back(false);
  const XNode* wb = nullptr;
  for (auto& k : root->kids) if (k->tag == "worldbody") wb = k.get();
  if (!wb) throw std::runtime_error("MJCF: missing <worldbody>");
  // world geoms first (MuJoCo numbers world geoms before child bodies' geoms)
  for (auto& k : wb->kids) if (k->tag == "geom") add_geom(*k, 0, "main", false);
  for (auto& k : wb->kids) if (k->tag == "site") add_site(*k, 0, "main");
  for (auto& k : wb->kids) {
    if (k->tag == "body") body(*k, 0, "main", false);
    else if (k->tag == "geom" || k->tag == "site" || k->tag == "light" || k->tag == "camera") continue;
    else throw std::runtime_error("MJCF: unsupported element <" + k->tag + "> in worldbody");
  }
  finish_inertia();
  for (auto& k : root->kids) {
    if (k->tag == "actuator") actuators(*k);
    else if (k->tag == "contact") contact(*k);
    else if (k->tag == "sensor") continue;  // after the actuators, whose names sensors use
    else if (k->tag == "compiler" || k->tag == "option" || k->tag == "default" || k->tag == "worldbody" ||
             k->tag == "size" || k->tag == "visual" || k->tag == "statistic" || k->tag == "asset") continue;
    else throw std::runtime_error("MJCF: unsupported top-level element <" + k->tag + ">");
  }
  for (auto& k : root->kids) if (k->tag == "sensor") sensors(*k);
  // rootid / weldid
  m_.body_rootid.assign(m_.nbody, 0);
  m_.body_weldid.assign(m_.nbody, 0);
  for (int b = 1; b < m_.nbody; b++) {
    int p = m_.body_parentid[b];
    m_.body_rootid[b] = p == 0 ? b : m_.body_rootid[p];
    m_.body_weldid[b] = m_.body_jntnum[b] > 0 ? b : m_.body_weldid[p];
  }
  for (int b = 0; b < m_.nbody; b++) if (m_.body_dofadr[b] < 0) m_.body_dofadr[b] = 0;
  for (int b = 0; b < m_.nbody; b++) if (m_.body_jntadr[b] < 0) m_.body_jntadr[b] = 0;
  // subtree mass
  m_.body_subtreemass = m_.body_mass;
  for (int b = m_.nbody - 1; b > 0; b--) m_.body_subtreemass[m_.body_parentid[b]] += m_.body_subtreemass[b];
  for (int b = 1; b < m_.nbody; b++)
    if (m_.body_dofnum[b] > 0 && m_.body_subtreemass[b] < 1e-15)
      throw std::runtime_error("MJCF: moving body '" + m_.body_name[b] + "' has no mass in its subtree");
  if (m_.nv == 0 && m_.nbody > 1) {}  // static scenes are allowed
  set_const(m_);
  return m_;
}

int Model::name2id(int objtype, const std::string& nm) const {
  const std::vector<std::string>* v = nullptr;
  switch (objtype) {
    case 1: case 2: v = &body_name; break;
    case 3: v = &jnt_name; break;
    case 5: v = &geom_name; break;
    case 6: v = &site_name; break;
    case 19: v = &actuator_name; break;
    case 20: v = &sensor_name; break;
    default: return -1;
  }
  if (nm.empty()) return -1;
  for (size_t i = 0; i < v->size(); i++) if ((*v)[i] == nm) return (int)i;
  return -1;
}

const char* Model::id2name(int objtype, int id) const {
  const std::vector<std::string>* v = nullptr;
  switch (objtype) {
    case 1: case 2: v = &body_name; break;
    case 3: v = &jnt_name; break;
    case 5: v = &geom_name; break;
    case 6: v = &site_name; break;
    case 19: v = &actuator_name; break;
    case 20: v = &sensor_name; break;
    default: return nullptr;
  }
  if (id < 0 || id >= (int)v->size() || (*v)[id].empty()) return nullptr;
  return (*v)[id].c_str();
}

Model compile_mjcf(const std::string& xml, const LoadOptions& opt) {
  Compiler c(xml, opt);
  return c.run();
}

// ---------------------------------------------------------------- pair table
std::vector<CollisionPair> build_pairs(const Model& m) {
  std::vector<CollisionPair> out;
  auto excluded = [&](int b1, int b2) {
    int sig = (std::min(b1, b2) << 16) | std::max(b1, b2);
    for (int s : m.exclude_signature) if (s == sig) return true;
    return false;
  };
  if (m.disableflags & (DSBL_CONTACT | DSBL_CONSTRAINT)) return out;
  for (int i = 0; i < m.ngeom; i++) {
    for (int j = i + 1; j < m.ngeom; j++) {
      int g1 = i, g2 = j;
      if (m.geom_type[g1] > m.geom_type[g2]) std::swap(g1, g2);
      int mask = (m.geom_contype[g1] & m.geom_conaffinity[g2]) | (m.geom_contype[g2] & m.geom_conaffinity[g1]);
      if (!mask) continue;
      int b1 = m.geom_bodyid[g1], b2 = m.geom_bodyid[g2];
      int w1 = m.body_weldid[b1], w2 = m.body_weldid[b2];
      if (w1 == w2) continue;
      if (!(m.disableflags & DSBL_FILTERPARENT)) {
        int w1p = m.body_weldid[m.body_parentid[w1]], w2p = m.body_weldid[m.body_parentid[w2]];
        if (w1 != 0 && w2 != 0 && (w1 == w2p || w2 == w1p)) continue;
      }
      if (excluded(b1, b2)) continue;
      // planes only collide with non-plane geoms
      if (m.geom_type[g1] == GEOM_PLANE && m.geom_type[g2] == GEOM_PLANE) continue;
      CollisionPair p{};
      p.g1 = g1; p.g2 = g2; p.b1 = b1; p.b2 = b2;
      p.margin = std::max(m.geom_margin[g1], m.geom_margin[g2]);
      p.gap = std::max(m.geom_gap[g1], m.geom_gap[g2]);
      int pr1 = m.geom_priority[g1], pr2 = m.geom_priority[g2];
      double fr[3];
      if (pr1 != pr2) {
        int gp = pr1 > pr2 ? g1 : g2;
        p.condim = m.geom_condim[gp];
        for (int k = 0; k < 3; k++) fr[k] = m.geom_friction[3 * gp + k];
        for (int k = 0; k < 2; k++) p.solref[k] = m.geom_solref[2 * gp + k];
        for (int k = 0; k < 5; k++) p.solimp[k] = m.geom_solimp[5 * gp + k];
      } else {
        p.condim = std::max(m.geom_condim[g1], m.geom_condim[g2]);
        double s1 = m.geom_solmix[g1], s2 = m.geom_solmix[g2], mix;
        const double mn = 1e-15;
        if (s1 >= mn && s2 >= mn) mix = s1 / (s1 + s2);
        else if (s1 < mn && s2 < mn) mix = 0.5;
        else mix = s1 < mn ? 0.0 : 1.0;
        const double* r1 = &m.geom_solref[2 * g1]; const double* r2 = &m.geom_solref[2 * g2];
        if (r1[0] > 0 && r2[0] > 0) for (int k = 0; k < 2; k++) p.solref[k] = mix * r1[k] + (1 - mix) * r2[k];
        else for (int k = 0; k < 2; k++) p.solref[k] = std::min(r1[k], r2[k]);
        for (int k = 0; k < 5; k++) p.solimp[k] = mix * m.geom_solimp[5 * g1 + k] + (1 - mix) * m.geom_solimp[5 * g2 + k];
        for (int k = 0; k < 3; k++) fr[k] = std::max(m.geom_friction[3 * g1 + k], m.geom_friction[3 * g2 + k]);
      }
      p.friction[0] = fr[0]; p.friction[1] = fr[0]; p.friction[2] = fr[1]; p.friction[3] = fr[2]; p.friction[4] = fr[2];
      out.push_back(p);
    }
  }
  // explicit pairs (kept even when contype/conaffinity would filter them)
  for (size_t k = 0; k < m.pair_geom1.size(); k++) {
    int g1 = m.pair_geom1[k], g2 = m.pair_geom2[k];
    if (m.geom_type[g1] > m.geom_type[g2]) std::swap(g1, g2);
    bool dup = false;
    for (auto& e : out) if (e.g1 == g1 && e.g2 == g2) { dup = true; break; }
    CollisionPair p{};
    p.g1 = g1; p.g2 = g2; p.b1 = m.geom_bodyid[g1]; p.b2 = m.geom_bodyid[g2];
    p.condim = m.pair_dim[k]; p.margin = m.pair_margin[k]; p.gap = m.pair_gap[k];
    for (int i = 0; i < 5; i++) { p.friction[i] = m.pair_friction[5 * k + i]; p.solimp[i] = m.pair_solimp[5 * k + i]; }
    for (int i = 0; i < 2; i++) p.solref[i] = m.pair_solref[2 * k + i];
    if (dup) { for (auto& e : out) if (e.g1 == g1 && e.g2 == g2) e = p; }
    else out.push_back(p);
  }
  return out;
}

// ---------------------------------------------------------------- mj_setConst subset
// Kinematics + composite rigid body at qpos0 in FP64, then M^-1 by dense
// Cholesky; body invweight0 = mean diagonal of J M^-1 J^T (translation and
// rotation blocks of the body-COM Jacobian), dof invweight0 from diag(M^-1),
// meaninertia = trace(M)/nv.
void set_const(Model& m) {
  int nb = m.nbody, nv = m.nv;
  m.body_invweight0.assign(2 * nb, 0.0);
  m.dof_invweight0.assign(nv, 0.0);
  if (nv == 0) { m.meaninertia = 1.0; return; }
  std::vector<double> xpos(3 * nb, 0), xquat(4 * nb, 0), xmat(9 * nb, 0), xipos(3 * nb, 0), ximat(9 * nb, 0);
  std::vector<double> xanchor(3 * m.njnt), xaxis(3 * m.njnt);
  xquat[0] = 1; xmat[0] = xmat[4] = xmat[8] = 1;
  const double* qpos = m.qpos0.data();
  for (int b = 1; b < nb; b++) {
    int p = m.body_parentid[b];
    double pos[3], q[4];
    int ja = m.body_jntadr[b], jn = m.body_jntnum[b];
    if (jn > 0 && m.jnt_type[ja] == JNT_FREE) {
      int a = m.jnt_qposadr[ja];
      for (int i = 0; i < 3; i++) pos[i] = qpos[a + i];
      for (int i = 0; i < 4; i++) q[i] = qpos[a + 3 + i];
      normquat(q);
      for (int i = 0; i < 3; i++) { xanchor[3 * ja + i] = pos[i]; xaxis[3 * ja + i] = m.jnt_axis[3 * ja + i]; }
    } else {
      double t[3]; rotvec(&xquat[4 * p], &m.body_pos[3 * b], t);
      for (int i = 0; i < 3; i++) pos[i] = xpos[3 * p + i] + t[i];
      mulquat(&xquat[4 * p], &m.body_quat[4 * b], q);
      for (int j = ja; j < ja + jn; j++) {
        int a = m.jnt_qposadr[j];
        double anc[3], ax[3];
        rotvec(q, &m.jnt_pos[3 * j], anc);
        for (int i = 0; i < 3; i++) anc[i] += pos[i];
        rotvec(q, &m.jnt_axis[3 * j], ax);
        for (int i = 0; i < 3; i++) { xanchor[3 * j + i] = anc[i]; xaxis[3 * j + i] = ax[i]; }
        if (m.jnt_type[j] == JNT_SLIDE) {
          for (int i = 0; i < 3; i++) pos[i] += ax[i] * (qpos[a] - m.qpos0[a]);
        } else if (m.jnt_type[j] == JNT_HINGE || m.jnt_type[j] == JNT_BALL) {
          double ql[4];
          if (m.jnt_type[j] == JNT_HINGE) axisangle2quat(&m.jnt_axis[3 * j], qpos[a] - m.qpos0[a], ql);
          else { for (int i = 0; i < 4; i++) ql[i] = qpos[a + i]; normquat(ql); }
          mulquat(q, ql, q);
          double r[3]; rotvec(q, &m.jnt_pos[3 * j], r);
          for (int i = 0; i < 3; i++) pos[i] = anc[i] - r[i];
        }
      }
      normquat(q);
    }
    for (int i = 0; i < 3; i++) xpos[3 * b + i] = pos[i];
    for (int i = 0; i < 4; i++) xquat[4 * b + i] = q[i];
    quat2mat(q, &xmat[9 * b]);
    double t[3]; rotvec(q, &m.body_ipos[3 * b], t);
    for (int i = 0; i < 3; i++) xipos[3 * b + i] = pos[i] + t[i];
    double qi[4]; mulquat(q, &m.body_iquat[4 * b], qi);
    quat2mat(qi, &ximat[9 * b]);
  }
  // subtree com
  std::vector<double> scom(3 * nb, 0), smass(nb, 0);
  for (int b = nb - 1; b >= 0; b--) {
    smass[b] += m.body_mass[b];
    for (int i = 0; i < 3; i++) scom[3 * b + i] += m.body_mass[b] * xipos[3 * b + i];
    if (b > 0) { int p = m.body_parentid[b]; smass[p] += smass[b]; for (int i = 0; i < 3; i++) scom[3 * p + i] += scom[3 * b + i]; }
  }
  for (int b = 0; b < nb; b++) for (int i = 0; i < 3; i++) scom[3 * b + i] = smass[b] > 1e-15 ? scom[3 * b + i] / smass[b] : xipos[3 * b + i];
  // cdof (motion subspaces about the root subtree com, [ang; lin])
  std::vector<double> cdof(6 * nv, 0);
  for (int j = 0; j < m.njnt; j++) {
    int b = m.jnt_bodyid[j], d = m.jnt_dofadr[j];
    const double* c = &scom[3 * m.body_rootid[b]];
    const double* anc = &xanchor[3 * j];
    double off[3] = {c[0] - anc[0], c[1] - anc[1], c[2] - anc[2]};
    auto dofcom = [&](double* out, const double* ax) {
      out[0] = ax[0]; out[1] = ax[1]; out[2] = ax[2];
      out[3] = ax[1] * off[2] - ax[2] * off[1];
      out[4] = ax[2] * off[0] - ax[0] * off[2];
      out[5] = ax[0] * off[1] - ax[1] * off[0];
    };
    switch (m.jnt_type[j]) {
      case JNT_FREE:
        for (int i = 0; i < 3; i++) cdof[6 * (d + i) + 3 + i] = 1;
        for (int i = 0; i < 3; i++) { double ax[3] = {xmat[9 * b + i], xmat[9 * b + 3 + i], xmat[9 * b + 6 + i]}; dofcom(&cdof[6 * (d + 3 + i)], ax); }
        break;
      case JNT_BALL:
        for (int i = 0; i < 3; i++) { double ax[3] = {xmat[9 * b + i], xmat[9 * b + 3 + i], xmat[9 * b + 6 + i]}; dofcom(&cdof[6 * (d + i)], ax); }
        break;
      case JNT_SLIDE:
        for (int i = 0; i < 3; i++) cdof[6 * d + 3 + i] = xaxis[3 * j + i];
        break;
      default:
        dofcom(&cdof[6 * d], &xaxis[3 * j]);
    }
  }
  // body inertias about the root subtree com in world frame: [Ixx Iyy Izz Ixy Ixz Iyz, m*d, m]
  std::vector<double> crb(10 * nb, 0);
  for (int b = 1; b < nb; b++) {
    const double* R = &ximat[9 * b]; const double* I = &m.body_inertia[3 * b];
    double mass = m.body_mass[b];
    double d[3]; for (int i = 0; i < 3; i++) d[i] = xipos[3 * b + i] - scom[3 * m.body_rootid[b] + i];
    double A[9];
    for (int r = 0; r < 3; r++) for (int c = 0; c < 3; c++) A[3 * r + c] = R[3 * r] * I[0] * R[3 * c] + R[3 * r + 1] * I[1] * R[3 * c + 1] + R[3 * r + 2] * I[2] * R[3 * c + 2];
    double dd = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double* o = &crb[10 * b];
    o[0] = A[0] + mass * (dd - d[0] * d[0]); o[1] = A[4] + mass * (dd - d[1] * d[1]); o[2] = A[8] + mass * (dd - d[2] * d[2]);
    o[3] = A[1] - mass * d[0] * d[1]; o[4] = A[2] - mass * d[0] * d[2]; o[5] = A[5] - mass * d[1] * d[2];
    o[6] = mass * d[0]; o[7] = mass * d[1]; o[8] = mass * d[2]; o[9] = mass;
  }
  for (int b = nb - 1; b > 0; b--) { int p = m.body_parentid[b]; if (p > 0) for (int k = 0; k < 10; k++) crb[10 * p + k] += crb[10 * b + k]; }
  auto mulinert = [](const double* i, const double* v, double* r) {
    r[0] = i[0] * v[0] + i[3] * v[1] + i[4] * v[2] - i[8] * v[4] + i[7] * v[5];
    r[1] = i[3] * v[0] + i[1] * v[1] + i[5] * v[2] + i[8] * v[3] - i[6] * v[5];
    r[2] = i[4] * v[0] + i[5] * v[1] + i[2] * v[2] - i[7] * v[3] + i[6] * v[4];
    r[3] = i[8] * v[1] - i[7] * v[2] + i[9] * v[3];
    r[4] = i[6] * v[2] - i[8] * v[0] + i[9] * v[4];
    r[5] = i[7] * v[0] - i[6] * v[1] + i[9] * v[5];
  };
  std::vector<double> M(nv * nv, 0);
  for (int i = 0; i < nv; i++) {
    double buf[6];
    mulinert(&crb[10 * m.dof_bodyid[i]], &cdof[6 * i], buf);
    for (int j = i; j >= 0; j = m.dof_parentid[j]) {
      double s = 0; for (int k = 0; k < 6; k++) s += cdof[6 * j + k] * buf[k];
      M[i * nv + j] = M[j * nv + i] = s;
    }
    M[i * nv + i] += m.dof_armature[i];
  }
  double tr = 0; for (int i = 0; i < nv; i++) tr += M[i * nv + i];
  m.meaninertia = tr / nv;
  if (m.meaninertia < 1e-15) m.meaninertia = 1.0;
  // dense Cholesky M = L L^T and the inverse
  std::vector<double> L(M);
  for (int k = 0; k < nv; k++) {
    double s = L[k * nv + k];
    for (int p = 0; p < k; p++) s -= L[k * nv + p] * L[k * nv + p];
    if (s <= 0) throw std::runtime_error("MJCF: mass matrix is not positive definite at qpos0");
    L[k * nv + k] = std::sqrt(s);
    for (int i = k + 1; i < nv; i++) {
      double t = L[i * nv + k];
      for (int p = 0; p < k; p++) t -= L[i * nv + p] * L[k * nv + p];
      L[i * nv + k] = t / L[k * nv + k];
    }
    for (int j = k + 1; j < nv; j++) L[k * nv + j] = 0;
  }
  auto solve = [&](std::vector<double>& x) {
    for (int i = 0; i < nv; i++) { double s = x[i]; for (int p = 0; p < i; p++) s -= L[i * nv + p] * x[p]; x[i] = s / L[i * nv + i]; }
    for (int i = nv - 1; i >= 0; i--) { double s = x[i]; for (int p = i + 1; p < nv; p++) s -= L[p * nv + i] * x[p]; x[i] = s / L[i * nv + i]; }
  };
  std::vector<double> Minv(nv * nv);
  for (int c = 0; c < nv; c++) {
    std::vector<double> e(nv, 0); e[c] = 1; solve(e);
    for (int r = 0; r < nv; r++) Minv[r * nv + c] = e[r];
  }
  // body invweight0
  for (int b = 1; b < nb; b++) {
    if (m.body_weldid[b] == 0) continue;
    // Jacobian of the body com point: rows 0..2 translation, 3..5 rotation
    std::vector<double> J(6 * nv, 0);
    const double* pt = &xipos[3 * b];
    const double* c = &scom[3 * m.body_rootid[b]];
    double off[3] = {pt[0] - c[0], pt[1] - c[1], pt[2] - c[2]};
    // dofs that move body b: the dof chain ending at b's last dof (or its nearest moving ancestor)
    int bb = b;
    while (bb > 0 && m.body_dofnum[bb] == 0) bb = m.body_parentid[bb];
    if (bb == 0) continue;
    for (int d = m.body_dofadr[bb] + m.body_dofnum[bb] - 1; d >= 0; d = m.dof_parentid[d]) {
      const double* cd = &cdof[6 * d];
      J[0 * nv + d] = cd[3] + cd[1] * off[2] - cd[2] * off[1];
      J[1 * nv + d] = cd[4] + cd[2] * off[0] - cd[0] * off[2];
      J[2 * nv + d] = cd[5] + cd[0] * off[1] - cd[1] * off[0];
      J[3 * nv + d] = cd[0]; J[4 * nv + d] = cd[1]; J[5 * nv + d] = cd[2];
    }
    double diag[6];
    for (int r = 0; r < 6; r++) {
      double s = 0;
      for (int i = 0; i < nv; i++) { if (J[r * nv + i] == 0) continue; for (int k = 0; k < nv; k++) s += J[r * nv + i] * Minv[i * nv + k] * J[r * nv + k]; }
      diag[r] = s;
    }
    m.body_invweight0[2 * b] = (diag[0] + diag[1] + diag[2]) / 3.0;
    m.body_invweight0[2 * b + 1] = (diag[3] + diag[4] + diag[5]) / 3.0;
  }
  // dof invweight0
  for (int j = 0; j < m.njnt; j++) {
    int d = m.jnt_dofadr[j];
    switch (m.jnt_type[j]) {
      case JNT_FREE: {
        double t = (Minv[d * nv + d] + Minv[(d + 1) * nv + d + 1] + Minv[(d + 2) * nv + d + 2]) / 3.0;
        double r = (Minv[(d + 3) * nv + d + 3] + Minv[(d + 4) * nv + d + 4] + Minv[(d + 5) * nv + d + 5]) / 3.0;
        for (int i = 0; i < 3; i++) { m.dof_invweight0[d + i] = t; m.dof_invweight0[d + 3 + i] = r; }
        break;
      }
      case JNT_BALL: {
        double r = (Minv[d * nv + d] + Minv[(d + 1) * nv + d + 1] + Minv[(d + 2) * nv + d + 2]) / 3.0;
        for (int i = 0; i < 3; i++) m.dof_invweight0[d + i] = r;
        break;
      }
      default: m.dof_invweight0[d] = Minv[d * nv + d];
    }
  }
}

}  // namespace mjh
