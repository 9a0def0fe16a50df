// MJCF loader for the C ABI (pippin_hip_model_load_*): parses the subset of
// crates/pippin/src/mjcf.rs this backend supports (defaults/classes, compiler
// angle/eulerseq/autolimits, options, bodies with free/ball/slide/hinge
// joints, inertial or geom-derived mass for primitive geoms, cameras,
// motor/position/velocity/general actuators with joint or fixed-tendon
// transmission, joint equality, <contact><exclude>) into the raw descriptor.
// Arithmetic follows mjcf.rs operation for operation.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/pippin_hip.h"
#include "host_model.hpp"
#include "meshes.hpp"
#include "xml.hpp"

namespace {

using phx::XmlError;
using phx::XmlParser;
using phx::XNode;

// ---------------------------------------------------------------- math (math.rs order)
struct V {
  double x, y, z;
};
struct Q {
  double w, x, y, z;
};
typedef std::vector<double> Mat;  // 9, row-major

V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V vscale(V a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double vdot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V vcross(V a, V o) { return {a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x}; }
double vnorm(V a) { return std::sqrt(vdot(a, a)); }
V vnormalized(V a) {
  double n = vnorm(a);
  return n > 1e-15 ? vscale(a, 1.0 / n) : V{0, 0, 0};
}
Mat mmul(const Mat& a, const Mat& o) {
  Mat r(9);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = -0.0;
      for (int k = 0; k < 3; k++) s += a[3 * i + k] * o[3 * k + j];
      r[3 * i + j] = s;
    }
  return r;
}
Mat mT(const Mat& m) { return {m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]}; }
Mat mskew(V v) { return {0.0, -v.z, v.y, v.z, 0.0, -v.x, -v.y, v.x, 0.0}; }
Mat mscale(const Mat& m, double s) {
  Mat r(9);
  for (int i = 0; i < 9; i++) r[i] = m[i] * s;
  return r;
}
Mat madd(const Mat& a, const Mat& b) {
  Mat r(9);
  for (int i = 0; i < 9; i++) r[i] = a[i] + b[i];
  return r;
}
Mat mdiag(V d) { return {d.x, 0, 0, 0, d.y, 0, 0, 0, d.z}; }
Mat mrotate(const Mat& m, const Mat& r) { return mmul(mmul(r, m), mT(r)); }
V mulv(const Mat& r, V v) {
  return {r[0] * v.x + r[1] * v.y + r[2] * v.z, r[3] * v.x + r[4] * v.y + r[5] * v.z, r[6] * v.x + r[7] * v.y + r[8] * v.z};
}
Q qnormalized(Q q) {
  double s = -0.0;
  s += q.w * q.w;
  s += q.x * q.x;
  s += q.y * q.y;
  s += q.z * q.z;
  double n = std::sqrt(s);
  if (n < 1e-15) return {1, 0, 0, 0};
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}
Q qmul(Q p, Q o) {
  double a = p.w, b = p.x, c = p.y, d = p.z, e = o.w, f = o.x, g = o.y, h = o.z;
  return {a * e - b * f - c * g - d * h, a * f + b * e + c * h - d * g, a * g - b * h + c * e + d * f,
          a * h + b * g - c * f + d * e};
}
Q from_axis_angle(V axis, double angle) {
  V a = vnormalized(axis);
  double s = std::sin(0.5 * angle), c = std::cos(0.5 * angle);
  return {c, a.x * s, a.y * s, a.z * s};
}
Mat to_mat(Q q) {
  double w = q.w, x = q.x, y = q.y, z = q.z;
  double ww = w * w, xx = x * x, yy = y * y, zz = z * z;
  double xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
  return {ww + xx - yy - zz, 2.0 * (xy - wz), 2.0 * (xz + wy), 2.0 * (xy + wz), ww - xx + yy - zz,
          2.0 * (yz - wx),   2.0 * (xz - wy), 2.0 * (yz + wx), ww - xx - yy + zz};
}
Q from_mat(const Mat& m) {
  double tr = m[0] + m[4] + m[8];
  Q q;
  if (tr > 0.0) {
    double s = std::sqrt(tr + 1.0) * 2.0;
    q = {0.25 * s, (m[7] - m[5]) / s, (m[2] - m[6]) / s, (m[3] - m[1]) / s};
  } else if (m[0] > m[4] && m[0] > m[8]) {
    double s = std::sqrt(1.0 + m[0] - m[4] - m[8]) * 2.0;
    q = {(m[7] - m[5]) / s, 0.25 * s, (m[1] + m[3]) / s, (m[2] + m[6]) / s};
  } else if (m[4] > m[8]) {
    double s = std::sqrt(1.0 + m[4] - m[0] - m[8]) * 2.0;
    q = {(m[2] - m[6]) / s, (m[1] + m[3]) / s, 0.25 * s, (m[5] + m[7]) / s};
  } else {
    double s = std::sqrt(1.0 + m[8] - m[0] - m[4]) * 2.0;
    q = {(m[3] - m[1]) / s, (m[2] + m[6]) / s, (m[5] + m[7]) / s, 0.25 * s};
  }
  q = qnormalized(q);
  if (q.w < 0.0) q = {-q.w, -q.x, -q.y, -q.z};
  return q;
}
Q from_z_to(V dir) {
  V d = vnormalized(dir);
  double c = vdot(V{0, 0, 1}, d);
  if (c < -1.0 + 1e-12) return {0, 1, 0, 0};
  V ax = vcross(V{0, 0, 1}, d);
  return qnormalized({1.0 + c, ax.x, ax.y, ax.z});
}

struct LoadError {
  std::string msg;
};

bool parse_floats(const char* s, std::vector<double>& out) {
  out.clear();
  const char* p = s;
  for (;;) {
    while (*p && isspace((unsigned char)*p)) p++;
    if (!*p) return true;
    char* end;
    double v = strtod(p, &end);
    if (end == p || (*end && !isspace((unsigned char)*end))) throw LoadError{std::string("bad number in '") + s + "'"};
    out.push_back(v);
    p = end;
  }
}

typedef std::map<std::string, std::map<std::string, std::string>> DefaultClass;

struct Ctx {
  std::map<std::string, std::vector<float>> materials;
  std::vector<std::string> mesh_names;  // index = geom_dataid
  std::vector<phh::Hull> hulls;
  bool degrees = true;
  std::string eulerseq = "xyz";
  bool autolimits = true;
  std::map<std::string, DefaultClass> defaults;

  // mjcf.rs:45-62
  const char* get(const XNode& n, const std::string& cls, const char* name) const {
    if (const char* v = n.get(name)) return v;
    const char* c = n.get("class");
    std::string klass = c ? c : cls;
    auto it = defaults.find(klass);
    if (it == defaults.end()) return nullptr;
    auto t = it->second.find(n.tag);
    if (t != it->second.end()) {
      auto a = t->second.find(name);
      if (a != t->second.end()) return a->second.c_str();
    }
    if (n.tag == "motor" || n.tag == "position" || n.tag == "velocity") {
      auto g = it->second.find("general");
      if (g != it->second.end()) {
        auto a = g->second.find(name);
        if (a != g->second.end()) return a->second.c_str();
      }
    }
    return nullptr;
  }
  bool floats(const XNode& n, const std::string& cls, const char* name, std::vector<double>& out) const {
    const char* s = get(n, cls, name);
    if (!s) return false;
    parse_floats(s, out);
    return true;
  }
  double flt(const XNode& n, const std::string& cls, const char* name, double dflt) const {
    std::vector<double> v;
    if (floats(n, cls, name, v) && !v.empty()) return v[0];
    return dflt;
  }
  V vec3(const XNode& n, const std::string& cls, const char* name, V dflt) const {
    std::vector<double> v;
    if (!floats(n, cls, name, v)) return dflt;
    if (v.size() >= 3) return V{v[0], v[1], v[2]};
    throw LoadError{std::string("attribute '") + name + "' needs 3 numbers"};
  }
  double angle(double a) const { return degrees ? a * M_PI / 180.0 : a; }
  // mjcf.rs:91-139
  Q orientation(const XNode& n, const std::string& cls) const {
    std::vector<double> v;
    if (floats(n, cls, "quat", v)) {
      if (v.size() != 4) throw LoadError{"quat needs 4 numbers"};
      return qnormalized(Q{v[0], v[1], v[2], v[3]});
    }
    if (floats(n, cls, "axisangle", v)) {
      if (v.size() != 4) throw LoadError{"axisangle needs 4 numbers"};
      return from_axis_angle(V{v[0], v[1], v[2]}, angle(v[3]));
    }
    if (floats(n, cls, "euler", v)) {
      if (v.size() != 3) throw LoadError{"euler needs 3 numbers"};
      Q q{1, 0, 0, 0};
      for (int i = 0; i < 3; i++) {
        char c = eulerseq[i];
        char l = (char)tolower((unsigned char)c);
        V axis = l == 'x' ? V{1, 0, 0} : l == 'y' ? V{0, 1, 0} : l == 'z' ? V{0, 0, 1} : V{0, 0, 0};
        if (l != 'x' && l != 'y' && l != 'z') throw LoadError{"bad eulerseq"};
        Q r = from_axis_angle(axis, angle(v[i]));
        q = islower((unsigned char)c) ? qmul(q, r) : qmul(r, q);
      }
      return qnormalized(q);
    }
    if (floats(n, cls, "xyaxes", v)) {
      if (v.size() != 6) throw LoadError{"xyaxes needs 6 numbers"};
      V xa = vnormalized(V{v[0], v[1], v[2]});
      V ya{v[3], v[4], v[5]};
      ya = vnormalized(vsub(ya, vscale(xa, vdot(xa, ya))));
      V za = vcross(xa, ya);
      return from_mat({xa.x, ya.x, za.x, xa.y, ya.y, za.y, xa.z, ya.z, za.z});
    }
    if (floats(n, cls, "zaxis", v)) {
      if (v.size() != 3) throw LoadError{"zaxis needs 3 numbers"};
      return from_z_to(V{v[0], v[1], v[2]});
    }
    return Q{1, 0, 0, 0};
  }
};

struct Builder {
  phh::HostModel& m;
  std::vector<std::string> body_names, jnt_names, tendon_names;
  std::vector<std::vector<std::pair<int, double>>> tendon_joints;
  std::vector<double> gm_mass;
  std::vector<V> gm_mc;
  std::vector<Mat> gm_second;
  std::vector<bool> explicit_inertial;
  explicit Builder(phh::HostModel& hm) : m(hm) {}

  int add_body(const std::string& name, int parent, V pos, Q q) {
    body_names.push_back(name);
    m.body_parent.push_back(parent);
    m.body_pos.insert(m.body_pos.end(), {pos.x, pos.y, pos.z});
    m.body_quat.insert(m.body_quat.end(), {q.w, q.x, q.y, q.z});
    m.body_jntadr.push_back((int)m.jnt_type.size());
    m.body_jntnum.push_back(0);
    m.body_dofadr.push_back(m.d.nv);
    m.body_dofnum.push_back(0);
    m.body_mass.push_back(0.0);
    m.body_ipos.insert(m.body_ipos.end(), {0.0, 0.0, 0.0});
    m.body_inertia.insert(m.body_inertia.end(), 9, 0.0);
    gm_mass.push_back(0.0);
    gm_mc.push_back(V{0, 0, 0});
    gm_second.push_back(Mat(9, 0.0));
    explicit_inertial.push_back(false);
    return (int)body_names.size() - 1;
  }
};

void parse_default(Ctx& ctx, const XNode& node, const std::string& parent, bool top) {
  std::string name = top ? "main" : (node.get("class") ? node.get("class") : "main");
  DefaultClass dc = ctx.defaults.count(parent) ? ctx.defaults[parent] : DefaultClass();
  for (auto& c : node.kids) {
    if (c->tag == "default") continue;
    auto& entry = dc[c->tag];
    for (auto& a : c->attrs) entry[a.first] = a.second;
  }
  ctx.defaults[name] = dc;
  for (auto& c : node.kids)
    if (c->tag == "default") parse_default(ctx, *c, name, false);
}

// mjcf.rs:774-808
void geom_volume_inertia(int gt, V s, double* vol, V* unit) {
  const double PI = M_PI;
  if (gt == PH_GEOM_SPHERE) {
    double r = s.x;
    double v = 4.0 / 3.0 * PI * r * r * r;
    double i = 0.4 * v * r * r;
    *vol = v;
    *unit = V{i, i, i};
  } else if (gt == PH_GEOM_BOX) {
    double a = s.x, b = s.y, c = s.z;
    double v = 8.0 * a * b * c;
    *vol = v;
    *unit = V{v * (b * b + c * c) / 3.0, v * (a * a + c * c) / 3.0, v * (a * a + b * b) / 3.0};
  } else if (gt == PH_GEOM_CYLINDER) {
    double r = s.x, h = s.y;
    double v = PI * r * r * 2.0 * h;
    double ix = v * (3.0 * r * r + 4.0 * h * h) / 12.0;
    *vol = v;
    *unit = V{ix, ix, 0.5 * v * r * r};
  } else if (gt == PH_GEOM_CAPSULE) {
    double r = s.x, h = s.y;
    double vc = PI * r * r * 2.0 * h;
    double vs = 4.0 / 3.0 * PI * r * r * r;
    double v = vc + vs;
    double hs = vs * 0.5;
    double d = h + 3.0 * r / 8.0;
    double hs_ix = hs * (83.0 / 320.0) * r * r;
    double ix = vc * (3.0 * r * r + 4.0 * h * h) / 12.0 + 2.0 * (hs_ix + hs * d * d);
    double iz = 0.5 * vc * r * r + 0.4 * vs * r * r;
    *vol = v;
    *unit = V{ix, ix, iz};
  } else {
    *vol = 0.0;
    *unit = V{0, 0, 0};
  }
}

uint32_t as_u32(double x) {
  if (!(x > 0.0)) return 0;
  if (x >= 4294967295.0) return 4294967295u;
  return (uint32_t)x;
}
int32_t as_i32(double x) {
  if (x != x) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return -2147483647 - 1;
  return (int32_t)x;
}

int find_name(const std::vector<std::string>& v, const std::string& n) {
  for (size_t i = 0; i < v.size(); i++)
    if (v[i] == n) return (int)i;
  return -1;
}

void parse_joint(Ctx& ctx, Builder& b, const XNode& node, int body, const std::string& cls) {
  phh::HostModel& m = b.m;
  int jt;
  if (node.tag == "freejoint") jt = PH_JNT_FREE;
  else {
    const char* t = ctx.get(node, cls, "type");
    std::string ts = t ? t : "hinge";
    if (ts == "free") jt = PH_JNT_FREE;
    else if (ts == "ball") jt = PH_JNT_BALL;
    else if (ts == "slide") jt = PH_JNT_SLIDE;
    else if (ts == "hinge") jt = PH_JNT_HINGE;
    else throw LoadError{"unknown joint type '" + ts + "'"};
  }
  int jid = (int)m.jnt_type.size();
  const char* nm = node.get("name");
  std::string name = nm ? nm : "joint" + std::to_string(jid);
  V pos = jt == PH_JNT_FREE ? V{0, 0, 0} : ctx.vec3(node, cls, "pos", V{0, 0, 0});
  V axis = vnormalized(ctx.vec3(node, cls, "axis", V{0, 0, 1}));
  std::vector<double> rng;
  bool has_range = ctx.floats(node, cls, "range", rng);
  const char* lim = ctx.get(node, cls, "limited");
  bool limited;
  if (lim && std::string(lim) == "true") limited = true;
  else if (lim && std::string(lim) == "false") limited = false;
  else limited = ctx.autolimits && has_range;
  limited = limited && (jt == PH_JNT_HINGE || jt == PH_JNT_SLIDE);
  double r0 = 0.0, r1 = 0.0;
  if (has_range) {
    if (rng.size() < 2) throw LoadError{"range needs 2 numbers"};
    r0 = rng[0];
    r1 = rng[1];
  }
  if (jt == PH_JNT_HINGE) {
    r0 = ctx.angle(r0);
    r1 = ctx.angle(r1);
  }
  double refv = ctx.flt(node, cls, "ref", 0.0);
  double springref = ctx.flt(node, cls, "springref", 0.0);
  if (jt == PH_JNT_HINGE) {
    refv = ctx.angle(refv);
    springref = ctx.angle(springref);
  }
  double damping = ctx.flt(node, cls, "damping", 0.0);
  double armature = ctx.flt(node, cls, "armature", 0.0);
  double stiffness = ctx.flt(node, cls, "stiffness", 0.0);
  b.jnt_names.push_back(name);
  m.jnt_type.push_back(jt);
  m.jnt_body.push_back(body);
  m.jnt_qposadr.push_back(m.d.nq);
  m.jnt_dofadr.push_back(m.d.nv);
  m.jnt_pos.insert(m.jnt_pos.end(), {pos.x, pos.y, pos.z});
  m.jnt_axis.insert(m.jnt_axis.end(), {axis.x, axis.y, axis.z});
  m.jnt_limited.push_back(limited ? 1 : 0);
  m.jnt_range.insert(m.jnt_range.end(), {r0, r1});
  m.jnt_stiffness.push_back(stiffness);
  m.jnt_springref.push_back(springref);
  if (jt == PH_JNT_FREE) {
    const double* p = &m.body_pos[3 * body];
    const double* q = &m.body_quat[4 * body];
    m.qpos0.insert(m.qpos0.end(), {p[0], p[1], p[2], q[0], q[1], q[2], q[3]});
  } else if (jt == PH_JNT_BALL) {
    m.qpos0.insert(m.qpos0.end(), {1.0, 0.0, 0.0, 0.0});
  } else {
    m.qpos0.push_back(refv);
  }
  int nv = jt == PH_JNT_FREE ? 6 : (jt == PH_JNT_BALL ? 3 : 1);
  int nq = jt == PH_JNT_FREE ? 7 : (jt == PH_JNT_BALL ? 4 : 1);
  for (int k = 0; k < nv; k++) {
    m.dof_body.push_back(body);
    m.dof_jnt.push_back(jid);
    m.dof_damping.push_back(damping);
    m.dof_armature.push_back(armature);
  }
  m.d.nq += nq;
  m.d.nv += nv;
  m.body_jntnum[body] += 1;
  m.body_dofnum[body] += nv;
  if (m.body_jntnum[body] == 1) {
    m.body_jntadr[body] = jid;
    m.body_dofadr[body] = m.jnt_dofadr[jid];
  }
}

void parse_geom(Ctx& ctx, Builder& b, const XNode& node, int body, const std::string& cls) {
  phh::HostModel& m = b.m;
  const char* t = ctx.get(node, cls, "type");
  std::string ts = t ? t : "sphere";
  int gt;
  if (ts == "plane") gt = PH_GEOM_PLANE;
  else if (ts == "sphere") gt = PH_GEOM_SPHERE;
  else if (ts == "capsule") gt = PH_GEOM_CAPSULE;
  else if (ts == "box") gt = PH_GEOM_BOX;
  else if (ts == "cylinder") gt = PH_GEOM_CYLINDER;
  else if (ts == "mesh") gt = PH_GEOM_MESH;
  else return;  // unsupported geom types are skipped (mjcf.rs:645-648)
  int dataid = -1;
  if (gt == PH_GEOM_MESH) {
    const char* mn = ctx.get(node, cls, "mesh");
    if (!mn) throw LoadError{"mesh geom needs a 'mesh' attribute"};
    dataid = find_name(ctx.mesh_names, mn);
    if (dataid < 0) throw LoadError{std::string("unknown mesh '") + mn + "'"};
  }
  std::vector<double> sz;
  ctx.floats(node, cls, "size", sz);
  auto get = [&](size_t i) { return i < sz.size() ? sz[i] : 0.0; };
  V size{get(0), get(1), get(2)};
  V pos = ctx.vec3(node, cls, "pos", V{0, 0, 0});
  Q quat = ctx.orientation(node, cls);
  std::vector<double> ft;
  if (ctx.floats(node, cls, "fromto", ft)) {
    if (ft.size() != 6) throw LoadError{"fromto needs 6 numbers"};
    V a{ft[0], ft[1], ft[2]}, bb{ft[3], ft[4], ft[5]};
    V d = vsub(bb, a);
    pos = vscale(vadd(a, bb), 0.5);
    quat = from_z_to(d);
    if (gt == PH_GEOM_CAPSULE || gt == PH_GEOM_CYLINDER) size = V{get(0), vnorm(d) * 0.5, 0.0};
    else if (gt == PH_GEOM_BOX) size = V{get(0), get(1), vnorm(d) * 0.5};
    else throw LoadError{"fromto only valid for capsule, cylinder, box"};
  }
  std::vector<double> fr;
  double friction = (ctx.floats(node, cls, "friction", fr) && !fr.empty()) ? fr[0] : 1.0;
  uint32_t contype = as_u32(ctx.flt(node, cls, "contype", 1.0));
  uint32_t conaffinity = as_u32(ctx.flt(node, cls, "conaffinity", 1.0));
  std::vector<float> rgba;
  std::vector<double> rv;
  if (ctx.floats(node, cls, "rgba", rv) && rv.size() == 4) {
    rgba = {(float)rv[0], (float)rv[1], (float)rv[2], (float)rv[3]};
  } else if (const char* mat = ctx.get(node, cls, "material")) {
    auto it = ctx.materials.find(mat);
    if (it != ctx.materials.end()) rgba = it->second;
  }
  if (rgba.empty()) rgba = {0.5f, 0.5f, 0.5f, 1.0f};
  int32_t group = as_i32(ctx.flt(node, cls, "group", 0.0));
  if (gt == PH_GEOM_MESH) {  // mjcf.rs:686-703: the hull's solid mass properties
    double vol, com[3], icom[9];
    phh::mass_properties(ctx.hulls[dataid], &vol, com, icom);
    std::vector<double> mv;
    double mass = ctx.floats(node, cls, "mass", mv) && !mv.empty() ? mv[0] : ctx.flt(node, cls, "density", 1000.0) * vol;
    if (mass > 0.0 && vol > 0.0) {
      Mat r = to_mat(quat);
      V c = vadd(pos, mulv(r, V{com[0], com[1], com[2]}));
      Mat ig = mrotate(mscale(Mat(icom, icom + 9), mass / vol), r);
      Mat cc = mskew(c);
      Mat shifted = madd(ig, mscale(mmul(cc, mT(cc)), mass));
      b.gm_mass[body] += mass;
      b.gm_mc[body] = vadd(b.gm_mc[body], vscale(c, mass));
      b.gm_second[body] = madd(b.gm_second[body], shifted);
    }
  } else if (gt != PH_GEOM_PLANE) {
    double vol;
    V unit;
    geom_volume_inertia(gt, size, &vol, &unit);
    std::vector<double> mv;
    double mass = ctx.floats(node, cls, "mass", mv) && !mv.empty() ? mv[0] : ctx.flt(node, cls, "density", 1000.0) * vol;
    if (mass > 0.0 && vol > 0.0) {
      Mat ig = mrotate(mdiag(vscale(unit, mass / vol)), to_mat(quat));
      Mat pp = mskew(pos);
      Mat shifted = madd(ig, mscale(mmul(pp, mT(pp)), mass));
      b.gm_mass[body] += mass;
      b.gm_mc[body] = vadd(b.gm_mc[body], vscale(pos, mass));
      b.gm_second[body] = madd(b.gm_second[body], shifted);
    }
  }
  m.geom_type.push_back(gt);
  m.geom_body.push_back(body);
  m.geom_pos.insert(m.geom_pos.end(), {pos.x, pos.y, pos.z});
  m.geom_quat.insert(m.geom_quat.end(), {quat.w, quat.x, quat.y, quat.z});
  m.geom_size.insert(m.geom_size.end(), {size.x, size.y, size.z});
  m.geom_friction.push_back(friction);
  m.geom_contype.push_back(contype);
  m.geom_conaffinity.push_back(conaffinity);
  m.geom_rgba.insert(m.geom_rgba.end(), rgba.begin(), rgba.end());
  m.geom_group.push_back(group);
  m.geom_dataid.push_back(dataid);
}

void parse_inertial(Ctx& ctx, Builder& b, const XNode& node, int body, const std::string& cls) {
  phh::HostModel& m = b.m;
  V pos = ctx.vec3(node, cls, "pos", V{0, 0, 0});
  Q quat = ctx.orientation(node, cls);
  double mass = ctx.flt(node, cls, "mass", 0.0);
  std::vector<double> d;
  Mat inertia;
  if (ctx.floats(node, cls, "diaginertia", d)) {
    if (d.size() < 3) throw LoadError{"diaginertia needs 3 numbers"};
    inertia = mrotate(mdiag(V{d[0], d[1], d[2]}), to_mat(quat));
  } else if (ctx.floats(node, cls, "fullinertia", d)) {
    if (d.size() < 6) throw LoadError{"fullinertia needs 6 numbers"};
    inertia = mrotate({d[0], d[3], d[4], d[3], d[1], d[5], d[4], d[5], d[2]}, to_mat(quat));
  } else {
    throw LoadError{"<inertial> requires diaginertia or fullinertia"};
  }
  m.body_mass[body] = mass;
  m.body_ipos[3 * body] = pos.x;
  m.body_ipos[3 * body + 1] = pos.y;
  m.body_ipos[3 * body + 2] = pos.z;
  for (int k = 0; k < 9; k++) m.body_inertia[9 * body + k] = inertia[k];
  b.explicit_inertial[body] = true;
}

void parse_body_contents(Ctx& ctx, Builder& b, const XNode& node, int body, const std::string& cls) {
  for (auto& cp : node.kids) {
    const XNode& c = *cp;
    if (c.tag == "geom") parse_geom(ctx, b, c, body, cls);
    else if (c.tag == "joint" || c.tag == "freejoint") {
      if (body == 0) throw LoadError{"joints cannot be attached to the world body"};
      parse_joint(ctx, b, c, body, cls);
    } else if (c.tag == "inertial") parse_inertial(ctx, b, c, body, cls);
    else if (c.tag == "body") {
      const char* nm = c.get("name");
      std::string name = nm ? nm : "body" + std::to_string(b.body_names.size());
      V pos = ctx.vec3(c, cls, "pos", V{0, 0, 0});
      Q quat = ctx.orientation(c, cls);
      const char* cc = c.get("childclass");
      std::string child = cc ? cc : cls;
      int id = b.add_body(name, body, pos, quat);
      parse_body_contents(ctx, b, c, id, child);
    } else if (c.tag == "camera") {
      phh::HostModel& m = b.m;
      m.cam_body.push_back(body);
      V p = ctx.vec3(c, cls, "pos", V{0, 0, 0});
      Q q = ctx.orientation(c, cls);
      m.cam_pos.insert(m.cam_pos.end(), {p.x, p.y, p.z});
      m.cam_quat.insert(m.cam_quat.end(), {q.w, q.x, q.y, q.z});
      m.cam_fovy.push_back(ctx.flt(c, cls, "fovy", 45.0));
    }
  }
}

// mjcf.rs:830-850
void finalize_inertia(Builder& b) {
  phh::HostModel& m = b.m;
  for (size_t body = 1; body < b.body_names.size(); body++) {
    if (b.explicit_inertial[body]) continue;
    double mass = b.gm_mass[body];
    if (mass <= 0.0) {
      if (m.body_jntnum[body] > 0) throw LoadError{"moving body '" + b.body_names[body] + "' has no mass"};
      continue;
    }
    V com = vscale(b.gm_mc[body], 1.0 / mass);
    Mat cx = mskew(com);
    Mat icom = madd(b.gm_second[body], mscale(mmul(cx, mT(cx)), -mass));
    m.body_mass[body] = mass;
    m.body_ipos[3 * body] = com.x;
    m.body_ipos[3 * body + 1] = com.y;
    m.body_ipos[3 * body + 2] = com.z;
    for (int k = 0; k < 9; k++) m.body_inertia[9 * body + k] = icom[k];
  }
}

void parse_tendon(Ctx&, Builder& b, const XNode& node) {
  if (node.tag != "fixed") return;
  std::vector<std::pair<int, double>> joints;
  for (auto& jp : node.kids) {
    if (jp->tag != "joint") continue;
    const char* nm = jp->get("joint");
    std::string name = nm ? nm : "";
    int j = find_name(b.jnt_names, name);
    if (j < 0) throw LoadError{"tendon references unknown joint '" + name + "'"};
    int t = b.m.jnt_type[j];
    if (t != PH_JNT_HINGE && t != PH_JNT_SLIDE) throw LoadError{"fixed tendons need hinge or slide joints"};
    double coef = 1.0;
    if (const char* c = jp->get("coef")) {
      std::vector<double> v;
      parse_floats(c, v);
      if (!v.empty()) coef = v[0];
    }
    joints.push_back({j, coef});
  }
  if (joints.empty()) throw LoadError{"fixed tendon needs at least one joint"};
  const char* nm = node.get("name");
  b.tendon_names.push_back(nm ? nm : "tendon" + std::to_string(b.tendon_names.size()));
  b.tendon_joints.push_back(joints);
}

void parse_actuator(Ctx& ctx, Builder& b, const XNode& node) {
  phh::HostModel& m = b.m;
  const std::string& tag = node.tag;
  if (tag != "motor" && tag != "position" && tag != "velocity" && tag != "general")
    throw LoadError{"actuator <" + tag + "> not supported"};
  const std::string cls = "main";
  std::vector<std::pair<int, double>> moment;
  if (const char* jn = node.get("joint")) {
    int j = find_name(b.jnt_names, jn);
    if (j < 0) throw LoadError{std::string("actuator references unknown joint '") + jn + "'"};
    int t = m.jnt_type[j];
    if (t != PH_JNT_HINGE && t != PH_JNT_SLIDE) throw LoadError{"actuators on free/ball joints are not supported"};
    moment.push_back({j, 1.0});
  } else if (const char* tn = node.get("tendon")) {
    int t = find_name(b.tendon_names, tn);
    if (t < 0) throw LoadError{std::string("actuator references unknown tendon '") + tn + "'"};
    moment = b.tendon_joints[t];
  } else {
    throw LoadError{"only joint and fixed-tendon transmissions are supported"};
  }
  double gear = ctx.flt(node, cls, "gear", 1.0);
  double gain;
  double bias[3] = {0.0, 0.0, 0.0};
  if (tag == "motor") gain = 1.0;
  else if (tag == "position") {
    double kp = ctx.flt(node, cls, "kp", 1.0), kv = ctx.flt(node, cls, "kv", 0.0);
    gain = kp;
    bias[1] = -kp;
    bias[2] = -kv;
  } else if (tag == "velocity") {
    double kv = ctx.flt(node, cls, "kv", 1.0);
    gain = kv;
    bias[2] = -kv;
  } else {
    std::vector<double> gp, bp;
    gain = ctx.floats(node, cls, "gainprm", gp) && !gp.empty() ? gp[0] : 1.0;
    ctx.floats(node, cls, "biasprm", bp);
    const char* bt = ctx.get(node, cls, "biastype");
    if (bt && std::string(bt) == "affine")
      for (int i = 0; i < 3; i++) bias[i] = i < (int)bp.size() ? bp[i] : 0.0;
  }
  auto limited = [&](const char* flag, const char* range, double* r) {
    std::vector<double> v;
    bool has = ctx.floats(node, cls, range, v);
    const char* f = ctx.get(node, cls, flag);
    bool lim;
    if (f && std::string(f) == "true") lim = true;
    else if (f && std::string(f) == "false") lim = false;
    else lim = ctx.autolimits && has;
    r[0] = has && v.size() > 0 ? v[0] : 0.0;
    r[1] = has && v.size() > 1 ? v[1] : 0.0;
    return lim;
  };
  double cr[2], fr[2];
  bool cl = limited("ctrllimited", "ctrlrange", cr);
  bool fl = limited("forcelimited", "forcerange", fr);
  m.actuator_gear.push_back(gear);
  m.actuator_gain.push_back(gain);
  m.actuator_bias.insert(m.actuator_bias.end(), {bias[0], bias[1], bias[2]});
  m.actuator_ctrllimited.push_back(cl ? 1 : 0);
  m.actuator_ctrlrange.insert(m.actuator_ctrlrange.end(), {cr[0], cr[1]});
  m.actuator_forcelimited.push_back(fl ? 1 : 0);
  m.actuator_forcerange.insert(m.actuator_forcerange.end(), {fr[0], fr[1]});
  m.actuator_moment_adr.push_back((int)m.moment_jnt.size());
  m.actuator_moment_num.push_back((int)moment.size());
  for (auto& e : moment) {
    m.moment_jnt.push_back(e.first);
    m.moment_coef.push_back(e.second);
  }
}

void parse_equality(Ctx& ctx, Builder& b, const XNode& node) {
  phh::HostModel& m = b.m;
  if (node.tag != "joint") return;
  if (const char* a = node.get("active"))
    if (std::string(a) == "false") return;
  auto get = [&](const char* name) -> const char* {
    if (const char* v = node.get(name)) return v;
    const char* c = node.get("class");
    auto it = ctx.defaults.find(c ? c : "main");
    if (it == ctx.defaults.end()) return nullptr;
    auto t = it->second.find("equality");
    if (t == it->second.end()) return nullptr;
    auto a = t->second.find(name);
    return a == t->second.end() ? nullptr : a->second.c_str();
  };
  auto joint = [&](const char* attr) {
    const char* nm = node.get(attr);
    std::string name = nm ? nm : "";
    int j = find_name(b.jnt_names, name);
    if (j < 0) throw LoadError{"equality references unknown joint '" + name + "'"};
    int t = m.jnt_type[j];
    if (t != PH_JNT_HINGE && t != PH_JNT_SLIDE) throw LoadError{"joint equality needs hinge or slide joints"};
    return j;
  };
  int j1 = joint("joint1");
  int j2 = node.get("joint2") ? joint("joint2") : -1;
  double poly[5] = {0.0, 1.0, 0.0, 0.0, 0.0};
  std::vector<double> v;
  if (const char* s = get("polycoef")) {
    parse_floats(s, v);
    for (size_t k = 0; k < v.size() && k < 5; k++) poly[k] = v[k];
  }
  double solref[2] = {m.d.solref[0], m.d.solref[1]};
  if (const char* s = get("solref")) {
    parse_floats(s, v);
    if (v.empty()) throw LoadError{"solref needs numbers"};
    solref[0] = v[0];
    solref[1] = v.size() > 1 ? v[1] : 1.0;
  }
  double solimp[5];
  for (int k = 0; k < 5; k++) solimp[k] = m.d.solimp[k];
  if (const char* s = get("solimp")) {
    parse_floats(s, v);
    for (size_t k = 0; k < v.size() && k < 5; k++) solimp[k] = v[k];
  }
  m.eq_joint1.push_back(j1);
  m.eq_joint2.push_back(j2);
  m.eq_polycoef.insert(m.eq_polycoef.end(), poly, poly + 5);
  m.eq_solref.insert(m.eq_solref.end(), solref, solref + 2);
  m.eq_solimp.insert(m.eq_solimp.end(), solimp, solimp + 5);
}

// mjcf.rs:740-772: a mesh asset's points (OBJ or STL file, or the 'vertex'
// attribute), scaled per axis, and their convex hull.
void parse_mesh(Ctx& ctx, const XNode& node, const std::string& meshdir) {
  const char* file = node.get("file");
  std::string name;
  if (const char* nm = node.get("name")) name = nm;
  else if (file) {
    std::string f = file;
    size_t slash = f.find_last_of('/');
    f = slash == std::string::npos ? f : f.substr(slash + 1);
    size_t dot = f.find_last_of('.');
    name = dot == std::string::npos || dot == 0 ? f : f.substr(0, dot);
  } else
    throw LoadError{"mesh needs a name or file"};
  V sc = ctx.vec3(node, "main", "scale", V{1, 1, 1});
  const double scale[3] = {sc.x, sc.y, sc.z};
  std::vector<double> pts;
  if (file) {
    std::string path = (file[0] == '/' ? std::string() : meshdir + "/") + file;
    std::ifstream in(path, std::ios::binary);
    if (!in) throw LoadError{"cannot read mesh '" + path + "'"};
    std::stringstream ss;
    ss << in.rdbuf();
    std::string ext = path.size() > 4 ? path.substr(path.size() - 4) : "";
    for (char& ch : ext) ch = (char)tolower((unsigned char)ch);
    bool ok = ext == ".obj" ? phh::read_obj_vertices(ss.str(), scale, &pts)
                            : ext == ".stl" ? phh::read_stl_vertices(ss.str(), scale, &pts) : false;
    if (ext != ".obj" && ext != ".stl") throw LoadError{"unsupported mesh format (use .stl or .obj): " + path};
    if (!ok) throw LoadError{"cannot read mesh '" + path + "'"};
  } else {
    std::vector<double> v;
    if (!ctx.floats(node, "main", "vertex", v)) throw LoadError{"mesh '" + name + "' needs 'file' or 'vertex'"};
    for (size_t i = 0; i + 2 < v.size(); i += 3)
      for (int k = 0; k < 3; k++) pts.push_back(v[i + k] * scale[k]);
  }
  phh::Hull h;
  std::string err;
  if (!phh::convex_hull(pts, &h, &err)) throw LoadError{"mesh '" + name + "': " + err};
  ctx.mesh_names.push_back(name);
  ctx.hulls.push_back(h);
}

// mjcf.rs:178-209: each <include file="..."/> is replaced, in place, by the
// children of the included document's root element; files resolve against the
// including file's folder; nesting is limited to 16 levels.
void expand_includes(XNode& node, const std::string& dir, int depth) {
  std::vector<std::unique_ptr<XNode>> kids;
  for (auto& k : node.kids) {
    if (k->tag != "include") {
      expand_includes(*k, dir, depth);
      kids.push_back(std::move(k));
      continue;
    }
    const char* f = k->get("file");
    std::string file = f ? f : "";
    std::string path = (!file.empty() && file[0] == '/') ? file : dir + "/" + file;
    std::ifstream in(path, std::ios::binary);
    if (!in) throw LoadError{"<include file=\"" + file + "\">: cannot read " + path};
    // the included text is expanded one level deeper; deeper than 16 is an error
    if (depth + 1 > 16) throw LoadError{"<include> nested too deeply (cycle?)"};
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    XmlParser sub(text);
    std::unique_ptr<XNode> sroot = sub.parse();
    size_t slash = path.find_last_of('/');
    std::string sub_dir = slash == std::string::npos ? dir : path.substr(0, slash);
    expand_includes(*sroot, sub_dir, depth + 1);
    for (auto& sk : sroot->kids) kids.push_back(std::move(sk));
  }
  node.kids = std::move(kids);
}

// mjcf.rs:211-395
void load(const std::string& xml, const std::string& dir, phh::HostModel& m) {
  XmlParser parser(xml);
  std::unique_ptr<XNode> root = parser.parse();
  if (root->tag != "mujoco") throw LoadError{"root element must be <mujoco>"};
  expand_includes(*root, dir.empty() ? "." : dir, 0);
  Ctx ctx;
  std::string meshdir = dir.empty() ? "." : dir;
  for (auto& c : root->kids) {
    if (c->tag != "compiler") continue;
    const char* md = c->get("meshdir");
    if (!md) md = c->get("assetdir");
    if (md) meshdir = (md[0] == '/' ? std::string() : meshdir + "/") + md;
    if (const char* a = c->get("angle")) ctx.degrees = std::string(a) == "degree";
    if (const char* e = c->get("eulerseq")) {
      if (strlen(e) != 3) throw LoadError{"eulerseq must have 3 characters"};
      ctx.eulerseq = e;
    }
    if (const char* a = c->get("autolimits")) ctx.autolimits = std::string(a) == "true";
  }
  ctx.defaults["main"] = DefaultClass();
  for (auto& c : root->kids)
    if (c->tag == "default") parse_default(ctx, *c, "main", true);

  memset(&m.d, 0, sizeof(m.d));
  ph_model_desc& d = m.d;
  d.timestep = 0.002;
  d.gravity[0] = 0.0;
  d.gravity[1] = 0.0;
  d.gravity[2] = -9.81;
  d.integrator = PH_INTEGRATOR_EULER;
  // SolverOptions::default (model.rs:85-100)
  d.solver_kind = PH_SOLVER_NEWTON;
  d.solver_iterations = 30;
  d.tolerance = 1e-8;
  d.contact_margin = 1e-3;
  d.solref[0] = 0.02;
  d.solref[1] = 1.0;
  const double si[5] = {0.9, 0.95, 0.001, 0.5, 2.0};
  for (int k = 0; k < 5; k++) d.solimp[k] = si[k];
  d.impratio = 1.0;
  d.contact_hertz = 0.0;
  d.contact_damping_ratio = 1.0;
  d.contact_slop = 5e-4;
  std::vector<double> v;
  for (auto& o : root->kids) {
    if (o->tag != "option") continue;
    if (const char* s = o->get("timestep")) {
      parse_floats(s, v);
      d.timestep = v.at(0);
    }
    if (const char* s = o->get("gravity")) {
      parse_floats(s, v);
      if (v.size() < 3) throw LoadError{"gravity needs 3 numbers"};
      for (int k = 0; k < 3; k++) d.gravity[k] = v[k];
    }
    if (const char* s = o->get("solver")) d.solver_kind = std::string(s) == "PGS" ? PH_SOLVER_PGS : PH_SOLVER_NEWTON;
    if (const char* s = o->get("impratio")) {
      parse_floats(s, v);
      d.impratio = v.at(0);
    }
    if (const char* s = o->get("tolerance")) {
      parse_floats(s, v);
      d.tolerance = v.at(0);
    }
    if (const char* s = o->get("o_solref")) {
      parse_floats(s, v);
      if (v.size() < 2) throw LoadError{"o_solref needs 2 numbers"};
      d.solref[0] = v[0];
      d.solref[1] = v[1];
    }
    if (const char* s = o->get("iterations")) {
      parse_floats(s, v);
      d.solver_iterations = (int)v.at(0);
    }
    if (const char* s = o->get("integrator")) {
      std::string is = s;
      d.integrator = (is == "implicitfast" || is == "implicit") ? PH_INTEGRATOR_IMPLICITFAST : PH_INTEGRATOR_EULER;
    }
  }
  Builder b(m);
  for (auto& a : root->kids) {
    if (a->tag != "asset") continue;
    for (auto& c : a->kids) {
      if (c->tag == "mesh") {
        parse_mesh(ctx, *c, meshdir);
        continue;
      }
      if (c->tag != "material") continue;
      const char* k = c->get("class");
      std::string cls = k ? k : "main";
      std::vector<double> rgba;
      const char* name = c->get("name");
      if (name && ctx.floats(*c, cls, "rgba", rgba) && rgba.size() == 4)
        ctx.materials[name] = {(float)rgba[0], (float)rgba[1], (float)rgba[2], (float)rgba[3]};
    }
  }
  b.add_body("world", 0, V{0, 0, 0}, Q{1, 0, 0, 0});
  bool any_world = false;
  for (auto& w : root->kids)
    if (w->tag == "worldbody") {
      any_world = true;
      parse_body_contents(ctx, b, *w, 0, "main");
    }
  if (!any_world) throw LoadError{"missing <worldbody>"};
  finalize_inertia(b);
  for (auto& t : root->kids)
    if (t->tag == "tendon")
      for (auto& c : t->kids) parse_tendon(ctx, b, *c);
  for (auto& a : root->kids)
    if (a->tag == "actuator")
      for (auto& c : a->kids) parse_actuator(ctx, b, *c);
  for (auto& e : root->kids)
    if (e->tag == "equality")
      for (auto& c : e->kids) parse_equality(ctx, b, *c);
  for (auto& c : root->kids) {
    if (c->tag != "contact") continue;
    for (auto& x : c->kids) {
      if (x->tag != "exclude") continue;
      int ids[2];
      const char* at[2] = {"body1", "body2"};
      for (int k = 0; k < 2; k++) {
        const char* nm = x->get(at[k]);
        if (!nm) throw LoadError{std::string("<exclude> needs ") + at[k]};
        ids[k] = find_name(b.body_names, nm);
        if (ids[k] < 0) throw LoadError{std::string("<exclude> references unknown body '") + nm + "'"};
      }
      m.exclude_pairs.push_back(ids[0]);
      m.exclude_pairs.push_back(ids[1]);
    }
  }
  d.nbody = (int)b.body_names.size();
  d.njnt = (int)m.jnt_type.size();
  d.ngeom = (int)m.geom_type.size();
  d.ncam = (int)m.cam_body.size();
  d.nu = (int)m.actuator_gear.size();
  d.neq = (int)m.eq_joint1.size();
  d.nexclude = (int)m.exclude_pairs.size() / 2;
  d.nmoment = (int)m.moment_jnt.size();
  m.qpos_reset = m.qpos0;
  for (const phh::Hull& h : ctx.hulls) {  // the descriptor's hull arrays
    m.mesh_vertadr.push_back((int)m.mesh_vert.size() / 3);
    m.mesh_vertnum.push_back((int)h.vert.size() / 3);
    m.mesh_vert.insert(m.mesh_vert.end(), h.vert.begin(), h.vert.end());
    m.mesh_faceadr.push_back((int)m.face_vertadr.size());
    m.mesh_facenum.push_back((int)h.faces.size());
    for (const auto& f : h.faces) {
      m.face_vertadr.push_back((int)m.face_vert.size());
      m.face_vertnum.push_back((int)f.size());
      m.face_vert.insert(m.face_vert.end(), f.begin(), f.end());
    }
  }
  d.nmesh = (int)ctx.hulls.size();
  d.nmeshvert = (int)m.mesh_vert.size() / 3;
  d.nmeshface = (int)m.face_vertadr.size();
  d.nmeshfacevert = (int)m.face_vert.size();
  phh::retarget(m);
}

}  // namespace

void ph_set_error(const std::string& msg);  // runtime.hip

struct ph_model {
  phh::HostModel hm;
};

extern "C" int pippin_hip_model_load_xml(const char* xml, const char* dir, ph_model** out) {
  if (!xml || !out) {
    ph_set_error("null argument");
    return PH_ERR_ARG;
  }
  std::unique_ptr<ph_model> m(new ph_model());
  try {
    load(xml, dir ? dir : ".", m->hm);
  } catch (const XmlError& e) {
    ph_set_error("xml parse error: " + e.msg);
    return PH_ERR_MODEL;
  } catch (const LoadError& e) {
    ph_set_error("invalid model: " + e.msg);
    return PH_ERR_MODEL;
  } catch (const std::exception& e) {
    ph_set_error(std::string("invalid model: ") + e.what());
    return PH_ERR_MODEL;
  }
  ph_set_error("");
  *out = m.release();
  return PH_OK;
}

extern "C" int pippin_hip_model_load_file(const char* path, ph_model** out) {
  if (!path) {
    ph_set_error("null path");
    return PH_ERR_ARG;
  }
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    ph_set_error(std::string("io error: cannot open ") + path);
    return PH_ERR_MODEL;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  std::string dir = path;
  size_t slash = dir.find_last_of('/');
  dir = slash == std::string::npos ? "." : dir.substr(0, slash);
  return pippin_hip_model_load_xml(ss.str().c_str(), dir.c_str(), out);
}

extern "C" const ph_model_desc* pippin_hip_model_desc(const ph_model* m) { return m ? &m->hm.d : nullptr; }

extern "C" void pippin_hip_model_free(ph_model* m) { delete m; }
