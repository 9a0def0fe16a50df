// URDF import for the C ABI (pippin_hip_model_load_urdf,
// pippin_hip_urdf_to_mjcf). As crates/pippin/src/urdf.rs does, the URDF is
// rewritten as an MJCF document and loaded by the MJCF loader:
//   - links become bodies; revolute and continuous joints become hinges,
//     prismatic joints slides, floating joints free joints, fixed joints
//     nothing; the root link is welded to the world unless floating_base;
//   - <collision> geometry collides; <visual> geometry is massless, does not
//     collide (contype = conaffinity = 0) and sits in group 1; collision
//     geometry of a link that also has visuals sits in group 3;
//   - <inertial> is taken as given (mass > 0); other links get their mass from
//     their collision geometry;
//   - mesh files (STL/OBJ) may be named by package://, file:// or a path
//     relative to the URDF file.
// Numbers are written back with 17 significant digits, so the loader reads
// exactly the doubles this file computed.
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include <sys/stat.h>

#include "../../include/pippin_hip.h"
#include "xml.hpp"

void ph_set_error(const std::string& msg);  // runtime.hip

namespace {

using phx::XmlError;
using phx::XmlParser;
using phx::XNode;

struct UrdfError {
  std::string msg;
};

struct Options {
  bool floating_base = false;
  std::vector<std::string> package_dirs;  // searched for package://<pkg>/... (urdf.rs:22-28)
};

std::string num(double v) {
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", v);
  return b;
}

// urdf.rs:34-36: whitespace-separated numbers
std::vector<double> floats(const char* s) {
  std::vector<double> out;
  std::istringstream in(s ? s : "");
  std::string t;
  while (in >> t) {
    char* end = nullptr;
    double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') throw UrdfError{"bad number '" + t + "' in URDF"};
    out.push_back(v);
  }
  return out;
}

double first(const char* s) {
  std::vector<double> v = floats(s);
  if (v.empty()) throw UrdfError{"expected a number in URDF"};
  return v[0];
}

// urdf.rs:38-49
void attr3(const XNode* n, const char* name, const double* def, double* out) {
  const char* s = n ? n->get(name) : nullptr;
  if (!s) {
    for (int k = 0; k < 3; k++) out[k] = def[k];
    return;
  }
  std::vector<double> v = floats(s);
  if (v.size() != 3) throw UrdfError{std::string("'") + name + "' needs 3 numbers"};
  for (int k = 0; k < 3; k++) out[k] = v[k];
}

const XNode* child(const XNode* n, const char* tag) {
  if (!n) return nullptr;
  for (auto& k : n->kids)
    if (k->tag == tag) return k.get();
  return nullptr;
}

std::vector<const XNode*> children(const XNode* n, const char* tag) {
  std::vector<const XNode*> out;
  for (auto& k : n->kids)
    if (k->tag == tag) out.push_back(k.get());
  return out;
}

// Quaternions in math.rs's operation order (math.rs:214-238)
struct Qt {
  double w, x, y, z;
};
Qt axis_angle(int axis, double angle) {  // unit axis: normalising leaves it as is
  const double s = std::sin(0.5 * angle), c = std::cos(0.5 * angle);
  const double a[3] = {axis == 0 ? 1.0 : 0.0, axis == 1 ? 1.0 : 0.0, axis == 2 ? 1.0 : 0.0};
  return {c, a[0] * s, a[1] * s, a[2] * s};
}
Qt qmul(Qt p, Qt o) {
  const double a = p.w, b = p.x, c = p.y, d = p.z, e = o.w, f = o.x, g = o.y, h = o.z;
  return {a * e - b * f - c * g - d * h, a * f + b * e + c * h - d * g, a * g - b * h + c * e + d * f,
          a * h + b * g - c * f + d * e};
}
Qt qnormalized(Qt q) {
  double s = -0.0;
  s += q.w * q.w;
  s += q.x * q.x;
  s += q.y * q.y;
  s += q.z * q.z;
  const double n = std::sqrt(s);
  if (n < 1e-15) return {1, 0, 0, 0};
  return {q.w / n, q.x / n, q.y / n, q.z / n};
}

// urdf.rs:60-64: fixed-axis roll, pitch, yaw, R = Rz Ry Rx
Qt rpy_quat(const double* rpy) {
  return qnormalized(qmul(qmul(axis_angle(2, rpy[2]), axis_angle(1, rpy[1])), axis_angle(0, rpy[0])));
}

// urdf.rs:66-73: ` pos="x y z" quat="w x y z"` of an <origin>
std::string origin_attrs(const XNode* origin) {
  const double zero[3] = {0, 0, 0};
  double xyz[3], rpy[3];
  attr3(origin, "xyz", zero, xyz);
  attr3(origin, "rpy", zero, rpy);
  const Qt q = rpy_quat(rpy);
  return " pos=\"" + num(xyz[0]) + " " + num(xyz[1]) + " " + num(xyz[2]) + "\" quat=\"" + num(q.w) + " " + num(q.x) +
         " " + num(q.y) + " " + num(q.z) + "\"";
}

std::string escape(const std::string& s) {  // urdf.rs:200-202
  std::string o;
  for (char c : s) {
    if (c == '&') o += "&amp;";
    else if (c == '"') o += "&quot;";
    else if (c == '<') o += "&lt;";
    else if (c == '>') o += "&gt;";
    else o += c;
  }
  return o;
}

bool exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

std::string canonical(const std::string& p) {
  char buf[PATH_MAX];
  return realpath(p.c_str(), buf) ? std::string(buf) : std::string();
}

std::string join(const std::string& a, const std::string& b) {
  if (a.empty()) return b;
  return a.back() == '/' ? a + b : a + "/" + b;
}

// Path::pop: drop the last component ("" once none is left)
std::string parent_dir(const std::string& d) {
  if (d.empty()) return d;
  std::string s = d;
  while (s.size() > 1 && s.back() == '/') s.pop_back();
  const size_t k = s.find_last_of('/');
  if (k == std::string::npos) return "";
  return k == 0 ? "/" : s.substr(0, k);
}

struct Converter {
  std::string base;
  const Options& opts;
  std::map<std::string, std::vector<double>> materials;
  std::vector<std::pair<std::string, std::pair<std::string, std::vector<double>>>> meshes;  // name, (path, scale)
  std::map<std::pair<std::string, std::vector<uint64_t>>, std::string> mesh_ids;
  std::vector<std::string> warnings;

  // urdf.rs:85-113
  std::string resolve(const std::string& filename) const {
    const std::string pkg_prefix = "package://", file_prefix = "file://";
    if (filename.compare(0, pkg_prefix.size(), pkg_prefix) == 0) {
      const std::string rest = filename.substr(pkg_prefix.size());
      const size_t slash = rest.find('/');
      if (slash == std::string::npos) return "";
      const std::string pkg = rest.substr(0, slash), rel = rest.substr(slash + 1);
      std::vector<std::string> dirs = opts.package_dirs;
      if (const char* env = std::getenv("ROS_PACKAGE_PATH")) {
        std::string e = env;
        size_t a = 0;
        for (;;) {
          const size_t b = e.find(':', a);
          dirs.push_back(e.substr(a, b == std::string::npos ? std::string::npos : b - a));
          if (b == std::string::npos) break;
          a = b + 1;
        }
      }
      for (int up = 0; up < 4; up++) {  // conventional layouts relative to the URDF file
        std::string d = base;
        for (int k = 0; k < up; k++) d = parent_dir(d);
        dirs.push_back(d);
      }
      for (const std::string& d : dirs)
        for (const std::string& cand : {join(join(d, pkg), rel), join(d, rel)})
          if (exists(cand)) return canonical(cand);
      return "";
    }
    std::string p = filename.compare(0, file_prefix.size(), file_prefix) == 0 ? filename.substr(file_prefix.size())
                                                                              : filename;
    if (p.empty() || p[0] != '/') p = join(base, p);
    return canonical(p);  // absolute, so the MJCF does not depend on where it is loaded from
  }

  // urdf.rs:115-137
  bool mesh(const std::string& filename, const double* scale, std::string* name) {
    const std::string path = resolve(filename);
    if (path.empty()) {
      warnings.push_back("mesh '" + filename + "' not found; geom skipped");
      return false;
    }
    const size_t dot = path.find_last_of('.'), slash = path.find_last_of('/');
    std::string ext = dot == std::string::npos || (slash != std::string::npos && dot < slash) ? "" : path.substr(dot + 1);
    for (char& c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext != "stl" && ext != "obj") {
      warnings.push_back("mesh '" + filename + "': ." + ext + " not supported (use STL/OBJ); geom skipped");
      return false;
    }
    std::vector<uint64_t> bits(3);
    for (int k = 0; k < 3; k++) std::memcpy(&bits[k], &scale[k], 8);
    auto key = std::make_pair(path, bits);
    auto it = mesh_ids.find(key);
    if (it != mesh_ids.end()) {
      *name = it->second;
      return true;
    }
    *name = "urdf_mesh" + std::to_string(meshes.size());
    meshes.push_back({*name, {path, std::vector<double>(scale, scale + 3)}});
    mesh_ids[key] = *name;
    return true;
  }

  // urdf.rs:186-193
  bool material_rgba(const XNode* visual, std::vector<double>* rgba) const {
    const XNode* mat = child(visual, "material");
    if (!mat) return false;
    const XNode* color = child(mat, "color");
    if (color && color->get("rgba")) {
      std::vector<double> v;
      try {
        v = floats(color->get("rgba"));
      } catch (const UrdfError&) {
        return false;
      }
      if (v.size() != 4) return false;
      *rgba = v;
      return true;
    }
    const char* nm = mat->get("name");
    if (!nm) return false;
    auto it = materials.find(nm);
    if (it == materials.end()) return false;
    *rgba = it->second;
    return true;
  }

  // urdf.rs:139-184: a <geom .../> for a visual or collision element, or "" if skipped
  std::string geom(const XNode* el, bool visual, bool hide) {
    const XNode* geometry = child(el, "geometry");
    if (!geometry || geometry->kids.empty()) return "";
    const XNode* shape = geometry->kids[0].get();
    std::string g = "<geom";
    const std::string& t = shape->tag;
    if (t == "box") {
      const double zero[3] = {0, 0, 0};
      double s[3];
      attr3(shape, "size", zero, s);
      g += " type=\"box\" size=\"" + num(s[0] / 2.0) + " " + num(s[1] / 2.0) + " " + num(s[2] / 2.0) + "\"";
    } else if (t == "cylinder" || t == "capsule") {
      const double r = first(shape->get("radius") ? shape->get("radius") : "0");
      const double l = first(shape->get("length") ? shape->get("length") : "0");
      g += " type=\"" + t + "\" size=\"" + num(r) + " " + num(l / 2.0) + "\"";
    } else if (t == "sphere") {
      const double r = first(shape->get("radius") ? shape->get("radius") : "0");
      g += " type=\"sphere\" size=\"" + num(r) + "\"";
    } else if (t == "mesh") {
      const char* file = shape->get("filename");
      const double one[3] = {1, 1, 1};
      double scale[3];
      attr3(shape, "scale", one, scale);
      std::string name;
      if (!mesh(file ? file : "", scale, &name)) return "";
      g += " type=\"mesh\" mesh=\"" + name + "\"";
    } else {
      warnings.push_back("URDF geometry <" + t + "> not supported; skipped");
      return "";
    }
    g += origin_attrs(child(el, "origin"));
    if (visual) {
      g += " contype=\"0\" conaffinity=\"0\" group=\"1\" mass=\"0\"";
      std::vector<double> rgba;
      if (material_rgba(el, &rgba))
        g += " rgba=\"" + num(rgba[0]) + " " + num(rgba[1]) + " " + num(rgba[2]) + " " + num(rgba[3]) + "\"";
    } else if (hide) {
      g += " group=\"3\"";
    }
    return g + "/>";
  }
};

struct Tree {
  std::map<std::string, const XNode*> links;
  std::map<std::string, std::vector<const XNode*>> kids;  // joints by parent link, in document order
};

double attr_or(const XNode* n, const char* name, double def) {
  const char* s = n ? n->get(name) : nullptr;
  return s ? first(s) : def;
}

// urdf.rs:251-311
void emit_link(Converter& cv, const Tree& tr, const std::string& link, std::string& out) {
  const XNode* l = tr.links.at(link);
  if (const XNode* inertial = child(l, "inertial")) {
    const double mass = attr_or(child(inertial, "mass"), "value", 0.0);
    const XNode* i = child(inertial, "inertia");
    const double ixx = attr_or(i, "ixx", 0.0), iyy = attr_or(i, "iyy", 0.0), izz = attr_or(i, "izz", 0.0);
    const double ixy = attr_or(i, "ixy", 0.0), ixz = attr_or(i, "ixz", 0.0), iyz = attr_or(i, "iyz", 0.0);
    if (mass > 0.0)
      out += "<inertial" + origin_attrs(child(inertial, "origin")) + " mass=\"" + num(mass) + "\" fullinertia=\"" +
             num(ixx) + " " + num(iyy) + " " + num(izz) + " " + num(ixy) + " " + num(ixz) + " " + num(iyz) + "\"/>";
  }
  const bool has_visual = !children(l, "visual").empty();
  for (const XNode* c : children(l, "collision")) out += cv.geom(c, false, has_visual);
  for (const XNode* v : children(l, "visual")) out += cv.geom(v, true, false);
  auto it = tr.kids.find(link);
  if (it == tr.kids.end()) return;
  for (const XNode* j : it->second) {
    const char* jn = j->get("name");
    const std::string name = escape(jn ? jn : "joint");
    const std::string ch = child(j, "child")->get("link");
    out += "<body name=\"" + escape(ch) + "\"" + origin_attrs(child(j, "origin")) + ">";
    const double xaxis[3] = {1, 0, 0};
    double axis[3];
    attr3(child(j, "axis"), "xyz", xaxis, axis);
    const double damping = attr_or(child(j, "dynamics"), "damping", 0.0);
    const XNode* limit = child(j, "limit");
    const double lo = attr_or(limit, "lower", 0.0), hi = attr_or(limit, "upper", 0.0);
    const std::string range = lo < hi ? " range=\"" + num(lo) + " " + num(hi) + "\"" : "";
    const std::string common = "name=\"" + name + "\" axis=\"" + num(axis[0]) + " " + num(axis[1]) + " " +
                               num(axis[2]) + "\" damping=\"" + num(damping) + "\"";
    const char* tp = j->get("type");
    const std::string type = tp ? tp : "fixed";
    if (type == "revolute") out += "<joint type=\"hinge\" " + common + range + "/>";
    else if (type == "continuous") out += "<joint type=\"hinge\" " + common + "/>";
    else if (type == "prismatic") out += "<joint type=\"slide\" " + common + range + "/>";
    else if (type == "fixed") {
    } else if (type == "floating") out += "<freejoint name=\"" + name + "\"/>";
    else cv.warnings.push_back("joint type '" + type + "' not supported; treated as fixed");
    emit_link(cv, tr, ch, out);
    out += "</body>";
  }
}

// urdf.rs:205-327
std::string to_mjcf(const std::string& urdf, const std::string& base, const Options& opts,
                    std::vector<std::string>* warnings) {
  XmlParser parser(urdf);
  std::unique_ptr<XNode> robot = parser.parse();
  if (robot->tag != "robot") throw UrdfError{"URDF root element must be <robot>"};
  Converter cv{base, opts, {}, {}, {}, {}};
  for (const XNode* m : children(robot.get(), "material")) {
    const XNode* color = child(m, "color");
    if (m->get("name") && color && color->get("rgba")) {
      std::vector<double> v = floats(color->get("rgba"));
      if (v.size() == 4) cv.materials[m->get("name")] = v;
    }
  }
  Tree tr;
  std::vector<std::string> order;
  for (const XNode* l : children(robot.get(), "link"))
    if (const char* nm = l->get("name")) {
      tr.links[nm] = l;  // a repeated name keeps the last link, as collecting into a map does
      order.push_back(nm);
    }
  std::set<std::string> is_child;
  for (const XNode* j : children(robot.get(), "joint")) {
    const XNode* pn = child(j, "parent");
    const XNode* cn = child(j, "child");
    const char* p = pn ? pn->get("link") : nullptr;
    const char* c = cn ? cn->get("link") : nullptr;
    if (!p || !c) throw UrdfError{"joint missing parent/child"};
    if (!tr.links.count(p) || !tr.links.count(c)) {
      const char* jn = j->get("name");
      throw UrdfError{std::string("joint '") + (jn ? jn : "?") + "' references unknown link"};
    }
    if (!is_child.insert(c).second) throw UrdfError{std::string("link '") + c + "' has two parent joints (URDF must be a tree)"};
    tr.kids[p].push_back(j);
  }
  std::vector<std::string> roots;
  for (const std::string& l : order)
    if (!is_child.count(l)) roots.push_back(l);
  if (roots.size() != 1) {
    std::string list;
    for (const std::string& r : roots) list += (list.empty() ? "" : ", ") + r;
    throw UrdfError{"URDF must have exactly one root link, found [" + list + "]"};
  }
  std::string body = "<body name=\"" + escape(roots[0]) + "\">";
  if (opts.floating_base) body += "<freejoint name=\"root\"/>";
  emit_link(cv, tr, roots[0], body);
  body += "</body>";
  std::string assets;
  for (auto& m : cv.meshes)
    assets += "<mesh name=\"" + m.first + "\" file=\"" + escape(m.second.first) + "\" scale=\"" +
              num(m.second.second[0]) + " " + num(m.second.second[1]) + " " + num(m.second.second[2]) + "\"/>";
  const char* nm = robot->get("name");
  *warnings = cv.warnings;
  return "<mujoco model=\"" + escape(nm ? nm : "robot") +
         "\"><compiler angle=\"radian\" autolimits=\"true\"/><asset>" + assets + "</asset><worldbody>" + body +
         "</worldbody></mujoco>";
}

Options make_options(int32_t floating_base, const char* package_dirs) {
  Options o;
  o.floating_base = floating_base != 0;
  if (package_dirs && *package_dirs) {
    std::string s = package_dirs;
    size_t a = 0;
    for (;;) {
      const size_t b = s.find(':', a);
      o.package_dirs.push_back(s.substr(a, b == std::string::npos ? std::string::npos : b - a));
      if (b == std::string::npos) break;
      a = b + 1;
    }
  }
  return o;
}

}  // namespace

extern "C" int pippin_hip_urdf_to_mjcf(const char* urdf, const char* base, int32_t floating_base,
                                       const char* package_dirs, char* out, int64_t cap, int64_t* len) {
  if (!urdf || !len) {
    ph_set_error("null argument");
    return PH_ERR_ARG;
  }
  std::string mjcf;
  try {
    std::vector<std::string> warnings;
    mjcf = to_mjcf(urdf, base && *base ? base : ".", make_options(floating_base, package_dirs), &warnings);
  } catch (const XmlError& e) {
    ph_set_error("xml parse error: " + e.msg);
    return PH_ERR_MODEL;
  } catch (const UrdfError& e) {
    ph_set_error("invalid model: " + e.msg);
    return PH_ERR_MODEL;
  }
  *len = (int64_t)mjcf.size();
  if (!out || cap < (int64_t)mjcf.size() + 1) {
    ph_set_error("output buffer holds " + std::to_string(cap) + " bytes, the document needs " +
                 std::to_string(mjcf.size() + 1));
    return PH_ERR_SIZE;
  }
  std::memcpy(out, mjcf.c_str(), mjcf.size() + 1);
  ph_set_error("");
  return PH_OK;
}

extern "C" int pippin_hip_model_load_urdf(const char* path, int32_t floating_base, const char* package_dirs,
                                          ph_model** out) {
  if (!path || !out) {
    ph_set_error("null argument");
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
  const size_t slash = dir.find_last_of('/');
  dir = slash == std::string::npos ? "." : (slash == 0 ? "/" : dir.substr(0, slash));
  std::string mjcf;
  try {
    std::vector<std::string> warnings;
    mjcf = to_mjcf(ss.str(), dir, make_options(floating_base, package_dirs), &warnings);
    for (const std::string& w : warnings) std::fprintf(stderr, "pippin urdf warning: %s\n", w.c_str());
  } catch (const XmlError& e) {
    ph_set_error("xml parse error: " + e.msg);
    return PH_ERR_MODEL;
  } catch (const UrdfError& e) {
    ph_set_error("invalid model: " + e.msg);
    return PH_ERR_MODEL;
  }
  return pippin_hip_model_load_xml(mjcf.c_str(), dir.c_str(), out);
}
