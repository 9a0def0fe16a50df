// Mesh assets for the library's MJCF loader: vertex files, convex hulls and
// solid mass properties (crates/pippin/src/mesh.rs). Meshes collide through
// their convex hull and keep their file coordinates: the geom frame is the mesh
// frame (mesh.rs:1-6).
#include "meshes.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <sstream>

namespace phh {

namespace {

using V3 = std::array<double, 3>;

V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 cross(const V3& a, const V3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Strict convex hull (monotone chain) of 2-D points, counter-clockwise,
// collinear points dropped; returns indices into `ids`.
std::vector<int> hull2d(const std::vector<std::array<double, 2>>& p, const std::vector<int>& ids) {
  std::vector<int> order(p.size());
  for (size_t k = 0; k < p.size(); k++) order[k] = (int)k;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return p[a][0] < p[b][0] || (p[a][0] == p[b][0] && p[a][1] < p[b][1]);
  });
  auto turn = [&](int o, int a, int b) {
    return (p[a][0] - p[o][0]) * (p[b][1] - p[o][1]) - (p[a][1] - p[o][1]) * (p[b][0] - p[o][0]);
  };
  std::vector<int> lo, up;
  for (int k : order) {
    while (lo.size() >= 2 && turn(lo[lo.size() - 2], lo.back(), k) <= 0) lo.pop_back();
    lo.push_back(k);
  }
  for (auto it = order.rbegin(); it != order.rend(); ++it) {
    while (up.size() >= 2 && turn(up[up.size() - 2], up.back(), *it) <= 0) up.pop_back();
    up.push_back(*it);
  }
  std::vector<int> out;
  for (size_t k = 0; k + 1 < lo.size(); k++) out.push_back(ids[lo[k]]);
  for (size_t k = 0; k + 1 < up.size(); k++) out.push_back(ids[up[k]]);
  return out;
}

}  // namespace

bool read_obj_vertices(const std::string& text, const double* scale, std::vector<double>* out) {
  std::istringstream in(text);
  std::string line;
  out->clear();
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string tag;
    if (!(ls >> tag) || tag != "v") continue;
    double c[3];
    for (int k = 0; k < 3; k++) {
      std::string tok;
      if (!(ls >> tok)) return false;
      // OBJ coordinates arrive as f32 (the reference reads them with tobj)
      c[k] = (double)std::strtof(tok.c_str(), nullptr);
    }
    for (int k = 0; k < 3; k++) out->push_back(c[k] * scale[k]);
  }
  return true;
}

bool read_stl_vertices(const std::string& bytes, const double* scale, std::vector<double>* out) {
  out->clear();
  auto push = [&](const float* v) {
    for (int k = 0; k < 3; k++) out->push_back((double)v[k] * scale[k]);
  };
  if (bytes.size() >= 84) {
    uint32_t n = 0;
    std::memcpy(&n, bytes.data() + 80, 4);
    if (bytes.size() == 84 + (size_t)n * 50) {  // binary: 50 bytes per triangle
      for (uint32_t t = 0; t < n; t++) {
        const char* rec = bytes.data() + 84 + (size_t)t * 50;
        for (int v = 0; v < 3; v++) {
          float f[3];
          std::memcpy(f, rec + 12 + 12 * v, 12);
          push(f);
        }
      }
      return true;
    }
  }
  std::istringstream in(bytes);  // ASCII: "vertex x y z" lines
  std::string tok;
  bool any = false;
  while (in >> tok) {
    if (tok != "vertex") continue;
    float f[3];
    for (int k = 0; k < 3; k++) {
      std::string s;
      if (!(in >> s)) return false;
      f[k] = std::strtof(s.c_str(), nullptr);
    }
    push(f);
    any = true;
  }
  return any;
}

bool convex_hull(const std::vector<double>& pts, Hull* out, std::string* err) {
  std::vector<V3> P;
  for (size_t i = 0; 3 * i + 2 < pts.size(); i++) {
    V3 p = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    if (std::find(P.begin(), P.end(), p) == P.end()) P.push_back(p);
  }
  const int n = (int)P.size();
  if (n > kMaxHullPoints) {
    *err = "meshes with more than " + std::to_string(kMaxHullPoints) + " distinct points are not supported";
    return false;
  }
  double scale = 0.0;
  for (const V3& p : P)
    for (double c : p) scale = std::max(scale, std::fabs(c));
  scale = std::max(scale, 1e-300);
  const double tol = 1e-10 * scale;
  // every plane through three points with all points on one side; keyed by
  // the set of points lying on it, first normal found kept
  std::vector<std::vector<int>> on_sets;
  std::vector<V3> normals;
  std::map<std::vector<int>, int> seen;
  for (int i = 0; i < n; i++)
    for (int j = i + 1; j < n; j++)
      for (int k = j + 1; k < n; k++) {
        V3 nrm = cross(sub(P[j], P[i]), sub(P[k], P[i]));
        const double ln = std::sqrt(dot(nrm, nrm));
        if (ln <= tol * scale) continue;
        for (double& c : nrm) c /= ln;
        std::vector<double> d(n);
        bool all_le = true, all_ge = true;
        for (int t = 0; t < n; t++) {
          d[t] = dot(sub(P[t], P[i]), nrm);
          all_le = all_le && d[t] <= tol;
          all_ge = all_ge && d[t] >= -tol;
        }
        if (!all_le) {
          if (!all_ge) continue;
          for (double& c : nrm) c = -c;
          for (double& x : d) x = -x;
        }
        std::vector<int> on;
        for (int t = 0; t < n; t++)
          if (std::fabs(d[t]) <= tol) on.push_back(t);
        if (seen.count(on)) continue;
        seen[on] = (int)on_sets.size();
        on_sets.push_back(on);
        normals.push_back(nrm);
      }
  bool flat = on_sets.empty();  // no plane, or every point on one plane
  for (const auto& on : on_sets) flat = flat || (int)on.size() == n;
  if (flat) {
    *err = "the mesh is degenerate (no volume)";
    return false;
  }
  std::vector<std::vector<int>> polys;
  for (size_t f = 0; f < on_sets.size(); f++) {
    const V3& nrm = normals[f];
    const V3 a = std::fabs(nrm[0]) < 0.57 ? V3{1.0, 0.0, 0.0} : V3{0.0, 1.0, 0.0};
    V3 u = cross(nrm, a);
    const double lu = std::sqrt(dot(u, u));
    for (double& c : u) c /= lu;
    const V3 v = cross(nrm, u);
    std::vector<std::array<double, 2>> p2;
    for (int t : on_sets[f]) p2.push_back({dot(P[t], u), dot(P[t], v)});
    std::vector<int> poly = hull2d(p2, on_sets[f]);
    if (poly.size() >= 3) polys.push_back(poly);
  }
  std::vector<int> used;
  for (const auto& poly : polys)
    for (int t : poly)
      if (std::find(used.begin(), used.end(), t) == used.end()) used.push_back(t);
  std::sort(used.begin(), used.end(), [&](int a, int b) { return P[a] < P[b]; });
  std::vector<int> remap(n, -1);
  for (size_t r = 0; r < used.size(); r++) remap[used[r]] = (int)r;
  out->vert.clear();
  for (int t : used)
    for (double c : P[t]) out->vert.push_back(c);
  out->faces.clear();
  for (const auto& poly : polys) {
    std::vector<int> f;
    for (int t : poly) f.push_back(remap[t]);
    const auto m = std::min_element(f.begin(), f.end()) - f.begin();
    std::rotate(f.begin(), f.begin() + m, f.end());
    out->faces.push_back(f);
  }
  std::sort(out->faces.begin(), out->faces.end());
  return true;
}

// mesh.rs:72-105 over the fan triangles (f0, fi, fi+1) of each face: signed
// tetrahedra from the origin, sums in triangle order, no fused multiply-adds.
void mass_properties(const Hull& h, double* vol_out, double* com, double* icom) {
  auto V = [&](int i) { return V3{h.vert[3 * i], h.vert[3 * i + 1], h.vert[3 * i + 2]}; };
  double vol = 0.0;
  double first[3] = {0.0, 0.0, 0.0};
  double c[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (const auto& f : h.faces)
    for (size_t i = 1; i + 1 < f.size(); i++) {
      const V3 a = V(f[0]), b = V(f[i]), d = V(f[i + 1]);
      const double v6 = dot(a, cross(b, d));
      vol += v6 / 6.0;
      const double fq = v6 / 24.0;
      for (int k = 0; k < 3; k++) first[k] = first[k] + ((a[k] + b[k]) + d[k]) * fq;
      for (int r = 0; r < 3; r++)
        for (int s = 0; s < 3; s++) {
          const double ss = a[r] * a[s] + b[r] * b[s] + d[r] * d[s];
          const double pp = (a[r] + b[r] + d[r]) * (a[s] + b[s] + d[s]);
          c[r][s] += v6 / 120.0 * (ss + pp);
        }
    }
  *vol_out = vol;
  if (std::fabs(vol) < 1e-18) {
    for (int k = 0; k < 3; k++) com[k] = 0.0;
    for (int k = 0; k < 9; k++) icom[k] = 0.0;
    *vol_out = 0.0;
    return;
  }
  const double inv = 1.0 / vol;
  for (int k = 0; k < 3; k++) com[k] = first[k] * inv;
  const double tr = c[0][0] + c[1][1] + c[2][2];
  double io[9];
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) io[3 * r + s] = (r == s ? tr : 0.0) - c[r][s];
  // parallel axis to the centre of mass: I_com = I_o + vol * skew(com) skew(com)^T * -1
  const double cx[9] = {0.0, -com[2], com[1], com[2], 0.0, -com[0], -com[1], com[0], 0.0};
  const double cxt[9] = {cx[0], cx[3], cx[6], cx[1], cx[4], cx[7], cx[2], cx[5], cx[8]};
  for (int r = 0; r < 3; r++)
    for (int s = 0; s < 3; s++) {
      double acc = -0.0;
      for (int k = 0; k < 3; k++) acc = acc + cx[3 * r + k] * cxt[3 * k + s];
      icom[3 * r + s] = io[3 * r + s] + acc * -vol;
    }
}

}  // namespace phh
