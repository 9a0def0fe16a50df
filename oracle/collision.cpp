// ORACLE — test infrastructure only (see math.hpp).
// Primitive narrow phase: crates/pippin/src/collision.rs. Normals point from
// the first geom toward the second; depth > 0 means penetration.
#include <algorithm>

#include "engine.hpp"

namespace orc {

static Hit mkhit(Vec3 pos, Vec3 n, Real depth) {
  Hit h;
  h.pos = pos;
  h.normal = n;
  h.depth = depth;
  h.tangent = v3zero();
  return h;
}

// collision.rs:47-55
void contact_frame(const Contact& c, Vec3* t1, Vec3* t2) {
  Vec3 t = c.tangent - c.normal * dot(c.normal, c.tangent);
  if (norm2(t) > 1e-12) {
    *t1 = normalized(t);
    *t2 = cross(c.normal, *t1);
  } else {
    tangents(c.normal, t1, t2);
  }
}

bool pair_supported(int a, int b) { return !(a == Plane && b == Plane); }
// collision.rs:104-107
bool needs_general(int a, int b) {
  return a == Mesh || b == Mesh || (a == Cylinder && b != Plane) || (b == Cylinder && a != Plane);
}

static void segment(const GeomPose& g, Vec3* p0, Vec3* p1) {
  Vec3 ax = col(g.mat, 2) * g.size[1];
  *p0 = g.pos - ax;
  *p1 = g.pos + ax;
}

// collision.rs:154-162
static Vec3 closest_on_segment(Vec3 a, Vec3 b, Vec3 p) {
  Vec3 d = b - a;
  Real l2 = norm2(d);
  if (l2 < 1e-18) return a;
  Real t = rclamp(dot(p - a, d) / l2, 0.0, 1.0);
  return a + d * t;
}

// collision.rs:165-201
static void closest_segments(Vec3 p0, Vec3 p1, Vec3 q0, Vec3 q1, Vec3* x, Vec3* y) {
  Vec3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
  Real a = norm2(d1), e = norm2(d2), f = dot(d2, r);
  Real s, t;
  if (a < 1e-18 && e < 1e-18) {
    *x = p0;
    *y = q0;
    return;
  }
  if (a < 1e-18) {
    s = 0.0;
    t = rclamp(f / e, 0.0, 1.0);
  } else {
    Real c = dot(d1, r);
    if (e < 1e-18) {
      t = 0.0;
      s = rclamp(-c / a, 0.0, 1.0);
    } else {
      Real b = dot(d1, d2);
      Real denom = a * e - b * b;
      Real s0 = denom > 1e-18 ? rclamp((b * f - c * e) / denom, 0.0, 1.0) : 0.0;
      Real t0 = (b * s0 + f) / e;
      if (t0 < 0.0) {
        t0 = 0.0;
        s0 = rclamp(-c / a, 0.0, 1.0);
      } else if (t0 > 1.0) {
        t0 = 1.0;
        s0 = rclamp((b - c) / a, 0.0, 1.0);
      }
      s = s0;
      t = t0;
    }
  }
  *x = p0 + d1 * s;
  *y = q0 + d2 * t;
}

// collision.rs:203-212
static void plane_sphere(const GeomPose& plane, Vec3 c, Real r, Real margin, std::vector<Hit>& out) {
  Vec3 n = col(plane.mat, 2);
  Real dist = dot(c - plane.pos, n);
  Real depth = r - dist;
  if (depth > -margin) {
    Vec3 pos = c - n * (dist + r) * 0.5;
    out.push_back(mkhit(pos, n, depth));
  }
}

// Stable sort by distance (Rust sort_by is stable), then the first four.
static void emit_sorted(std::vector<std::pair<Real, Vec3>>& hits, Vec3 n, std::vector<Hit>& out) {
  std::stable_sort(hits.begin(), hits.end(),
                   [](const std::pair<Real, Vec3>& a, const std::pair<Real, Vec3>& b) { return a.first < b.first; });
  for (size_t i = 0; i < hits.size() && i < 4; i++) {
    Real dist = hits[i].first;
    out.push_back(mkhit(hits[i].second - n * (dist * 0.5), n, -dist));
  }
}

// collision.rs:214-236
static void plane_box(const GeomPose& plane, const GeomPose& b, Real margin, std::vector<Hit>& out) {
  Vec3 n = col(plane.mat, 2);
  std::vector<std::pair<Real, Vec3>> hits;
  for (int i = 0; i < 8; i++) {
    Vec3 s = V3((i & 1) == 0 ? -1.0 : 1.0, (i & 2) == 0 ? -1.0 : 1.0, (i & 4) == 0 ? -1.0 : 1.0);
    Vec3 corner = b.pos + mul_vec(b.mat, mul_elem(b.size, s));
    Real dist = dot(corner - plane.pos, n);
    if (dist < margin) hits.push_back({dist, corner});
  }
  emit_sorted(hits, n, out);
}

// collision.rs:238-267
static void plane_cylinder(const GeomPose& plane, const GeomPose& c, Real margin, std::vector<Hit>& out) {
  Vec3 n = col(plane.mat, 2);
  Vec3 axis = col(c.mat, 2);
  Real r = c.size[0], h = c.size[1];
  Vec3 radial = n - axis * dot(axis, n);
  Real rl = norm(radial);
  std::vector<Vec3> pts;
  const Real caps[2] = {-1.0, 1.0};
  for (Real cap : caps) {
    Vec3 cc = c.pos + axis * (h * cap);
    if (rl > 0.05) {
      Vec3 d = radial * (-1.0 / rl);
      pts.push_back(cc + d * r);
    } else {
      Vec3 t1, t2;
      tangents(axis, &t1, &t2);
      Vec3 vs[4] = {t1, -t1, t2, -t2};
      for (Vec3 v : vs) pts.push_back(cc + v * r);
    }
  }
  std::vector<std::pair<Real, Vec3>> hits;
  for (Vec3 p : pts) {
    Real dd = dot(p - plane.pos, n);
    if (dd < margin) hits.push_back({dd, p});
  }
  emit_sorted(hits, n, out);
}

// collision.rs:269-279
static void sphere_sphere(Vec3 c1, Real r1, Vec3 c2, Real r2, Real margin, std::vector<Hit>& out) {
  Vec3 d = c2 - c1;
  Real dist = norm(d);
  Real depth = r1 + r2 - dist;
  if (depth <= -margin) return;
  Vec3 n = dist > 1e-12 ? d * (1.0 / dist) : V3(0, 0, 1);
  Vec3 pos = c1 + n * (r1 - depth * 0.5);
  out.push_back(mkhit(pos, n, depth));
}

// collision.rs:282-316
static void sphere_box(Vec3 c, Real r, const GeomPose& b, Real margin, std::vector<Hit>& out) {
  Vec3 local = tmul_vec(b.mat, c - b.pos);
  Vec3 h = b.size;
  Vec3 clamped = V3(rclamp(local[0], -h[0], h[0]), rclamp(local[1], -h[1], h[1]), rclamp(local[2], -h[2], h[2]));
  Vec3 diff = local - clamped;
  Real dist = norm(diff);
  if (dist > 1e-12) {
    Real depth = r - dist;
    if (depth <= -margin) return;
    Vec3 n_local = diff * (-1.0 / dist);
    Vec3 n = mul_vec(b.mat, n_local);
    Vec3 surf = b.pos + mul_vec(b.mat, clamped);
    out.push_back(mkhit(surf - n * (depth * 0.5), n, depth));
  } else {
    Real bd = INFINITY;
    int bi = 0;
    Real bs = 1.0;
    for (int i = 0; i < 3; i++) {
      const Real ss[2] = {-1.0, 1.0};
      for (Real s : ss) {
        Real dd = h[i] - s * local[i];
        if (dd < bd) {
          bd = dd;
          bi = i;
          bs = s;
        }
      }
    }
    Vec3 nl = v3zero();
    nl[bi] = -bs;
    Vec3 n = mul_vec(b.mat, nl);
    Real depth = r + bd;
    out.push_back(mkhit(c + n * (r - depth * 0.5), n, depth));
  }
}

// collision.rs:318-335
static void capsule_capsule(const GeomPose& a, const GeomPose& b, Real margin, std::vector<Hit>& out) {
  Vec3 p0, p1, q0, q1;
  segment(a, &p0, &p1);
  segment(b, &q0, &q1);
  Real ra = a.size[0], rb = b.size[0];
  Vec3 x, y;
  closest_segments(p0, p1, q0, q1, &x, &y);
  sphere_sphere(x, ra, y, rb, margin, out);
  Vec3 da = normalized(p1 - p0);
  Vec3 db = normalized(q1 - q0);
  if (norm(cross(da, db)) < 0.05) {
    Vec3 ends[2] = {p0, p1};
    for (Vec3 p : ends) {
      Vec3 q = closest_on_segment(q0, q1, p);
      if (norm(q - y) > 0.25 * (ra + rb) || norm(p - x) > 0.25 * (ra + rb)) sphere_sphere(p, ra, q, rb, margin, out);
    }
  }
}

// collision.rs:338-344
static Real box_sdf(const GeomPose& b, Vec3 p) {
  Vec3 l = tmul_vec(b.mat, p - b.pos);
  Vec3 q = vabs(l) - b.size;
  Real outside = norm(vmax(q, v3zero()));
  Real inside = rmin(rmax(rmax(q[0], q[1]), q[2]), 0.0);
  return outside + inside;
}

// collision.rs:346-380
static void capsule_box(const GeomPose& a, const GeomPose& b, Real margin, std::vector<Hit>& out) {
  Vec3 p0, p1;
  segment(a, &p0, &p1);
  Real r = a.size[0];
  auto f = [&](Real t) { return box_sdf(b, p0 + (p1 - p0) * t); };
  const Real g = 0.618033988750;
  Real lo = 0.0, hi = 1.0;
  Real x1 = hi - g * (hi - lo), x2 = lo + g * (hi - lo);
  Real f1 = f(x1), f2 = f(x2);
  for (int it = 0; it < 30; it++) {
    if (f1 < f2) {
      hi = x2;
      x2 = x1;
      f2 = f1;
      x1 = hi - g * (hi - lo);
      f1 = f(x1);
    } else {
      lo = x1;
      x1 = x2;
      f1 = f2;
      x2 = lo + g * (hi - lo);
      f2 = f(x2);
    }
  }
  Real tmin = 0.5 * (lo + hi);
  std::vector<Real> ts = {tmin};
  const Real ends[2] = {0.0, 1.0};
  for (Real t : ends)
    if (r_fabs(t - tmin) > 0.1 && f(t) < r + margin) ts.push_back(t);
  for (Real t : ts) sphere_box(p0 + (p1 - p0) * t, r, b, margin, out);
}

// collision.rs:496-511
static std::vector<Vec3> clip(const std::vector<Vec3>& poly, Vec3 pn, Real off) {
  std::vector<Vec3> out;
  size_t n = poly.size();
  for (size_t i = 0; i < n; i++) {
    Vec3 a = poly[i], b = poly[(i + 1) % n];
    Real da = dot(pn, a) - off, db = dot(pn, b) - off;
    if (da <= 0.0) out.push_back(a);
    if ((da < 0.0) != (db < 0.0) && r_fabs(da - db) > 1e-15) out.push_back(a + (b - a) * (da / (da - db)));
  }
  return out;
}

// collision.rs:514-531. Rust's min_by keeps the first minimum, max_by the last maximum.
void reduce_to_four(std::vector<std::pair<Real, Vec3>>& pts) {
  if (pts.size() <= 4) return;
  std::vector<std::pair<Real, Vec3>> keep;
  size_t di = 0;
  for (size_t i = 1; i < pts.size(); i++)
    if (pts[i].first < pts[di].first) di = i;
  keep.push_back(pts[di]);
  auto argmax_last = [&](auto key) {
    size_t bi = 0;
    Real bk = key(pts[0].second);
    for (size_t i = 1; i < pts.size(); i++) {
      Real k = key(pts[i].second);
      if (!(k < bk)) {  // ties move to the later element
        bk = k;
        bi = i;
      }
    }
    return bi;
  };
  Vec3 deep = keep[0].second;
  keep.push_back(pts[argmax_last([&](Vec3 p) { return norm2(p - deep); })]);
  auto area = [&](Vec3 p) { return norm(cross(keep[1].second - keep[0].second, p - keep[0].second)); };
  keep.push_back(pts[argmax_last(area)]);
  auto spread = [&](Vec3 p) {
    Real s = INFINITY;
    for (const auto& k : keep) s = rmin(s, norm(k.second - p));
    return s;
  };
  keep.push_back(pts[argmax_last(spread)]);
  pts = keep;
}

// collision.rs:383-493
static void box_box(const GeomPose& a, const GeomPose& b, Real margin, std::vector<Hit>& out) {
  Vec3 d = b.pos - a.pos;
  Vec3 ax[3] = {col(a.mat, 0), col(a.mat, 1), col(a.mat, 2)};
  Vec3 bx[3] = {col(b.mat, 0), col(b.mat, 1), col(b.mat, 2)};
  auto proj = [](const Vec3* axes, Vec3 h, Vec3 l) {
    return h[0] * r_fabs(dot(axes[0], l)) + h[1] * r_fabs(dot(axes[1], l)) + h[2] * r_fabs(dot(axes[2], l));
  };
  bool has_best = false, has_face = false;
  Real best_o = 0, face_o = 0;
  Vec3 best_l = v3zero(), face_l = v3zero();
  int best_k = 0, face_k = 0;
  for (int k = 0; k < 15; k++) {
    Vec3 l;
    if (k < 3) l = ax[k];
    else if (k < 6) l = bx[k - 3];
    else {
      Vec3 c = cross(ax[(k - 6) / 3], bx[(k - 6) % 3]);
      Real n = norm(c);
      if (n < 1e-6) continue;
      l = c * (1.0 / n);
    }
    Real overlap = proj(ax, a.size, l) + proj(bx, b.size, l) - r_fabs(dot(d, l));
    if (overlap < -margin) return;
    if (dot(d, l) < 0.0) l = -l;
    if (!has_best || overlap < best_o) {
      has_best = true;
      best_o = overlap;
      best_l = l;
      best_k = k;
    }
    if (k < 6 && (!has_face || overlap < face_o)) {
      has_face = true;
      face_o = overlap;
      face_l = l;
      face_k = k;
    }
  }
  if (!has_best || !has_face) return;
  if (best_k >= 6 && face_o < best_o * 1.05 + 1e-4) {
    best_o = face_o;
    best_l = face_l;
    best_k = face_k;
  }
  Real overlap = best_o;
  Vec3 n = best_l;
  int k = best_k;
  if (k >= 6) {
    int i = (k - 6) / 3, j = (k - 6) % 3;
    Vec3 pa = a.pos;
    for (int mm = 0; mm < 3; mm++)
      if (mm != i) pa += ax[mm] * (a.size[mm] * rsignum(dot(ax[mm], n)));
    Vec3 pb = b.pos;
    for (int mm = 0; mm < 3; mm++)
      if (mm != j) pb -= bx[mm] * (b.size[mm] * rsignum(dot(bx[mm], n)));
    Vec3 x, y;
    closest_segments(pa - ax[i] * a.size[i], pa + ax[i] * a.size[i], pb - bx[j] * b.size[j], pb + bx[j] * b.size[j],
                     &x, &y);
    out.push_back(mkhit((x + y) * 0.5, n, overlap));
    return;
  }
  const GeomPose* rf = k < 3 ? &a : &b;
  const GeomPose* inc = k < 3 ? &b : &a;
  const Vec3* raxes = k < 3 ? ax : bx;
  const Vec3* iaxes = k < 3 ? bx : ax;
  Vec3 nref = k < 3 ? n : -n;
  int fi = k < 3 ? k : k - 3;
  Vec3 ref_center = rf->pos + nref * rf->size[fi];
  int u = (fi + 1) % 3, v = (fi + 2) % 3;
  int j = 0;
  Real bestdot = INFINITY;
  for (int mm = 0; mm < 3; mm++) {
    Real dt = dot(iaxes[mm], nref);
    if (-r_fabs(dt) < bestdot) {
      bestdot = -r_fabs(dt);
      j = mm;
    }
  }
  Real s = -rsignum(dot(iaxes[j], nref));
  Vec3 ic = inc->pos + iaxes[j] * (s * inc->size[j]);
  int ku = (j + 1) % 3, kv = (j + 2) % 3;
  Vec3 eu = iaxes[ku] * inc->size[ku];
  Vec3 ev = iaxes[kv] * inc->size[kv];
  std::vector<Vec3> poly = {ic + eu + ev, ic - eu + ev, ic - eu - ev, ic + eu - ev};
  Vec3 caxes[2] = {raxes[u], raxes[v]};
  Real cext[2] = {rf->size[u], rf->size[v]};
  for (int c = 0; c < 2; c++) {
    const Real signs[2] = {1.0, -1.0};
    for (Real sign : signs) {
      Vec3 pn = caxes[c] * sign;
      Real off = dot(pn, ref_center) + cext[c];
      poly = clip(poly, pn, off);
      if (poly.empty()) return;
    }
  }
  std::vector<std::pair<Real, Vec3>> pts;
  for (Vec3 p : poly) {
    Real sep = dot(p - ref_center, nref);
    if (sep < margin) pts.push_back({sep, p});
  }
  reduce_to_four(pts);
  for (const auto& sp : pts) out.push_back(mkhit(sp.second - nref * (sp.first * 0.5), n, -sp.first));
}

// collision.rs:59-96
void collide_pair(const GeomPose& a, const GeomPose& b, Real margin, std::vector<Hit>& out) {
  if (a.typ > b.typ) {
    size_t start = out.size();
    collide_pair(b, a, margin, out);
    for (size_t i = start; i < out.size(); i++) out[i].normal = -out[i].normal;
    return;
  }
  if (a.typ == Plane && b.typ == Sphere) plane_sphere(a, b.pos, b.size[0], margin, out);
  else if (a.typ == Plane && b.typ == Capsule) {
    Vec3 p0, p1;
    segment(b, &p0, &p1);
    size_t start = out.size();
    plane_sphere(a, p0, b.size[0], margin, out);
    plane_sphere(a, p1, b.size[0], margin, out);
    for (size_t i = start; i < out.size(); i++) out[i].tangent = col(b.mat, 2);
  } else if (a.typ == Plane && b.typ == Box) plane_box(a, b, margin, out);
  else if (a.typ == Plane && b.typ == Cylinder) plane_cylinder(a, b, margin, out);
  else if (a.typ == Sphere && b.typ == Sphere) sphere_sphere(a.pos, a.size[0], b.pos, b.size[0], margin, out);
  else if (a.typ == Sphere && b.typ == Capsule) {
    Vec3 p0, p1;
    segment(b, &p0, &p1);
    Vec3 q = closest_on_segment(p0, p1, a.pos);
    sphere_sphere(a.pos, a.size[0], q, b.size[0], margin, out);
  } else if (a.typ == Sphere && b.typ == Box) sphere_box(a.pos, a.size[0], b, margin, out);
  else if (a.typ == Capsule && b.typ == Capsule) capsule_capsule(a, b, margin, out);
  else if (a.typ == Capsule && b.typ == Box) capsule_box(a, b, margin, out);
  else if (a.typ == Box && b.typ == Box) box_box(a, b, margin, out);
}

}  // namespace orc
