// Primitive narrow phase for the device (one lane per geom pair).
// Follows crates/pippin/src/collision.rs routine by routine; normals point from
// the first geom to the second and depth > 0 means penetration. Hits go to a
// fixed sink of MAXHIT records (no pair can produce more).
#pragma once
#include "dmodel.hpp"

namespace phk {

// A mesh geom's convex hull as the general narrow phase reads it.
struct HullRef {
  gp<const Real> vert;          // this hull's vertices (geom frame), 3 per vertex
  gp<const int> vfadr, vfnum;   // per vertex of this hull: its faces in vface
  gp<const int> vface;          // face ids (all hulls), increasing per vertex
  gp<const Real> fnormal;       // unit face normals (all hulls), 3 per face
  gp<const int> fvadr, fvnum, fvert;  // face vertex lists, ids within the face's hull
  int nvert;
};

struct GPose {
  int typ;
  V3 pos;
  M3 mat;
  V3 size;
  HullRef hull;  // mesh geoms only
};

struct Sink {
  gp<Real> out;  // MAXHIT x CW, in global memory
  int n;
  PHD void push(V3 pos, V3 nrm, Real depth) {
    gp<Real> h = out + n * CW;
    st3(h, pos);
    st3(h + 3, nrm);
    st3(h + 6, v3z());
    h[9] = depth;
    n++;
  }
};

PHD void segment(const GPose& g, V3* p0, V3* p1) {
  V3 ax = col(g.mat, 2) * g.size.y;
  *p0 = g.pos - ax;
  *p1 = g.pos + ax;
}

// collision.rs:154-162
PHD V3 closest_on_segment(V3 a, V3 b, V3 p) {
  V3 d = b - a;
  Real l2 = norm2(d);
  if (l2 < 1e-18) return a;
  Real t = rclamp(dot(p - a, d) / l2, 0.0, 1.0);
  return a + d * t;
}

// collision.rs:165-201
PHD void closest_segments(V3 p0, V3 p1, V3 q0, V3 q1, V3* x, V3* y) {
  V3 d1 = p1 - p0, d2 = q1 - q0, r = p0 - q0;
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
PHD void plane_sphere(const GPose& plane, V3 c, Real r, Real margin, Sink& out) {
  V3 n = col(plane.mat, 2);
  Real dist = dot(c - plane.pos, n);
  Real depth = r - dist;
  if (depth > -margin) out.push(c - n * (dist + r) * 0.5, n, depth);
}

// Stable insertion sort of up to 8 (dist, point) by dist, then the first four
// (Rust's sort_by is stable), collision.rs:231-235 and 263-266.
PHD void emit_sorted(Real* dist, V3* pts, int k, V3 n, Sink& out) {
  for (int i = 1; i < k; i++) {
    Real d = dist[i];
    V3 p = pts[i];
    int j = i - 1;
    while (j >= 0 && dist[j] > d) {
      dist[j + 1] = dist[j];
      pts[j + 1] = pts[j];
      j--;
    }
    dist[j + 1] = d;
    pts[j + 1] = p;
  }
  for (int i = 0; i < k && i < 4; i++) out.push(pts[i] - n * (dist[i] * 0.5), n, -dist[i]);
}

// collision.rs:214-236
PHD void plane_box(const GPose& plane, const GPose& b, Real margin, Sink& out) {
  V3 n = col(plane.mat, 2);
  Real dist[8];
  V3 pts[8];
  int k = 0;
  for (int i = 0; i < 8; i++) {
    V3 s = v3((i & 1) == 0 ? -1.0 : 1.0, (i & 2) == 0 ? -1.0 : 1.0, (i & 4) == 0 ? -1.0 : 1.0);
    V3 corner = b.pos + mul_vec(b.mat, mul_elem(b.size, s));
    Real d = dot(corner - plane.pos, n);
    if (d < margin) {
      dist[k] = d;
      pts[k] = corner;
      k++;
    }
  }
  emit_sorted(dist, pts, k, n, out);
}

// collision.rs:238-267
PHD void plane_cylinder(const GPose& plane, const GPose& c, Real margin, Sink& out) {
  V3 n = col(plane.mat, 2);
  V3 axis = col(c.mat, 2);
  Real r = c.size.x, h = c.size.y;
  V3 radial = n - axis * dot(axis, n);
  Real rl = norm(radial);
  V3 cand[8];
  int np = 0;
  for (int ci = 0; ci < 2; ci++) {
    Real cap = ci == 0 ? -1.0 : 1.0;
    V3 cc = c.pos + axis * (h * cap);
    if (rl > 0.05) {
      V3 d = radial * (-1.0 / rl);
      cand[np++] = cc + d * r;
    } else {
      V3 t1, t2;
      tangents(axis, &t1, &t2);
      cand[np++] = cc + t1 * r;
      cand[np++] = cc + (-t1) * r;
      cand[np++] = cc + t2 * r;
      cand[np++] = cc + (-t2) * r;
    }
  }
  Real dist[8];
  V3 pts[8];
  int k = 0;
  for (int i = 0; i < np; i++) {
    Real d = dot(cand[i] - plane.pos, n);
    if (d < margin) {
      dist[k] = d;
      pts[k] = cand[i];
      k++;
    }
  }
  emit_sorted(dist, pts, k, n, out);
}

// collision.rs:269-279
PHD void sphere_sphere(V3 c1, Real r1, V3 c2, Real r2, Real margin, Sink& out) {
  V3 d = c2 - c1;
  Real dist = norm(d);
  Real depth = r1 + r2 - dist;
  if (depth <= -margin) return;
  V3 n = dist > 1e-12 ? d * (1.0 / dist) : v3(0, 0, 1);
  out.push(c1 + n * (r1 - depth * 0.5), n, depth);
}

// collision.rs:282-316
PHD void sphere_box(V3 c, Real r, const GPose& b, Real margin, Sink& out) {
  V3 local = tmul_vec(b.mat, c - b.pos);
  V3 h = b.size;
  V3 clamped = v3(rclamp(local.x, -h.x, h.x), rclamp(local.y, -h.y, h.y), rclamp(local.z, -h.z, h.z));
  V3 diff = local - clamped;
  Real dist = norm(diff);
  if (dist > 1e-12) {
    Real depth = r - dist;
    if (depth <= -margin) return;
    V3 n_local = diff * (-1.0 / dist);
    V3 n = mul_vec(b.mat, n_local);
    V3 surf = b.pos + mul_vec(b.mat, clamped);
    out.push(surf - n * (depth * 0.5), n, depth);
  } else {
    Real bd = INFINITY, bs = 1.0;
    int bi = 0;
    for (int i = 0; i < 3; i++) {
      for (int si = 0; si < 2; si++) {
        Real s = si == 0 ? -1.0 : 1.0;
        Real d = h[i] - s * local[i];
        if (d < bd) {
          bd = d;
          bi = i;
          bs = s;
        }
      }
    }
    V3 nl = v3(bi == 0 ? -bs : 0.0, bi == 1 ? -bs : 0.0, bi == 2 ? -bs : 0.0);
    V3 n = mul_vec(b.mat, nl);
    Real depth = r + bd;
    out.push(c + n * (r - depth * 0.5), n, depth);
  }
}

// collision.rs:318-335
PHD void capsule_capsule(const GPose& a, const GPose& b, Real margin, Sink& out) {
  V3 p0, p1, q0, q1;
  segment(a, &p0, &p1);
  segment(b, &q0, &q1);
  Real ra = a.size.x, rb = b.size.x;
  V3 x, y;
  closest_segments(p0, p1, q0, q1, &x, &y);
  sphere_sphere(x, ra, y, rb, margin, out);
  V3 da = normalized(p1 - p0);
  V3 db = normalized(q1 - q0);
  if (norm(cross(da, db)) < 0.05) {
    for (int e = 0; e < 2; e++) {
      V3 p = e == 0 ? p0 : p1;
      V3 q = closest_on_segment(q0, q1, p);
      if (norm(q - y) > 0.25 * (ra + rb) || norm(p - x) > 0.25 * (ra + rb)) sphere_sphere(p, ra, q, rb, margin, out);
    }
  }
}

// collision.rs:338-344
PHD Real box_sdf(const GPose& b, V3 p) {
  V3 l = tmul_vec(b.mat, p - b.pos);
  V3 q = vabs(l) - b.size;
  Real outside = norm(vmax(q, v3z()));
  Real inside = rmin(rmax(rmax(q.x, q.y), q.z), 0.0);
  return outside + inside;
}

// collision.rs:346-380
PHD void capsule_box(const GPose& a, const GPose& b, Real margin, Sink& out) {
  V3 p0, p1;
  segment(a, &p0, &p1);
  Real r = a.size.x;
  V3 d = p1 - p0;
  const Real g = 0.618033988750;
  Real lo = 0.0, hi = 1.0;
  Real x1 = hi - g * (hi - lo), x2 = lo + g * (hi - lo);
  Real f1 = box_sdf(b, p0 + d * x1), f2 = box_sdf(b, p0 + d * x2);
  for (int it = 0; it < 30; it++) {
    if (f1 < f2) {
      hi = x2;
      x2 = x1;
      f2 = f1;
      x1 = hi - g * (hi - lo);
      f1 = box_sdf(b, p0 + d * x1);
    } else {
      lo = x1;
      x1 = x2;
      f1 = f2;
      x2 = lo + g * (hi - lo);
      f2 = box_sdf(b, p0 + d * x2);
    }
  }
  Real tmin = 0.5 * (lo + hi);
  sphere_box(p0 + d * tmin, r, b, margin, out);
  for (int e = 0; e < 2; e++) {
    Real t = e == 0 ? 0.0 : 1.0;
    if (fabs(t - tmin) > 0.1 && box_sdf(b, p0 + d * t) < r + margin) sphere_box(p0 + d * t, r, b, margin, out);
  }
}

// collision.rs:496-511: keep the part of poly with pn.p <= off.
// Capacity of the clipping polygon. One clip keeps the vertices on the inner
// side (distance <= 0) and adds one point per sign change; for a convex
// polygon there are at most two, so a clip adds at most two vertices and the
// face quad has at most 12 after its four clips (vertices on a clip line
// really do produce them: 9 to 12 occur when two box faces are parallel).
// The reference's vectors are unbounded; here a vertex past the capacity
// would be dropped rather than written out of bounds.
constexpr int CLIP_CAP = 24;
PHD int clip(const V3* poly, int n, V3 pn, Real off, V3* out) {
  int m = 0;
  for (int i = 0; i < n; i++) {
    V3 a = poly[i], b = poly[(i + 1) % n];
    Real da = dot(pn, a) - off, db = dot(pn, b) - off;
    if (da <= 0.0 && m < CLIP_CAP) out[m++] = a;
    if ((da < 0.0) != (db < 0.0) && fabs(da - db) > 1e-15 && m < CLIP_CAP) out[m++] = a + (b - a) * (da / (da - db));
  }
  return m;
}

// collision.rs:383-493
PHD void box_box(const GPose& a, const GPose& b, Real margin, Sink& out) {
  V3 d = b.pos - a.pos;
  V3 ax[3] = {col(a.mat, 0), col(a.mat, 1), col(a.mat, 2)};
  V3 bx[3] = {col(b.mat, 0), col(b.mat, 1), col(b.mat, 2)};
  bool has_best = false, has_face = false;
  Real best_o = 0, face_o = 0;
  V3 best_l = v3z(), face_l = v3z();
  int best_k = 0, face_k = 0;
  for (int k = 0; k < 15; k++) {
    V3 l;
    if (k < 3) l = ax[k];
    else if (k < 6) l = bx[k - 3];
    else {
      V3 c = cross(ax[(k - 6) / 3], bx[(k - 6) % 3]);
      Real n = norm(c);
      if (n < 1e-6) continue;
      l = c * (1.0 / n);
    }
    Real pa = a.size.x * fabs(dot(ax[0], l)) + a.size.y * fabs(dot(ax[1], l)) + a.size.z * fabs(dot(ax[2], l));
    Real pb = b.size.x * fabs(dot(bx[0], l)) + b.size.y * fabs(dot(bx[1], l)) + b.size.z * fabs(dot(bx[2], l));
    Real overlap = pa + pb - fabs(dot(d, l));
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
  V3 n = best_l;
  int k = best_k;
  if (k >= 6) {
    int i = (k - 6) / 3, j = (k - 6) % 3;
    V3 pa = a.pos;
    for (int m = 0; m < 3; m++)
      if (m != i) pa = pa + ax[m] * (a.size[m] * rsignum(dot(ax[m], n)));
    V3 pb = b.pos;
    for (int m = 0; m < 3; m++)
      if (m != j) pb = pb - bx[m] * (b.size[m] * rsignum(dot(bx[m], n)));
    V3 x, y;
    closest_segments(pa - ax[i] * a.size[i], pa + ax[i] * a.size[i], pb - bx[j] * b.size[j], pb + bx[j] * b.size[j],
                     &x, &y);
    out.push((x + y) * 0.5, n, overlap);
    return;
  }
  const GPose& rf = k < 3 ? a : b;
  const GPose& inc = k < 3 ? b : a;
  const V3* raxes = k < 3 ? ax : bx;
  const V3* iaxes = k < 3 ? bx : ax;
  V3 nref = k < 3 ? n : -n;
  int fi = k < 3 ? k : k - 3;
  V3 ref_center = rf.pos + nref * rf.size[fi];
  int u = (fi + 1) % 3, v = (fi + 2) % 3;
  int j = 0;
  Real bestdot = INFINITY;
  for (int m = 0; m < 3; m++) {
    Real dt = dot(iaxes[m], nref);
    if (-fabs(dt) < bestdot) {
      bestdot = -fabs(dt);
      j = m;
    }
  }
  Real s = -rsignum(dot(iaxes[j], nref));
  V3 ic = inc.pos + iaxes[j] * (s * inc.size[j]);
  int ku = (j + 1) % 3, kv = (j + 2) % 3;
  V3 eu = iaxes[ku] * inc.size[ku];
  V3 ev = iaxes[kv] * inc.size[kv];
  V3 poly[CLIP_CAP], tmp[CLIP_CAP];
  poly[0] = ic + eu + ev;
  poly[1] = ic - eu + ev;
  poly[2] = ic - eu - ev;
  poly[3] = ic + eu - ev;
  int np = 4;
  for (int c = 0; c < 2; c++) {
    V3 axis = c == 0 ? raxes[u] : raxes[v];
    Real ext = c == 0 ? rf.size[u] : rf.size[v];
    for (int si = 0; si < 2; si++) {
      V3 pn = axis * (si == 0 ? 1.0 : -1.0);
      Real off = dot(pn, ref_center) + ext;
      np = clip(poly, np, pn, off, tmp);
      for (int q = 0; q < np; q++) poly[q] = tmp[q];
      if (np == 0) return;
    }
  }
  Real sep[CLIP_CAP];
  V3 pts[CLIP_CAP];
  int npt = 0;
  for (int q = 0; q < np; q++) {
    Real sp_ = dot(poly[q] - ref_center, nref);
    if (sp_ < margin) {
      sep[npt] = sp_;
      pts[npt] = poly[q];
      npt++;
    }
  }
  // reduce_to_four, collision.rs:514-531 (min_by keeps the first minimum, max_by the last maximum)
  if (npt > 4) {
    int keep[4];
    int di = 0;
    for (int q = 1; q < npt; q++)
      if (sep[q] < sep[di]) di = q;
    keep[0] = di;
    V3 deep = pts[di];
    int fa = 0;
    Real fk = norm2(pts[0] - deep);
    for (int q = 1; q < npt; q++) {
      Real kk = norm2(pts[q] - deep);
      if (!(kk < fk)) {
        fk = kk;
        fa = q;
      }
    }
    keep[1] = fa;
    V3 e01 = pts[keep[1]] - pts[keep[0]];
    int th = 0;
    Real tk = norm(cross(e01, pts[0] - pts[keep[0]]));
    for (int q = 1; q < npt; q++) {
      Real kk = norm(cross(e01, pts[q] - pts[keep[0]]));
      if (!(kk < tk)) {
        tk = kk;
        th = q;
      }
    }
    keep[2] = th;
    int fo = 0;
    Real fok = 0;
    for (int q = 0; q < npt; q++) {
      Real sprd = INFINITY;
      for (int t = 0; t < 3; t++) sprd = rmin(sprd, norm(pts[keep[t]] - pts[q]));
      if (q == 0 || !(sprd < fok)) {
        fok = sprd;
        fo = q;
      }
    }
    keep[3] = fo;
    Real s4[4];
    V3 p4[4];
    for (int t = 0; t < 4; t++) {
      s4[t] = sep[keep[t]];
      p4[t] = pts[keep[t]];
    }
    for (int t = 0; t < 4; t++) {
      sep[t] = s4[t];
      pts[t] = p4[t];
    }
    npt = 4;
  }
  for (int q = 0; q < npt; q++) out.push(pts[q] - nref * (sep[q] * 0.5), n, -sep[q]);
}

// collision.rs:59-96 for the ordered pair type(a) <= type(b).
PHD void collide_ordered(const GPose& a, const GPose& b, Real margin, Sink& out) {
  int ta = a.typ, tb = b.typ;
  if (ta == G_PLANE && tb == G_SPHERE) plane_sphere(a, b.pos, b.size.x, margin, out);
  else if (ta == G_PLANE && tb == G_CAPSULE) {
    V3 p0, p1;
    segment(b, &p0, &p1);
    int start = out.n;
    plane_sphere(a, p0, b.size.x, margin, out);
    plane_sphere(a, p1, b.size.x, margin, out);
    V3 t = col(b.mat, 2);
    for (int i = start; i < out.n; i++) st3(out.out + i * CW + 6, t);
  } else if (ta == G_PLANE && tb == G_BOX) plane_box(a, b, margin, out);
  else if (ta == G_PLANE && tb == G_CYLINDER) plane_cylinder(a, b, margin, out);
  else if (ta == G_SPHERE && tb == G_SPHERE) sphere_sphere(a.pos, a.size.x, b.pos, b.size.x, margin, out);
  else if (ta == G_SPHERE && tb == G_CAPSULE) {
    V3 p0, p1;
    segment(b, &p0, &p1);
    V3 q = closest_on_segment(p0, p1, a.pos);
    sphere_sphere(a.pos, a.size.x, q, b.size.x, margin, out);
  } else if (ta == G_SPHERE && tb == G_BOX) sphere_box(a.pos, a.size.x, b, margin, out);
  else if (ta == G_CAPSULE && tb == G_CAPSULE) capsule_capsule(a, b, margin, out);
  else if (ta == G_CAPSULE && tb == G_BOX) capsule_box(a, b, margin, out);
  else if (ta == G_BOX && tb == G_BOX) box_box(a, b, margin, out);
}

PHD void collide_pair(const GPose& a, const GPose& b, Real margin, Sink& out) {
  if (a.typ > b.typ) {
    int start = out.n;
    collide_ordered(b, a, margin, out);
    for (int i = start; i < out.n; i++) {
      gp<Real> h = out.out + (i * CW + 3);
      h[0] = -h[0];
      h[1] = -h[1];
      h[2] = -h[2];
    }
    return;
  }
  collide_ordered(a, b, margin, out);
}

// Upper bound on the hits a pair of these types produces (used to size buffers).
inline int pair_max_hits(int a, int b) {
  if (a == G_MESH || b == G_MESH || (a == G_CYLINDER && b != G_PLANE) || (b == G_CYLINDER && a != G_PLANE)) return 4;
  if (a > b) {
    int t = a;
    a = b;
    b = t;
  }
  if (a == G_PLANE && b == G_CAPSULE) return 2;
  if ((a == G_PLANE && b == G_BOX) || (a == G_PLANE && b == G_CYLINDER) || (a == G_BOX && b == G_BOX)) return 4;
  if ((a == G_CAPSULE && b == G_CAPSULE) || (a == G_CAPSULE && b == G_BOX)) return 3;
  return 1;
}

}  // namespace phk
