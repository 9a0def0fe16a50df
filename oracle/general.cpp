// ORACLE — test infrastructure only (see math.hpp).
//
// General convex narrow phase: the pairs that crates/pippin/src/collision.rs:102-107
// (needs_general) sends to parry3d-f64 0.31.1 through collide_general
// (collision.rs:124-147): a cylinder against a sphere, capsule, box or cylinder,
// and a mesh (its convex hull) against a plane or any other geom.
//
// parry is a crates.io dependency that is not in the reference tree, so this is a
// restatement of its published method, not of its code:
//   1. GJK distance between the two convex cores (sphere = its centre, capsule =
//      its segment, box, cylinder), the radii of sphere and capsule added after;
//   2. EPA (expanding polytope) for the depth and normal when the cores overlap;
//   3. the support feature of each shape along the contact normal, as parry's
//      PolygonalFeatureMap gives it (box: the face most aligned; cylinder: the cap
//      approximated by a square when |n_axis| >= 0.5, else the side segment;
//      capsule: its segment; sphere: its centre; hull: the face most aligned
//      among those around the support vertex, at most its first four vertices,
//      as FastHull::local_support_feature, fasthull.rs:125-153), and their overlap in the plane
//      normal to the contact normal (polygon clip, segment clip, parallel
//      segments), each point lifted back onto both features;
//   4. points whose separation is below the margin become hits; more than four
//      are reduced with the box-box rule (collision.rs:514-531), and when the
//      features give none the GJK/EPA witness pair is the hit.
// The hit is the midpoint of the two surface points, the normal points from the
// first geom to the second and depth = -separation, as collide_general converts
// parry's manifold (collision.rs:139-146).
// A plane against a hull is parry's half-space manifold: every vertex of the
// hull's support feature along the plane's inward normal whose distance to the
// plane is at most the margin is a hit, at the midpoint between the vertex and
// its projection on the plane.
// A hull's support point is the first of its vertices with the largest
// projection (a scan: FastHull scans hulls under 64 vertices, fasthull.rs:85-96;
// larger ones hill-climb to the same maximum, ties aside).
//
// Parity against parry is unpinned (its arithmetic is not available here). The
// tests pin this restatement with the reference's own physical tests of the path
// (crates/pippin/tests/mesh.rs:66-83: cylinders resting upright, on their side
// and stacked) and with closed forms (sphere and box against a cylinder at known
// separations and depths).
#include <algorithm>

#include "engine.hpp"

namespace orc {

namespace {

constexpr int GJK_ITERS = 64;
constexpr int EPA_ITERS = 28;
constexpr int EPA_V = 32;
constexpr int EPA_F = 64;
constexpr int EPA_E = 64;
constexpr int GEN_CAP = 16;
constexpr double EPA_TOL = 1e-9;

struct Core {
  int typ;
  Vec3 pos;
  Mat3 mat;
  Vec3 size;
  Real rad;
  const HullShape* hull;
};

Core core_of(const GeomPose& g) {
  Core c;
  c.typ = g.typ;
  c.pos = g.pos;
  c.mat = g.mat;
  c.size = g.size;
  c.rad = (g.typ == Sphere || g.typ == Capsule) ? g.size[0] : Real(0.0);
  c.hull = g.hull;
  return c;
}

// Index of the first hull vertex with the largest projection on d (local frame).
size_t hull_support(const HullShape& h, Vec3 d) {
  size_t v = 0;
  Real best = -INFINITY;
  for (size_t i = 0; i < h.vert.size(); i++) {
    Real x = dot(h.vert[i], d);
    if (x > best) {
      best = x;
      v = i;
    }
  }
  return v;
}

// Extreme point of the core along d, local frame.
Vec3 support_local(const Core& c, Vec3 d) {
  if (c.typ == Mesh) return c.hull->vert[hull_support(*c.hull, d)];
  if (c.typ == Sphere) return v3zero();
  if (c.typ == Capsule) return V3(0.0, 0.0, d[2] >= 0.0 ? c.size[1] : -c.size[1]);
  if (c.typ == Box)
    return V3(d[0] >= 0.0 ? c.size[0] : -c.size[0], d[1] >= 0.0 ? c.size[1] : -c.size[1],
              d[2] >= 0.0 ? c.size[2] : -c.size[2]);
  Real z = d[2] >= 0.0 ? c.size[1] : -c.size[1];
  Real rho2 = d[0] * d[0] + d[1] * d[1];
  if (rho2 > 1e-30) {
    Real s = c.size[0] / r_sqrt(rho2);
    return V3(d[0] * s, d[1] * s, z);
  }
  return V3(0.0, 0.0, z);
}

Vec3 support(const Core& c, Vec3 d) { return c.pos + mul_vec(c.mat, support_local(c, tmul_vec(c.mat, d))); }

// A point of the Minkowski difference A - B with the two points it came from.
struct SV {
  Vec3 w, a, b;
};

SV msup(const Core& A, const Core& B, Vec3 d) {
  SV s;
  s.a = support(A, d);
  s.b = support(B, -d);
  s.w = s.a - s.b;
  return s;
}

struct Simplex {
  SV p[4];
  Real lam[4];
  int n;
};

Vec3 simplex_point(const Simplex& s) {
  Vec3 v = s.p[0].w * s.lam[0];
  for (int i = 1; i < s.n; i++) v += s.p[i].w * s.lam[i];
  return v;
}

void keep1(Simplex& s, int i) {
  s.p[0] = s.p[i];
  s.lam[0] = 1.0;
  s.n = 1;
}

void keep2(Simplex& s, int i, int j, Real t) {
  SV a = s.p[i], b = s.p[j];
  s.p[0] = a;
  s.p[1] = b;
  s.lam[0] = 1.0 - t;
  s.lam[1] = t;
  s.n = 2;
}

// Closest point of segment (0, 1) of s to the origin; reduces s to its support.
void closest_segment(Simplex& s) {
  Vec3 a = s.p[0].w, ab = s.p[1].w - s.p[0].w;
  Real t = -dot(a, ab);
  if (t <= 0.0) {
    keep1(s, 0);
    return;
  }
  Real den = norm2(ab);
  if (t >= den) {
    keep1(s, 1);
    return;
  }
  keep2(s, 0, 1, t / den);
}

// Closest point of triangle (i, j, k) of s to the origin by Voronoi regions
// (Ericson, Real-Time Collision Detection, 5.1.5); writes the reduced simplex to
// r and returns the squared distance.
Real closest_triangle(const Simplex& s, int i, int j, int k, Simplex& r) {
  Vec3 a = s.p[i].w, b = s.p[j].w, c = s.p[k].w;
  Vec3 ab = b - a, ac = c - a;
  Real d1 = -dot(ab, a), d2 = -dot(ac, a);
  r = s;
  if (d1 <= 0.0 && d2 <= 0.0) {
    keep1(r, i);
    return norm2(a);
  }
  Real d3 = -dot(ab, b), d4 = -dot(ac, b);
  if (d3 >= 0.0 && d4 <= d3) {
    keep1(r, j);
    return norm2(b);
  }
  Real vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    keep2(r, i, j, d1 / (d1 - d3));
    return norm2(simplex_point(r));
  }
  Real d5 = -dot(ab, c), d6 = -dot(ac, c);
  if (d6 >= 0.0 && d5 <= d6) {
    keep1(r, k);
    return norm2(c);
  }
  Real vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    keep2(r, i, k, d2 / (d2 - d6));
    return norm2(simplex_point(r));
  }
  Real va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
    keep2(r, j, k, (d4 - d3) / ((d4 - d3) + (d5 - d6)));
    return norm2(simplex_point(r));
  }
  Real den = 1.0 / (va + vb + vc);
  Real v = vb * den, w = vc * den;
  SV pa = s.p[i], pb = s.p[j], pc = s.p[k];
  r.p[0] = pa;
  r.p[1] = pb;
  r.p[2] = pc;
  r.lam[0] = 1.0 - v - w;
  r.lam[1] = v;
  r.lam[2] = w;
  r.n = 3;
  return norm2(simplex_point(r));
}

// Reduces s to the sub-simplex nearest the origin. Returns 1 when the origin is
// inside the tetrahedron, -1 when the tetrahedron is flat, else 0.
int closest(Simplex& s) {
  if (s.n == 2) {
    closest_segment(s);
    return 0;
  }
  if (s.n == 3) {
    Simplex r;
    closest_triangle(s, 0, 1, 2, r);
    s = r;
    return 0;
  }
  // tetrahedron (Ericson 5.1.6): faces whose plane separates the origin from the
  // opposite vertex are candidates; none means the origin is inside
  static const int F[4][4] = {{0, 1, 2, 3}, {0, 2, 3, 1}, {0, 3, 1, 2}, {1, 3, 2, 0}};
  Vec3 a = s.p[0].w;
  Real vol = dot(s.p[3].w - a, cross(s.p[1].w - a, s.p[2].w - a));
  if (r_fabs(vol) <= 1e-30) return -1;
  bool any = false;
  Real best = 0.0;
  Simplex bs;
  for (int f = 0; f < 4; f++) {
    Vec3 p0 = s.p[F[f][0]].w;
    Vec3 nrm = cross(s.p[F[f][1]].w - p0, s.p[F[f][2]].w - p0);
    Real so = -dot(p0, nrm), sd = dot(s.p[F[f][3]].w - p0, nrm);
    if (!(so * sd < 0.0)) continue;
    Simplex r;
    Real d = closest_triangle(s, F[f][0], F[f][1], F[f][2], r);
    if (!any || d < best) {
      any = true;
      best = d;
      bs = r;
    }
  }
  if (!any) return 1;
  s = bs;
  return 0;
}

struct Face {
  int i[3];
  Vec3 n;
  Real d;
};

bool make_face(Face& f, const SV* V, int a, int b, int c) {
  Vec3 nrm = cross(V[b].w - V[a].w, V[c].w - V[a].w);
  Real l = norm(nrm);
  if (!(l > 1e-30)) return false;
  f.i[0] = a;
  f.i[1] = b;
  f.i[2] = c;
  f.n = nrm * (1.0 / l);
  f.d = dot(f.n, V[a].w);
  return true;
}

struct Penetration {
  Vec3 n, pa, pb;
  Real depth;
};

// Witness points of the origin's projection n * d onto face f.
Penetration face_result(const Face& f, const SV* V) {
  Penetration r;
  r.n = f.n;
  r.depth = f.d;
  Vec3 a = V[f.i[0]].w;
  Vec3 v0 = V[f.i[1]].w - a, v1 = V[f.i[2]].w - a, v2 = f.n * f.d - a;
  Real d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1), d20 = dot(v2, v0), d21 = dot(v2, v1);
  Real den = d00 * d11 - d01 * d01;
  Real u = 1.0, v = 0.0, w = 0.0;
  if (r_fabs(den) > 1e-30) {
    v = (d11 * d20 - d01 * d21) / den;
    w = (d00 * d21 - d01 * d20) / den;
    u = 1.0 - v - w;
  }
  r.pa = V[f.i[0]].a * u + V[f.i[1]].a * v + V[f.i[2]].a * w;
  r.pb = V[f.i[0]].b * u + V[f.i[1]].b * v + V[f.i[2]].b * w;
  return r;
}

// EPA from the simplex GJK ended with. False when no tetrahedron can be built.
bool epa(const Core& A, const Core& B, const Simplex& s, Penetration* out) {
  SV V[EPA_V];
  int nv = s.n;
  for (int k = 0; k < nv; k++) V[k] = s.p[k];
  static const Real AX[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  if (nv == 1) {
    for (int k = 0; k < 6 && nv == 1; k++) {
      SV w = msup(A, B, V3(AX[k][0], AX[k][1], AX[k][2]));
      if (norm2(w.w - V[0].w) > 1e-18) V[nv++] = w;
    }
  }
  if (nv == 2) {
    Vec3 e = V[1].w - V[0].w, ea = vabs(e);
    int k = 0;
    if (ea[1] < ea[k]) k = 1;
    if (ea[2] < ea[k]) k = 2;
    Vec3 d = cross(e, V3(AX[2 * k][0], AX[2 * k][1], AX[2 * k][2]));
    for (int sg = 0; sg < 2 && nv == 2; sg++) {
      SV w = msup(A, B, sg == 0 ? d : -d);
      if (norm2(cross(w.w - V[0].w, e)) > 1e-18 * norm2(e)) V[nv++] = w;
    }
  }
  if (nv == 3) {
    Vec3 nrm = cross(V[1].w - V[0].w, V[2].w - V[0].w);
    for (int sg = 0; sg < 2 && nv == 3; sg++) {
      SV w = msup(A, B, sg == 0 ? nrm : -nrm);
      if (r_fabs(dot(w.w - V[0].w, nrm)) > 1e-9 * norm(nrm)) V[nv++] = w;
    }
  }
  if (nv < 4) return false;
  if (dot(cross(V[1].w - V[0].w, V[2].w - V[0].w), V[3].w - V[0].w) > 0.0) {
    SV t = V[1];
    V[1] = V[2];
    V[2] = t;
  }
  Face F[EPA_F];
  int nf = 0;
  static const int T[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {1, 3, 2}};
  for (int k = 0; k < 4; k++)
    if (!make_face(F[nf++], V, T[k][0], T[k][1], T[k][2])) return false;
  int E[EPA_E][2];
  for (int it = 0; it < EPA_ITERS; it++) {
    int bi = 0;
    for (int k = 1; k < nf; k++)
      if (F[k].d < F[bi].d) bi = k;
    *out = face_result(F[bi], V);
    if (nv == EPA_V) return true;
    SV w = msup(A, B, F[bi].n);
    if (dot(w.w, F[bi].n) - F[bi].d <= EPA_TOL) return true;
    // faces that see the new point go; their unshared edges form the horizon
    int ne = 0;
    for (int k = 0; k < nf;) {
      if (dot(F[k].n, w.w - V[F[k].i[0]].w) > 0.0) {
        for (int q = 0; q < 3; q++) {
          int ea = F[k].i[q], eb = F[k].i[(q + 1) % 3];
          int found = -1;
          for (int m = 0; m < ne; m++)
            if (E[m][0] == eb && E[m][1] == ea) found = m;
          if (found >= 0) {
            E[found][0] = E[ne - 1][0];
            E[found][1] = E[ne - 1][1];
            ne--;
          } else {
            if (ne == EPA_E) return true;
            E[ne][0] = ea;
            E[ne][1] = eb;
            ne++;
          }
        }
        F[k] = F[nf - 1];
        nf--;
      } else {
        k++;
      }
    }
    V[nv] = w;
    for (int m = 0; m < ne; m++) {
      if (nf == EPA_F) return true;
      if (!make_face(F[nf], V, E[m][0], E[m][1], nv)) return true;
      nf++;
    }
    nv++;
  }
  int bi = 0;
  for (int k = 1; k < nf; k++)
    if (F[k].d < F[bi].d) bi = k;
  *out = face_result(F[bi], V);
  return true;
}

// Support feature of the core along the world direction dir (world points).
int feature(const Core& c, Vec3 dir, Vec3* out) {
  Vec3 d = tmul_vec(c.mat, dir);
  Vec3 loc[4];
  int n;
  if (c.typ == Mesh) {
    // the face most aligned with d among those around the support vertex
    // (Iterator::max_by: the last of equal maxima)
    const HullShape& h = *c.hull;
    size_t best = h.vert_faces[hull_support(h, d)][0];
    Real bd = dot(h.normal[best], d);
    for (size_t f : h.vert_faces[hull_support(h, d)]) {
      Real x = dot(h.normal[f], d);
      if (x >= bd) {
        bd = x;
        best = f;
      }
    }
    const std::vector<size_t>& face = h.faces[best];
    n = face.size() < 4 ? (int)face.size() : 4;
    for (int q = 0; q < n; q++) loc[q] = h.vert[face[(size_t)q]];
  } else if (c.typ == Sphere) {
    loc[0] = v3zero();
    n = 1;
  } else if (c.typ == Capsule) {
    loc[0] = V3(0.0, 0.0, -c.size[1]);
    loc[1] = V3(0.0, 0.0, c.size[1]);
    n = 2;
  } else if (c.typ == Box) {
    Vec3 da = vabs(d);
    int k = 0;
    if (da[1] > da[k]) k = 1;
    if (da[2] > da[k]) k = 2;
    int u = (k + 1) % 3, v = (k + 2) % 3;
    Real su = c.size[u], sv = c.size[v];
    const Real cu[4] = {su, -su, -su, su};
    const Real cv[4] = {sv, sv, -sv, -sv};
    for (int q = 0; q < 4; q++) {
      Vec3 p = v3zero();
      p[k] = d[k] >= 0.0 ? c.size[k] : -c.size[k];
      p[u] = cu[q];
      p[v] = cv[q];
      loc[q] = p;
    }
    n = 4;
  } else {
    Real r = c.size[0], h = c.size[1];
    Real rho2 = d[0] * d[0] + d[1] * d[1];
    Real cx = 1.0, cy = 0.0;
    if (rho2 > 1e-30) {
      Real s = 1.0 / r_sqrt(rho2);
      cx = d[0] * s;
      cy = d[1] * s;
    }
    if (r_fabs(d[2]) < 0.5) {
      loc[0] = V3(r * cx, r * cy, -h);
      loc[1] = V3(r * cx, r * cy, h);
      n = 2;
    } else {
      Real z = d[2] >= 0.0 ? h : -h;
      loc[0] = V3(r * cx, r * cy, z);
      loc[1] = V3(-(r * cy), r * cx, z);
      loc[2] = V3(-(r * cx), -(r * cy), z);
      loc[3] = V3(r * cy, -(r * cx), z);
      n = 4;
    }
  }
  for (int q = 0; q < n; q++) out[q] = c.pos + mul_vec(c.mat, loc[q]);
  return n;
}

struct P2 {
  Real x, y;
};

P2 proj2(Vec3 p, Vec3 t1, Vec3 t2) {
  P2 r;
  r.x = dot(p, t1);
  r.y = dot(p, t2);
  return r;
}

Real cross2(P2 a, P2 b) { return a.x * b.y - a.y * b.x; }
P2 sub2(P2 a, P2 b) {
  P2 r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}

struct Cands {
  Real dist[GEN_CAP];
  Vec3 pos[GEN_CAP];
  int n;
};

void emit(Cands& c, Vec3 pa, Vec3 pb, Vec3 n, Real ra, Real rb, Real margin) {
  Real dist = dot(pb - pa, n) - ra - rb;
  if (dist < margin && c.n < GEN_CAP) {
    c.dist[c.n] = dist;
    c.pos[c.n] = ((pa + n * ra) + (pb - n * rb)) * 0.5;
    c.n++;
  }
}

// Overlap of segment s[0..2) (on shape S) and convex polygon p[0..np) (on shape
// P) seen along n; points go out as (on A, on B) with A = S when s_is_a.
void seg_poly(const Vec3* s, const Vec3* p, int np, bool s_is_a, Vec3 n, Vec3 t1, Vec3 t2, Real ra, Real rb,
              Real margin, Cands& c) {
  P2 s0 = proj2(s[0], t1, t2), s1 = proj2(s[1], t1, t2);
  P2 e = sub2(s1, s0);
  if (!(e.x * e.x + e.y * e.y > 1e-18)) return;
  P2 q[4];
  for (int k = 0; k < np; k++) q[k] = proj2(p[k], t1, t2);
  Real area = 0.0;
  for (int k = 0; k < np; k++) area = area + cross2(q[k], q[(k + 1) % np]);
  Real orient = area >= 0.0 ? 1.0 : -1.0;
  Real tlo = 0.0, thi = 1.0;
  for (int k = 0; k < np; k++) {
    P2 ed = sub2(q[(k + 1) % np], q[k]);
    Real num = orient * cross2(ed, sub2(s0, q[k]));
    Real den = orient * cross2(ed, e);
    if (den == 0.0) {
      if (num < 0.0) return;
      continue;
    }
    Real t = -num / den;
    if (den > 0.0) tlo = rmax(tlo, t);
    else thi = rmin(thi, t);
  }
  if (tlo > thi) return;
  Vec3 np3 = cross(p[1] - p[0], p[2] - p[0]);
  Real den = dot(n, np3);
  if (!(r_fabs(den) > 1e-12 * norm(np3))) return;
  int m = thi - tlo > 1e-12 ? 2 : 1;
  for (int k = 0; k < m; k++) {
    Real t = k == 0 ? tlo : thi;
    Vec3 ps = s[0] + (s[1] - s[0]) * t;
    Vec3 pp = ps + n * (dot(p[0] - ps, np3) / den);
    if (s_is_a) emit(c, ps, pp, n, ra, rb, margin);
    else emit(c, pp, ps, n, ra, rb, margin);
  }
}

// Parallel segments a[0..2), b[0..2) seen along n: the two ends of their overlap.
void seg_seg(const Vec3* a, const Vec3* b, Vec3 n, Vec3 t1, Vec3 t2, Real ra, Real rb, Real margin, Cands& c) {
  P2 a0 = proj2(a[0], t1, t2), a1 = proj2(a[1], t1, t2), b0 = proj2(b[0], t1, t2), b1 = proj2(b[1], t1, t2);
  P2 e = sub2(a1, a0), f = sub2(b1, b0);
  Real ee = e.x * e.x + e.y * e.y, ff = f.x * f.x + f.y * f.y;
  if (!(ee > 1e-18) || !(ff > 1e-18)) return;
  Real cr = cross2(e, f);
  if (cr * cr > 1e-12 * ee * ff) return;  // crossing: the witness pair is the contact
  P2 g0 = sub2(b0, a0), g1 = sub2(b1, a0);
  Real u0 = (g0.x * e.x + g0.y * e.y) / ee, u1 = (g1.x * e.x + g1.y * e.y) / ee;
  Real lo = rmax(0.0, rmin(u0, u1)), hi = rmin(1.0, rmax(u0, u1));
  if (lo > hi) return;
  int m = hi - lo > 1e-12 ? 2 : 1;
  for (int k = 0; k < m; k++) {
    Real t = k == 0 ? lo : hi;
    Vec3 pa = a[0] + (a[1] - a[0]) * t;
    P2 q = proj2(pa, t1, t2);
    P2 h = sub2(q, b0);
    Real s = (h.x * f.x + h.y * f.y) / ff;
    Vec3 pb = b[0] + (b[1] - b[0]) * s;
    emit(c, pa, pb, n, ra, rb, margin);
  }
}

// Overlap of two convex polygons seen along n (Sutherland-Hodgman in the plane
// normal to n), each point lifted onto both polygons' planes.
void poly_poly(const Vec3* a, int na, const Vec3* b, int nb, Vec3 n, Vec3 t1, Vec3 t2, Real ra, Real rb,
               Real margin, Cands& c) {
  P2 qa[4], buf[2][GEN_CAP];
  for (int k = 0; k < na; k++) qa[k] = proj2(a[k], t1, t2);
  int m = nb;
  for (int k = 0; k < nb; k++) buf[0][k] = proj2(b[k], t1, t2);
  Real area = 0.0;
  for (int k = 0; k < na; k++) area = area + cross2(qa[k], qa[(k + 1) % na]);
  Real orient = area >= 0.0 ? 1.0 : -1.0;
  int cur = 0;
  for (int k = 0; k < na && m > 0; k++) {
    P2 p0 = qa[k], ed = sub2(qa[(k + 1) % na], qa[k]);
    const P2* in = buf[cur];
    P2* o = buf[1 - cur];
    int mo = 0;
    for (int q = 0; q < m; q++) {
      P2 s = in[(q + m - 1) % m], e = in[q];
      Real ds = orient * cross2(ed, sub2(s, p0)), de = orient * cross2(ed, sub2(e, p0));
      if (de >= 0.0) {
        if (ds < 0.0 && mo < GEN_CAP) {
          Real t = ds / (ds - de);
          P2 x;
          x.x = s.x + (e.x - s.x) * t;
          x.y = s.y + (e.y - s.y) * t;
          o[mo++] = x;
        }
        if (mo < GEN_CAP) o[mo++] = e;
      } else if (ds >= 0.0 && mo < GEN_CAP) {
        Real t = ds / (ds - de);
        P2 x;
        x.x = s.x + (e.x - s.x) * t;
        x.y = s.y + (e.y - s.y) * t;
        o[mo++] = x;
      }
    }
    m = mo;
    cur = 1 - cur;
  }
  if (m == 0) return;
  Vec3 nA = cross(a[1] - a[0], a[2] - a[0]), nB = cross(b[1] - b[0], b[2] - b[0]);
  Real dA = dot(n, nA), dB = dot(n, nB);
  if (!(r_fabs(dA) > 1e-12 * norm(nA)) || !(r_fabs(dB) > 1e-12 * norm(nB))) return;
  for (int k = 0; k < m; k++) {
    Vec3 x = t1 * buf[cur][k].x + t2 * buf[cur][k].y;
    Vec3 pa = x + n * (dot(a[0] - x, nA) / dA);
    Vec3 pb = x + n * (dot(b[0] - x, nB) / dB);
    emit(c, pa, pb, n, ra, rb, margin);
  }
}

// A plane against a hull: the hull's support feature along the plane's inward
// normal, each vertex within the margin a hit. `plane_first`: the plane is the
// pair's first geom (normal along the plane's normal), else its negation.
void plane_hull(const GeomPose& pl, const Core& H, bool plane_first, Real margin, std::vector<Hit>& out) {
  Vec3 n = col(pl.mat, 2);
  Vec3 fv[4];
  int nv = feature(H, -n, fv);
  for (int k = 0; k < nv; k++) {
    Real dist = dot(fv[k] - pl.pos, n);
    if (dist <= margin) {
      Hit h;
      h.pos = fv[k] - n * (dist * 0.5);
      h.normal = plane_first ? n : -n;
      h.depth = -dist;
      h.tangent = v3zero();
      out.push_back(h);
    }
  }
}

}  // namespace

// collision.rs:124-147 (parry contact manifold), restated as described above.
void collide_general(const GeomPose& ga, const GeomPose& gb, Real margin, std::vector<Hit>& out) {
  if (ga.typ == Plane && gb.typ == Mesh) {
    plane_hull(ga, core_of(gb), true, margin, out);
    return;
  }
  if (gb.typ == Plane && ga.typ == Mesh) {
    plane_hull(gb, core_of(ga), false, margin, out);
    return;
  }
  if (ga.typ == Plane || gb.typ == Plane) return;
  Core A = core_of(ga), B = core_of(gb);
  Simplex s;
  Vec3 d0 = B.pos - A.pos;
  if (!(norm2(d0) > 1e-30)) d0 = V3(1.0, 0.0, 0.0);
  s.p[0] = msup(A, B, d0);
  s.lam[0] = 1.0;
  s.n = 1;
  Vec3 v = s.p[0].w;
  Real vv = norm2(v);
  bool overlap = false;
  for (int it = 0; it < GJK_ITERS; it++) {
    if (vv <= 1e-24) {
      overlap = true;
      break;
    }
    SV w = msup(A, B, -v);
    if (vv - dot(v, w.w) <= 1e-12 * vv) break;
    Simplex prev = s;
    s.p[s.n] = w;
    s.n++;
    int st = closest(s);
    if (st == 1) {
      overlap = true;
      break;
    }
    if (st == -1) {
      s = prev;
      break;
    }
    Vec3 vn = simplex_point(s);
    Real vvn = norm2(vn);
    if (!(vvn < vv)) {
      s = prev;
      break;
    }
    v = vn;
    vv = vvn;
  }
  Vec3 n, wa, wb;
  Real dcore;
  if (overlap) {
    Penetration p;
    if (!epa(A, B, s, &p)) return;
    n = p.n;
    dcore = -p.depth;
    wa = p.pa;
    wb = p.pb;
  } else {
    dcore = r_sqrt(vv);
    n = v * (-1.0 / dcore);
    wa = s.p[0].a * s.lam[0];
    wb = s.p[0].b * s.lam[0];
    for (int i = 1; i < s.n; i++) {
      wa += s.p[i].a * s.lam[i];
      wb += s.p[i].b * s.lam[i];
    }
  }
  if (!(dcore - A.rad - B.rad < margin)) return;
  Vec3 fa[4], fb[4];
  int na = feature(A, n, fa), nb = feature(B, -n, fb);
  Vec3 t1, t2;
  tangents(n, &t1, &t2);
  Cands c;
  c.n = 0;
  if (na == 2 && nb == 2) seg_seg(fa, fb, n, t1, t2, A.rad, B.rad, margin, c);
  else if (na == 2 && nb >= 3) seg_poly(fa, fb, nb, true, n, t1, t2, A.rad, B.rad, margin, c);
  else if (na >= 3 && nb == 2) seg_poly(fb, fa, na, false, n, t1, t2, A.rad, B.rad, margin, c);
  else if (na >= 3 && nb >= 3) poly_poly(fa, na, fb, nb, n, t1, t2, A.rad, B.rad, margin, c);
  if (c.n == 0) emit(c, wa, wb, n, A.rad, B.rad, margin);
  std::vector<std::pair<Real, Vec3>> pts;
  for (int k = 0; k < c.n; k++) pts.push_back({c.dist[k], c.pos[k]});
  reduce_to_four(pts);
  for (const auto& sp : pts) {
    Hit h;
    h.pos = sp.second;
    h.normal = n;
    h.depth = -sp.first;
    h.tangent = v3zero();
    out.push_back(h);
  }
}

}  // namespace orc
