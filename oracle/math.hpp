// ORACLE — test infrastructure only. Nothing in the product (pippin_amd/, include/)
// may include, link or call this code; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, as the checker.
//
// Plain-old-data 3-D and spatial algebra restating crates/pippin/src/math.rs and
// crates/pippin/src/simd.rs (x86 branch) operation for operation. Build with
// -ffp-contract=off and without fast-math: the Rust reference never contracts
// a*b+c into an FMA, and the order of every sum below is the reference's.
#pragma once
#include <cmath>
#include <cstddef>

namespace orc {

#ifdef ORC_OPCOUNT
}  // namespace orc
#include "opcount.hpp"
namespace orc {
#else
typedef double Real;
static inline Real r_sqrt(Real x) { return std::sqrt(x); }
static inline Real r_fabs(Real x) { return std::fabs(x); }
static inline Real r_pow(Real a, Real b) { return std::pow(a, b); }
static inline bool r_isfinite(Real x) { return std::isfinite(x); }
static inline bool r_isnan(Real x) { return std::isnan(x); }
static inline Real r_copysign(Real a, Real b) { return std::copysign(a, b); }
static inline Real r_fmax(Real a, Real b) { return std::fmax(a, b); }
static inline Real r_fmin(Real a, Real b) { return std::fmin(a, b); }
static inline double r_val(Real x) { return x; }
static inline void r_count(int) {}
#endif

// Rust's `impl Sum for f64` folds from -0.0; `rsum_start()` keeps that sign rule.
static inline Real rsum_start() { return -0.0; }

// Rust f64::clamp: `if x < lo {x = lo}; if x > hi {x = hi}` (NaN passes through).
static inline Real rclamp(Real x, Real lo, Real hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
// Rust f64::max / f64::min (IEEE maxNum/minNum: a NaN operand yields the other).
static inline Real rmax(Real a, Real b) { return r_fmax(a, b); }
static inline Real rmin(Real a, Real b) { return r_fmin(a, b); }
// Rust f64::signum: +1 for +0.0 and positives, -1 for -0.0 and negatives.
static inline Real rsignum(Real x) { return r_isnan(x) ? x : r_copysign(1.0, x); }

struct Vec3 {
  Real v[3];
  Real& operator[](int i) { return v[i]; }
  Real operator[](int i) const { return v[i]; }
};
static inline Vec3 V3(Real x, Real y, Real z) { Vec3 r = {{x, y, z}}; return r; }
static inline Vec3 v3zero() { return V3(0.0, 0.0, 0.0); }
static inline Vec3 operator+(Vec3 a, Vec3 b) { return V3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
static inline Vec3 operator-(Vec3 a, Vec3 b) { return V3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
static inline Vec3 operator-(Vec3 a) { return V3(-a[0], -a[1], -a[2]); }
static inline Vec3 operator*(Vec3 a, Real s) { return V3(a[0] * s, a[1] * s, a[2] * s); }
static inline Vec3& operator+=(Vec3& a, Vec3 b) { a = a + b; return a; }
static inline Vec3& operator-=(Vec3& a, Vec3 b) { a = a - b; return a; }
// math.rs:22-24
static inline Real dot(Vec3 a, Vec3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// math.rs:26-30
static inline Vec3 cross(Vec3 a, Vec3 o) {
  return V3(a[1] * o[2] - a[2] * o[1], a[2] * o[0] - a[0] * o[2], a[0] * o[1] - a[1] * o[0]);
}
static inline Real norm2(Vec3 a) { return dot(a, a); }
static inline Real norm(Vec3 a) { return r_sqrt(norm2(a)); }
// math.rs:40-47
static inline Vec3 normalized(Vec3 a) {
  Real n = norm(a);
  if (n > 1e-15) return a * (1.0 / n);
  return v3zero();
}
static inline Vec3 vabs(Vec3 a) { return V3(r_fabs(a[0]), r_fabs(a[1]), r_fabs(a[2])); }
static inline Vec3 mul_elem(Vec3 a, Vec3 o) { return V3(a[0] * o[0], a[1] * o[1], a[2] * o[2]); }
static inline Vec3 vmin(Vec3 a, Vec3 o) { return V3(rmin(a[0], o[0]), rmin(a[1], o[1]), rmin(a[2], o[2])); }
static inline Vec3 vmax(Vec3 a, Vec3 o) { return V3(rmax(a[0], o[0]), rmax(a[1], o[1]), rmax(a[2], o[2])); }
// math.rs:65-70
static inline void tangents(Vec3 n, Vec3* t1, Vec3* t2) {
  Vec3 a = r_fabs(n[0]) < 0.57 ? V3(1, 0, 0) : V3(0, 1, 0);
  *t1 = normalized(cross(n, a));
  *t2 = cross(n, *t1);
}

// Row-major 3x3 (math.rs:127-199).
struct Mat3 {
  Real m[9];
  Real& operator[](int i) { return m[i]; }
  Real operator[](int i) const { return m[i]; }
};
static inline Mat3 m3identity() { Mat3 r = {{1, 0, 0, 0, 1, 0, 0, 0, 1}}; return r; }
static inline Mat3 m3zero() { Mat3 r = {{0, 0, 0, 0, 0, 0, 0, 0, 0}}; return r; }
static inline Mat3 m3diag(Vec3 d) { Mat3 r = {{d[0], 0, 0, 0, d[1], 0, 0, 0, d[2]}}; return r; }
static inline Mat3 from_cols(Vec3 a, Vec3 b, Vec3 c) {
  Mat3 r = {{a[0], b[0], c[0], a[1], b[1], c[1], a[2], b[2], c[2]}};
  return r;
}
static inline Vec3 col(const Mat3& m, int j) { return V3(m[j], m[3 + j], m[6 + j]); }
static inline Vec3 mul_vec(const Mat3& m, Vec3 v) {
  return V3(m[0] * v[0] + m[1] * v[1] + m[2] * v[2], m[3] * v[0] + m[4] * v[1] + m[5] * v[2],
            m[6] * v[0] + m[7] * v[1] + m[8] * v[2]);
}
static inline Vec3 tmul_vec(const Mat3& m, Vec3 v) {
  return V3(m[0] * v[0] + m[3] * v[1] + m[6] * v[2], m[1] * v[0] + m[4] * v[1] + m[7] * v[2],
            m[2] * v[0] + m[5] * v[1] + m[8] * v[2]);
}
// math.rs:169-177: each entry is an iterator sum (starts at -0.0, then k = 0, 1, 2).
static inline Mat3 mul_mat(const Mat3& a, const Mat3& o) {
  Mat3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Real s = rsum_start();
      for (int k = 0; k < 3; k++) s += a[3 * i + k] * o[3 * k + j];
      r[3 * i + j] = s;
    }
  return r;
}
static inline Mat3 transpose(const Mat3& m) {
  Mat3 r = {{m[0], m[3], m[6], m[1], m[4], m[7], m[2], m[5], m[8]}};
  return r;
}
static inline Mat3 skew(Vec3 v) {
  Mat3 r = {{0.0, -v[2], v[1], v[2], 0.0, -v[0], -v[1], v[0], 0.0}};
  return r;
}
static inline Mat3 mscale(const Mat3& m, Real s) {
  Mat3 r;
  for (int i = 0; i < 9; i++) r[i] = m[i] * s;
  return r;
}
static inline Mat3 madd(const Mat3& a, const Mat3& o) {
  Mat3 r;
  for (int i = 0; i < 9; i++) r[i] = a[i] + o[i];
  return r;
}
// math.rs:196-198: R * self * R^T
static inline Mat3 mrotate(const Mat3& self, const Mat3& r) { return mul_mat(mul_mat(r, self), transpose(r)); }

// ---- correctly rounded sin and cos (see the note below) ----
// The reference calls the platform libm (Rust f64::sin_cos); libms differ from
// each other by up to one ulp, and through contact one ulp grows into visible
// trajectory differences. Both the device and the oracle therefore evaluate
// the correctly rounded sin/cos: argument reduction by a three-double pi/2
// (Cody-Waite with exact products) and Taylor series in double-double
// arithmetic (terms 0..8) with a double tail (terms 9..14). The result is
// within ~2^-100 relative of the exact value before the final rounding, so it
// equals the exactly rounded value except at ties closer than that.
struct DD {
  double hi, lo;
};
static inline DD dd_fast_two_sum(double a, double b) {
  double s = a + b;
  DD r;
  r.hi = s;
  r.lo = b - (s - a);
  return r;
}
static inline DD dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  DD r;
  r.hi = s;
  r.lo = (a - (s - bb)) + (b - bb);
  return r;
}
static inline DD dd_add(DD a, DD b) {
  DD s = dd_two_sum(a.hi, b.hi);
  DD t = dd_two_sum(a.lo, b.lo);
  double e = s.lo + t.hi;
  DD u = dd_fast_two_sum(s.hi, e);
  e = u.lo + t.lo;
  return dd_fast_two_sum(u.hi, e);
}
static inline DD dd_mul(DD a, DD b) {
  double p = a.hi * b.hi;
  double e = std::fma(a.hi, b.hi, -p);
  e = e + (a.hi * b.lo + a.lo * b.hi);
  return dd_fast_two_sum(p, e);
}
static inline DD dd_from(double hi, double lo) {
  DD r;
  r.hi = hi;
  r.lo = lo;
  return r;
}
static inline void cr_sincos(double x, double* sout, double* cout) {
  if (!std::isfinite(x)) {
    *sout = x - x;
    *cout = x - x;
    return;
  }
  if (x == 0.0) {  // sin(-0) = -0
    *sout = x;
    *cout = 1.0;
    return;
  }
  const double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
  double k = std::rint(x * 0.6366197723675814);
  DD r = dd_from(x, 0.0);
  double h = k * P1;
  r = dd_add(r, dd_from(-h, -std::fma(k, P1, -h)));
  h = k * P2;
  r = dd_add(r, dd_from(-h, -std::fma(k, P2, -h)));
  r = dd_add(r, dd_from(-(k * P3), 0.0));
  DD r2 = dd_mul(r, r);
  const double S[15][2] = {{1.0, 0.0},
                         {-0.16666666666666666, -9.25185853854297e-18},
                         {0.008333333333333333, 1.1564823173178714e-19},
                         {-0.0001984126984126984, -1.7209558293420705e-22},
                         {2.7557319223985893e-06, -1.858393274046472e-22},
                         {-2.505210838544172e-08, 1.448814070935912e-24},
                         {1.6059043836821613e-10, 1.2585294588752098e-26},
                         {-7.647163731819816e-13, -7.03872877733453e-30},
                         {2.8114572543455206e-15, 1.6508842730861433e-31},
                         {-8.22063524662433e-18, 0.0},
                         {1.9572941063391263e-20, 0.0},
                         {-3.868170170630684e-23, 0.0},
                         {6.446950284384474e-26, 0.0},
                         {-9.183689863795546e-29, 0.0},
                         {1.1309962886447716e-31, 0.0}};
  const double Cc[15][2] = {{1.0, 0.0},
                          {-0.5, 0.0},
                          {0.041666666666666664, 2.3129646346357427e-18},
                          {-0.001388888888888889, 5.300543954373577e-20},
                          {2.48015873015873e-05, 2.1511947866775882e-23},
                          {-2.755731922398589e-07, -2.3767714622250297e-23},
                          {2.08767569878681e-09, -1.20734505911326e-25},
                          {-1.1470745597729725e-11, -2.0655512752830745e-28},
                          {4.779477332387385e-14, 4.399205485834081e-31},
                          {-1.5619206968586225e-16, 0.0},
                          {4.110317623312165e-19, 0.0},
                          {-8.896791392450574e-22, 0.0},
                          {1.6117375710961184e-24, 0.0},
                          {-2.4795962632247976e-27, 0.0},
                          {3.279889237069838e-30, 0.0}};
  double ts = S[14][0], tc = Cc[14][0];
  for (int i = 13; i >= 9; i--) {
    ts = ts * r2.hi + S[i][0];
    tc = tc * r2.hi + Cc[i][0];
  }
  DD ps = dd_from(ts, 0.0), pc = dd_from(tc, 0.0);
  for (int i = 8; i >= 0; i--) {
    ps = dd_add(dd_mul(ps, r2), dd_from(S[i][0], S[i][1]));
    pc = dd_add(dd_mul(pc, r2), dd_from(Cc[i][0], Cc[i][1]));
  }
  DD s = dd_mul(ps, r);
  double sv = s.hi, cv = pc.hi;
  long q = (long)k;
  q = ((q % 4) + 4) % 4;
  if (q == 0) {
    *sout = sv;
    *cout = cv;
  } else if (q == 1) {
    *sout = cv;
    *cout = -sv;
  } else if (q == 2) {
    *sout = -sv;
    *cout = -cv;
  } else {
    *sout = -cv;
    *cout = sv;
  }
}

// Unit quaternion (w, x, y, z), math.rs:201-301.
struct Quat {
  Real q[4];
  Real& operator[](int i) { return q[i]; }
  Real operator[](int i) const { return q[i]; }
};
static inline Quat Q4(Real w, Real x, Real y, Real z) { Quat r = {{w, x, y, z}}; return r; }
static inline Quat qidentity() { return Q4(1, 0, 0, 0); }
static inline Quat from_axis_angle(Vec3 axis, Real angle) {
  Vec3 a = normalized(axis);
  double sd, cd;
  cr_sincos(r_val(0.5 * angle), &sd, &cd);
  r_count(2);  // one sin and one cos, as the reference's libm sin_cos
  Real s = sd, c = cd;
  return Q4(c, a[0] * s, a[1] * s, a[2] * s);
}
static inline Quat qnormalized(Quat q) {
  Real s = rsum_start();
  for (int i = 0; i < 4; i++) s += q[i] * q[i];
  Real n = r_sqrt(s);
  if (n < 1e-15) return qidentity();
  return Q4(q[0] / n, q[1] / n, q[2] / n, q[3] / n);
}
static inline Quat qmul(Quat p, Quat o) {
  Real a = p[0], b = p[1], c = p[2], d = p[3];
  Real e = o[0], f = o[1], g = o[2], h = o[3];
  return Q4(a * e - b * f - c * g - d * h, a * f + b * e + c * h - d * g, a * g - b * h + c * e + d * f,
            a * h + b * g - c * f + d * e);
}
static inline Mat3 to_mat(Quat q) {
  Real w = q[0], x = q[1], y = q[2], z = q[3];
  Real ww = w * w, xx = x * x, yy = y * y, zz = z * z;
  Real xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
  Mat3 r = {{ww + xx - yy - zz, 2.0 * (xy - wz), 2.0 * (xz + wy), 2.0 * (xy + wz), ww - xx + yy - zz,
             2.0 * (yz - wx), 2.0 * (xz - wy), 2.0 * (yz + wx), ww - xx - yy + zz}};
  return r;
}
static inline Vec3 qrotate(Quat q, Vec3 v) { return mul_vec(to_mat(q), v); }
// Shepperd's method, math.rs:260-282
static inline Quat from_mat(const Mat3& m) {
  Real tr = m[0] + m[4] + m[8];
  Quat q;
  if (tr > 0.0) {
    Real s = r_sqrt(tr + 1.0) * 2.0;
    q = Q4(0.25 * s, (m[7] - m[5]) / s, (m[2] - m[6]) / s, (m[3] - m[1]) / s);
  } else if (m[0] > m[4] && m[0] > m[8]) {
    Real s = r_sqrt(1.0 + m[0] - m[4] - m[8]) * 2.0;
    q = Q4((m[7] - m[5]) / s, 0.25 * s, (m[1] + m[3]) / s, (m[2] + m[6]) / s);
  } else if (m[4] > m[8]) {
    Real s = r_sqrt(1.0 + m[4] - m[0] - m[8]) * 2.0;
    q = Q4((m[2] - m[6]) / s, (m[1] + m[3]) / s, 0.25 * s, (m[5] + m[7]) / s);
  } else {
    Real s = r_sqrt(1.0 + m[8] - m[0] - m[4]) * 2.0;
    q = Q4((m[3] - m[1]) / s, (m[2] + m[6]) / s, (m[5] + m[7]) / s, 0.25 * s);
  }
  q = qnormalized(q);
  if (q[0] < 0.0) q = Q4(-q[0], -q[1], -q[2], -q[3]);
  return q;
}
// math.rs:284-292
static inline Quat from_z_to(Vec3 dir) {
  Vec3 d = normalized(dir);
  Real c = dot(V3(0, 0, 1), d);
  if (c < -1.0 + 1e-12) return Q4(0.0, 1.0, 0.0, 0.0);
  Vec3 axis = cross(V3(0, 0, 1), d);
  return qnormalized(Q4(1.0 + c, axis[0], axis[1], axis[2]));
}
// math.rs:294-300
static inline Quat integrate_local(Quat self, Vec3 omega, Real h) {
  Real n = norm(omega);
  if (n * h < 1e-15) return self;
  return qnormalized(qmul(self, from_axis_angle(omega * (1.0 / n), n * h)));
}

// Spatial motion/force [ang; lin] about the world origin, math.rs:303-369.
struct Spatial {
  Vec3 ang, lin;
};
static inline Spatial SP(Vec3 a, Vec3 l) { Spatial s; s.ang = a; s.lin = l; return s; }
static inline Spatial spzero() { return SP(v3zero(), v3zero()); }
static inline Real sdot(const Spatial& a, const Spatial& o) { return dot(a.ang, o.ang) + dot(a.lin, o.lin); }
static inline Spatial sscale(const Spatial& a, Real s) { return SP(a.ang * s, a.lin * s); }
static inline Spatial cross_motion(const Spatial& s, const Spatial& m) {
  return SP(cross(s.ang, m.ang), cross(s.ang, m.lin) + cross(s.lin, m.ang));
}
static inline Spatial cross_force(const Spatial& s, const Spatial& f) {
  return SP(cross(s.ang, f.ang) + cross(s.lin, f.lin), cross(s.ang, f.lin));
}
static inline Vec3 point_velocity(const Spatial& s, Vec3 p) { return s.lin + cross(s.ang, p); }
static inline Spatial operator+(const Spatial& a, const Spatial& o) { return SP(a.ang + o.ang, a.lin + o.lin); }
static inline Spatial& operator+=(Spatial& a, const Spatial& o) {
  a.ang += o.ang;
  a.lin += o.lin;
  return a;
}

// math.rs:371-410
struct SpatialInertia {
  Real mass;
  Vec3 mc;
  Mat3 inertia_o;
};
static inline SpatialInertia sizero() {
  SpatialInertia s;
  s.mass = 0.0;
  s.mc = v3zero();
  s.inertia_o = m3zero();
  return s;
}
static inline SpatialInertia from_com(Real mass, Vec3 com, const Mat3& inertia_com) {
  Mat3 cx = skew(com);
  SpatialInertia s;
  s.inertia_o = madd(inertia_com, mscale(mul_mat(cx, transpose(cx)), mass));
  s.mass = mass;
  s.mc = com * mass;
  return s;
}
static inline SpatialInertia siadd(const SpatialInertia& a, const SpatialInertia& o) {
  SpatialInertia s;
  s.mass = a.mass + o.mass;
  s.mc = a.mc + o.mc;
  s.inertia_o = madd(a.inertia_o, o.inertia_o);
  return s;
}
static inline Spatial mul_motion(const SpatialInertia& I, const Spatial& v) {
  return SP(mul_vec(I.inertia_o, v.ang) + cross(I.mc, v.lin), v.lin * I.mass - cross(I.mc, v.ang));
}

// simd.rs:14-64, x86 branch: four strided accumulators when n >= 8.
static inline Real sdot_n(const Real* a, const Real* b, size_t n) {
  if (n < 8) {
    Real s = 0.0;
    for (size_t k = 0; k < n; k++) s += a[k] * b[k];
    return s;
  }
  Real s[4] = {0.0, 0.0, 0.0, 0.0};
  size_t k = 0;
  while (k + 4 <= n) {
    for (int j = 0; j < 4; j++) s[j] += a[k + j] * b[k + j];
    k += 4;
  }
  Real t = (s[0] + s[1]) + (s[2] + s[3]);
  while (k < n) {
    t += a[k] * b[k];
    k++;
  }
  return t;
}
// simd.rs:68-72
static inline void axpy(Real alpha, const Real* x, Real* y, size_t n) {
  for (size_t k = 0; k < n; k++) y[k] += alpha * x[k];
}

// math.rs:425-447: in-place lower Cholesky of the diagonal block [s0, e) of a
// row-major matrix with row stride `stride`.
static inline bool cholesky_block(Real* a, size_t stride, size_t s0, size_t e) {
  for (size_t j = s0; j < e; j++) {
    Real* rj = a + j * stride + s0;
    size_t len = j - s0;
    Real d = a[j * stride + j] - sdot_n(rj, rj, len);
    if (d <= 0.0) return false;
    d = r_sqrt(d);
    a[j * stride + j] = d;
    Real inv = 1.0 / d;
    for (size_t i = j + 1; i < e; i++) {
      Real* ri = a + i * stride + s0;
      size_t k = j - s0;
      ri[k] = (ri[k] - sdot_n(ri, rj, k)) * inv;
    }
  }
  return true;
}
// math.rs:451-463
static inline void cholesky_solve_block(const Real* l, size_t stride, size_t s0, size_t e, Real* x) {
  for (size_t i = s0; i < e; i++)
    x[i] = (x[i] - sdot_n(l + i * stride + s0, x + s0, i - s0)) / l[i * stride + i];
  for (size_t ii = e; ii > s0; ii--) {
    size_t i = ii - 1;
    x[i] /= l[i * stride + i];
    Real xi = x[i];
    axpy(-xi, l + i * stride + s0, x + s0, i - s0);
  }
}
static inline bool cholesky(Real* a, size_t n) { return cholesky_block(a, n, 0, n); }
static inline void cholesky_solve(const Real* l, size_t n, Real* x) { cholesky_solve_block(l, n, 0, n, x); }

}  // namespace orc
