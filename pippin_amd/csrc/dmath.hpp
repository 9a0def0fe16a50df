// FP64 3-D and spatial algebra for the device kernels.
//
// Operation order mirrors crates/pippin/src/math.rs and simd.rs (x86 branch)
// exactly, and every file including this is compiled with -ffp-contract=off:
// together that makes the device arithmetic identical to the reference's
// wherever the same operands meet: sqrt and division are IEEE on both sides and
// sin/cos are evaluated correctly rounded (cr_sincos).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PHD __host__ __device__ __forceinline__

namespace phk {

typedef double Real;

PHD Real rsum0() { return -0.0; }  // Rust's f64 Sum folds from -0.0
PHD Real rclamp(Real x, Real lo, Real hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
PHD Real rmax(Real a, Real b) { return fmax(a, b); }
PHD Real rmin(Real a, Real b) { return fmin(a, b); }
// f64::is_finite without the libm call (x - x is 0 exactly for finite x, NaN otherwise).
PHD bool rfinite(Real x) { return x - x == 0.0; }
PHD Real rsignum(Real x) { return x != x ? x : copysign(1.0, x); }

// IEEE (correctly rounded) square root. On the device, for 2^-767 <= x <=
// DBL_MAX this is the compiler's own f64 sqrt sequence (rsq estimate, one
// Goldschmidt refinement of the root and half-reciprocal, two residual
// corrections) without the input scaling and special-value selects that the
// full sequence applies outside that range, where both give the same bits;
// other inputs take the full sequence.
PHD Real sqrt_cr(Real x) {
#ifdef __HIP_DEVICE_COMPILE__
  if (x >= 0x1p-767 && x <= 0x1.fffffffffffffp+1023) {
    const Real y = __builtin_amdgcn_rsq(x);
    Real g = x * y, h = y * 0.5;
    const Real r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    Real d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
  }
  return __builtin_sqrt(x);
#else
  return sqrt(x);
#endif
}

// 1.0 / s, correctly rounded, for s a square root of a positive finite double
// (so 2^-537 <= s <= 2^512). On the device this is the compiler's f64 division
// sequence with numerator 1 (reciprocal estimate, two refinements, quotient,
// residual correction); for such s its operand scaling and special-value fixup
// leave every value unchanged, so they are left out. A NaN s gives NaN.
PHD Real recip_of_root(Real s) {
#ifdef __HIP_DEVICE_COMPILE__
  Real r = __builtin_amdgcn_rcp(s);
  Real e = fma(-s, r, 1.0);
  r = fma(r, e, r);
  e = fma(-s, r, 1.0);
  r = fma(r, e, r);
  const Real q = 1.0 * r;
  e = fma(-s, q, 1.0);
  return fma(e, r, q);
#else
  return 1.0 / s;
#endif
}

// ---- correctly rounded sin and cos ----
// The reference calls the platform libm (Rust f64::sin_cos); libms differ from
// each other by up to one ulp, and through contact one ulp grows into visible
// trajectory differences. Both the device and the oracle therefore evaluate
// the correctly rounded sin/cos: argument reduction by a three-double pi/2
// (Cody-Waite with exact products) and Taylor series in double-double
// arithmetic (terms 0..8) with a double tail (terms 9..14). The result is
// within ~2^-100 relative of the exact value before the final rounding, so it
// equals the exactly rounded value except at ties closer than that.
struct DD {
  Real hi, lo;
};
PHD DD dd_fast_two_sum(Real a, Real b) {
  Real s = a + b;
  DD r;
  r.hi = s;
  r.lo = b - (s - a);
  return r;
}
PHD DD dd_two_sum(Real a, Real b) {
  Real s = a + b;
  Real bb = s - a;
  DD r;
  r.hi = s;
  r.lo = (a - (s - bb)) + (b - bb);
  return r;
}
PHD DD dd_add(DD a, DD b) {
  DD s = dd_two_sum(a.hi, b.hi);
  DD t = dd_two_sum(a.lo, b.lo);
  Real e = s.lo + t.hi;
  DD u = dd_fast_two_sum(s.hi, e);
  e = u.lo + t.lo;
  return dd_fast_two_sum(u.hi, e);
}
PHD DD dd_mul(DD a, DD b) {
  Real p = a.hi * b.hi;
  Real e = fma(a.hi, b.hi, -p);
  e = e + (a.hi * b.lo + a.lo * b.hi);
  return dd_fast_two_sum(p, e);
}
PHD DD dd_from(Real hi, Real lo) {
  DD r;
  r.hi = hi;
  r.lo = lo;
  return r;
}
PHD void cr_sincos(Real x, Real* sout, Real* cout) {
  if (!rfinite(x)) {
    *sout = x - x;
    *cout = x - x;
    return;
  }
  if (x == 0.0) {  // sin(-0) = -0
    *sout = x;
    *cout = 1.0;
    return;
  }
  const Real P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
  Real k = rint(x * 0.6366197723675814);
  DD r = dd_from(x, 0.0);
  Real h = k * P1;
  r = dd_add(r, dd_from(-h, -fma(k, P1, -h)));
  h = k * P2;
  r = dd_add(r, dd_from(-h, -fma(k, P2, -h)));
  r = dd_add(r, dd_from(-(k * P3), 0.0));
  DD r2 = dd_mul(r, r);
  const Real S[15][2] = {{1.0, 0.0},
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
  const Real Cc[15][2] = {{1.0, 0.0},
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
  Real ts = S[14][0], tc = Cc[14][0];
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
  Real sv = s.hi, cv = pc.hi;
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

struct V3 {
  Real x, y, z;
  PHD Real operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
PHD V3 v3(Real x, Real y, Real z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
PHD V3 v3z() { return v3(0.0, 0.0, 0.0); }
PHD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
PHD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
PHD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
PHD V3 operator*(V3 a, Real s) { return v3(a.x * s, a.y * s, a.z * s); }
PHD Real dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PHD V3 cross(V3 a, V3 o) { return v3(a.y * o.z - a.z * o.y, a.z * o.x - a.x * o.z, a.x * o.y - a.y * o.x); }
PHD Real norm2(V3 a) { return dot(a, a); }
PHD Real norm(V3 a) { return sqrt_cr(norm2(a)); }
PHD V3 normalized(V3 a) {
  Real n = norm(a);
  if (n > 1e-15) return a * (1.0 / n);
  return v3z();
}
PHD V3 vabs(V3 a) { return v3(fabs(a.x), fabs(a.y), fabs(a.z)); }
PHD V3 mul_elem(V3 a, V3 o) { return v3(a.x * o.x, a.y * o.y, a.z * o.z); }
PHD V3 vmax(V3 a, V3 o) { return v3(rmax(a.x, o.x), rmax(a.y, o.y), rmax(a.z, o.z)); }
template <class P>
PHD V3 ld3(P p) { return v3(p[0], p[1], p[2]); }
template <class P>
PHD void st3(P p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
PHD void tangents(V3 n, V3* t1, V3* t2) {
  V3 a = fabs(n.x) < 0.57 ? v3(1, 0, 0) : v3(0, 1, 0);
  *t1 = normalized(cross(n, a));
  *t2 = cross(n, *t1);
}

// Row-major 3x3.
struct M3 {
  Real m[9];
};
PHD M3 m3id() { M3 r = {{1, 0, 0, 0, 1, 0, 0, 0, 1}}; return r; }
template <class P>
PHD M3 ldm(P p) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = p[i];
  return r;
}
template <class P>
PHD void stm(P p, const M3& r) {
  for (int i = 0; i < 9; i++) p[i] = r.m[i];
}
PHD V3 col(const M3& a, int j) { return v3(a.m[j], a.m[3 + j], a.m[6 + j]); }
PHD V3 mul_vec(const M3& a, V3 v) {
  return v3(a.m[0] * v.x + a.m[1] * v.y + a.m[2] * v.z, a.m[3] * v.x + a.m[4] * v.y + a.m[5] * v.z,
            a.m[6] * v.x + a.m[7] * v.y + a.m[8] * v.z);
}
PHD V3 tmul_vec(const M3& a, V3 v) {
  return v3(a.m[0] * v.x + a.m[3] * v.y + a.m[6] * v.z, a.m[1] * v.x + a.m[4] * v.y + a.m[7] * v.z,
            a.m[2] * v.x + a.m[5] * v.y + a.m[8] * v.z);
}
// math.rs:169-177 (each entry an iterator sum from -0.0)
PHD M3 mul_mat(const M3& a, const M3& o) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Real s = rsum0();
      for (int k = 0; k < 3; k++) s += a.m[3 * i + k] * o.m[3 * k + j];
      r.m[3 * i + j] = s;
    }
  return r;
}
PHD M3 transpose(const M3& a) {
  M3 r = {{a.m[0], a.m[3], a.m[6], a.m[1], a.m[4], a.m[7], a.m[2], a.m[5], a.m[8]}};
  return r;
}
PHD M3 skew(V3 v) {
  M3 r = {{0.0, -v.z, v.y, v.z, 0.0, -v.x, -v.y, v.x, 0.0}};
  return r;
}
PHD M3 mscale(const M3& a, Real s) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = a.m[i] * s;
  return r;
}
PHD M3 madd(const M3& a, const M3& o) {
  M3 r;
  for (int i = 0; i < 9; i++) r.m[i] = a.m[i] + o.m[i];
  return r;
}
PHD M3 mrotate(const M3& self, const M3& r) { return mul_mat(mul_mat(r, self), transpose(r)); }

// Quaternion (w, x, y, z).
struct Q4 {
  Real w, x, y, z;
};
PHD Q4 q4(Real w, Real x, Real y, Real z) { Q4 r; r.w = w; r.x = x; r.y = y; r.z = z; return r; }
template <class P>
PHD Q4 ldq(P p) { return q4(p[0], p[1], p[2], p[3]); }
template <class P>
PHD void stq(P p, Q4 q) { p[0] = q.w; p[1] = q.x; p[2] = q.y; p[3] = q.z; }
PHD Q4 qnormalized(Q4 q) {
  Real s = rsum0();
  s += q.w * q.w;
  s += q.x * q.x;
  s += q.y * q.y;
  s += q.z * q.z;
  Real n = sqrt_cr(s);
  if (n < 1e-15) return q4(1, 0, 0, 0);
  return q4(q.w / n, q.x / n, q.y / n, q.z / n);
}
PHD Q4 qmul(Q4 p, Q4 o) {
  Real a = p.w, b = p.x, c = p.y, d = p.z, e = o.w, f = o.x, g = o.y, h = o.z;
  return q4(a * e - b * f - c * g - d * h, a * f + b * e + c * h - d * g, a * g - b * h + c * e + d * f,
            a * h + b * g - c * f + d * e);
}
PHD Q4 from_axis_angle(V3 axis, Real angle) {
  V3 a = normalized(axis);
  Real s, c;
  cr_sincos(0.5 * angle, &s, &c);
  return q4(c, a.x * s, a.y * s, a.z * s);
}
PHD M3 to_mat(Q4 q) {
  Real w = q.w, x = q.x, y = q.y, z = q.z;
  Real ww = w * w, xx = x * x, yy = y * y, zz = z * z;
  Real xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
  M3 r = {{ww + xx - yy - zz, 2.0 * (xy - wz), 2.0 * (xz + wy), 2.0 * (xy + wz), ww - xx + yy - zz, 2.0 * (yz - wx),
           2.0 * (xz - wy), 2.0 * (yz + wx), ww - xx - yy + zz}};
  return r;
}
PHD V3 qrotate(Q4 q, V3 v) { return mul_vec(to_mat(q), v); }
PHD Q4 integrate_local(Q4 self, V3 omega, Real h) {
  Real n = norm(omega);
  if (n * h < 1e-15) return self;
  return qnormalized(qmul(self, from_axis_angle(omega * (1.0 / n), n * h)));
}

// Spatial vector [ang; lin] stored as 6 consecutive doubles.
struct SP {
  V3 a, l;
};
PHD SP sp(V3 a, V3 l) { SP s; s.a = a; s.l = l; return s; }
template <class P>
PHD SP ldsp(P p) { return sp(ld3(p), ld3(p + 3)); }
template <class P>
PHD void stsp(P p, const SP& s) { st3(p, s.a); st3(p + 3, s.l); }
PHD Real sdot(const SP& a, const SP& o) { return dot(a.a, o.a) + dot(a.l, o.l); }
PHD SP sscale(const SP& a, Real s) { return sp(a.a * s, a.l * s); }
PHD SP spadd(const SP& a, const SP& o) { return sp(a.a + o.a, a.l + o.l); }
PHD SP cross_motion(const SP& s, const SP& m) { return sp(cross(s.a, m.a), cross(s.a, m.l) + cross(s.l, m.a)); }
PHD SP cross_force(const SP& s, const SP& f) { return sp(cross(s.a, f.a) + cross(s.l, f.l), cross(s.a, f.l)); }
PHD V3 point_velocity(const SP& s, V3 p) { return s.l + cross(s.a, p); }

// Spatial inertia {mass, m*c, I_O} stored as 13 doubles: [mass, mc(3), I(9)].
template <class P>
PHD SP mul_motion(P I, const SP& v) {
  M3 io = ldm(I + 4);
  V3 mc = ld3(I + 1);
  Real mass = I[0];
  return sp(mul_vec(io, v.a) + cross(mc, v.l), v.l * mass - cross(mc, v.a));
}

// simd.rs:14-64 (x86 branch), over a strided pointer.
template <class PA, class PB>
PHD Real sdot_n(PA a, PB b, int n) {
  if (n < 8) {
    Real s = 0.0;
    for (int k = 0; k < n; k++) s += a[k] * b[k];
    return s;
  }
  Real s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int k = 0;
  while (k + 4 <= n) {
    s0 += a[k] * b[k];
    s1 += a[k + 1] * b[k + 1];
    s2 += a[k + 2] * b[k + 2];
    s3 += a[k + 3] * b[k + 3];
    k += 4;
  }
  Real t = (s0 + s1) + (s2 + s3);
  while (k < n) {
    t += a[k] * b[k];
    k++;
  }
  return t;
}

}  // namespace phk
