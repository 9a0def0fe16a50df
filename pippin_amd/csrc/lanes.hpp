// Per-lane values of a 16-lane environment group.
//
// Small dense linear algebra (nv <= 16) runs with one matrix row per lane and
// the rows exchanged by broadcasts inside the group. The same source serves
// two policies:
//   device: DLV holds this lane's value; the group is one 16-lane DPP row and
//           bc<J> is a row_newbcast from lane J of the row;
//   host:   HLV holds all 16 lanes' values and every operation is applied
//           lane by lane, which is the serial emulation of the device code.
// Arithmetic is elementwise IEEE on both, so results agree bit for bit.
#pragma once
#include <type_traits>

#include "dmath.hpp"

namespace phk {

constexpr int GL = 16;  // lanes per environment group

// compile-time loop: f(integral_constant<int, I>) for I in [B, E) while f returns true
template <int B, int E, class F>
PHD void sfor(F&& f) {
  if constexpr (B < E) {
    if (f(std::integral_constant<int, B>{})) sfor<B + 1, E>(f);
  }
}
// descending: I = E-1 down to B
template <int B, int E, class F>
PHD void sfor_down(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, E - 1>{});
    sfor_down<B, E - 1>(f);
  }
}

// ---------------- device lanes ----------------
struct DLV {
  Real v;
};
__device__ __forceinline__ DLV operator+(DLV a, DLV b) { return {a.v + b.v}; }
__device__ __forceinline__ DLV operator-(DLV a, DLV b) { return {a.v - b.v}; }
__device__ __forceinline__ DLV operator*(DLV a, DLV b) { return {a.v * b.v}; }
__device__ __forceinline__ DLV operator/(DLV a, DLV b) { return {a.v / b.v}; }
__device__ __forceinline__ DLV operator/(Real a, DLV b) { return {a / b.v}; }
__device__ __forceinline__ DLV operator-(DLV a) { return {-a.v}; }

struct DevLanes {
  using LV = DLV;
  __device__ __forceinline__ static int lane() { return (int)(threadIdx.x & (GL - 1)); }
  __device__ __forceinline__ static DLV splat(Real c) { return {c}; }
  template <class F>
  __device__ __forceinline__ static DLV gen(F&& f) {
    return {f(lane())};
  }
  template <class F>
  __device__ __forceinline__ static void each(F&& f) {
    f(lane());
  }
  __device__ __forceinline__ static Real at(DLV x, int) { return x.v; }
  // lane l takes a if pred(l) else b
  template <class P>
  __device__ __forceinline__ static DLV sel(P&& pred, DLV a, DLV b) {
    return pred(lane()) ? a : b;
  }
  __device__ __forceinline__ static DLV sqrt_(DLV a) { return {sqrt(a.v)}; }
  // the value of a group-uniform vector
  __device__ __forceinline__ static Real uni(DLV a) { return a.v; }
  template <int J>
  __device__ __forceinline__ static DLV bc(DLV x) {
    int lo = __double2loint(x.v), hi = __double2hiint(x.v);
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x150 + J, 0xF, 0xF, false);  // row_newbcast:J
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x150 + J, 0xF, 0xF, false);
    return {__hiloint2double(hi, lo)};
  }
};

// ---------------- host lanes (serial emulation) ----------------
struct HLV {
  Real v[GL];
};
#define PH_HLV_BIN(op)                                          \
  inline HLV operator op(const HLV& a, const HLV& b) {          \
    HLV r;                                                      \
    for (int l = 0; l < GL; l++) r.v[l] = a.v[l] op b.v[l];     \
    return r;                                                   \
  }
PH_HLV_BIN(+)
PH_HLV_BIN(-)
PH_HLV_BIN(*)
PH_HLV_BIN(/)
#undef PH_HLV_BIN
inline HLV operator/(Real a, const HLV& b) {
  HLV r;
  for (int l = 0; l < GL; l++) r.v[l] = a / b.v[l];
  return r;
}
inline HLV operator-(const HLV& a) {
  HLV r;
  for (int l = 0; l < GL; l++) r.v[l] = -a.v[l];
  return r;
}

struct HostLanes {
  using LV = HLV;
  static HLV splat(Real c) {
    HLV r;
    for (int l = 0; l < GL; l++) r.v[l] = c;
    return r;
  }
  template <class F>
  static HLV gen(F&& f) {
    HLV r;
    for (int l = 0; l < GL; l++) r.v[l] = f(l);
    return r;
  }
  template <class F>
  static void each(F&& f) {
    for (int l = 0; l < GL; l++) f(l);
  }
  static Real at(const HLV& x, int l) { return x.v[l]; }
  template <class P>
  static HLV sel(P&& pred, const HLV& a, const HLV& b) {
    HLV r;
    for (int l = 0; l < GL; l++) r.v[l] = pred(l) ? a.v[l] : b.v[l];
    return r;
  }
  static HLV sqrt_(const HLV& a) {
    HLV r;
    for (int l = 0; l < GL; l++) r.v[l] = sqrt(a.v[l]);
    return r;
  }
  static Real uni(const HLV& a) { return a.v[0]; }
  template <int J>
  static HLV bc(const HLV& x) {
    return splat(x.v[J]);
  }
};

// simd.rs:14-64 (x86 branch) over lanes: sum_k a[k] * b[k] for a compile-time
// length N, in the reference's order (sequential below 8, else four strided
// accumulators combined as (s0 + s1) + (s2 + s3), then the tail).
template <class X, int N, class LV>
PHD LV lsdot(const LV* a, const LV* b) {
  if constexpr (N < 8) {
    LV s = X::splat(0.0);
#pragma unroll
    for (int k = 0; k < N; k++) s = s + a[k] * b[k];
    return s;
  } else {
    LV s0 = X::splat(0.0), s1 = s0, s2 = s0, s3 = s0;
    constexpr int NB = N / 4 * 4;
#pragma unroll
    for (int k = 0; k < NB; k += 4) {
      s0 = s0 + a[k] * b[k];
      s1 = s1 + a[k + 1] * b[k + 1];
      s2 = s2 + a[k + 2] * b[k + 2];
      s3 = s3 + a[k + 3] * b[k + 3];
    }
    LV t = (s0 + s1) + (s2 + s3);
#pragma unroll
    for (int k = NB; k < N; k++) t = t + a[k] * b[k];
    return t;
  }
}

// math.rs:425-447 with one row per lane: in-place lower Cholesky of the
// diagonal block [s0, s0 + n) (n <= 16) of the row-major stride-nv matrix A.
// Column j: lanes receive row j (entries < j) by broadcast, each lane forms its
// dot with row j, lane j's diagonal is broadcast, square-rooted and inverted,
// and rows below scale their entry j. Only the lower triangle is written, as in
// the reference. Returns false when a pivot is not positive.
template <class X>
PHD bool chol_lanes(Real* A, int nv, int s0, int n) {
  using LV = typename X::LV;
  LV r[GL];
#pragma unroll
  for (int k = 0; k < GL; k++)
    r[k] = X::gen([&](int l) { return (l < n && k < n) ? A[(s0 + l) * nv + s0 + k] : 0.0; });
  bool ok = true;
  sfor<0, GL>([&](auto JC) {
    constexpr int j = decltype(JC)::value;
    if (j >= n) return false;
    LV b[j > 0 ? j : 1];
#pragma unroll
    for (int k = 0; k < j; k++) b[k] = X::template bc<j>(r[k]);
    LV d = r[j] - lsdot<X, j>(r, b);
    LV dj = X::template bc<j>(d);
    if (X::uni(dj) <= 0.0) {  // a NaN pivot passes, as in the reference
      ok = false;
      return false;
    }
    LV ljj = X::sqrt_(dj);
    LV inv = 1.0 / ljj;
    r[j] = X::sel([](int l) { return l == j; }, ljj, X::sel([](int l) { return l > j; }, d * inv, r[j]));
    return true;
  });
  if (!ok) return false;
  X::each([&](int l) {
    if (l < n) {
#pragma unroll
      for (int k = 0; k < GL; k++)
        if (k <= l) A[(s0 + l) * nv + s0 + k] = X::at(r[k], l);
    }
  });
  return true;
}

// math.rs:449-463 with one row per lane: forward substitution (lane i forms
// its dot with the solved prefix, divides, and broadcasts), then backward
// substitution (lane i divides and broadcasts; lanes above update with
// column entries they loaded up front). x is overwritten with the solution.
template <class X>
PHD void solve_lanes(const Real* L, int nv, int s0, int n, Real* x) {
  using LV = typename X::LV;
  LV r[GL], c[GL], xs[GL];
#pragma unroll
  for (int k = 0; k < GL; k++) {
    r[k] = X::gen([&](int l) { return (l < n && k < l) ? L[(s0 + l) * nv + s0 + k] : 0.0; });
    c[k] = X::gen([&](int l) { return (k < n && k > l) ? L[(s0 + k) * nv + s0 + l] : 0.0; });
  }
  LV lii = X::gen([&](int l) { return l < n ? L[(s0 + l) * nv + s0 + l] : 1.0; });
  LV xv = X::gen([&](int l) { return l < n ? x[s0 + l] : 0.0; });
  sfor<0, GL>([&](auto IC) {
    constexpr int i = decltype(IC)::value;
    if (i >= n) return false;
    LV t = (xv - lsdot<X, i>(r, xs)) / lii;
    xs[i] = X::template bc<i>(t);
    xv = X::sel([](int l) { return l == i; }, t, xv);
    return true;
  });
  sfor_down<0, GL>([&](auto IC) {
    constexpr int i = decltype(IC)::value;
    if (i >= n) return;
    LV q = xv / lii;
    LV xi = X::template bc<i>(q);
    xv = X::sel([](int l) { return l == i; }, q, X::sel([](int l) { return l < i; }, xv + (-xi) * c[i], xv));
  });
  X::each([&](int l) {
    if (l < n) x[s0 + l] = X::at(xv, l);
  });
}

}  // namespace phk
