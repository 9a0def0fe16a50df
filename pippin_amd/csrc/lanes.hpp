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

// Clamp an index into [0, n): lets a lane load unconditionally (every address
// in bounds, no branch around the load) and select the value afterwards.
PHD int clampi(int i, int n) { return i < n ? i : n - 1; }

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
__device__ __forceinline__ DLV operator*(Real a, DLV b) { return {a * b.v}; }
__device__ __forceinline__ DLV operator*(DLV a, Real b) { return {a.v * b}; }
__device__ __forceinline__ DLV operator+(DLV a, Real b) { return {a.v + b}; }

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
  __device__ __forceinline__ static DLV sqrt_(DLV a) { return {sqrt_cr(a.v)}; }
  __device__ __forceinline__ static DLV recip_root_(DLV a) { return {recip_of_root(a.v)}; }
  // the value of a group-uniform vector
  __device__ __forceinline__ static Real uni(DLV a) { return a.v; }
  // bit l set when pred holds on lane l of this lane's 16-lane row
  template <class P>
  __device__ __forceinline__ static uint32_t ballot(P&& pred) {
    uint64_t m = __ballot(pred(lane()) ? 1 : 0);
    return (uint32_t)((m >> (threadIdx.x & 48u)) & 0xFFFFu);
  }
  // lane J's value in every lane of this lane's 16-lane row: one 64-bit DPP
  // move (v_mov_b64_dpp row_newbcast:J; every source lane is valid)
  template <int J>
  __device__ __forceinline__ static DLV bc(DLV x) {
    long v = __builtin_bit_cast(long, x.v);
    v = __builtin_amdgcn_mov_dpp(v, 0x150 + J, 0xF, 0xF, true);
    return {__builtin_bit_cast(double, v)};
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
inline HLV operator*(Real a, const HLV& b) {
  HLV r;
  for (int l = 0; l < GL; l++) r.v[l] = a * b.v[l];
  return r;
}
inline HLV operator*(const HLV& a, Real b) {
  HLV r;
  for (int l = 0; l < GL; l++) r.v[l] = a.v[l] * b;
  return r;
}
inline HLV operator+(const HLV& a, Real b) {
  HLV r;
  for (int l = 0; l < GL; l++) r.v[l] = a.v[l] + b;
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
  static HLV recip_root_(HLV a) {
    HLV r;
    for (int l = 0; l < GL; l++) r.v[l] = 1.0 / a.v[l];
    return r;
  }
  static Real uni(const HLV& a) { return a.v[0]; }
  template <class P>
  static uint32_t ballot(P&& pred) {
    uint32_t m = 0;
    for (int l = 0; l < GL; l++)
      if (pred(l)) m |= 1u << l;
    return m;
  }
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

// out[k] = lane k's x, in every lane
template <class X, class LV>
PHD void bcast_all(LV x, LV* out) {
  sfor<0, GL>([&](auto KC) {
    constexpr int k = decltype(KC)::value;
    out[k] = X::template bc<k>(x);
    return true;
  });
}

// sdot_n (simd.rs:14-64) of a runtime length n <= 16 over register arrays;
// the structure switches on n (uniform), the unrolled terms stay in order.
template <class X, class LV>
PHD LV lsdot_rt(const LV* a, const LV* b, int n) {
  if (n < 8) {
    LV s = X::splat(0.0);
#pragma unroll
    for (int k = 0; k < 7; k++)
      if (k < n) s = s + a[k] * b[k];
    return s;
  }
  LV s0 = X::splat(0.0), s1 = s0, s2 = s0, s3 = s0;
  const int nb = n / 4 * 4;
#pragma unroll
  for (int k = 0; k < GL; k += 4)
    if (k + 4 <= nb) {
      s0 = s0 + a[k] * b[k];
      s1 = s1 + a[k + 1] * b[k + 1];
      s2 = s2 + a[k + 2] * b[k + 2];
      s3 = s3 + a[k + 3] * b[k + 3];
    }
  LV t = (s0 + s1) + (s2 + s3);
#pragma unroll
  for (int k = 8; k < GL; k++)
    if (k >= nb && k < n) t = t + a[k] * b[k];
  return t;
}

// Sequential sum of lane values 0..n-1 starting from `start` (Rust's
// iterator sum folds from -0.0), computed redundantly in every lane.
template <class X, class LV>
PHD LV lsum_lanes(LV x, int n, Real start) {
  LV xb[GL];
  bcast_all<X>(x, xb);
  LV s = X::splat(start);
#pragma unroll
  for (int k = 0; k < GL; k++)
    if (k < n) s = s + xb[k];
  return s;
}

// math.rs:425-447 with one row per lane: in-place lower Cholesky of the
// diagonal block [s0, s0 + n) (n <= 16) of the row-major stride-nv matrix A.
// Column j: lanes receive row j (entries < j) by broadcast, each lane forms its
// dot with row j, lane j's diagonal is broadcast, square-rooted and inverted,
// and rows below scale their entry j. Only the lower triangle is written, as in
// the reference. Returns false when a pivot is not positive.
template <class X, class LV>
PHD bool chol_rows(LV* r, int n) {
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
    LV inv = X::recip_root_(ljj);  // 1.0 / ljj
    r[j] = X::sel([](int l) { return l == j; }, ljj, X::sel([](int l) { return l > j; }, d * inv, r[j]));
    return true;
  });
  return ok;
}

// Store the lower triangle (with the diagonal) of register rows into the
// diagonal block [s0, s0 + n) of A.
template <class X, class LV>
PHD void store_lower(const LV* r, Real* A, int nv, int s0, int n) {
  X::each([&](int l) {
    if (l < n) {
#pragma unroll
      for (int k = 0; k < GL; k++)
        if (k <= l) A[(s0 + l) * nv + s0 + k] = X::at(r[k], l);
    }
  });
}

template <class X>
PHD bool chol_lanes(Real* A, int nv, int s0, int n) {
  using LV = typename X::LV;
  LV r[GL];
#pragma unroll
  for (int k = 0; k < GL; k++)
    r[k] = X::gen([&](int l) {
      Real v = A[(s0 + clampi(l, n)) * nv + s0 + clampi(k, n)];
      return (l < n && k < n) ? v : 0.0;
    });
  if (!chol_rows<X>(r, n)) return false;
  store_lower<X>(r, A, nv, s0, n);
  return true;
}

// math.rs:449-463 with one row per lane: forward substitution (lane i forms
// its dot with the solved prefix, divides, and broadcasts), then backward
// substitution (lane i divides and broadcasts; lanes above update with
// column entries they loaded up front). x is overwritten with the solution.
template <class X, class LV>
PHD LV solve_rows(const LV* r, const LV* c, LV lii, LV xv, int n) {
  LV xs[GL];
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
  return xv;
}

// Load, for lane l, row l of the strict lower triangle (r), column l below the
// diagonal (c) and the diagonal entry of the factor in block [s0, s0 + n).
template <class X, class LV>
PHD void load_factor(const Real* L, int nv, int s0, int n, LV* r, LV* c, LV* lii) {
#pragma unroll
  for (int k = 0; k < GL; k++) {
    r[k] = X::gen([&](int l) {
      Real v = L[(s0 + clampi(l, n)) * nv + s0 + clampi(k, n)];
      return (l < n && k < l) ? v : 0.0;
    });
    c[k] = X::gen([&](int l) {
      Real v = L[(s0 + clampi(k, n)) * nv + s0 + clampi(l, n)];
      return (k < n && k > l) ? v : 0.0;
    });
  }
  *lii = X::gen([&](int l) {
    int m = clampi(l, n);
    Real v = L[(s0 + m) * nv + s0 + m];
    return l < n ? v : 1.0;
  });
}

template <class X>
PHD void solve_lanes(const Real* L, int nv, int s0, int n, Real* x) {
  using LV = typename X::LV;
  LV r[GL], c[GL], lii;
  load_factor<X>(L, nv, s0, n, r, c, &lii);
  LV xv = X::gen([&](int l) {
    Real v = x[s0 + clampi(l, n)];
    return l < n ? v : 0.0;
  });
  xv = solve_rows<X>(r, c, lii, xv, n);
  X::each([&](int l) {
    if (l < n) x[s0 + l] = X::at(xv, l);
  });
}

}  // namespace phk
