// ORACLE — test infrastructure only. Included by math.hpp when the oracle is
// built with -DORC_OPCOUNT: `Real` becomes a double that counts its FP64
// arithmetic (+ - * / sqrt pow: 1 each, sin and cos: 1 each; comparisons,
// negation, fabs, min/max and copies: 0). Running the restated engine in this
// build gives the algorithmic FP64 operation count of Pippin's step on a given
// trajectory, the numerator of the kernels' FLOP/s.
#pragma once
#include <cmath>

namespace orc {

extern thread_local unsigned long long g_ops;

struct CReal {
  double v;
  CReal() = default;
  CReal(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator float() const { return (float)v; }
  CReal& operator+=(CReal o) {
    g_ops++;
    v += o.v;
    return *this;
  }
  CReal& operator-=(CReal o) {
    g_ops++;
    v -= o.v;
    return *this;
  }
  CReal& operator*=(CReal o) {
    g_ops++;
    v *= o.v;
    return *this;
  }
  CReal& operator/=(CReal o) {
    g_ops++;
    v /= o.v;
    return *this;
  }
};
typedef CReal Real;
inline CReal operator+(CReal a, CReal b) { g_ops++; return CReal(a.v + b.v); }
inline CReal operator-(CReal a, CReal b) { g_ops++; return CReal(a.v - b.v); }
inline CReal operator*(CReal a, CReal b) { g_ops++; return CReal(a.v * b.v); }
inline CReal operator/(CReal a, CReal b) { g_ops++; return CReal(a.v / b.v); }
inline CReal operator-(CReal a) { return CReal(-a.v); }
inline bool operator<(CReal a, CReal b) { return a.v < b.v; }
inline bool operator>(CReal a, CReal b) { return a.v > b.v; }
inline bool operator<=(CReal a, CReal b) { return a.v <= b.v; }
inline bool operator>=(CReal a, CReal b) { return a.v >= b.v; }
inline bool operator==(CReal a, CReal b) { return a.v == b.v; }
inline bool operator!=(CReal a, CReal b) { return a.v != b.v; }
static inline Real r_sqrt(Real x) { g_ops++; return std::sqrt(x.v); }
static inline Real r_fabs(Real x) { return std::fabs(x.v); }
static inline Real r_pow(Real a, Real b) { g_ops++; return std::pow(a.v, b.v); }
static inline bool r_isfinite(Real x) { return std::isfinite(x.v); }
static inline bool r_isnan(Real x) { return std::isnan(x.v); }
static inline Real r_copysign(Real a, Real b) { return std::copysign(a.v, b.v); }
static inline Real r_fmax(Real a, Real b) { return std::fmax(a.v, b.v); }
static inline Real r_fmin(Real a, Real b) { return std::fmin(a.v, b.v); }
static inline double r_val(Real x) { return x.v; }
static inline void r_count(int n) { g_ops += (unsigned long long)n; }

}  // namespace orc
