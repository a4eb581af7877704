// oracle_dual.hpp -- forward-mode dual numbers for the oracle (test infrastructure only).
//
// Instantiating the oracle's residual on ODual gives the CPU tangent R_U dU + R_p dp of the restated
// algorithm: every value carries one directional derivative, and comparisons (hence max / min / clamp
// and every branch) look at the value only.  The derivative rules:
//   (a b)' = a' b + a b',  (a / b)' = (a' - (a/b) b') / b,  sqrt(a)' = a' / (2 sqrt a) (0 at 0),
//   |a|' = -a' for a < 0 else a'.
#pragma once

#include <cmath>

#include "aero_oracle.hpp"

namespace aero_oracle {

struct ODual {
  double v = 0.0, d = 0.0;
  constexpr ODual() = default;
  constexpr ODual(double x) : v(x), d(0.0) {}  // NOLINT: constants promote implicitly
  constexpr ODual(double x, double dx) : v(x), d(dx) {}
  ODual& operator+=(const ODual& b) {
    v += b.v;
    d += b.d;
    return *this;
  }
  ODual& operator-=(const ODual& b) {
    v -= b.v;
    d -= b.d;
    return *this;
  }
  ODual& operator*=(const ODual& b) {
    d = d * b.v + v * b.d;
    v = v * b.v;
    return *this;
  }
};

inline ODual operator-(const ODual& a) { return ODual(-a.v, -a.d); }
inline ODual operator+(const ODual& a, const ODual& b) { return ODual(a.v + b.v, a.d + b.d); }
inline ODual operator-(const ODual& a, const ODual& b) { return ODual(a.v - b.v, a.d - b.d); }
inline ODual operator*(const ODual& a, const ODual& b) { return ODual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline ODual operator/(const ODual& a, const ODual& b) {
  const double q = a.v / b.v;
  return ODual(q, (a.d - q * b.d) / b.v);
}
inline bool operator<(const ODual& a, const ODual& b) { return a.v < b.v; }
inline bool operator>(const ODual& a, const ODual& b) { return a.v > b.v; }
inline bool operator<=(const ODual& a, const ODual& b) { return a.v <= b.v; }
inline bool operator>=(const ODual& a, const ODual& b) { return a.v >= b.v; }
inline bool operator==(const ODual& a, const ODual& b) { return a.v == b.v; }
inline bool operator!=(const ODual& a, const ODual& b) { return a.v != b.v; }

inline ODual ao_sqrt(const ODual& a) {
  const double s = std::sqrt(a.v);
  return ODual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
inline ODual ao_abs(const ODual& a) { return a.v < 0.0 ? -a : a; }
inline bool ao_isfinite(const ODual& a) { return std::isfinite(a.v); }
inline double ao_value(const ODual& a) { return a.v; }

template <> constexpr ODual kc<ODual>(float, double d) { return ODual(d); }

}  // namespace aero_oracle
