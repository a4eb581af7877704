// dual.hpp -- forward-mode dual numbers (value + one directional derivative) in FP64.
//
// Arithmetic follows the calculus rules, written out so that host and device builds round the same
// way (no contraction: the library is built with -ffp-contract=off):
//   (a b)' = a' b + a b'        (a / b)' = (a' - (a/b) b') / b
//   sqrt(a)' = a' / (2 sqrt(a)), taken as 0 where sqrt(a) == 0
//   |a|' = -a' for a < 0, else a'
// Comparisons look at the value only, so max / min / clamp and every branch pick one operand and
// carry its derivative: the derivative of the branch the value takes.
#pragma once

#include <hip/hip_runtime.h>

namespace aero_dev {

struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants promote implicitly
  __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
  __host__ __device__ Dual& operator+=(const Dual& b) {
    v += b.v;
    d += b.d;
    return *this;
  }
  __host__ __device__ Dual& operator-=(const Dual& b) {
    v -= b.v;
    d -= b.d;
    return *this;
  }
};

__host__ __device__ __forceinline__ Dual operator-(const Dual& a) { return Dual(-a.v, -a.d); }
__host__ __device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
// Mixed forms with a constant: no derivative term for the constant.
__host__ __device__ __forceinline__ Dual operator+(const Dual& a, double b) { return Dual(a.v + b, a.d); }
__host__ __device__ __forceinline__ Dual operator+(double a, const Dual& b) { return Dual(a + b.v, b.d); }
__host__ __device__ __forceinline__ Dual operator-(const Dual& a, double b) { return Dual(a.v - b, a.d); }
__host__ __device__ __forceinline__ Dual operator-(double a, const Dual& b) { return Dual(a - b.v, -b.d); }
__host__ __device__ __forceinline__ Dual operator*(const Dual& a, double b) { return Dual(a.v * b, a.d * b); }
__host__ __device__ __forceinline__ Dual operator*(double a, const Dual& b) { return Dual(a * b.v, a * b.d); }
__host__ __device__ __forceinline__ Dual operator/(const Dual& a, double b) { return Dual(a.v / b, a.d / b); }
__host__ __device__ __forceinline__ Dual operator/(double a, const Dual& b) {
  const double q = a / b.v;
  return Dual(q, (-(q * b.d)) / b.v);
}

__host__ __device__ __forceinline__ bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
__host__ __device__ __forceinline__ bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
__host__ __device__ __forceinline__ bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
__host__ __device__ __forceinline__ bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
__host__ __device__ __forceinline__ bool operator<(const Dual& a, double b) { return a.v < b; }
__host__ __device__ __forceinline__ bool operator>(const Dual& a, double b) { return a.v > b; }
__host__ __device__ __forceinline__ bool operator<=(const Dual& a, double b) { return a.v <= b; }
__host__ __device__ __forceinline__ bool operator>=(const Dual& a, double b) { return a.v >= b; }

__device__ __forceinline__ Dual vsqrt(const Dual& a) {
  const double s = sqrt(a.v);
  return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
__device__ __forceinline__ Dual vabs(const Dual& a) { return a.v < 0.0 ? -a : a; }

}  // namespace aero_dev
