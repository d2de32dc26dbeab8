// Scalar types and precision policies of the per-cell chain.
//
// The chain's code is written once over two types: R for the bulk of the
// arithmetic and W ("wide") for the parts that keep FP64 in the mixed build:
// the analytic H2SO4 update exp(-k dt), the Vehkamaki ln J polynomials,
// renaming's erfc tail fractions, calcsize's lognormal moment conversions, the
// SOA sub-stepping and the mass sums that form the grid-cell result.
//
//   Fp64    R = W = double          the drop-in
//   Mixed   R = F32, W = double     FP32 bulk
//   Tangent R = W = Dual            forward-mode directional derivatives
//
// F32 wraps a float so that the double literals and double constants the
// chain is written with are converted to float on use instead of promoting
// the expression to double.  Its division is a reciprocal times the
// numerator (<= 2 ulp) and its transcendentals are the single-precision
// device functions.  Dual carries a value and one derivative.
#pragma once

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#else
#include <cmath>
#endif

#include "chain_consts.hpp"

namespace mam4x {
namespace num {

// ------------------------------------------------------------------ F32
struct F32 {
  float v;
  F32() = default;
  __host__ __device__ constexpr F32(double x) : v((float)x) {}
  __host__ __device__ constexpr F32(float x) : v(x) {}
  __host__ __device__ constexpr F32(int x) : v((float)x) {}
  __host__ __device__ explicit constexpr operator double() const { return v; }
};

__host__ __device__ __forceinline__ float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

__host__ __device__ __forceinline__ F32 operator+(F32 a, F32 b) { return F32(a.v + b.v); }
__host__ __device__ __forceinline__ F32 operator-(F32 a, F32 b) { return F32(a.v - b.v); }
__host__ __device__ __forceinline__ F32 operator*(F32 a, F32 b) { return F32(a.v * b.v); }
__host__ __device__ __forceinline__ F32 operator/(F32 a, F32 b) { return F32(a.v * frcp(b.v)); }
__host__ __device__ __forceinline__ F32 operator-(F32 a) { return F32(-a.v); }
__host__ __device__ __forceinline__ F32 &operator+=(F32 &a, F32 b) { return a = a + b; }
__host__ __device__ __forceinline__ F32 &operator-=(F32 &a, F32 b) { return a = a - b; }
__host__ __device__ __forceinline__ F32 &operator*=(F32 &a, F32 b) { return a = a * b; }
__host__ __device__ __forceinline__ F32 &operator/=(F32 &a, F32 b) { return a = a / b; }
__host__ __device__ __forceinline__ bool operator<(F32 a, F32 b) { return a.v < b.v; }
__host__ __device__ __forceinline__ bool operator>(F32 a, F32 b) { return a.v > b.v; }
__host__ __device__ __forceinline__ bool operator<=(F32 a, F32 b) { return a.v <= b.v; }
__host__ __device__ __forceinline__ bool operator>=(F32 a, F32 b) { return a.v >= b.v; }
__host__ __device__ __forceinline__ bool operator==(F32 a, F32 b) { return a.v == b.v; }
__host__ __device__ __forceinline__ bool operator!=(F32 a, F32 b) { return a.v != b.v; }
__host__ __device__ __forceinline__ F32 exp(F32 a) { return F32(expf(a.v)); }
__host__ __device__ __forceinline__ F32 log(F32 a) { return F32(logf(a.v)); }
__host__ __device__ __forceinline__ F32 pow(F32 a, F32 b) { return F32(powf(a.v, b.v)); }
__host__ __device__ __forceinline__ F32 sqrt(F32 a) { return F32(sqrtf(a.v)); }
__host__ __device__ __forceinline__ F32 cbrt(F32 a) { return F32(cbrtf(a.v)); }
__host__ __device__ __forceinline__ F32 erfc(F32 a) { return F32(erfcf(a.v)); }
__host__ __device__ __forceinline__ F32 fabs(F32 a) { return F32(fabsf(a.v)); }
__host__ __device__ __forceinline__ F32 round(F32 a) { return F32(roundf(a.v)); }

// ------------------------------------------------------------------ Dual
struct Dual {
  double v, d;
  Dual() = default;
  __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
  __host__ __device__ constexpr Dual(int x) : v((double)x), d(0.0) {}
  __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
  __host__ __device__ explicit constexpr operator double() const { return v; }
};

__host__ __device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return Dual(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__host__ __device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__host__ __device__ __forceinline__ Dual &operator+=(Dual &a, Dual b) { return a = a + b; }
__host__ __device__ __forceinline__ Dual &operator-=(Dual &a, Dual b) { return a = a - b; }
__host__ __device__ __forceinline__ Dual &operator*=(Dual &a, Dual b) { return a = a * b; }
__host__ __device__ __forceinline__ Dual &operator/=(Dual &a, Dual b) { return a = a / b; }
__host__ __device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__host__ __device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__host__ __device__ __forceinline__ bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
__host__ __device__ __forceinline__ bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
__host__ __device__ __forceinline__ bool operator==(Dual a, Dual b) { return a.v == b.v; }
__host__ __device__ __forceinline__ bool operator!=(Dual a, Dual b) { return a.v != b.v; }
__host__ __device__ __forceinline__ Dual exp(Dual a) {
  const double e = ::exp(a.v);
  return Dual(e, a.d * e);
}
__host__ __device__ __forceinline__ Dual log(Dual a) { return Dual(::log(a.v), a.d / a.v); }
__host__ __device__ __forceinline__ Dual pow(Dual a, Dual b) {
  const double r = ::pow(a.v, b.v);
  // d(a^b) = a^b (b' ln a + b a'/a); the b' term only when b varies
  const double db = b.d != 0.0 ? b.d * ::log(a.v) : 0.0;
  return Dual(r, r * (db + b.v * a.d / a.v));
}
__host__ __device__ __forceinline__ Dual sqrt(Dual a) {
  const double s = ::sqrt(a.v);
  return Dual(s, a.d / (2.0 * s));
}
__host__ __device__ __forceinline__ Dual cbrt(Dual a) {
  const double c = ::cbrt(a.v);
  return Dual(c, a.d / (3.0 * c * c));
}
__host__ __device__ __forceinline__ Dual erfc(Dual a) {
  // d erfc(x)/dx = -2/sqrt(pi) exp(-x^2)
  return Dual(::erfc(a.v), -1.1283791670955126 * ::exp(-a.v * a.v) * a.d);
}
__host__ __device__ __forceinline__ Dual fabs(Dual a) { return a.v < 0 ? -a : a; }
__host__ __device__ __forceinline__ Dual round(Dual a) { return Dual(::round(a.v), 0.0); }

// ------------------------------------------------------------------ division
// Correctly rounded FP64 division without the range scaling and special-case
// fix-up of the general IEEE sequence: reciprocal, two Newton steps, then
// q + r (a - b q) (Markstein).  These are the same operations the general
// sequence performs when its scaling is the identity, which holds whenever a,
// b and a/b are normal numbers away from the ends of the exponent range
// (|a| >= 2^-968, 2^-1021 < |b| < 2^1021, |a/b| >= 2^-1021); a +0 numerator
// gives +0, a -0 numerator +0 as well (IEEE: -0).  Below 2^-968 the residual
// a - b q is not representable and the quotient can be one unit in the last
// place off; b = 0 or inf gives NaN.  Callers use it only where they know
// the operands to be in range (the SOA sub-step, whose loop checks that it
// is: soa_fast_exact in chain_device.hpp).
//
// rcp(b) is the reciprocal with its two Newton steps and div_r(a, b, r) the
// quotient from it, so a divisor used more than once is inverted once:
// div_fast(a, b) = div_r(a, b, rcp(b)).  Host builds divide with a / b.
__host__ __device__ __forceinline__ double rcp(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  return __builtin_fma(r, e, r);
#else
  return 1.0 / b;
#endif
}
__host__ __device__ __forceinline__ double div_r(double a, double b, double r) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double q = a * r;
  const double res = __builtin_fma(-b, q, a);
  return __builtin_fma(res, r, q);
#else
  (void)r;
  return a / b;
#endif
}
__host__ __device__ __forceinline__ double div_fast(double a, double b) {
  return div_r(a, b, rcp(b));
}
__host__ __device__ __forceinline__ F32 div_fast(F32 a, F32 b) { return a / b; }
__host__ __device__ __forceinline__ Dual div_fast(Dual a, Dual b) { return a / b; }

// ------------------------------------------------------------------ helpers
// primal value as a double (branch decisions, integer indices, outputs)
__host__ __device__ __forceinline__ double val(double x) { return x; }
__host__ __device__ __forceinline__ double val(F32 x) { return x.v; }
__host__ __device__ __forceinline__ double val(Dual x) { return x.v; }
// derivative part (zero for the plain types)
__host__ __device__ __forceinline__ double tan(double) { return 0.0; }
__host__ __device__ __forceinline__ double tan(F32) { return 0.0; }
__host__ __device__ __forceinline__ double tan(Dual x) { return x.d; }

// conversion between the bulk and the wide type of a policy
template <class T>
__host__ __device__ __forceinline__ T to(double x) { return T(x); }
template <class T>
__host__ __device__ __forceinline__ T to(F32 x) { return T((double)x.v); }
template <class T>
__host__ __device__ __forceinline__ T to(Dual x) { return x; }
template <>
__host__ __device__ __forceinline__ double to<double>(Dual x) { return x.v; }

}  // namespace num

// precision policies; diag: the step also forms the per-process tendency
// diagnostics (a separate instantiation, so the plain step carries none of
// their bookkeeping)
struct Fp64 {
  using R = double;
  using W = double;
  static constexpr bool diag = false;
};
struct Fp64Diag {
  using R = double;
  using W = double;
  static constexpr bool diag = true;
};
struct Mixed {
  using R = num::F32;
  using W = double;
  static constexpr bool diag = false;
};
struct Tangent {
  using R = num::Dual;
  using W = num::Dual;
  static constexpr bool diag = false;
};

}  // namespace mam4x
