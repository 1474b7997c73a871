// mpcd_dual.h -- forward-mode dual numbers for the tangent of the solvent step.
//
// A Dual carries a value and its derivative along one direction in parameter
// space (body force, activity).  The value part is computed by exactly the
// operations of the plain FP64 code, so a step run in Dual arithmetic gives
// the FP64 step's values bit for bit; comparisons and integer conversions
// look at the value only, which freezes every discrete decision of the step
// (cell assignment, wall events, Metropolis acceptances, activity side): the
// derivative part is the pathwise derivative.
//
// The type and its functions live in their own namespace and are found by
// argument-dependent lookup, so `sqrt(x)` on a double inside namespace mpcd
// still calls the device library's sqrt.
#pragma once

#include <hip/hip_runtime.h>

#define MPCD_AD __host__ __device__ __forceinline__

namespace mpcd_ad {

struct Dual {
  double v, d;
  MPCD_AD Dual() : v(0.0), d(0.0) {}
  MPCD_AD Dual(double x) : v(x), d(0.0) {}  // NOLINT: constants promote implicitly
  MPCD_AD Dual(double x, double y) : v(x), d(y) {}
  MPCD_AD Dual &operator+=(const Dual &b) { v += b.v; d += b.d; return *this; }
  MPCD_AD Dual &operator-=(const Dual &b) { v -= b.v; d -= b.d; return *this; }
  MPCD_AD Dual &operator*=(const Dual &b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
  MPCD_AD Dual &operator/=(const Dual &b) {
    const double q = v / b.v;
    d = (d - q * b.d) / b.v;
    v = q;
    return *this;
  }
};

MPCD_AD double val(const Dual &a) { return a.v; }
MPCD_AD double der(const Dual &a) { return a.d; }

MPCD_AD Dual operator-(const Dual &a) { return Dual(-a.v, -a.d); }
MPCD_AD Dual operator+(const Dual &a, const Dual &b) { return Dual(a.v + b.v, a.d + b.d); }
MPCD_AD Dual operator-(const Dual &a, const Dual &b) { return Dual(a.v - b.v, a.d - b.d); }
MPCD_AD Dual operator*(const Dual &a, const Dual &b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
MPCD_AD Dual operator/(const Dual &a, const Dual &b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
MPCD_AD Dual operator+(const Dual &a, double b) { return Dual(a.v + b, a.d); }
MPCD_AD Dual operator+(double a, const Dual &b) { return Dual(a + b.v, b.d); }
MPCD_AD Dual operator-(const Dual &a, double b) { return Dual(a.v - b, a.d); }
MPCD_AD Dual operator-(double a, const Dual &b) { return Dual(a - b.v, -b.d); }
MPCD_AD Dual operator*(const Dual &a, double b) { return Dual(a.v * b, a.d * b); }
MPCD_AD Dual operator*(double a, const Dual &b) { return Dual(a * b.v, a * b.d); }
MPCD_AD Dual operator/(const Dual &a, double b) { return Dual(a.v / b, a.d / b); }
MPCD_AD Dual operator/(double a, const Dual &b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}

#define MPCD_AD_CMP(op)                                                            \
  MPCD_AD bool operator op(const Dual &a, const Dual &b) { return a.v op b.v; }    \
  MPCD_AD bool operator op(const Dual &a, double b) { return a.v op b; }           \
  MPCD_AD bool operator op(double a, const Dual &b) { return a op b.v; }
MPCD_AD_CMP(<)
MPCD_AD_CMP(>)
MPCD_AD_CMP(<=)
MPCD_AD_CMP(>=)
MPCD_AD_CMP(==)
MPCD_AD_CMP(!=)
#undef MPCD_AD_CMP

// the reference's tolerance tests (mtools.c:61, 74) on values
MPCD_AD bool feq(const Dual &x, const Dual &y) { return ::fabs(x.v - y.v) <= 1e-8; }
MPCD_AD bool feq(const Dual &x, double y) { return ::fabs(x.v - y) <= 1e-8; }
MPCD_AD bool fneq(const Dual &x, const Dual &y) { return ::fabs(x.v - y.v) >= 1e-8; }
MPCD_AD bool fneq(const Dual &x, double y) { return ::fabs(x.v - y) >= 1e-8; }

// elementary functions: value by the library routine, derivative by the chain rule
MPCD_AD Dual sqrt(const Dual &a) {
  const double s = ::sqrt(a.v);
  return Dual(s, s > 0.0 ? 0.5 * a.d / s : 0.0);
}
MPCD_AD Dual sin(const Dual &a) { return Dual(::sin(a.v), ::cos(a.v) * a.d); }
MPCD_AD Dual cos(const Dual &a) { return Dual(::cos(a.v), -::sin(a.v) * a.d); }
MPCD_AD Dual exp(const Dual &a) {
  const double e = ::exp(a.v);
  return Dual(e, e * a.d);
}
MPCD_AD Dual acos(const Dual &a) {
  const double s = 1.0 - a.v * a.v;
  return Dual(::acos(a.v), s > 0.0 ? -a.d / ::sqrt(s) : 0.0);
}
MPCD_AD Dual atan(const Dual &a) { return Dual(::atan(a.v), a.d / (1.0 + a.v * a.v)); }
MPCD_AD Dual tanh(const Dual &a) {
  const double t = ::tanh(a.v);
  return Dual(t, (1.0 - t * t) * a.d);
}
MPCD_AD Dual pow(const Dual &a, double p) {
  const double r = ::pow(a.v, p);
  return Dual(r, a.v != 0.0 ? p * r / a.v * a.d : 0.0);
}
MPCD_AD Dual fabs(const Dual &a) { return a.v < 0.0 ? -a : a; }

}  // namespace mpcd_ad
