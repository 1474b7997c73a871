// dual_check.cpp -- host-side check of the forward-mode dual numbers and of the
// geometry helpers the tangent runs them through (test infrastructure).
//
// For each function: the value part must equal the plain double result bit
// for bit, and the derivative part must match a central finite difference of
// the double function.  Prints one JSON object with the worst errors.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "mpcd_math.h"
#include "mpcd_dual.h"

using mpcd_ad::Dual;

static double worst_fd = 0.0;
static int value_mismatch = 0;
static int checks = 0;

static void check(const char *what, double val, double dval, double f_plus, double f_minus, double h, double ref_val) {
  checks++;
  if (memcmp(&val, &ref_val, sizeof(double)) != 0) {
    value_mismatch++;
    fprintf(stderr, "value mismatch in %s: %.17g vs %.17g\n", what, val, ref_val);
  }
  const double fd = (f_plus - f_minus) / (2.0 * h);
  const double err = fabs(dval - fd) / fmax(1.0, fabs(fd));
  if (err > worst_fd) worst_fd = err;
  if (err > 1e-6) fprintf(stderr, "derivative of %s: %.12g vs finite difference %.12g\n", what, dval, fd);
}

template <typename Fn>
static void scalar(const char *what, Fn f, double x) {
  const double h = 1e-6 * fmax(1.0, fabs(x));
  const Dual y = f(Dual(x, 1.0));
  check(what, y.v, y.d, f(x + h), f(x - h), h, f(x));
}

// a symmetric 2x2 matrix with eigenvalues a > b and first eigenvector at
// angle th: its director must turn with th at unit rate
template <typename F>
static F director_angle(const F &th) {
  const double a = 0.8, b = -0.3;
  F c = cos(th), s = sin(th);
  F m[2][2] = {{a * c * c + b * s * s, (a - b) * c * s}, {(a - b) * c * s, a * s * s + b * c * c}};
  F v[2];
  mpcd::eig2_top(m, v);
  return atan(v[1] / v[0]);
}

template <typename F>
static F rodrigues_x(const F &th) {
  F vec[3] = {0.3, -0.7, 0.2}, ax[3] = {0.1, 0.5, 0.8};
  mpcd::rodrigues(vec, ax, th);
  return vec[0] + 2.0 * vec[1] - vec[2];
}

template <typename F>
static F normalised(const F &x) {
  F v[3] = {x, 0.4 * x + 1.0, -0.3};
  mpcd::normn(v, 3);
  return v[0] - v[2];
}

template <typename F>
static F angle_between(const F &x) {
  F a[3] = {1.0, x, 0.0}, b[3] = {-0.2, 1.0, 0.0};
  return mpcd::signed_angle(a, b, 2) + mpcd::abs_angle(a, b, 2);
}

template <typename F>
static F eig3_top_value(const F &x) {
  F m[3][3] = {{1.0 + x, 0.2, 0.1}, {0.2, 0.5, -0.3 * x}, {0.1, -0.3 * x, -0.4}};
  F ev[3], v0[3];
  mpcd::eig3_top(m, ev, v0);
  return ev[0] + 0.5 * v0[0] - v0[2];
}

int main() {
  const double xs[] = {0.3, 0.77, 1.9};
  for (double x : xs) {
    scalar("sqrt", [](auto v) { return sqrt(v); }, x);
    scalar("sin", [](auto v) { return sin(v); }, x);
    scalar("cos", [](auto v) { return cos(v); }, x);
    scalar("exp", [](auto v) { return exp(v); }, x);
    scalar("atan", [](auto v) { return atan(v); }, x);
    scalar("tanh", [](auto v) { return tanh(v); }, x);
    scalar("pow", [](auto v) { return pow(v, 1.0 / 3.0); }, x);
    scalar("fabs", [](auto v) { return fabs(v - 1.0); }, x);
    scalar("quotient", [](auto v) { return (v * v + 1.0) / (2.0 - v * 0.3); }, x);
    scalar("compound", [](auto v) {
      auto w = v;
      w *= v;
      w /= (v + 3.0);
      w -= 0.5 * v;
      w += 1.0;
      return w;
    }, x);
    scalar("eig2 director", [](auto v) { return director_angle(v); }, 0.1 * x);
    scalar("rodrigues", [](auto v) { return rodrigues_x(v); }, x);
    scalar("normn", [](auto v) { return normalised(v); }, x);
    scalar("angles", [](auto v) { return angle_between(v); }, 0.5 * x);
    scalar("eig3 top", [](auto v) { return eig3_top_value(v); }, 0.2 * x);
  }
  scalar("acos", [](auto v) { return acos(v); }, 0.35);
  scalar("acos", [](auto v) { return acos(v); }, -0.8);
  // the director of eig2_top turns at unit rate with the eigenbasis
  const Dual d = director_angle(Dual(0.25, 1.0));
  const double rate_err = fabs(d.d - 1.0);
  printf("{\"checks\": %d, \"value_mismatch\": %d, \"worst_fd_rel_err\": %.3e, \"director_rate_err\": %.3e}\n", checks,
         value_mismatch, worst_fd, rate_err);
  return 0;
}
