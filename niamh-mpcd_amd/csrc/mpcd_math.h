// mpcd_math.h -- per-cell and per-particle arithmetic of the MPCD solvent step
// for gfx950.  Every helper keeps the operation order of the reference routine
// it names, so that FP64 results agree with the reference to round-off (bit
// for bit where only +,-,*,/,sqrt are involved; the library is compiled with
// -ffp-contract=off for that reason).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpcd {

constexpr double kTOL = 1e-8;          // definitions.h:57
constexpr int kNBOUNCE = 50;           // definitions.h:63
constexpr double kPI = 3.14159265358979323846;
constexpr int kMCINT = 5;              // definitions.h:355
constexpr double kMCSLOPE = 0.1;       // definitions.h:353
constexpr double kBUSMIN = 0.5;        // definitions.h:179
constexpr double kBUSMAX = 5.0;        // definitions.h:181

#define MPCD_HD __host__ __device__ __forceinline__

MPCD_HD bool feq(double x, double y) { return fabs(x - y) <= kTOL; }   // mtools.c:61
MPCD_HD bool fneq(double x, double y) { return fabs(x - y) >= kTOL; }  // mtools.c:74

// The geometric helpers are templates on the arithmetic type F: double for
// the step, mpcd_ad::Dual (mpcd_dual.h) for its tangent.
MPCD_HD double val(double x) { return x; }
MPCD_HD float val(float x) { return x; }

template <typename F>
MPCD_HD F dotn(const F *x, const F *y, int n) {                        // mtools.c:111
  F r = 0.;
  for (int i = 0; i < n; i++) r += x[i] * y[i];
  return r;
}

template <typename F>
MPCD_HD void cross3(const F *x, const F *y, F *r) {                    // mtools.c:177
  r[0] = x[1] * y[2] - x[2] * y[1];
  r[1] = x[2] * y[0] - x[0] * y[2];
  r[2] = x[0] * y[1] - x[1] * y[0];
}

template <typename F>
MPCD_HD F lengthn(const F *x, int n) {                                 // mtools.c:227
  F r = 0.;
  for (int i = 0; i < n; i++) r += x[i] * x[i];
  return sqrt(r);
}

template <typename F>
MPCD_HD void normn(F *x, int n) {                                      // mtools.c:246
  F l = lengthn(x, n);
  if (fneq(l, 0.0))
    for (int i = 0; i < n; i++) x[i] = x[i] / l;
}

template <typename F>
MPCD_HD void projn(const F *v, const F *nr, F *vn, int n) {            // mtools.c:293
  F x = dotn(v, nr, n);
  for (int i = 0; i < n; i++) vn[i] = x * nr[i];
}

// the reference's atan2 (mtools.c:342-356), which overrides libm's
template <typename F>
MPCD_HD F ref_atan2(const F &y, const F &x) {
  F at = 0.0;
  if (x > 0.0) at = atan(y / x);
  else if (x < 0.0 && y >= 0.0) at = atan(y / x) + kPI;
  else if (x < 0.0 && y < 0.0) at = atan(y / x) - kPI;
  else if (feq(x, 0.0) && y > 0.0) at = 0.5 * kPI;
  else if (feq(x, 0.0) && y < 0.0) at = -0.5 * kPI;
  return at;
}

template <typename F>
MPCD_HD F abs_angle(const F *a, const F *b, int n) {                   // mtools.c:325,365
  F c = dotn(a, b, n);
  c /= lengthn(a, n);
  c /= lengthn(b, n);
  return acos(c);
}

template <typename F>
MPCD_HD F signed_angle(const F *v1, const F *v2, int n) {              // mtools.c:380
  F A[3] = {0., 0., 0.}, B[3] = {0., 0., 0.}, c3[3];
  for (int i = 0; i < n; i++) { A[i] = v1[i]; B[i] = v2[i]; }
  cross3(A, B, c3);
  return ref_atan2(lengthn(c3, 3), dotn(A, B, n));
}

// rodriguesRotation, mtools.c:1572-1584 (renormalises the axis in place)
template <typename F>
MPCD_HD void rodrigues(F *vec, F *ax, const F &theta) {
  F cp[3];
  normn(ax, 3);
  F oldl = lengthn(vec, 3);
  cross3(ax, vec, cp);
  F dp = dotn(ax, vec, 3);
  for (int i = 0; i < 3; i++) vec[i] = vec[i] * cos(theta) + cp[i] * sin(theta) + ax[i] * dp * (1.0 - cos(theta));
  F newl = lengthn(vec, 3);
  for (int i = 0; i < 3; i++) vec[i] *= (oldl / newl);
}

template <typename F>
MPCD_HD F det3(const F m[3][3]) {                                      // mtools.c:857
  F c1 = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
  F c2 = m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
  F c3 = m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  return (c1 - c2 + c3);
}

// the ORTHAXIS rotation about axis CA (0 x, 1 y, 2 z) by the angle whose
// cosine and sine are Cc and Sn (rotate, mpc.c:1206-1237, with the x-axis
// entry R[0][2][1] = -.1 S as written there).  The entries are selected, not
// branched on: an if-else chain over CA in a cell loop was compiled to code
// that left the z-axis matrix's entry [1][0] at zero on gfx950.
MPCD_HD void orth_matrix(int CA, double Cc, double Sn, double M[9]) {
  const bool a0 = CA == 0, a1 = CA == 1, a2 = !a0 && !a1;
  M[0] = a0 ? 1. : Cc;
  M[1] = a2 ? Sn : 0.;
  M[2] = a1 ? Sn : 0.;
  M[3] = a2 ? -1. * Sn : 0.;
  M[4] = a1 ? 1. : Cc;
  M[5] = a0 ? Sn : 0.;
  M[6] = a1 ? -1. * Sn : 0.;
  M[7] = a0 ? -.1 * Sn : 0.;
  M[8] = a2 ? 1. : Cc;
}

// invert3x3 via cofactors, mtools.c:972-1037
template <typename F>
MPCD_HD void inv3(F inv[3][3], const F m[3][3]) {
  F det = det3(m);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      int r0 = (i == 0) ? 1 : 0, r1 = (i == 2) ? 1 : 2;
      int c0 = (j == 0) ? 1 : 0, c1 = (j == 2) ? 1 : 2;
      int sg = (2 * ((i + j) / 2) == (i + j)) ? 1 : -1;
      F cof = (double)sg * (m[r0][c0] * m[r1][c1] - m[r0][c1] * m[r1][c0]);
      inv[j][i] = cof / det;
    }
}

// eigenvalues2x2 + eigenvectors2x2, mtools.c:1230-1305.  Returns the largest
// eigenvalue and writes the first eigenvector (the cell director).
// *br (when given) receives the branch taken: 0 m10 != 0, 1 m01 != 0,
// 2 diagonal with equal entries, 3 other diagonal
template <typename F>
MPCD_HD F eig2_top(const F m[2][2], F v0[2], int *br = nullptr) {
  F det = m[0][0] * m[1][1] - m[1][0] * m[0][1];
  F tr = m[0][0] + m[1][1];
  F sq = sqrt(tr * tr * 0.25 - det);
  F e0 = tr * 0.5 + sq;
  v0[0] = 0.0;
  v0[1] = 0.0;
  if (br) *br = fneq(m[1][0], 0.0) ? 0 : fneq(m[0][1], 0.0) ? 1 : feq(m[0][0], m[1][1]) ? 2 : 3;
  if (fneq(m[1][0], 0.0)) {
    v0[0] = e0 - m[1][1];
    v0[1] = m[1][0];
    normn(v0, 2);
  } else if (fneq(m[0][1], 0.0)) {
    v0[0] = m[0][1];
    v0[1] = e0 - m[0][0];
    normn(v0, 2);
  } else {
    if (feq(m[0][0], m[1][1])) {
      v0[0] = sqrt(2.);
      v0[1] = v0[0];
    } else if (feq(m[0][0], e0)) {
      v0[0] = 1.0;
      v0[1] = 0.0;
    } else if (feq(m[1][1], e0)) {
      v0[0] = 1.0;
      v0[1] = 0.0;
    }
  }
  return e0;
}

// eigenvalues3x3 (Smith's trigonometric method) + eigenvectors3x3
// (Cayley-Hamilton column products), mtools.c:1307-1430.  Writes all three
// eigenvalues (descending by construction) and the first eigenvector.
// *br (when given) receives the branches taken: bits 0-1 the column of the
// eigenvector product used (3: none), bit 2 the diagonal branch, bit 3 the
// clamp of r
template <typename F>
MPCD_HD void eig3_top(const F m[3][3], F ev[3], F v0[3], int *br = nullptr) {
  F p1 = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
  int b = 0;
  if (feq(p1, 0.0)) {
    b |= 4;
    F q;
    for (int i = 0; i < 3; i++) ev[i] = m[i][i];
    if (ev[1] > ev[0]) { q = ev[0]; ev[0] = ev[1]; ev[1] = q; }
    if (ev[2] > ev[0]) { q = ev[0]; ev[0] = ev[2]; ev[2] = q; }
    if (ev[2] > ev[1]) { q = ev[1]; ev[1] = ev[2]; ev[2] = q; }
  } else {
    F p2 = 0.;
    F q = (m[0][0] + m[1][1] + m[2][2]) / 3.;
    for (int i = 0; i < 3; i++) p2 += (m[i][i] - q) * (m[i][i] - q);
    p2 += 2. * p1;
    F p = sqrt(p2 / 6.);
    F ip = 1. / p;
    F B[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) B[i][j] = m[i][j];
    for (int i = 0; i < 3; i++) B[i][i] -= q;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) B[i][j] *= ip;
    F r = det3(B) * 0.5, phi;
    if (r <= -1. || r >= 1.) b |= 8;
    if (r <= -1.) phi = kPI / 3.;
    else if (r >= 1.) phi = 0.;
    else phi = acos(r) / 3.;
    ev[0] = q + 2. * p * cos(phi);
    ev[2] = q + 2. * p * cos(phi + (2. * kPI / 3.));
    ev[1] = 3. * q - ev[0] - ev[2];
  }
  // first eigenvector: product (m - e1)(m - e2), first non-zero column
  int used = 3;
  for (int col = 0; col < 3; col++) {
    for (int row = 0; row < 3; row++) v0[row] = 0.;
    for (int row = 0; row < 3; row++)
      for (int i = 0; i < 3; i++) {
        F a = m[row][i];
        if (i == row) a -= ev[1];
        F bb = m[i][col];
        if (i == col) bb -= ev[2];
        v0[row] += a * bb;
      }
    if (!(feq(v0[0], 0.0) && feq(v0[1], 0.0) && feq(v0[2], 0.0))) {
      used = col;
      break;
    }
  }
  normn(v0, 3);
  if (br) *br = b | used;
}

// ---------------------------------------------------------------------------
// Counter-based random numbers: Philox-4x32-10 (Salmon et al., SC11).
// A variate is a pure function of (seed, replica, step, purpose, key, k):
// counter = {key, purpose << 24 | block, step, replica}, key = seed.
// ---------------------------------------------------------------------------
MPCD_HD void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
  }
}

MPCD_HD double u53(uint32_t a, uint32_t b) {
  return (double)(((uint64_t)a << 21) | (uint64_t)(b >> 11)) * (1.0 / 9007199254740992.0);
}

// sine and cosine of 2 pi u for u = n 2^-53 in [0, 1): the octant of u and
// the position in it are exact in FP64 (8u, its fraction f and 1 - f), the
// angle in the octant is g pi/4 with g = f or 1 - f, and Taylor polynomials of
// degree 13 (sine) and 14 (cosine) on [0, pi/4] leave truncation errors below
// 3e-14.  Each operation is a plain FP64 add or multiply in a fixed order, so
// a C build without contraction gives the same bits (no library sin or cos,
// whose last bits differ between implementations).
MPCD_HD void sincos_turn(double u, double *sn, double *cs) {
  const double t = 8.0 * u;
  const int o = (int)t;
  const double f = t - (double)o;
  const double g = (o & 1) ? 1.0 - f : f;
  const double th = g * 0.78539816339744830962;
  const double x = th * th;
  double p = 1.0 / 6227020800.0;
  p = p * x + -1.0 / 39916800.0;
  p = p * x + 1.0 / 362880.0;
  p = p * x + -1.0 / 5040.0;
  p = p * x + 1.0 / 120.0;
  p = p * x + -1.0 / 6.0;
  const double s = th + (th * x) * p;
  double q = -1.0 / 87178291200.0;
  q = q * x + 1.0 / 479001600.0;
  q = q * x + -1.0 / 3628800.0;
  q = q * x + 1.0 / 40320.0;
  q = q * x + -1.0 / 720.0;
  q = q * x + 1.0 / 24.0;
  q = q * x + -0.5;
  const double c = 1.0 + x * q;
  // octants 1, 2, 5, 6 swap sine and cosine; 2-5 negate the cosine, 4-7 the sine
  const bool swap = (((o + 1) >> 1) & 1) != 0;
  const double a = swap ? s : c, b = swap ? c : s;
  *cs = ((o + 2) & 4) ? -a : a;
  *sn = (o & 4) ? -b : b;
}

// natural log of u in (0, 1] without a library log: u = m 2^e with m in
// [1/sqrt 2, sqrt 2) (frexp and a doubling, both exact), log m = 2 atanh s
// with s = (m - 1) / (m + 1), |s| < 0.172, from its series to s^19
// (truncation below 2e-16 relative), plus e ln 2.  Plain FP64 operations in
// a fixed order, as in sincos_turn.
MPCD_HD double log_unit(double u) {
  int e;
  double m = frexp(u, &e);
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    e = e - 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double x = s * s;
  double p = 1.0 / 19.0;
  p = p * x + 1.0 / 17.0;
  p = p * x + 1.0 / 15.0;
  p = p * x + 1.0 / 13.0;
  p = p * x + 1.0 / 11.0;
  p = p * x + 1.0 / 9.0;
  p = p * x + 1.0 / 7.0;
  p = p * x + 1.0 / 5.0;
  p = p * x + 1.0 / 3.0;
  p = p * x + 1.0;
  return (double)e * 0.69314718055994530942 + 2.0 * (s * p);
}

struct Rng {
  uint64_t seed;
  uint32_t step;
  MPCD_HD void block(uint32_t replica, int purpose, uint32_t key, uint32_t blk, uint32_t w[4]) const {
    w[0] = key;
    w[1] = ((uint32_t)purpose << 24) | (blk & 0xFFFFFFu);
    w[2] = step;
    w[3] = replica;
    philox4x32_10(w, (uint32_t)seed, (uint32_t)(seed >> 32));
  }
  MPCD_HD double uniform(uint32_t rep, int purpose, uint32_t key, uint32_t k) const {
    uint32_t w[4];
    block(rep, purpose, key, k, w);
    return u53(w[0], w[1]);
  }
  MPCD_HD double normal(uint32_t rep, int purpose, uint32_t key, uint32_t k) const {
    uint32_t w[4];
    block(rep, purpose, key, k >> 1, w);
    double u1 = 1.0 - u53(w[0], w[1]);
    double u2 = u53(w[2], w[3]);
    double r = sqrt(-2.0 * log(u1));
    return (k & 1) ? r * sin(6.283185307179586 * u2) : r * cos(6.283185307179586 * u2);
  }
  // the two normals of one block, normal(.., 2 * pair) and normal(.., 2 * pair + 1),
  // by the same expressions: one counter block, one logarithm and one root for both
  MPCD_HD void normal_pair(uint32_t rep, int purpose, uint32_t key, uint32_t pair, double out[2]) const {
    uint32_t w[4];
    block(rep, purpose, key, pair, w);
    double u1 = 1.0 - u53(w[0], w[1]);
    double u2 = u53(w[2], w[3]);
    double r = sqrt(-2.0 * log(u1));
    out[0] = r * cos(6.283185307179586 * u2);
    out[1] = r * sin(6.283185307179586 * u2);
  }
  // the same pair with log_unit and sincos_turn (the exact-sum form)
  MPCD_HD void normal_pair_turn(uint32_t rep, int purpose, uint32_t key, uint32_t pair, double out[2]) const {
    uint32_t w[4];
    block(rep, purpose, key, pair, w);
    double u1 = 1.0 - u53(w[0], w[1]);
    double r = sqrt(-2.0 * log_unit(u1));
    double sn, cs;
    sincos_turn(u53(w[2], w[3]), &sn, &cs);
    out[0] = r * cs;
    out[1] = r * sn;
  }
  MPCD_HD void unitvec(uint32_t rep, int purpose, uint32_t key, uint32_t k, double v[3]) const {
    uint32_t w[4];
    block(rep, purpose, key, k, w);
    double z = 1.0 - 2.0 * u53(w[0], w[1]);
    double phi = 6.283185307179586 * u53(w[2], w[3]);
    double s = sqrt((1.0 - z) * (1.0 + z));
    v[0] = s * cos(phi);
    v[1] = s * sin(phi);
    v[2] = z;
  }
  // the same axis with the sine and cosine of sincos_turn (the exact-sum form)
  MPCD_HD void unitvec_turn(uint32_t rep, int purpose, uint32_t key, uint32_t k, double v[3]) const {
    uint32_t w[4];
    block(rep, purpose, key, k, w);
    double z = 1.0 - 2.0 * u53(w[0], w[1]);
    double sn, cs;
    sincos_turn(u53(w[2], w[3]), &sn, &cs);
    double s = sqrt((1.0 - z) * (1.0 + z));
    v[0] = s * cs;
    v[1] = s * sn;
    v[2] = z;
  }
};

// P_ESCAPE: escape repair in the step's shifted bin; P_ESCAPE2: in its
// unshifted bin after streaming
enum Purpose { P_SHIFT = 1, P_CELL = 2, P_GHOST = 3, P_NORMAL = 4, P_LC = 5, P_ESCAPE = 6, P_RETHERM = 7,
               P_ESCAPE2 = 8 };

// Keyed layout of the rare draws (the same numbers as MPCD_ESC_STRIDE and
// MPCD_RETHERM_STRIDE of the C-ABI).  Escape repair, per particle and
// binning: [0..2] Maxwell normals, [3..4] orientation uniforms, then DIM
// position uniforms per try from kEscPos0.  The generator serves index k of
// a normal as normal(k) (counter blocks 0 and 1) and of a uniform as
// uniform(k), k >= 3, so no counter block feeds two variates.
constexpr int kEscStride = 32, kEscPos0 = 5, kEscMaxTries = 65536;
constexpr int kRethStride = 24;

// Draw source seen by the kernels: the counter-based generator or injected
// arrays in the keyed layout of include/mpcd_gpu.h.
struct Draws {
  Rng rng;
  int injected;
  const double *shift, *cell, *ghost, *part, *lc;
  int lc_stride;
  // escape and re-thermalisation arrays (null: the generator serves them
  // even when the rest is injected) and the particles of the context
  const double *esc, *reth;
  int ntot;
  // graph replays: the step number lives in device memory (step_dev) and a
  // launch is step_off steps after it; eager launches carry rng.step
  const uint32_t *step_dev;
  uint32_t step_off;

  __device__ void resolve() {
    if (step_dev) rng.step = *step_dev + step_off;
  }

  MPCD_HD double cell_u(uint32_t rep, uint32_t c, int k) const {
    return injected ? cell[(size_t)c * 4 + k] : rng.uniform(rep, P_CELL, c, k);
  }
  MPCD_HD void cell_axis(uint32_t rep, uint32_t c, double v[3]) const {
    if (injected) { v[0] = cell[(size_t)c * 4]; v[1] = cell[(size_t)c * 4 + 1]; v[2] = cell[(size_t)c * 4 + 2]; }
    else rng.unitvec(rep, P_CELL, c, 0, v);
  }
  MPCD_HD double ghost_n(uint32_t rep, uint32_t c, int k) const {
    return injected ? ghost[(size_t)c * 3 + k] : rng.normal(rep, P_GHOST, c, k);
  }
  MPCD_HD double part_n(uint32_t rep, uint32_t slot, int k) const {
    return injected ? part[(size_t)slot * 3 + k] : rng.normal(rep, P_NORMAL, slot, k);
  }
  // part_n(rep, slot, k) for k < dim, the normals of a block computed once
  MPCD_HD void part_n_all(uint32_t rep, uint32_t slot, int dim, double out[3]) const {
    if (injected) {
      for (int k = 0; k < dim; k++) out[k] = part[(size_t)slot * 3 + k];
      return;
    }
    rng.normal_pair(rep, P_NORMAL, slot, 0, out);
    if (dim == 3) {
      double t[2];
      rng.normal_pair(rep, P_NORMAL, slot, 1, t);
      out[2] = t[0];
    }
  }
  MPCD_HD double lc_u(uint32_t rep, uint32_t slot, int k) const {
    return injected ? lc[(size_t)slot * lc_stride + k] : rng.uniform(rep, P_LC, slot, k);
  }
  MPCD_HD double lc_n(uint32_t rep, uint32_t slot, int k) const {
    return injected ? lc[(size_t)slot * lc_stride + k] : rng.normal(rep, P_LC, slot, k);
  }
  // escape draws of particle pid (global index gp = replica offset + pid) in
  // binning b (0 shifted, 1 unshifted); k < 3 normals, k >= 3 uniforms
  MPCD_HD bool esc_keyed() const { return injected && esc; }
  MPCD_HD double esc_n(uint32_t rep, int b, size_t gp, uint32_t pid, int k) const {
    return esc_keyed() ? esc[((size_t)b * ntot + gp) * kEscStride + k]
                       : rng.normal(rep, b ? P_ESCAPE2 : P_ESCAPE, pid, k);
  }
  MPCD_HD double esc_u(uint32_t rep, int b, size_t gp, uint32_t pid, int k) const {
    return esc_keyed() ? esc[((size_t)b * ntot + gp) * kEscStride + k]
                       : rng.uniform(rep, b ? P_ESCAPE2 : P_ESCAPE, pid, k);
  }
  // variate k = 3r + j of the particle's r-th wall re-thermalisation
  MPCD_HD bool reth_keyed() const { return injected && reth; }
  MPCD_HD double reth_n(uint32_t rep, size_t gp, uint32_t pid, int k) const {
    return reth_keyed() ? reth[gp * kRethStride + k] : rng.normal(rep, P_RETHERM, pid, k);
  }
  MPCD_HD double reth_u(uint32_t rep, size_t gp, uint32_t pid, int k) const {
    return reth_keyed() ? reth[gp * kRethStride + k] : rng.uniform(rep, P_RETHERM, pid, k);
  }
};

}  // namespace mpcd
