// mpcd_walls.h -- the general wall engine of the streaming step for gfx950:
// surfaces that are not axis-aligned planes (ellipses, circles, cylinders,
// superellipses, pipes, their inverted forms, |.| surfaces and wavy planes)
// for immobile walls with the four-fold rotational symmetry (ROTSYMM 4 4).
//
// A wall is the level set W(r) = 0 of
//     W(r) = sum_i (A_i (r_i - Q_i))^P_i - R^P_3  [+ wavy term]  [* -1 if INV]
// (calcW, bc.c:55-80; smrtPow, mtools.c:38-50; calcWavyW, bc.c:94-150), with
// W < 0 inside the wall.  A particle that ends a streaming step inside a wall
// is taken back along its velocity to the crossing (crosstimeReverse,
// bc.c:251-298: the plane formula, the quadratic of a P = 2 surface, or the
// secant iteration of bc.c:329-373 for the rest; chooseT, bc.c:849-880), the
// wall's rule-based collision is applied at the surface normal there
// (normal / normalWavy, bc.c:1516-1640), and it streams on for the time left.
// Curved walls act through the periodic image nearest the particle (shiftBC,
// bc.c:389-422).
//
// Only FP64 arithmetic in the operation order of the reference: the library is
// compiled with -ffp-contract=off, so +, -, *, / and sqrt round as the host's
// do; pow, cos, sin and atan may differ from glibc's in the last bit.
#pragma once

#include "mpcd_math.h"
#include "../../include/mpcd_gpu.h"

namespace mpcd {

// smrtPow (mtools.c:38-50): integer exponents up to 10 by repeated products
// (a negative integer exponent gives 1, as the reference's loop does)
__device__ __forceinline__ double smrt_pow(double x, double y) {
  if ((y - (int)y == 0) && (y <= 10)) {
    double r = 1;
    for (int i = 0; i < y; i++) r *= x;
    return r;
  }
  return pow(x, y);
}

// the reference's PLANAR flag (read.c:1423-1434): an axis-aligned plane that
// is not wavy
__host__ __device__ __forceinline__ bool wall_is_plane(const mpcd_gpu_wall &W) {
  if (!feq(W.B[0], 0.0)) return false;
  if (!(feq(W.P[0], 1.0) && feq(W.P[1], 1.0) && feq(W.P[2], 1.0))) return false;
  const bool a0 = fneq(W.A[0], 0.0), a1 = fneq(W.A[1], 0.0), a2 = fneq(W.A[2], 0.0);
  const bool z0 = feq(W.A[0], 0.0), z1 = feq(W.A[1], 0.0), z2 = feq(W.A[2], 0.0);
  return (a0 && z1 && z2) || (z0 && a1 && z2) || (z0 && z1 && a2);
}

// the wavy term of calcW (calcWavyW, bc.c:94-150) for the wall at wq
__device__ __forceinline__ double wavy_w(const mpcd_gpu_wall &W, const double *wq, const double *pos, double Wv,
                                         int dim) {
  double W1 = 0.0, W2 = 0.0;
  int flag = 0;
  for (int i = 0; i < dim; i++)
    if (!feq(W.P[i], 1.0)) flag += 1;
  if (!flag) {
    // a wavy plane: the phase runs along the plane
    if (!feq(W.B[1], 0.0)) W1 = (W.A[1] * pos[0] - W.A[0] * pos[1]) / sqrt(W.A[0] * W.A[0] + W.A[1] * W.A[1]);
    if (dim > 2 && !feq(W.B[2], 0.0)) {
      W2 = W.A[0] * W.A[2] * pos[0] + W.A[1] * W.A[2] * pos[1] - (W.A[0] * W.A[0] + W.A[1] * W.A[1]) * pos[2];
      W2 /= sqrt(pow(W.A[0], 4) + pow(W.A[1], 4) + pow(W.A[0], 2) * pow(W.A[2], 2) +
                 2.0 * pow(W.A[0], 2) * pow(W.A[1], 2) + pow(W.A[1], 2) * pow(W.A[2], 2));
    }
  } else {
    flag = 0;
    for (int i = 0; i < dim; i++)
      if (!feq(W.P[i], 2.0) || feq(W.A[i], 0.0)) flag += 1;
    if (!flag) {
      // a wavy ellipse or ellipsoid: the phase runs with the polar angles
      if (!feq(W.B[1], 0.0)) W1 = ref_atan2(W.A[1] * (pos[1] - wq[1]), W.A[0] * (pos[0] - wq[0]));
      if (dim > 2 && !feq(W.B[2], 0.0))
        W2 = ref_atan2(sqrt(pow(W.A[0] * (pos[0] - wq[0]), 2) + pow(W.A[1] * (pos[1] - wq[1]), 2)),
                       W.A[2] * (pos[2] - wq[2]));
    } else if (dim > 2 && flag == 1) {
      // a wavy cylinder: the axis is the component with A = 0
      int f = -1;
      for (int i = 0; i < dim; i++)
        if (feq(W.A[i], 0.0)) f = i;
      if (f != -1) {
        const int j = f, k = (f + 1) % 3, l = (f + 2) % 3;
        if (!feq(W.B[1], 0.0)) W1 = ref_atan2(W.A[l] * (pos[l] - wq[l]), W.A[k] * (pos[k] - wq[k]));
        if (!feq(W.B[2], 0.0)) W2 = ref_atan2(sqrt(Wv + pow(W.R, W.P[3])), sqrt(pos[j]));
      }
    }
  }
  return W.B[0] * cos(W.B[1] * W1) * cos(W.B[2] * W2);
}

// calcW (bc.c:55-80) of the wall at wq, four-fold symmetric surfaces
__device__ __forceinline__ double calc_w(const mpcd_gpu_wall &W, const double *wq, const double *pos, int dim) {
  double terms, Wv = 0.0;
  for (int i = 0; i < dim; i++) {
    terms = W.A[i] * (pos[i] - wq[i]);
    if (W.ABS) terms = fabs(terms);
    terms = smrt_pow(terms, W.P[i]);
    Wv += terms;
  }
  terms = W.R;
  if (W.ABS) terms = fabs(terms);
  terms = smrt_pow(terms, W.P[3]);
  Wv -= terms;
  if (!feq(W.B[0], 0.0)) Wv += wavy_w(W, wq, pos, Wv, dim);
  if (W.INV) Wv *= -1.0;
  return Wv;
}

// walls the planar fast path cannot take: not PLANAR, or a plane whose calcW
// takes absolute values (its calcW then differs from calcW_PLANE)
__host__ __device__ __forceinline__ bool wall_is_general(const mpcd_gpu_wall &W) {
  return !wall_is_plane(W) || W.ABS != 0;
}

// calcW_PLANE (bc.c:306-317)
__device__ __forceinline__ double calc_w_plane(const mpcd_gpu_wall &W, const double *wq, const double *pos,
                                               int dim) {
  double Wv = 0.;
  for (int i = 0; i < dim; i++) Wv += W.A[i] * (pos[i] - wq[i]);
  Wv -= W.R;
  if (W.INV) Wv *= -1.;
  return Wv;
}

// secant_time (bc.c:329-373): the time back along -V at which the surface was
// crossed, from the end point Q and the start point Q - t_step V
__device__ __forceinline__ double secant_time(const mpcd_gpu_wall &W, const double *wq, const double *Q,
                                              const double *V, int dim, double t_step) {
  double Qi[3] = {0, 0, 0}, QiM1[3] = {0, 0, 0}, ti, tiM1, root, fi, fiM1;
  int iter = 0;
  for (int i = 0; i < dim; i++) {
    QiM1[i] = Q[i];
    Qi[i] = Q[i] + t_step * (-V[i]);
  }
  tiM1 = 0.0;
  ti = t_step;
  root = ti;
  do {
    iter++;
    fi = calc_w(W, wq, Qi, dim);
    fiM1 = calc_w(W, wq, QiM1, dim);
    if (!feq(fi, fiM1)) root = ti - fi * (ti - tiM1) / (fi - fiM1);
    if (root < kTOL / 100.) {
      ti = tiM1;
      tiM1 = root;
    } else {
      tiM1 = ti;
      ti = root;
    }
    for (int i = 0; i < dim; i++) {
      QiM1[i] = Q[i] + tiM1 * (-V[i]);
      Qi[i] = Q[i] + ti * (-V[i]);
    }
  } while (iter < 10 && fabs(ti - tiM1) > kTOL / 100. && fabs(fi - fiM1) > kTOL / 100.);
  return root;
}

// crosstimeReverse (bc.c:251-298): the two candidate times back to the
// surface (equal unless the surface is a quadric with P = 2)
__device__ __forceinline__ void crosstime_reverse(const mpcd_gpu_wall &W, bool plane, const double *wq,
                                                  const double *Q, const double *V, int dim, double t_step,
                                                  double *tc_pos, double *tc_neg) {
  if (plane || (feq(W.P[0], 1.0) && feq(W.P[1], 1.0) && feq(W.P[2], 1.0) && feq(W.P[3], 1.0) &&
                feq(W.B[0], 0.0))) {
    double t = W.R;
    for (int i = 0; i < dim; i++) t += W.A[i] * (wq[i] - Q[i]);
    t /= (W.A[0] * (-V[0]) + W.A[1] * (-V[1]) + W.A[2] * (-V[2]));
    *tc_pos = t;
    *tc_neg = t;
  } else if ((dim == 2 && feq(W.P[0], 2.0) && feq(W.P[1], 2.0) && feq(W.B[0], 0.0)) ||
             (dim > 2 && feq(W.P[0], 2.0) && feq(W.P[1], 2.0) && feq(W.P[2], 2.0) && feq(W.B[0], 0.0))) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < dim; i++) {
      a += W.A[i] * W.A[i] * (-V[i]) * (-V[i]);
      b += W.A[i] * W.A[i] * (-V[i]) * (Q[i] - wq[i]);
      c += W.A[i] * W.A[i] * (Q[i] * Q[i] - 2. * Q[i] * wq[i] + wq[i] * wq[i]);
    }
    b *= 2.0;
    c -= smrt_pow(W.R, W.P[3]);
    double tn = -b - sqrt(b * b - 4. * a * c);
    tn /= (2. * a);
    double tp = -b + sqrt(b * b - 4. * a * c);
    tp /= (2. * a);
    *tc_neg = tn;
    *tc_pos = tp;
  } else {
    double t = secant_time(W, wq, Q, V, dim, t_step);
    if (t < 0 || t > t_step) t = t_step;
    *tc_pos = t;
    *tc_neg = t;
  }
}

// chooseT (bc.c:849-880): the collision time out of the two candidates
__device__ __forceinline__ double choose_t(double tstep, double tp, double tn) {
  double tc = tstep;
  const double zero = -kTOL, step = tstep + kTOL;
  if (tp > step && tn > step) tc = 0.;
  else if (tp < zero && tn < zero) tc = 0.;
  else if (tp < zero) tc = tn;
  else if (tn < zero) tc = tp;
  else if (tn < tp) tc = tn;
  else tc = tp;
  if (tc < zero) tc = 0.;
  if (tc > step) tc = 0.;
  return tc;
}

// shiftBC (bc.c:389-405): the position of the periodic image of a curved
// wall nearest the point q (the wall itself when it is a plane)
__device__ __forceinline__ void wall_image(const mpcd_gpu_wall &W, const int *L, int dim, const double *wq,
                                           const double *q, double *wqi) {
  for (int k = 0; k < 3; k++) wqi[k] = wq[k];
  if (fneq(W.P[0], 1.0) && fneq(W.P[1], 1.0)) {
    for (int k = 0; k < dim; k++) {
      double s = 0.0;
      if (q[k] - wq[k] >= 0.5 * (double)L[k]) s = (double)L[k];
      else if (q[k] - wq[k] <= -0.5 * (double)L[k]) s = -1.0 * (double)L[k];
      wqi[k] = wq[k] + s;
    }
  }
}

// the wavy part of the surface normal (normalWavy, bc.c:1568-1640)
__device__ __forceinline__ void wavy_normal(const mpcd_gpu_wall &W, const double *wq, const double *point,
                                            int dimension, double *n) {
  double W1 = 0.0, W2 = 0.0, div, dw1[3] = {0, 0, 0}, dw2[3] = {0, 0, 0};
  if (feq(W.P[0], 1.0) && feq(W.P[1], 1.0) && feq(W.P[2], 1.0)) {
    if (!feq(W.B[1], 0.0)) {
      const double h = sqrt(W.A[0] * W.A[0] + W.A[1] * W.A[1]);
      W1 = (W.A[1] * point[0] - W.A[0] * point[1]) / h;
      dw1[0] = W.A[1] / h;
      dw1[1] = -W.A[0] / h;
      dw1[2] = 0;
    }
    if (dimension > 2 && !feq(W.B[2], 0.0)) {
      div = sqrt(pow(W.A[0], 4) + pow(W.A[1], 4) + pow(W.A[0], 2) * pow(W.A[2], 2) +
                 2.0 * pow(W.A[0], 2) * pow(W.A[1], 2) + pow(W.A[1], 2) * pow(W.A[2], 2));
      W2 = (W.A[0] * W.A[2] * point[0] + W.A[1] * W.A[2] * point[1] -
            (W.A[0] * W.A[0] + W.A[1] * W.A[1]) * point[2]) / div;
      dw2[0] = (W.A[0] * W.A[2]) / div;
      dw2[1] = (W.A[1] * W.A[2]) / div;
      dw2[2] = -(W.A[0] * W.A[0] + W.A[1] * W.A[1]) / div;
    }
  } else {
    int flag = 0;
    for (int i = 0; i < dimension; i++)
      if (!feq(W.P[i], 2.0) || feq(W.A[i], 0.0)) flag += 1;
    if (!flag) {
      div = pow(W.A[0] * (point[0] - wq[0]), 2) + pow(W.A[1] * (point[1] - wq[1]), 2);
      W1 = ref_atan2(W.A[1] * (-wq[1] + point[1]), W.A[0] * (-wq[0] + point[0]));
      if (!feq(W.B[1], 0.0)) {
        dw1[0] = W.A[0] * W.A[1] * (wq[1] - point[1]) / div;
        dw1[1] = W.A[0] * W.A[1] * (point[0] - wq[0]) / div;
        dw1[2] = 0;
      }
      if (dimension > 2 && !feq(W.B[2], 0.0)) {
        const double dz = pow(W.A[2] * (point[2] - wq[2]), 2);
        dw2[0] = W.A[2] * W.A[0] * (point[0] - wq[0]) * (point[2] - wq[2]) / (sqrt(div) * (div + dz));
        dw2[1] = W.A[2] * W.A[1] * (point[1] - wq[1]) * (point[2] - wq[2]) / (sqrt(div) * (div + dz));
        dw2[2] = -W.A[2] * sqrt(div) / (div + dz);
      }
    } else if (dimension > 2 && flag == 1) {
      int f = -1;
      for (int i = 0; i < dimension; i++)
        if (feq(W.A[i], 0.0)) f = i;
      if (f != -1) {
        const int j = f, k = (f + 1) % 3, l = (f + 2) % 3;
        div = pow(W.A[k] * (point[k] - wq[k]), 2) + pow(W.A[l] * (point[l] - wq[l]), 2);
        if (!feq(W.B[1], 0.0)) {
          dw1[k] = W.A[k] * W.A[l] * (wq[l] - point[l]) / div;
          dw1[l] = W.A[k] * W.A[l] * (point[k] - wq[k]) / div;
          dw1[j] = 0;
        }
        if (!feq(W.B[2], 0.0)) {
          dw2[k] = W.A[k] * W.A[k] * (point[k] - wq[k]) * point[j] / (sqrt(div) * (div + point[k]));
          dw2[l] = W.A[l] * W.A[l] * (point[l] - wq[l]) * point[j] / (sqrt(div) * (div + point[l]));
          dw2[j] = -sqrt(div) / (div + point[j]);
        }
      }
    }
  }
  for (int i = 0; i < 3; i++)
    n[i] -= W.B[0] * (W.B[1] * dw1[i] * sin(W.B[1] * W1) * cos(W.B[2] * W2) +
                      W.B[2] * dw2[i] * cos(W.B[1] * W1) * sin(W.B[2] * W2));
}

// the surface normal at point (normal, bc.c:1516-1556), not normalised
__device__ __forceinline__ void wall_normal(const mpcd_gpu_wall &W, const double *wq, const double *point,
                                            int dimension, double *n) {
  if (feq(W.P[0], 1.0) && feq(W.P[1], 1.0) && feq(W.P[2], 1.0)) {
    for (int i = 0; i < dimension; i++) n[i] = W.A[i];
  } else if ((dimension == 2 && feq(W.P[0], 2.0) && feq(W.P[1], 2.0)) ||
             (dimension > 2 && feq(W.P[0], 2.0) && feq(W.P[1], 2.0) && feq(W.P[2], 2.0))) {
    for (int i = 0; i < dimension; i++) n[i] = 2.0 * W.A[i] * W.A[i] * (point[i] - wq[i]);
  } else if ((dimension == 2 && feq(W.P[0], 4.0) && feq(W.P[1], 4.0)) ||
             (dimension > 2 && feq(W.P[0], 4.0) && feq(W.P[1], 4.0) && feq(W.P[2], 4.0))) {
    for (int i = 0; i < dimension; i++) {
      const double d = point[i] - wq[i];
      n[i] = 4.0 * W.A[i] * W.A[i] * W.A[i] * W.A[i] * d * d * d;
    }
  } else if ((dimension == 2 && feq(W.P[0], 6.0) && feq(W.P[1], 6.0)) ||
             (dimension > 2 && feq(W.P[0], 6.0) && feq(W.P[1], 6.0) && feq(W.P[2], 6.0))) {
    // the reference's coefficient for P = 6 is 4 A^6 (not 6 A^6), kept
    for (int i = 0; i < dimension; i++) {
      const double d = point[i] - wq[i];
      n[i] = 4.0 * W.A[i] * W.A[i] * W.A[i] * W.A[i] * W.A[i] * W.A[i];
      n[i] *= d * d * d * d * d;
    }
  } else if (W.ABS) {
    for (int i = 0; i < dimension; i++)
      n[i] = W.A[i] * W.A[i] * W.P[i] * (point[i] - wq[i]) * smrt_pow(fabs(W.A[i] * (point[i] - wq[i])), W.P[i] - 2.0);
  } else {
    for (int i = 0; i < dimension; i++)
      n[i] = W.P[i] * smrt_pow(W.A[i], W.P[i]) * smrt_pow(point[i] - wq[i], W.P[i] - 1.0);
  }
  if (!feq(W.B[0], 0.0)) wavy_normal(W, wq, point, dimension, n);
}

}  // namespace mpcd
