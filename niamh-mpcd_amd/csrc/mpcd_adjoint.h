// mpcd_adjoint.h -- the pathwise adjoint of the isotropic step (SRD and
// MPC-AT) with cell assignments, random draws and wall events held fixed.
//
// With those discrete choices frozen, one step maps the particle velocities
// linearly: v1 = v + dt g (acc_all, mpc.c:697), the collision mixes the
// velocities of each cell (SRD: v' = VCM' + R (v1 - VCM'), mpc.c:1181-1333;
// MPC-AT: v' = VCM' (1 - DAMP/POP) + RV - RS, mpc.c:1353-1444; VCM' is the
// ghost-corrected centre-of-mass velocity, mpc.c:549-557), and streaming
// multiplies each component by the wall factors MVN / MVT of the bounces it
// made (bc.c:656-672).  Positions do not feed back into velocities for these
// operators, so the adjoint carries lambda = dJ/dv only:
//
//   lambda' = wmul (.) (lambda + dJ_t/dv)                      (walls, observable)
//   SRD:   lambda1_i = R^T lambda'_i + a (m_i / M)(I - R)^T sum_cell lambda'
//   MPC-AT: lambda1_i = a (m_i / M)(1 - DAMP/POP) sum_cell lambda'
//   dJ/dg += dt sum_i lambda1_i,   dJ/d(angle) += sum_i lambda'_i . (dR/dangle)(v1_i - VCM')
//
// The forward pass records, per step, the tape below (TapeDev); the backward
// pass replays it in reverse.
#pragma once

#include "mpcd_kernels.h"

namespace mpcd {

struct TapeDev {
  int32_t *id, *sp, *start;   // slot -> particle index, species; cell offsets [ncell + 1]
  double *v1;                 // [3 * n] pre-collision velocity (after acc) in slot order
  double *vcm, *a, *mass;     // per cell: ghost-corrected VCM [3], VCM scale, mass
  double *M, *dM;             // per cell SRD rotation [9] and its derivative in the angle [9]
  float *wmul;                // [3 * n] product of the wall factors of the step's bounces (NaN: not smooth)
};

// the SRD matrix for given (cos, sin) values; rows of M in M[0..8]
// (rotate(), mpc.c:1181-1257, including the ORTHAXIS x-axis entry -.1 S)
__device__ __forceinline__ void srd_matrix(int dim, int rtech, int CA, const double L[3], double Cc, double Sn,
                                           double M[9]) {
  if (dim == 3) {
    if (rtech == MPCD_ARBAXIS) {
      M[0] = L[0] * L[0] + (1. - L[0] * L[0]) * Cc;
      M[1] = L[0] * L[1] * (1. - Cc) - L[2] * Sn;
      M[2] = L[0] * L[2] * (1. - Cc) + L[1] * Sn;
      M[3] = L[0] * L[1] * (1. - Cc) + L[2] * Sn;
      M[4] = L[1] * L[1] + (1. - L[1] * L[1]) * Cc;
      M[5] = L[1] * L[2] * (1. - Cc) - L[0] * Sn;
      M[6] = L[0] * L[2] * (1. - Cc) - L[1] * Sn;
      M[7] = L[1] * L[2] * (1. - Cc) + L[0] * Sn;
      M[8] = L[2] * L[2] + (1. - L[2] * L[2]) * Cc;
    } else if (CA == 0) {
      M[0] = 1.; M[1] = 0.; M[2] = 0.; M[3] = 0.; M[4] = Cc; M[5] = Sn; M[6] = 0.; M[7] = -.1 * Sn; M[8] = Cc;
    } else if (CA == 1) {
      M[0] = Cc; M[1] = 0.; M[2] = Sn; M[3] = 0.; M[4] = 1.; M[5] = 0.; M[6] = -1. * Sn; M[7] = 0.; M[8] = Cc;
    } else {
      M[0] = Cc; M[1] = Sn; M[2] = 0.; M[3] = -1. * Sn; M[4] = Cc; M[5] = 0.; M[6] = 0.; M[7] = 0.; M[8] = 1.;
    }
  } else {
    M[0] = Cc; M[1] = -1. * Sn; M[2] = 0.;
    M[3] = Sn; M[4] = Cc; M[5] = 0.;
    M[6] = 0.; M[7] = 0.; M[8] = 0.;
  }
}

// one lane per cell, after the cell properties, ghost particles and gravity
// of the step and before its collision: record the cell's linear map and
// its particles' pre-collision velocities
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_tape_cells(Ctx C, State<Real> S, const BlkDev *cblk, Draws Dr, TapeDev T) {
  Dr.resolve();
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  const int dim = R.dim;
  const int b = C.start[c], e = C.start[c + 1], n = e - b;
  T.start[c] = b;
  if (c == C.nctot - 1) T.start[C.nctot] = e;
  for (int t = b; t < e; t++) {
    T.id[t] = R.part_off + S.id[t];  // particle index over the concatenated replicas
    T.sp[t] = S.sp[t];
    T.v1[3 * t] = (double)S.vx[t];
    T.v1[3 * t + 1] = (double)S.vy[t];
    T.v1[3 * t + 2] = (double)S.vz[t];
  }
  for (int k = 0; k < 3; k++) T.vcm[3 * c + k] = C.cell.vcm[3 * c + k];
  T.mass[c] = C.cell.mass[c];
  // the ghost scale of VCM (ghostPart, mpc.c:549-557): same test as k_ghost
  double a = 1.0;
  if (R.anyphantom && n > 0 && n < R.gndnst) {
    const int nco = (dim == 3) ? 8 : 4;
    int lc_ = c - R.cell_off;
    int ia = lc_ / (R.ny * R.nz), ib = (lc_ / R.nz) % R.ny, ic = lc_ % R.nz;
    int flagW = 0;
    for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
      const mpcd_gpu_wall &W = C.wall[w];
      if (!W.PHANTOM) continue;
      for (int j = 0; j < nco; j++) {
        double tq[3] = {(double)ia + (double)(j & 1), (double)ib + (double)((j >> 1) & 1),
                        (double)ic + (double)((j >> 2) & 1)};
        if (wall_W(W, &C.wq_s[w * 3], tq, dim) < 0.0) flagW = 1;
      }
    }
    if (flagW) a = 1.0 / R.gndnst;
  }
  T.a[c] = a;
  double *M = &T.M[9 * c], *dM = &T.dM[9 * c];
  for (int k = 0; k < 9; k++) M[k] = dM[k] = 0.0;
  if (n <= 1) return;  // not collided (mpc.c:4713)
  if (R.rtech == MPCD_ARBAXIS || R.rtech == MPCD_ORTHAXIS) {
    double RV[3] = {0., 0., 0.};
    int CA = 0;
    if (R.rtech == MPCD_ARBAXIS) {
      if (dim == 3) Dr.cell_axis(B.rep, c, RV);
      else { double x = Dr.cell_u(B.rep, c, 0); RV[0] = x; RV[1] = sqrt(1. - x * x); }
    } else {
      CA = (int)((int)dim * Dr.cell_u(B.rep, c, 0));
    }
    const double sgn = Dr.cell_u(B.rep, c, 3);
    const double s = (R.true_sign ? (sgn >= 0.5) : ((long)sgn > 0.5)) ? 1.0 : -1.0;
    double M0[9], Mc[9], Ms[9];
    srd_matrix(dim, R.rtech, CA, RV, 0.0, 0.0, M0);
    srd_matrix(dim, R.rtech, CA, RV, 1.0, 0.0, Mc);
    srd_matrix(dim, R.rtech, CA, RV, 0.0, 1.0, Ms);
    srd_matrix(dim, R.rtech, CA, RV, R.cosA, s * R.sinA, M);
    // M is affine in (cos, sin): dM/dangle = -sin Mc' + s cos Ms' with Mc' = Mc - M0, Ms' = Ms - M0
    for (int k = 0; k < 9; k++) dM[k] = -R.sinA * (Mc[k] - M0[k]) + s * R.cosA * (Ms[k] - M0[k]);
  }
}

// lambda' in slot order: walls and the observable's contribution
__global__ void k_adj_gather(TapeDev T, const double *lam, const double *adj_in, double *lam_s, int ntot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntot) return;
  const int i = T.id[t];
  for (int k = 0; k < 3; k++) {
    double l = lam[3 * i + k];
    if (adj_in) l += adj_in[3 * i + k];
    lam_s[3 * t + k] = (double)T.wmul[3 * t + k] * l;
  }
}

// one lane per cell: the transposed collision, the parameter derivatives
__global__ __launch_bounds__(kBlock) void k_adj_cells(Ctx C, TapeDev T, const BlkDev *cblk, const double *lam_s, double *lam1_s,
                            double *dtheta) {
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  double dg[3] = {0, 0, 0}, da = 0.0;
  if (c < B.end) {
    const RepDev &R = C.rep[B.rep];
    const int dim = R.dim;
    const int b = T.start[c], e = T.start[c + 1], n = e - b;
    if (n == 1) {
      for (int k = 0; k < 3; k++) lam1_s[3 * b + k] = lam_s[3 * b + k];  // a single particle is not collided
    } else if (n > 1) {
      double Lc[3] = {0, 0, 0};
      for (int t = b; t < e; t++)
        for (int k = 0; k < dim; k++) Lc[k] += lam_s[3 * t + k];
      const double a = T.a[c], Mc = T.mass[c];
      const double *M = &T.M[9 * c], *dM = &T.dM[9 * c];
      const bool srd = (R.rtech == MPCD_ARBAXIS || R.rtech == MPCD_ORTHAXIS);
      double G[3] = {0, 0, 0};  // (I - M)^T Lc  or  f Lc
      if (srd) {
        for (int j = 0; j < dim; j++) {
          double s = Lc[j];
          for (int i = 0; i < dim; i++) s -= M[3 * i + j] * Lc[i];
          G[j] = s;
        }
      } else {
        // MPC-AT: v'_t = VCM' (1 - DAMP_t / POP) + RV_t - RS, so G = sum_t (1 - DAMP_t / POP) lambda'_t
        for (int t = b; t < e; t++) {
          const double ft = 1.0 - C.sp[T.sp[t]].damp / (double)n;
          for (int j = 0; j < dim; j++) G[j] += ft * lam_s[3 * t + j];
        }
      }
      const double *vcm = &T.vcm[3 * c];
      for (int t = b; t < e; t++) {
        const double m = C.sp[T.sp[t]].mass, w = a * m / Mc;
        const double *l = &lam_s[3 * t];
        double o[3] = {0, 0, 0};
        for (int j = 0; j < dim; j++) {
          double s = w * G[j];
          if (srd)
            for (int i = 0; i < dim; i++) s += M[3 * i + j] * l[i];
          o[j] = s;
        }
        if (srd) {
          const double *v = &T.v1[3 * t];
          for (int i = 0; i < dim; i++) {
            double r = 0.0;
            for (int j = 0; j < dim; j++) r += dM[3 * i + j] * (v[j] - vcm[j]);
            da += l[i] * r;
          }
        }
        for (int k = 0; k < 3; k++) lam1_s[3 * t + k] = (k < dim) ? o[k] : l[k];
      }
    }
    if (n > 0)
      for (int t = b; t < e; t++)
        for (int k = 0; k < dim; k++) dg[k] += R.dt * lam1_s[3 * t + k];
  }
  // per-replica accumulation: dtheta[4 r + 0..2] = dJ/dg, dtheta[4 r + 3] = dJ/dangle
  __shared__ double sh[kBlock][4];
  sh[threadIdx.x][0] = dg[0];
  sh[threadIdx.x][1] = dg[1];
  sh[threadIdx.x][2] = dg[2];
  sh[threadIdx.x][3] = da;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 4; k++) sh[threadIdx.x][k] += sh[threadIdx.x + off][k];
    __syncthreads();
  }
  if (threadIdx.x < 4) atomicAdd(&dtheta[4 * B.rep + threadIdx.x], sh[0][threadIdx.x]);
}

__global__ void k_adj_scatter(TapeDev T, const double *lam1_s, double *lam, int ntot) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntot) return;
  const int i = T.id[t];
  for (int k = 0; k < 3; k++) lam[3 * i + k] = lam1_s[3 * t + k];
}

}  // namespace mpcd
