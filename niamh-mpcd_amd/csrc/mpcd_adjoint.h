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
    } else {
      orth_matrix(CA, Cc, Sn, M);
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

// ---------------------------------------------------------------------------
// The reverse mode of every operator of the unfused step (mpcd_rev.h): the
// step's kernels instantiated with Rv, one launch per kernel, each on the
// state that kernel saw in the forward run.  The adjoint arrays A follow the
// layout of the tangent's (TanDev); A.dth receives the parameters' adjoints.
// ---------------------------------------------------------------------------
namespace mpcd {

using mpcd_ad::RevTape;
static_assert(kBlock == mpcd_ad::kRevBlock, "reverse launches use the step's workgroup size");
static_assert(kNTheta == MPCD_NTHETA, "parameter adjoints per replica");

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_rev_cellprops(Ctx C, State<Real> S, const BlkDev *cblk, int do_acc,
                                                          TanDev A, RevTape tape) {
  mpcd_ad::rev_begin(tape);
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c < B.end) cellprops_cell<Real, Rv>(C, S, B.rep, c, do_acc, A);
  mpcd_ad::rev_finish();
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_rev_ghost(Ctx C, State<Real> S, const BlkDev *cblk, Draws Dr, TanDev A,
                                                      RevTape tape) {
  Dr.resolve();
  mpcd_ad::rev_begin(tape);
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c < B.end) ghost_cell<Real, Rv>(C, S, B.rep, c, Dr, A);
  mpcd_ad::rev_finish();
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_rev_lc(Ctx C, State<Real> S, const BlkDev *pblk, Draws Dr, TanDev A,
                                                   RevTape tape) {
  Dr.resolve();
  mpcd_ad::rev_begin(tape);
  const BlkDev B = pblk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t < B.end) lc_particle<Real, Rv>(C, S, B.rep, t, Dr, A);
  mpcd_ad::rev_finish();
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_rev_collide(Ctx C, State<Real> S, const BlkDev *cblk, Draws Dr, TanDev A,
                                                        RevTape tape) {
  Dr.resolve();
  mpcd_ad::rev_begin(tape);
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c < B.end) (void)collide_cell<Real, Rv>(C, S, B.rep, c, Dr, A);
  mpcd_ad::rev_finish();
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_rev_stream(Ctx C, State<Real> S, const BlkDev *pblk, Draws Dr, TanDev A,
                                                       const double *boxes, RevTape tape) {
  Dr.resolve();
  mpcd_ad::rev_begin(tape);
  const BlkDev B = pblk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t < B.end) stream_particle<Real, Rv>(C, S, B.rep, t, Dr, nullptr, A, boxes);
  mpcd_ad::rev_finish();
}

// the transposed velocity-gradient stencil (velGradD2Q9 / velGradD3Q15,
// lc.c:2070-2162): E_c[i][j] = 3 sum_k w_k e_k[j] VCM_{c + e_k}[i] gives
// lambda_VCM[c + e_k][i] += 3 w_k e_k[j] lambda_E_c[i][j]; lambda_E_c is
// then spent
__global__ void k_velgrad_bulk_adj(Ctx C, const BlkDev *cblk, double *lvcm, double *lE) {
  const BlkDev B = cblk[blockIdx.x];
  int c = B.beg + threadIdx.x;
  if (c >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  int lc_ = c - R.cell_off;
  int a = lc_ / (R.ny * R.nz), b = (lc_ / R.nz) % R.ny, cz = lc_ % R.nz;
  double *E = &lE[(size_t)c * 9];
  auto idx = [&](int aa, int bb, int cc) { return R.cell_off + (aa * R.ny + bb) * R.nz + cc; };
  if (R.dim == 2) {
    if (!(a >= 1 && a < R.L[0] && b >= 1 && b < R.L[1])) return;
    const int uv[8][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
    const double w1 = 1. / 9., w5 = 1. / 36.;
    for (int k = 0; k < 8; k++) {
      const double w = 3. * ((k < 4) ? w1 : w5);
      const int nb = idx(a + uv[k][0], b + uv[k][1], 0);
      for (int i = 0; i < 2; i++) {
        double s = 0.0;
        for (int j = 0; j < 2; j++) s += uv[k][j] * E[i * 3 + j];
        if (s != 0.0) atomicAdd(&lvcm[(size_t)nb * 3 + i], w * s);
      }
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) E[i * 3 + j] = 0.0;
  } else {
    if (!(a >= 1 && a < R.L[0] && b >= 1 && b < R.L[1] && cz >= 1 && cz < R.L[2])) return;
    const int uv[14][3] = {{1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0},  {0, 0, 1},
                           {0, 0, -1}, {1, 1, 1},  {-1, 1, 1}, {1, -1, 1},  {-1, -1, 1},
                           {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
    const double w1 = 1. / 9., w7 = 1. / 72.;
    for (int k = 0; k < 14; k++) {
      const double w = 3. * ((k < 6) ? w1 : w7);
      const int nb = idx(a + uv[k][0], b + uv[k][1], cz + uv[k][2]);
      for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int j = 0; j < 3; j++) s += uv[k][j] * E[i * 3 + j];
        if (s != 0.0) atomicAdd(&lvcm[(size_t)nb * 3 + i], w * s);
      }
    }
    for (int k = 0; k < 9; k++) E[k] = 0.0;
  }
}

// the edge copies of k_velgrad_edges in reverse order (z faces, y faces, x
// faces): each copy dst = src adds lambda_dst to lambda_src and spends it
__global__ void k_velgrad_edges_adj(Ctx C, double *lE) {
  const RepDev &R = C.rep[blockIdx.x];
  const int X = R.L[0], Y = R.L[1], dim = R.dim;
  const int Z = (dim == 3) ? R.L[2] : 1;
  auto idx = [&](int aa, int bb, int cc) { return R.cell_off + (aa * R.ny + bb) * R.nz + cc; };
  const int zr = (dim == 3) ? Z : 1;
  auto back = [&](int dst, int src, int ij) {
    lE[(size_t)src * 9 + ij] += lE[(size_t)dst * 9 + ij];
    lE[(size_t)dst * 9 + ij] = 0.0;
  };
  if (dim == 3) {
    for (int k = threadIdx.x; k < X * Y; k += blockDim.x) {
      int a = k / Y, b = k % Y;
      for (int ij = 0; ij < 9; ij++) {
        back(idx(a, b, Z), idx(a, b, Z - 1), ij);
        back(idx(a, b, 0), idx(a, b, 1), ij);
      }
    }
    __syncthreads();
  }
  for (int k = threadIdx.x; k < X * zr; k += blockDim.x) {
    int a = k / zr, c = k % zr;
    for (int ij = 0; ij < 9; ij++) {
      if ((ij / 3) >= dim || (ij % 3) >= dim) continue;
      back(idx(a, Y, c), idx(a, Y - 1, c), ij);
      back(idx(a, 0, c), idx(a, 1, c), ij);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < Y * zr; k += blockDim.x) {
    int b = k / zr, c = k % zr;
    for (int ij = 0; ij < 9; ij++) {
      if ((ij / 3) >= dim || (ij % 3) >= dim) continue;
      back(idx(X, b, c), idx(X - 1, b, c), ij);
      back(idx(0, b, c), idx(1, b, c), ij);
    }
  }
}

// adjoints in particle order (replicas concatenated, 9 per particle: q, v,
// u) <-> slot order of the state S (the A arrays)
template <typename Real>
__global__ void k_rev_to_slots(State<Real> S, const BlkDev *blk, const RepDev *reps, const double *lam, TanDev A) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t >= B.end) return;
  const size_t i = (size_t)reps[B.rep].part_off + S.id[t];
  for (int k = 0; k < 3; k++) {
    A.q[k][t] = lam[9 * i + k];
    A.v[k][t] = lam[9 * i + 3 + k];
    A.u[k][t] = lam[9 * i + 6 + k];
  }
}

// after the reverse of a step: slot t of the step's shifted bin came from
// slot skey[t] of the state before it, re-placed if it had escaped (then
// nothing before the bin reaches it)
template <typename Real>
__global__ void k_rev_from_slots(Ctx C, State<Real> Sold, const BlkDev *blk, TanDev A, double *lam) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  const int src = (int)(C.skey[t] & 0x7FFFFFFFull);
  double q[3];
  load_q(Sold, src, q);
  shift_fwd(R, &C.shift[B.rep * 3], q);
  const bool fresh = cell_of(R, q, 1) < 0;
  const size_t i = (size_t)R.part_off + Sold.id[src];
  for (int k = 0; k < 3; k++) {
    lam[9 * i + k] = fresh ? 0.0 : A.q[k][t];
    lam[9 * i + 3 + k] = fresh ? 0.0 : A.v[k][t];
    lam[9 * i + 6 + k] = (fresh && R.lc != MPCD_ISOF) ? 0.0 : A.u[k][t];
  }
}

// the state's slots back to particle order (no bin in between)
template <typename Real>
__global__ void k_rev_unslot(State<Real> S, const BlkDev *blk, const RepDev *reps, TanDev A, double *lam) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t >= B.end) return;
  const size_t i = (size_t)reps[B.rep].part_off + S.id[t];
  for (int k = 0; k < 3; k++) {
    lam[9 * i + k] = A.q[k][t];
    lam[9 * i + 3 + k] = A.v[k][t];
    lam[9 * i + 6 + k] = A.u[k][t];
  }
}

__global__ void k_add_f64(double *a, const double *b, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] += b[i];
}

__global__ void k_fill_f64(double *a, double v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// observables of a state in its unshifted binning, per replica
// band velocity: the mean of v[comp] over the particles with
// lo <= (int)q[axis] < hi; sums[2 r] = count, sums[2 r + 1] = sum
template <typename Real>
__global__ void k_obs_band(Ctx C, State<Real> S, const BlkDev *blk, int comp, int axis, int lo, int hi, double *sums) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  double n = 0.0, s = 0.0;
  if (t < B.end) {
    const Real *qa = axis == 0 ? S.x : axis == 1 ? S.y : S.z;
    const Real *va = comp == 0 ? S.vx : comp == 1 ? S.vy : S.vz;
    const int y = (int)qa[t];
    if (y >= lo && y < hi) {
      n = 1.0;
      s = (double)va[t];
    }
  }
  __shared__ double sh[2][kBlock];
  sh[0][threadIdx.x] = n;
  sh[1][threadIdx.x] = s;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&sums[2 * B.rep], sh[0][0]);
    atomicAdd(&sums[2 * B.rep + 1], sh[1][0]);
  }
}

// its gradient: 1 / count at v[comp] of the particles in the band
template <typename Real>
__global__ void k_obs_band_adj(Ctx C, State<Real> S, const BlkDev *blk, int comp, int axis, int lo, int hi,
                               const double *sums, double *lam) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  const size_t i = (size_t)R.part_off + S.id[t];
  for (int k = 0; k < 9; k++) lam[9 * i + k] = 0.0;
  const Real *qa = axis == 0 ? S.x : axis == 1 ? S.y : S.z;
  const int y = (int)qa[t];
  const double cnt = sums[2 * B.rep];
  if (y >= lo && y < hi && cnt > 0.0) lam[9 * i + 3 + comp] = 1.0 / cnt;
}

// mean S over the cells holding more than one particle: sums[2 r] = count,
// sums[2 r + 1] = sum of S; with w != null, w[c] = 1 / count for those cells
__global__ void k_obs_meanS(Ctx C, const BlkDev *cblk, double *sums) {
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  double n = 0.0, s = 0.0;
  if (c < B.end && C.cell.pop[c] > 1) {
    n = 1.0;
    s = C.cell.S[c];
  }
  __shared__ double sh[2][kBlock];
  sh[0][threadIdx.x] = n;
  sh[1][threadIdx.x] = s;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&sums[2 * B.rep], sh[0][0]);
    atomicAdd(&sums[2 * B.rep + 1], sh[1][0]);
  }
}

__global__ void k_obs_meanS_seed(Ctx C, const BlkDev *cblk, const double *sums, double *lS) {
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c >= B.end) return;
  const double cnt = sums[2 * B.rep];
  lS[c] = (C.cell.pop[c] > 1 && cnt > 0.0) ? 1.0 / cnt : 0.0;
}

}  // namespace mpcd
