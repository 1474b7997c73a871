// mpcd_fused.h -- the fused, LDS-tiled step kernel for the isotropic
// collision operators (SRD, MPC-AT, MPC-AT+a).
//
// One wavefront owns a tile of consecutive cells.  It
//   1. loads the tile's sorted 64-bit keys and ranks them inside each cell
//      (the composite key reproduces the reference's linked-list order),
//   2. gathers the particles' r, v (and id, species) from the previous
//      buffer into LDS, one lane per particle, applying the grid shift and
//      the body force and drawing the thermal velocities of MPC-AT(+a),
//   3. runs one lane per cell over its particles in list order for the FP64
//      cell sums (mass, centre-of-mass velocity and position, inertia
//      tensor, random-velocity sums, angular momentum) -- the reference's own
//      summation order, so FP64 results match it to round-off,
//   4. applies the collision, the kinetic-energy partial sum, the shift
//      back, streaming and the planar walls with one lane per particle; a
//      particle strictly inside the step's wall-free box (computed with the
//      wall positions) skips the wall search, which would find nothing,
//   5. writes the new state coalesced, and -- when the next step's grid
//      shift is known -- the next step's cell, in-cell rank and key plus its
//      histogram (global atomics), so the next step starts at the scan.
// Tiles larger than the LDS capacity are processed in cell-aligned chunks.
#pragma once

#include "mpcd_kernels.h"

namespace mpcd {

constexpr int kTileThreads = 64;  // one wavefront per tile: no cross-wave barriers
constexpr int kTileCells = 32;    // cells per tile at most

enum { OP_SRD = 0, OP_AT = 1, OP_RAT = 2 };

// particles staged in LDS at once (sized for about a dozen workgroups per CU)
__host__ __device__ constexpr int tile_cap(bool fp64, int op) { return (fp64 || op != OP_SRD) ? 192 : 256; }

// cell_of / shift_fwd / shift_back with the dimension known at compile time
// (same arithmetic as the runtime-dimension versions)
template <int DIM>
__device__ __forceinline__ int cell_of_d(const RepDev &R, const double q[3], int shifted) {
  int a = (int)q[0], b = (int)q[1], c = (int)q[2];
  const int Lz = (DIM == 3) ? R.L[2] : 1;
  if (shifted) {
    if (R.pbc[0] && a == R.L[0]) a = 0;
    if (R.pbc[1] && b == R.L[1]) b = 0;
    if (R.pbc[2] && c == Lz) c = 0;
  }
  if (a > R.L[0] || b > R.L[1] || c > Lz || a < 0 || b < 0 || c < 0) return -1;
  if (DIM < 3) c = 0;
  return R.cell_off + (a * R.ny + b) * R.nz + c;
}

template <int DIM>
__device__ __forceinline__ void shift_fwd_d(const RepDev &R, const double *s, double q[3]) {
#pragma unroll
  for (int j = 0; j < DIM; j++) {
    q[j] += s[j];
    if (q[j] > R.L[j] && q[j] <= R.L[j] + s[j]) q[j] -= R.L[j];
  }
}

template <int DIM>
__device__ __forceinline__ void shift_back_d(const RepDev &R, const double *s, double q[3]) {
#pragma unroll
  for (int j = 0; j < DIM; j++) {
    double sg = -1.0 * s[j];
    q[j] += sg;
    if (q[j] < 0.0 && q[j] >= sg) q[j] += R.L[j];
  }
}

// MPC_BCcollision for planar walls (bc.c:1363-1496) on an already drifted
// particle: search the walls, bounce, repeat until no wall is crossed
template <int DIM>
__device__ __forceinline__ void wall_loop(const Ctx &C, const RepDev &R, const mpcd_gpu_species &sp,
                                       const mpcd_gpu_wall *__restrict__ walls, const WallHot *__restrict__ wh,
                                       const double *__restrict__ wq_b, const Draws &Dr, int rep, int spid, int pid,
                                       double Q[3], double V[3], double U[3], bool lc) {
  const int dim = DIM;
  double t_step = R.dt, t_left = t_step, t_coll = 0.0;
  int chosen = 0, cnt = 0, cnt2 = 0, flagCNT = 0, rk = 0;
  double n[3] = {0, 0, 0};
  while (R.nwall > 0) {
    double tmin = t_left, chW = 1.0;
    for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
      const WallHot &H = wh[w];
      if (!((H.mask >> spid) & 1)) continue;
      double tw = wall_W_hot(H, Q, H.wqb);
      if (tw < 0.0) {
        const mpcd_gpu_wall &W = walls[w];
        const double *wq = &wq_b[w * 3];
        double tc = W.R;
        for (int i = 0; i < dim; i++) tc += W.A[i] * (wq[i] - Q[i]);
        tc /= (W.A[0] * (-V[0]) + W.A[1] * (-V[1]) + W.A[2] * (-V[2]));
        double t1 = t_left - tc;
        if (t1 < tmin) { tmin = t1; chosen = w; chW = tw; }
      }
    }
    t_coll = tmin;
    if (chW > -kTOL) break;
    cnt++;
    const mpcd_gpu_wall &W = walls[chosen];
    const double *wq = &wq_b[chosen * 3];
    if (t_coll > -kTOL) {
      for (int i = 0; i < dim; i++) Q[i] = Q[i] - (t_left - t_coll) * V[i];
      for (int i = 0; i < dim; i++) n[i] = W.A[i];
      normn(n, dim);
      double Rr[3] = {0, 0, 0}, VR[3], Vr[3] = {0, 0, 0}, VN[3] = {0, 0, 0}, VT[3] = {0, 0, 0};
      for (int i = 0; i < dim; i++) Rr[i] = Q[i] - wq[i];
      cross3(W.L, Rr, VR);
      for (int i = 0; i < dim; i++) Vr[i] = V[i] - W.V[i] - VR[i];
      projn(Vr, n, VN, dim);
      for (int i = 0; i < dim; i++) VT[i] = Vr[i] - VN[i];
      double IMpart = 1.0 / sp.mass;
      for (int i = 0; i < dim; i++) {
        VN[i] *= W.MVN;
        VN[i] += W.DVN * n[i];
        VT[i] *= W.MVT;
        VT[i] += W.DVT;
      }
      for (int i = 0; i < dim; i++) Vr[i] = VN[i] + VT[i];
      for (int i = 0; i < dim; i++) Vr[i] += W.DVxyz[i];
      for (int i = 0; i < dim; i++) Vr[i] += W.V[i] + VR[i];
      for (int i = 0; i < dim; i++) {
        Vr[i] -= V[i];
        Vr[i] *= sp.mass;
      }
      for (int i = 0; i < dim; i++) V[i] += Vr[i] * 1.0 * IMpart;
      double PN[3] = {0, 0, 0}, PT[3] = {0, 0, 0};
      projn(Q, n, PN, dim);
      for (int i = 0; i < dim; i++) PT[i] = Q[i] - PN[i];
      for (int i = 0; i < dim; i++) {
        PN[i] += W.DN * n[i] * (1. - 2. * kTOL);
        PT[i] += W.DT;
      }
      for (int i = 0; i < dim; i++) Q[i] = PN[i] + PT[i];
      if (lc) ori_bc(U, W, n, dim);
      t_left -= t_coll;
      for (int i = 0; i < dim; i++) Q[i] = Q[i] + t_left * V[i];
      if (wall_W(W, wq, Q, dim) < -kTOL)
        for (int i = 0; i < dim; i++) Q[i] = Q[i] - t_left * V[i];
    } else {
      for (int i = 0; i < dim; i++) Q[i] = Q[i] - (t_left - t_coll) * V[i];
    }
    if (cnt > kNBOUNCE) {
      flagCNT = 1;
      if (cnt < 2 * kNBOUNCE && cnt2 == 0) {
        atomicAdd(&C.events[2], 1ull);
        if (t_left > kTOL) t_left *= 1.1;
        else {
          t_left = 0.9 * t_step;
          t_coll = 0.1 * t_step;
        }
      } else {
        atomicAdd(&C.events[3], 1ull);
        if (Dr.reth_keyed() && rk + DIM > kRethStride) {
          atomicOr(&C.flags[0], kFlagPushes);
          break;
        }
        push_velocity(sp, R.kbt, DIM, Dr, rep, (size_t)R.part_off + pid, pid, rk, V);
        rk += 3;
        t_left = t_step;
        t_coll = 0.0;
        cnt2++;
        cnt = 0;
      }
      if (cnt2 > 1000) {  // a particle the walls leave no room for: the reference would loop for ever
        atomicOr(&C.flags[0], kFlagStuck);
        break;
      }
    }
  }
  if (flagCNT) atomicAdd(&C.events[1], 1ull);
}

// the unshifted bin after streaming (mpc.c:4936): escaped particles are
// re-placed and stay in the list they were in (mpc.c:970-982)
template <int DIM>
__device__ __forceinline__ int unshifted_cell(const Ctx &C, const RepDev &R, const mpcd_gpu_species &sp,
                                              const WallHot *wh, const Draws &Dr, int rep, int pid, int cell_this,
                                              double Q[3], double V[3], double U[3], bool lc) {
  int cu = cell_of_d<DIM>(R, Q, 0);
  if (cu < 0) {
    atomicAdd(&C.events[0], 1ull);
    if (!escape_repair(R, wh, 1, Dr, rep, pid, sp.mass, lc, Q, V, U)) atomicOr(&C.flags[0], kFlagEscape);
    cu = cell_this;
  }
  return cu;
}

// stream_all (mpc.c:680) + walls + unshifted cell of one particle.  `box`
// holds the step's wall-free box [lo0, hi0, lo1, hi1, lo2, hi2]: a drifted
// particle strictly inside it crosses no wall, so the search is skipped.
template <int DIM>
__device__ __forceinline__ int stream_walls(const Ctx &C, const RepDev &R, const mpcd_gpu_species *__restrict__ sps,
                                            const mpcd_gpu_wall *__restrict__ walls,
                                            const WallHot *__restrict__ wh, const double *__restrict__ wq_b,
                                            const double *__restrict__ box, const Draws &Dr, int rep, int spid,
                                            int pid, int cell_this, double Q[3], double V[3], double U[3], bool lc) {
  const mpcd_gpu_species &sp = sps[spid];
#pragma unroll
  for (int i = 0; i < DIM; i++) Q[i] = Q[i] + R.dt * V[i];
  bool inside = true;
#pragma unroll
  for (int i = 0; i < DIM; i++) inside = inside && (Q[i] > box[2 * i]) && (Q[i] < box[2 * i + 1]);
  if (!inside) wall_loop<DIM>(C, R, sp, walls, wh, wq_b, Dr, rep, spid, pid, Q, V, U, lc);
  return unshifted_cell<DIM>(C, R, sp, wh, Dr, rep, pid, cell_this, Q, V, U, lc);
}

// next-step shifted cell of a streamed particle (bin(shifted=1), mpc.c:948);
// an escaped particle stays in its list
template <int DIM>
__device__ __forceinline__ int next_cell(const RepDev &R, const double *shift_next, int cu, const double Q[3]) {
  double q[3] = {Q[0], Q[1], Q[2]};
  shift_fwd_d<DIM>(R, shift_next, q);
  int cn = cell_of_d<DIM>(R, q, 1);
  return cn < 0 ? cu : cn;
}

__device__ __forceinline__ uint64_t bin_key(int cn, int cu, int cell_this, int slot) {
  uint32_t a1 = (cu != cell_this) ? 1u : 0u, a2 = (cn != cu) ? 1u : 0u;
  return ((uint64_t)a2 << 63) | ((uint64_t)(uint32_t)cu << 32) | ((uint64_t)a1 << 31) | (uint64_t)(uint32_t)slot;
}

// next-step binning of one particle: histogram + rank + composite key
template <int DIM>
__device__ __forceinline__ void next_bin(const Ctx &C, const RepDev &R, const double *shift_next, int32_t *hist_next,
                                         int slot, int cell_this, int cu, const double Q[3]) {
  const int cn = next_cell<DIM>(R, shift_next, cu, Q);
  const int l = atomicAdd(&hist_next[cn], 1);
  C.cnew[slot] = cn;
  C.loc[slot] = l;
  C.key[slot] = bin_key(cn, cu, cell_this, slot);
}

// the terms of one cell whose particles sit at positions [b, e) of the LDS
// arrays (component d of particle p at [d * stride + p]): FP64 sums in list
// order (localPROP, mpc.c:64-216), ghost particles at phantom walls, and the
// collision's cell data cd[12]: vcm[3], then the SRD matrix[9] or the
// thermal-draw sum RS[3], W[3] (angular velocity) and CM[3].  GHOST = false
// leaves out the ghost particles (no replica has phantom walls).
template <typename Real, int OP, bool ONESP, int DIM, bool GHOST = true>
__device__ __forceinline__ void cell_terms(const RepDev &R, const mpcd_gpu_species *__restrict__ sps,
                                           const WallHot *__restrict__ wh, const Draws &Dr, int rep, int c, int b,
                                           int e, const Real *sv, const Real *sq, const double *srv,
                                           const uint8_t *ssp, int stride, double *cd,
                                           const double *cdr_pre = nullptr) {
  const int np_ = e - b;
  double mass = 0.0, vcm[3] = {0, 0, 0}, cm[3] = {0, 0, 0};
  // the loads of several particles are in flight at once; the sums keep
  // their sequential order
#pragma unroll 4
  for (int p = b; p < e; p++) {
    const double m = sps[ONESP ? 0 : ssp[p]].mass;
    mass += m;
    for (int d = 0; d < DIM; d++) vcm[d] += (double)sv[d * stride + p] * m;
    if (OP == OP_RAT)
      for (int d = 0; d < DIM; d++) cm[d] += (double)sq[d * stride + p] * m;
  }
  if (mass > 0.0)
    for (int d = 0; d < DIM; d++) {
      vcm[d] /= mass;
      if (OP == OP_RAT) cm[d] /= mass;
    }
  // ghost particles at phantom walls (ghostPart, mpc.c:351-567)
  if (GHOST && R.anyphantom && np_ > 0) {
    int lc_ = c - R.cell_off;
    int a = lc_ / (R.ny * R.nz), bb = (lc_ / R.nz) % R.ny, cz = lc_ % R.nz;
    const int nco = (DIM == 3) ? 8 : 4;
    int flagW = 0;
    for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
      const WallHot &H = wh[w];
      if (!H.phantom) continue;
      for (int j = 0; j < nco; j++) {
        double tq[3] = {(double)a + (double)(j & 1), (double)bb + (double)((j >> 1) & 1),
                        (double)cz + (double)((j >> 2) & 1)};
        if (wall_W_hot(H, tq, H.wqs) < 0.0) flagW = 1;
      }
    }
    if (flagW && np_ < R.gndnst) {
      double invN = 1.0 / (R.gndnst - (double)np_);
      for (int d = 0; d < DIM; d++) {
        double sigma = sqrt(R.kbt / invN);
        vcm[d] += sigma * Dr.ghost_n(rep, c, d);
        vcm[d] /= R.gndnst;
      }
    }
  }
  cd[0] = vcm[0];
  cd[1] = vcm[1];
  cd[2] = vcm[2];
  if (np_ <= 1) return;
  if (OP == OP_SRD) {
    // the cell's draws in the keyed layout: injected, or written by
    // k_celldraws for this step (possibly already staged by the caller)
    const double *cdr = cdr_pre ? cdr_pre : &Dr.cell[(size_t)c * 4];
    double RV[3] = {0., 0., 0.};
    int CA = 0;
    if (R.rtech == MPCD_ARBAXIS) {
      if (DIM == 3) { RV[0] = cdr[0]; RV[1] = cdr[1]; RV[2] = cdr[2]; }
      else { double x = cdr[0]; RV[0] = x; RV[1] = sqrt(1. - x * x); }
    } else {
      CA = (int)((int)DIM * cdr[0]);
    }
    double sgn = cdr[3];
    double Cc = R.cosA, Sn_;
    if (R.true_sign ? (sgn >= 0.5) : ((long)sgn > 0.5)) Sn_ = R.sinA;
    else Sn_ = -1. * R.sinA;
    double *M = cd + 3;
    if (DIM == 3) {
      const double *L = RV;
      if (R.rtech == MPCD_ARBAXIS) {
        M[0] = L[0] * L[0] + (1. - L[0] * L[0]) * Cc;
        M[1] = L[0] * L[1] * (1. - Cc) - L[2] * Sn_;
        M[2] = L[0] * L[2] * (1. - Cc) + L[1] * Sn_;
        M[3] = L[0] * L[1] * (1. - Cc) + L[2] * Sn_;
        M[4] = L[1] * L[1] + (1. - L[1] * L[1]) * Cc;
        M[5] = L[1] * L[2] * (1. - Cc) - L[0] * Sn_;
        M[6] = L[0] * L[2] * (1. - Cc) - L[1] * Sn_;
        M[7] = L[1] * L[2] * (1. - Cc) + L[0] * Sn_;
        M[8] = L[2] * L[2] + (1. - L[2] * L[2]) * Cc;
      } else {
        orth_matrix(CA, Cc, Sn_, M);
      }
    } else {
      M[0] = Cc; M[1] = -1. * Sn_; M[2] = 0.;
      M[3] = Sn_; M[4] = Cc; M[5] = 0.;
      M[6] = 0.; M[7] = 0.; M[8] = 0.;
    }
  } else {
    // mass-weighted sum of the thermal draws, in list order
    double RS[3] = {0, 0, 0};
    for (int p = b; p < e; p++) {
      const double m = sps[ONESP ? 0 : ssp[p]].mass;
      for (int j = 0; j < DIM; j++) RS[j] += m * srv[j * stride + p];
    }
    for (int j = 0; j < DIM; j++) RS[j] /= mass;
    cd[3] = RS[0];
    cd[4] = RS[1];
    cd[5] = RS[2];
    if (OP == OP_RAT) {
      // inertia about the centre of mass (mpc.c:4209-4299) and angular momentum (mpc.c:1554-1618)
      double I[6] = {0, 0, 0, 0, 0, 0};
      for (int p = b; p < e; p++) {
        const double m = sps[ONESP ? 0 : ssp[p]].mass;
        double r[3] = {0, 0, 0};
        for (int d = 0; d < DIM; d++) r[d] = (double)sq[d * stride + p] - cm[d];
        I[0] += m * (r[1] * r[1] + r[2] * r[2]);
        I[1] += m * (r[0] * r[0] + r[2] * r[2]);
        I[2] += m * (r[0] * r[0] + r[1] * r[1]);
        I[3] += m * r[0] * r[1];
        I[4] += m * r[0] * r[2];
        I[5] += m * r[1] * r[2];
      }
      I[3] *= -1.; I[4] *= -1.; I[5] *= -1.;
      double II[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      if (DIM == 3) {
        if (np_ >= 3) {
          double Im[3][3] = {{I[0], I[3], I[4]}, {I[3], I[1], I[5]}, {I[4], I[5], I[2]}};
          inv3(II, Im);
        }
      } else {
        II[2][2] = 1. / I[2];
      }
      double L[3] = {0, 0, 0};
      for (int p = b; p < e; p++) {
        const double m = sps[ONESP ? 0 : ssp[p]].mass;
        double rq[3] = {0, 0, 0}, dv[3] = {0, 0, 0}, am[3];
        for (int j = 0; j < DIM; j++) {
          rq[j] = (double)sq[j * stride + p] - cm[j];
          dv[j] = m * ((double)sv[j * stride + p] - srv[j * stride + p]);
        }
        cross3(rq, dv, am);
        for (int j = 0; j < 3; j++) L[j] += am[j];
      }
      for (int i = 0; i < 3; i++) {
        double w = 0.;
        for (int j = 0; j < 3; j++) w += II[i][j] * L[j];
        cd[6 + i] = w;
        cd[9 + i] = cm[i];
      }
    }
  }
}

// the collision of one particle of a cell with more than one particle,
// given the cell data of cell_terms and, for MPC-AT(+a), its thermal draws rv
template <typename Real, int OP, int DIM>
__device__ __forceinline__ void collide_one(const mpcd_gpu_species &sp, const double *cd, int np_, const double rv[3],
                                            const double q[3], double v[3]) {
  const double vcm[3] = {cd[0], cd[1], cd[2]};
  if (OP == OP_SRD) {
    const double *M = cd + 3;
    double Vr[3] = {v[0] - vcm[0], v[1] - vcm[1], v[2] - vcm[2]};
    double TM[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
      for (int j = 0; j < DIM; j++) TM[i] += M[i * 3 + j] * Vr[j];
#pragma unroll
    for (int d = 0; d < DIM; d++) v[d] = vcm[d] + TM[d];
  } else if (OP == OP_AT) {
#pragma unroll
    for (int j = 0; j < DIM; j++) {
      double dv = (sp.damp) * (vcm[j]) / ((double)np_);
      v[j] = vcm[j] + rv[j] - cd[3 + j] - dv;
    }
  } else {
    double rq[3] = {0, 0, 0}, at[3];
#pragma unroll
    for (int j = 0; j < DIM; j++) rq[j] = q[j] - cd[9 + j];
    const double Wv[3] = {cd[6], cd[7], cd[8]};
    cross3(Wv, rq, at);
#pragma unroll
    for (int j = 0; j < DIM; j++) v[j] = vcm[j] + rv[j] - cd[3 + j] - (sp.damp) * (vcm[j]) + at[j];
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) v[d] = (double)(Real)v[d];  // the state keeps the collided velocity
}

template <typename Real, int OP, bool ONESP, int DIM>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(3))) void k_fused(
    Ctx C, State<Real> So, State<Real> Sn, const TileDev *__restrict__ tiles, const RepDev *__restrict__ reps,
    const mpcd_gpu_species *__restrict__ sps, const mpcd_gpu_wall *__restrict__ walls,
    const WallHot *__restrict__ wh, const double *__restrict__ boxes, Draws Dr,
    const double *__restrict__ shift_cur, const double *__restrict__ wq_b, const double *__restrict__ shift_next,
    int32_t *hist_next, int do_next, int do_stream) {
  Dr.resolve();
  // the uniform tables come in through restrict-qualified arguments so that
  // their loads can be scalar
  constexpr int kTileCap = tile_cap(sizeof(Real) == 8, OP);
  const TileDev T = tiles[blockIdx.x];
  const RepDev &R = reps[T.rep];
  const int tid = threadIdx.x;
  const double *shc = &shift_cur[T.rep * 3];
  const double *box = &boxes[T.rep * 6];

  __shared__ int32_t s_start[kTileCells + 1];
  __shared__ int32_t s_ce;
  __shared__ uint64_t s_key[kTileCap];
  __shared__ int32_t s_src[kTileCap];
  __shared__ uint8_t s_cell[kTileCap];
  __shared__ uint8_t s_sp[ONESP ? 1 : kTileCap];
  __shared__ int32_t s_id[kTileCap];
  __shared__ Real s_q[3][kTileCap];
  __shared__ Real s_v[3][kTileCap];
  __shared__ double s_rv[OP == OP_SRD ? 1 : 3][OP == OP_SRD ? 1 : kTileCap];
  __shared__ double s_cd[kTileCells][12];  // per cell: vcm[3], then SRD matrix[9] or RS[3], W[3], CM[3]
  __shared__ double s_ke[kTileThreads];

  double ke = 0.0;
  int cs = T.c0;
  while (cs < T.c1) {
    // cell-aligned chunk that fits the LDS capacity: the cell offsets of the
    // rest of the tile are loaded in parallel and the chunk end is the last
    // offset within capacity of the first (offsets are non-decreasing)
    const int nrest = T.c1 - cs;
    if (tid == 0) s_ce = cs + 1;
    for (int k = tid; k <= nrest; k += kTileThreads) s_start[k] = C.start[cs + k];
    __syncthreads();
    for (int k = 2 + tid; k <= nrest; k += kTileThreads)
      if (s_start[k] - s_start[0] <= kTileCap) atomicMax(&s_ce, cs + k);
    __syncthreads();
    const int ce = s_ce, ncl = ce - cs;
    const int base = s_start[0];
    int n = s_start[ncl] - base;
    if (n > kTileCap) {
      if (tid == 0) atomicOr(&C.flags[0], kFlagCellCap);  // one cell above capacity
      n = kTileCap;                            // never overrun LDS
    }
    __syncthreads();
    for (int k = tid; k <= ncl; k += kTileThreads) {
      int v = s_start[k] - base;
      s_start[k] = v > n ? n : v;
    }
    for (int e = tid; e < n; e += kTileThreads) s_key[e] = C.skey[base + e];
    __syncthreads();
    // rank inside the cell: the key order is the reference's list order
    for (int e = tid; e < n; e += kTileThreads) {
      int lo = 0, hi = ncl;  // last k with s_start[k] <= e
      while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (s_start[mid] <= e) lo = mid;
        else hi = mid;
      }
      const uint64_t ke_ = s_key[e];
      int r = 0;
      for (int f = s_start[lo]; f < s_start[lo + 1]; f++) r += (s_key[f] < ke_) ? 1 : 0;
      const int pos = s_start[lo] + r;
      s_src[pos] = (int)(ke_ & 0x7FFFFFFFull);
      s_cell[pos] = (uint8_t)lo;
    }
    __syncthreads();
    // gather with grid shift and body force; thermal draws of MPC-AT(+a)
    for (int p = tid; p < n; p += kTileThreads) {
      const int src = s_src[p];
      double q[3];
      load_q(So, src, q);
      double v[3] = {(double)So.vx[src], (double)So.vy[src], (double)So.vz[src]};
      const int pid = So.id[src];
      const int spid = ONESP ? 0 : So.sp[src];
      if (R.galinv) shift_fwd_d<DIM>(R, shc, q);
      if (cell_of_d<DIM>(R, q, 1) < 0) {
        // escape repair (mpc.c:970-982); the particle stays in its list
        atomicAdd(&C.events[0], 1ull);
        double u0[3] = {0.0, 0.0, 0.0};
        if (!escape_repair(R, wh, 0, Dr, T.rep, (uint32_t)pid, sps[spid].mass, false, q, v, u0))
          atomicOr(&C.flags[0], kFlagEscape);
      }
      s_q[0][p] = (Real)q[0];
      s_q[1][p] = (Real)q[1];
      s_q[2][p] = (Real)q[2];
      if (R.gravflag) {
        // acc_all (mpc.c:697), in the state precision
        Real vv[3] = {(Real)v[0], (Real)v[1], (Real)v[2]};
        for (int d = 0; d < DIM; d++) vv[d] = (Real)((double)vv[d] + R.dt * R.grav[d]);
        s_v[0][p] = vv[0];
        s_v[1][p] = vv[1];
        s_v[2][p] = vv[2];
      } else {
        s_v[0][p] = (Real)v[0];
        s_v[1][p] = (Real)v[1];
        s_v[2][p] = (Real)v[2];
      }
      Sn.id[base + p] = pid;
      s_id[p] = pid;
      if (!ONESP) {
        Sn.sp[base + p] = spid;
        s_sp[p] = (uint8_t)spid;
      }
      if (OP != OP_SRD) {
        // thermal velocities of the collision (mpc.c:1382, 1506), keyed by slot
        const double sig = sqrt(R.kbt / sps[spid].mass);
        double nrm[3];
        Dr.part_n_all(T.rep, base + p, DIM, nrm);
        for (int j = 0; j < DIM; j++) s_rv[j][p] = sig * nrm[j];
      }
    }
    __syncthreads();
    // one lane per cell: sums in list order and the collision's cell terms
    for (int k = tid; k < ncl; k += kTileThreads)
      cell_terms<Real, OP, ONESP, DIM>(R, sps, wh, Dr, T.rep, cs + k, s_start[k], s_start[k + 1], &s_v[0][0],
                                       &s_q[0][0], OP == OP_SRD ? nullptr : &s_rv[0][0], ONESP ? nullptr : s_sp,
                                       kTileCap, s_cd[k]);
    __syncthreads();
    // one lane per particle: collision, kinetic energy, shift back, stream,
    // walls, next bin
    for (int p = tid; p < n; p += kTileThreads) {
      const int k = s_cell[p], c = cs + k;
      const int np_ = s_start[k + 1] - s_start[k];
      const double *cd = s_cd[k];
      const int spid = ONESP ? 0 : s_sp[p];
      const mpcd_gpu_species &sp = sps[spid];
      double v[3] = {(double)s_v[0][p], (double)s_v[1][p], (double)s_v[2][p]};
      double q[3] = {(double)s_q[0][p], (double)s_q[1][p], (double)s_q[2][p]};
      if (np_ > 1) {
        double rv[3] = {0., 0., 0.};
        if (OP != OP_SRD)
          for (int j = 0; j < DIM; j++) rv[j] = s_rv[j][p];
        collide_one<Real, OP, DIM>(sp, cd, np_, rv, q, v);
      }
      {
        double v0 = v[0] - R.avv[0], v1 = v[1] - R.avv[1], v2 = v[2] - R.avv[2];
        double tmp = v0 * v0;
        tmp += v1 * v1;
        tmp += v2 * v2;
        ke += sp.mass * tmp;
      }
      if (R.galinv) {
        shift_back_d<DIM>(R, shc, q);
#pragma unroll
        for (int d = 0; d < DIM; d++) q[d] = (double)(Real)q[d];
      }
      const int slot = base + p;
      if (do_stream) {
        double U0[3] = {0, 0, 0};
        const int cu = stream_walls<DIM>(C, R, sps, walls, wh, wq_b, box, Dr, T.rep, spid, s_id[p], c, q, v, U0,
                                         false);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          q[d] = (double)(Real)q[d];
          v[d] = (double)(Real)v[d];
        }
        if (do_next) next_bin<DIM>(C, R, &shift_next[T.rep * 3], hist_next, slot, c, cu, q);
        else C.cu[slot] = cu;
      }
      Sn.x[slot] = (Real)q[0];
      Sn.y[slot] = (Real)q[1];
      Sn.z[slot] = (Real)q[2];
      Sn.vx[slot] = (Real)v[0];
      Sn.vy[slot] = (Real)v[1];
      Sn.vz[slot] = (Real)v[2];
      Sn.cst[slot] = c;
    }
    __syncthreads();
    cs = ce;
  }
  s_ke[tid] = ke;
  __syncthreads();
  for (int off = kTileThreads / 2; off > 0; off >>= 1) {
    if (tid < off) s_ke[tid] += s_ke[tid + off];
    __syncthreads();
  }
  if (tid == 0) C.tpart[blockIdx.x] = s_ke[0];
}

// the SRD cells' rotation draws of one step in the keyed layout of injected
// draws (cell [4 per cell]: ARBAXIS 3D unit axis in [0..2], otherwise a
// uniform in [0]; the rotation-sign uniform in [3]), one lane per cell
__global__ void k_celldraws(Ctx C, const BlkDev *__restrict__ cblk, Draws D, double *out) {
  D.resolve();
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  double *o = &out[(size_t)c * 4];
  if (R.rtech == MPCD_ARBAXIS && R.dim == 3) {
    double v[3];
    D.rng.unitvec(B.rep, P_CELL, c, 0, v);
    o[0] = v[0];
    o[1] = v[1];
    o[2] = v[2];
  } else {
    o[0] = D.rng.uniform(B.rep, P_CELL, c, 0);
  }
  o[3] = D.rng.uniform(B.rep, P_CELL, c, 3);
}

// stream after a global thermostat: scale, stream, walls, next bin
template <typename Real, int DIM>
__global__ __launch_bounds__(kBlock) void k_stream_next(Ctx C, State<Real> S, const BlkDev *__restrict__ blk, const RepDev *__restrict__ reps,
                              const mpcd_gpu_species *__restrict__ sps, const mpcd_gpu_wall *__restrict__ walls,
                              const WallHot *__restrict__ wh, const double *__restrict__ boxes, Draws Dr,
                              const double *__restrict__ wq_b, const double *__restrict__ shift_next,
                              int32_t *hist_next, int do_next) {
  Dr.resolve();
  const BlkDev B = blk[blockIdx.x];
  int t = B.beg + threadIdx.x;
  const RepDev &R = reps[B.rep];
  // the block's term of the temperature after the scaling (mpc.c:4760)
  double ke = 0.0;
  const double g = C.gscale[B.rep];
  double v[3] = {0.0, 0.0, 0.0};
  if (t < B.end) {
    v[0] = (double)S.vx[t];
    v[1] = (double)S.vy[t];
    v[2] = (double)S.vz[t];
    if (g != 1.0) {
      for (int d = 0; d < DIM; d++) v[d] = (double)(Real)(v[d] * g);
    }
    ke = temp_term(sps[S.sp[t]].mass, v, R.avv);
  }
  block_sum_to(ke, &C.tpart[blockIdx.x]);
  if (t >= B.end) return;
  const bool lc = R.lc != MPCD_ISOF;
  double q[3], U[3] = {0, 0, 0};
  load_q(S, t, q);
  if (lc) { U[0] = (double)S.ux[t]; U[1] = (double)S.uy[t]; U[2] = (double)S.uz[t]; }
  const int c = S.cst[t];
  int cu = stream_walls<DIM>(C, R, sps, walls, wh, wq_b, &boxes[B.rep * 6], Dr, B.rep, S.sp[t], S.id[t], c, q, v, U,
                             lc);
  for (int d = 0; d < 3; d++) {
    q[d] = (double)(Real)q[d];
    v[d] = (double)(Real)v[d];
  }
  if (do_next) next_bin<DIM>(C, R, &shift_next[B.rep * 3], hist_next, t, c, cu, q);
  else C.cu[t] = cu;
  S.x[t] = (Real)q[0];
  S.y[t] = (Real)q[1];
  S.z[t] = (Real)q[2];
  S.vx[t] = (Real)v[0];
  S.vy[t] = (Real)v[1];
  S.vz[t] = (Real)v[2];
  if (lc) {
    S.ux[t] = (Real)U[0];
    S.uy[t] = (Real)U[1];
    S.uz[t] = (Real)U[2];
  }
}

// TEMP (therm.c:209-235) from per-block kinetic-energy partials of one
// replica (blocks [beg, end)), with the thermostat factor (therm.c:52-143)
// and the active-matter guard (mpc.c:4766-4778); one workgroup per replica
constexpr int kTempThreads = 1024;

__global__ __launch_bounds__(kTempThreads) void k_temp_reduce(Ctx C, const int2 *ranges) {
  const int r = blockIdx.x;
  const RepDev &R = C.rep[r];
  const int2 rg = ranges[r];
  __shared__ double sh[kTempThreads];
  double acc = 0.0;
  for (int b = rg.x + threadIdx.x; b < rg.y; b += kTempThreads) acc += C.tpart[b];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = kTempThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double K = sh[0] / ((double)(R.npart * R.dim));
  double g = 1.0;
  if (R.tstech == MPCD_VSC) g = sqrt(R.kbt / K);
  else if (R.tstech == MPCD_BEREND) g = sqrt(1. + (R.kbt / K - 1.) * R.dt / R.tau);
  K = K * g * g;
  if (R.rtech >= 7 && R.rtech <= 17 && K > 50.0 * R.kbt) {
    double g2 = sqrt(R.kbt / K);
    g *= g2;
    K = K * g2 * g2;
  }
  C.gscale[r] = g;
  C.kbtnow[r] = K;
}

}  // namespace mpcd
