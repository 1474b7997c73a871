// mpcd_kernels.h -- HIP kernels of the MPCD solvent step (gfx950).
//
// Layout in HBM.  Particles live in structure-of-arrays form, sorted by
// collision cell ("slots"), concatenated over replicas: x,y,z,vx,vy,vz
// (+ux,uy,uz for nematics) in the state precision (double or float), plus
// the particle id, species id and the slot's cell.  Two such buffers are
// ping-ponged by the binning gather.  Cells are numbered x-major over
// (L0+1)(L1+1)(L2+1) (z extent 1 in 2D) and concatenated over replicas; all
// per-cell sums are kept in double.
//
// Binning.  The reference keeps one linked list per cell and relinks movers
// to the tail (mpc.c:948-992, addlink 1069-1087), which is a stable
// partition: in every cell the stayers keep their order and are followed by
// the arrivals in previous global order.  A step bins twice (shifted for the
// collision, unshifted after streaming); the two stable partitions compose
// into ONE counting sort by the new shifted cell with the in-cell key
//   (moved since the unshifted bin, unshifted cell, moved in the unshifted
//    bin, slot)
// so a step needs one histogram (global atomics), one scan, one scatter of
// 64-bit keys and one in-cell sort, and the unshifted bin is never
// materialised unless a caller asks for it.
#pragma once

#include "mpcd_math.h"
#include "mpcd_dual.h"
#include "mpcd_rev.h"
#include "mpcd_params.h"
#include "mpcd_walls.h"

namespace mpcd {

using mpcd_ad::Dual;
using mpcd_ad::Rv;

constexpr int kBlock = 256;

struct RepDev {
  int dim, L[3], nx, ny, nz, pbc[3];
  int cell_off, ncell, part_off, npart;
  int wall_off, nwall;
  int rtech, lc, tstech, galinv, true_sign, gravflag, anyphantom;
  int anycurved;  // a wall that is not an axis-aligned plane (mpcd_walls.h)
  int xwrap;      // exact-sum FP32 form: the walls are the box's periodic faces (mpcd_exact.h)
  double dt, kbt, cosA, sinA, tau, grav[3], avv[3], gndnst;
};

// a block of work inside one replica
struct BlkDev {
  int rep, beg, end, pad;
};

// a tile of consecutive cells [c0, c1) of one replica
struct TileDev {
  int rep, c0, c1, pad;
};

// tiles of the unfused path's cell kernels: slots staged in LDS per tile
constexpr int kLcThreads = 256, kLcPPT = 2, kLcCap = kLcThreads * kLcPPT, kLcTileCells = 64;

template <typename Real>
struct State {
  Real *x, *y, *z, *vx, *vy, *vz, *ux, *uy, *uz;
  int32_t *id;
  int32_t *sp;
  int32_t *cst;  // cell of the slot in this buffer's order
};

struct CellDev {
  int32_t *pop;
  double *mass, *vcm, *cm, *I, *S, *dir, *E;
  double *avm;  // nematic cells: mean mean-field potential (lc.c:495), cells of two or more
};

struct WallHot;

struct Ctx {
  const RepDev *rep;
  const mpcd_gpu_species *sp;
  const mpcd_gpu_wall *wall;
  double *wq_s, *wq_b;       // wall positions of this step: shifted, shifted back
  const WallHot *wh;         // the walls' hot fields of this step (escape repair)
  double *shift;             // [3 * nrep]
  int32_t *hist, *start, *cnew, *loc, *cu;
  uint64_t *key, *skey;
  CellDev cell;
  double *T;                 // LC torque per slot [3 * ntot]
  double *tpart;             // temperature partial sums per block
  double *kbtnow, *gscale, *sg; // per replica
  unsigned long long *events;   // [4]
  int *flags;                   // [0] error bits (kFlag*), [1] first overflowing step of a bucketed run
  int ntot, nctot, nrep;
  // nematic parity diagnostics (null unless recording): the discrete
  // decisions of the last step, per particle (particle-index order,
  // replicas concatenated) and per cell (lc_decision_word, cell_decision_word)
  uint32_t *pdec, *cdec;
};

// the decision words of the nematic path (also written by the oracle):
// particle: bit 0 collided (its cell holds more than one particle), bits 1-2
// sampler (0 2D Metropolis, 1 3D Gaussian, 2 3D cube root, 3 3D
// Metropolis), bit 3 the 2D angle redrawn at +-pi/2, bit 4 the sign pm = -1,
// bits 5-12 annealNum, bits 13-20 accepted moves, bit 21 the dipole force
// on the positive side;
// cell: bit 5 S and director computed (more than one particle), bits 0-3
// the eigen-solver's branches (eig2_top / eig3_top), bit 4 S clamped to 1
constexpr uint32_t kDecSide = 1u << 21;
MPCD_HD uint32_t lc_decision_word(int sampler, bool redraw, bool neg, int anneal, int accepted) {
  const uint32_t a = (uint32_t)min(max(anneal, 0), 255), k = (uint32_t)min(max(accepted, 0), 255);
  return 1u | ((uint32_t)sampler << 1) | (redraw ? 8u : 0u) | (neg ? 16u : 0u) | (a << 5) | (k << 13);
}
MPCD_HD uint32_t cell_decision_word(int br, bool clamped) { return 32u | (uint32_t)(br & 15) | (clamped ? 16u : 0u); }

// Tangent of the FP64 state along one direction in parameter space (forward
// mode, mpcd_gpu_tangent): per slot of the current buffer the derivatives of
// position, velocity and orientation; per cell those of the cell properties;
// and the direction itself (derivatives of the body force and of each
// species' activity).  The reverse mode (TA<Rv>) uses the same arrays for the
// adjoints of those values, and dth for the adjoints of the parameters.
struct TanDev {
  double *q[3], *v[3], *u[3];
  double *vcm, *cm, *I, *S, *dir, *E;  // per cell: 3, 3, 6, 1, 3, 9
  double *T;                           // LC torque per slot [3 * ntot]
  double dgrav[3];
  double dact[MPCD_MAXSPEC];
  double *dth;  // reverse mode: per replica kNTheta adjoints (MPCD_NTHETA layout), else null
};

// adjoints of the parameters per replica: GRAV[0..2], the rotation angle, ACT of each species
constexpr int kNTheta = 4 + MPCD_MAXSPEC;
constexpr int kThGrav = 0, kThAngle = 3, kThAct = 4;
__device__ __forceinline__ double *th_slot(const TanDev &T, int rep, int k) {
  return T.dth ? T.dth + (size_t)rep * kNTheta + k : nullptr;
}

// loads and stores of state values and their tangents: TA<double> touches
// the values only, TA<Dual> the values and the tangent arrays
template <typename F>
struct TA;
template <>
struct TA<double> {
  template <typename R>
  static MPCD_HD double ld(const R *a, const double *, size_t i) { return (double)a[i]; }
  template <typename R>
  static MPCD_HD void st(R *a, double *, size_t i, double x) { a[i] = (R)x; }
  static MPCD_HD double par(double x, double, double *) { return x; }
  static MPCD_HD bool moves(double) { return false; }
};
// FP32 arithmetic on the FP32 state (the orientation step of k_lc): values only
template <>
struct TA<float> {
  template <typename R>
  static MPCD_HD float ld(const R *a, const double *, size_t i) { return (float)a[i]; }
  template <typename R>
  static MPCD_HD void st(R *a, double *, size_t i, float x) { a[i] = (R)x; }
  static MPCD_HD float par(double x, double, double *) { return (float)x; }
  static MPCD_HD bool moves(float) { return false; }
};
template <>
struct TA<Dual> {
  template <typename R>
  static MPCD_HD Dual ld(const R *a, const double *d, size_t i) { return Dual((double)a[i], d[i]); }
  template <typename R>
  static MPCD_HD void st(R *a, double *d, size_t i, const Dual &x) { a[i] = (R)x.v; d[i] = x.d; }
  static MPCD_HD Dual par(double x, double dx, double *) { return Dual(x, dx); }
  static MPCD_HD bool moves(const Dual &x) { return x.d != 0.0; }
};
// reverse mode: loads are leaves (or the entry last stored to the same
// element), stores are outputs, parameters are leaves whose adjoints go to dth
template <>
struct TA<Rv> {
  template <typename R>
  static __device__ __forceinline__ Rv ld(const R *a, const double *d, size_t i) {
    const double v = (double)a[i];
    if (!d) return Rv(v);
    int var;
    if (mpcd_ad::rv_stored(d + i, var)) return Rv(v, var);
    return mpcd_ad::rv_leaf(v, const_cast<double *>(d) + i);
  }
  template <typename R>
  static __device__ __forceinline__ void st(R *a, double *d, size_t i, const Rv &x) {
    a[i] = (R)x.v;
    if (d) mpcd_ad::rv_store(d + i, x);
  }
  static __device__ __forceinline__ Rv par(double x, double, double *adj) { return mpcd_ad::rv_leaf(x, adj); }
  static __device__ __forceinline__ bool moves(const Rv &x) { return x.i >= 0; }
};

// bits of Ctx::flags[0]
constexpr int kFlagCellCap = 1;  // a cell above the fused kernel's LDS capacity
constexpr int kFlagOutside = 2;  // initial positions outside the box
constexpr int kFlagPoison = 4;   // a bucket overflowed: the bucketed steps from then on are no-ops
constexpr int kFlagBadCell = 8;  // a particle outside its bucket's tile (internal error)
constexpr int kFlagEscape = 16;  // escape repair ran out of position tries
constexpr int kFlagPushes = 32;  // a particle needed more re-thermalisations than its injected entry holds
constexpr int kFlagStuck = 64;   // a particle was re-thermalised 1000 times in one step (walls leave no room)

// hot fields of an axis-aligned planar wall for the per-particle wall test:
// the normal's one nonzero component A (0 when its axis is not a box axis),
// the offset R and the wall's shifted-back / shifted position along the axis
struct WallHot {
  double A, R, wqb, wqs;
  int axis, inv, mask, phantom;
};

// calcW_PLANE (bc.c:306) for an axis-aligned wall: the other axes add exact
// zeros to the reference's sum, so the sign tests agree with wall_W()
__device__ __forceinline__ double wall_W_hot(const WallHot &H, const double *q, double wq) {
  // select values, not addresses: a dynamic index would move q to scratch
  const double q0 = q[0], q1 = q[1], q2 = q[2];
  double qa = q2;
  if (H.axis == 0) qa = q0;
  if (H.axis == 1) qa = q1;
  double Wv = H.A * (qa - wq);
  Wv -= H.R;
  if (H.inv) Wv *= -1.;
  return Wv;
}

template <typename Real>
MPCD_HD void load_q(const State<Real> &s, int i, double q[3]) {
  q[0] = (double)s.x[i];
  q[1] = (double)s.y[i];
  q[2] = (double)s.z[i];
}

template <typename F, typename Real>
__device__ __forceinline__ void load_qt(const State<Real> &s, const TanDev &T, int i, F q[3]) {
  q[0] = TA<F>::ld(s.x, T.q[0], i);
  q[1] = TA<F>::ld(s.y, T.q[1], i);
  q[2] = TA<F>::ld(s.z, T.q[2], i);
}

// cell index of a position; shifted = 1 applies the periodic wrap of
// bin(shifted=1) (mpc.c:964-968).  Returns -1 for an escaped particle.
template <typename F>
__device__ __forceinline__ int cell_of(const RepDev &R, const F q[3], int shifted) {
  int a = (int)val(q[0]), b = (int)val(q[1]), c = (int)val(q[2]);
  int Lz = (R.dim == 3) ? R.L[2] : 1;
  if (shifted) {
    if (R.pbc[0] && a == R.L[0]) a = 0;
    if (R.pbc[1] && b == R.L[1]) b = 0;
    if (R.pbc[2] && c == Lz) c = 0;
  }
  if (a > R.L[0] || b > R.L[1] || c > Lz || a < 0 || b < 0 || c < 0) return -1;
  if (R.dim < 3) c = 0;
  return R.cell_off + (a * R.ny + b) * R.nz + c;
}

// gridShift_all(+SHIFT), mpc.c:728-736
__device__ __forceinline__ void shift_fwd(const RepDev &R, const double *s, double q[3]) {
  for (int j = 0; j < R.dim; j++) {
    q[j] += s[j];
    if (q[j] > R.L[j] && q[j] <= R.L[j] + s[j]) q[j] -= R.L[j];
  }
}

// gridShift_all(-SHIFT), mpc.c:728-736 with shiftBack = 1
template <typename F>
__device__ __forceinline__ void shift_back(const RepDev &R, const double *s, F q[3]) {
  for (int j = 0; j < R.dim; j++) {
    double sg = -1.0 * s[j];
    q[j] += sg;
    if (q[j] < 0.0 && q[j] >= sg) q[j] += R.L[j];
  }
}

// push() (init.c:1261-1293): a fresh velocity from the species' distribution
// (RANDVEL 0, HOMO 2, HOMOAXIS 3, GAUSS 4; READ is refused at creation),
// from variates k0 .. k0 + dim - 1 of the particle's re-thermalisation draws
__device__ __forceinline__ void push_velocity(const mpcd_gpu_species &sp, double kbt, int dim, const Draws &Dr,
                                              uint32_t rep, size_t gp, uint32_t pid, int k0, double V[3]) {
  if (sp.vdist == 0) {
    for (int d = 0; d < dim; d++) V[d] = sqrt(kbt / sp.mass) * (2. * Dr.reth_u(rep, gp, pid, k0 + d) - 1.);
  } else if (sp.vdist == 3) {
    for (int d = 0; d < dim; d++) {
      V[d] = sqrt((double)dim * kbt / sp.mass);
      V[d] *= (Dr.reth_u(rep, gp, pid, k0 + d) <= 0.5) ? -1.0 : 1.0;  // genrand_pmOne, rand.c:378
    }
  } else {
    for (int d = 0; d < dim; d++) V[d] = sqrt(kbt / sp.mass) * Dr.reth_n(rep, gp, pid, k0 + d);
    if (sp.vdist == 2) {
      normn(V, dim);
      const double nrm = sqrt((double)dim * kbt / sp.mass);
      for (int d = 0; d < dim; d++) V[d] *= nrm;
    }
  }
}

// bin()'s escape repair (mpc.c:970-982) of particle pid of replica R in
// binning b (0: the shifted bin, walls at their shifted positions wqs;
// 1: the unshifted bin after streaming, walls at wqb):
//  - replacePos_WithCheck (init.c:1462-1483): a uniform position in the box
//    (replace, init.c:1240), redrawn while any wall of the replica has
//    calcW <= TOL there;
//  - a Maxwell velocity, genrand_gaussMB whatever the species' VDIST;
//  - for nematics, an orientation genrand_coneNP(U, pi, DIM) (rand.c:397-425).
// The particle stays in the list it was in; the caller keeps its cell.
// Returns false when the tries ran out (the injected entry's, or
// kEscMaxTries: walls that leave no room in the box).  The bucketed step
// kernel does not call it (its preparation kernel re-places listed escapes).
__device__ __forceinline__ bool escape_repair(const RepDev &R, const WallHot *wh, int b, const Draws &Dr, uint32_t rep,
                                              uint32_t pid, double mass, bool lc, double q[3], double v[3],
                                              double u[3], const mpcd_gpu_wall *walls = nullptr,
                                              const double *wq3 = nullptr) {
  const size_t gp = (size_t)R.part_off + pid;
  const int dim = R.dim;
  const int tries = Dr.esc_keyed() ? (kEscStride - kEscPos0) / dim : kEscMaxTries;
  bool ok = false;
  for (int t = 0; t < tries && !ok; t++) {
    for (int j = 0; j < dim; j++) q[j] = (double)R.L[j] * Dr.esc_u(rep, b, gp, pid, kEscPos0 + dim * t + j);
    ok = true;
    for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
      if (R.anycurved) {
        // calcW of every wall at its current position (the general forms
        // equal calcW_PLANE on a plane)
        if (calc_w(walls[w], &wq3[w * 3], q, dim) <= kTOL) ok = false;
      } else {
        const WallHot &H = wh[w];
        if (wall_W_hot(H, q, b ? H.wqb : H.wqs) <= kTOL) ok = false;
      }
    }
  }
  for (int j = 0; j < dim; j++) v[j] = sqrt(R.kbt / mass) * Dr.esc_n(rep, b, gp, pid, j);
  if (lc) {
    if (dim == 3) {
      const double ct = cos(kPI);
      const double z = Dr.esc_u(rep, b, gp, pid, 3) * (1.0 - ct) + ct;
      const double phi = Dr.esc_u(rep, b, gp, pid, 4) * 2.0 * kPI;
      u[0] = z;
      u[1] = sqrt(1.0 - z * z) * cos(phi);
      u[2] = sqrt(1.0 - z * z) * sin(phi);
    } else {
      const double phi = kPI * (1. - 2. * Dr.esc_u(rep, b, gp, pid, 3));
      u[0] = cos(phi);
      u[1] = sin(phi);
    }
  }
  return ok;
}

// ---------------------------------------------------------------------------
// 1. per-replica grid shift draw and wall positions (ranshift, rand.c:678)
// ---------------------------------------------------------------------------
// the grid shift of replica r and the wall positions it implies: reads the
// previous step's shifted-back wall positions and writes this step's shift
// and wall positions, so a repeated launch is idempotent
__device__ __forceinline__ void shift_replica(const Ctx &C, const Draws &D, int r, int do_shift,
                                              const double *wq_prev_b, double *shift_out, double *wqs_out,
                                              double *wqb_out, WallHot *wh_out, double *box_out) {
  const RepDev &R = C.rep[r];
  double s[3] = {0.0, 0.0, 0.0};
  if (do_shift && R.galinv) {
    for (int j = 0; j < R.dim; j++) s[j] = D.injected ? D.shift[r * 3 + j] : D.rng.uniform(r, P_SHIFT, 0, j);
  }
  for (int j = 0; j < 3; j++) shift_out[r * 3 + j] = s[j];
  double lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
  double Lmax = 0.0;
  for (int j = 0; j < R.dim; j++) Lmax = fmax(Lmax, (double)R.L[j]);
  for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
    // the positions stay in registers (no store and reload through memory)
    double qs3[3], qb3[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double q = wq_prev_b[w * 3 + j];
      if (j < R.dim && do_shift) {
        double qs = q + s[j];
        qs3[j] = qs;
        qb3[j] = qs + (-1.0 * s[j]);
      } else {
        qs3[j] = q;
        qb3[j] = q;
      }
      wqs_out[w * 3 + j] = qs3[j];
      wqb_out[w * 3 + j] = qb3[j];
    }
    const mpcd_gpu_wall &W = C.wall[w];
    int k = 0;
    for (int j = 0; j < 3; j++)
      if (W.A[j] != 0.0) k = j;
    WallHot H;
    H.A = (k < R.dim) ? W.A[k] : 0.0;
    H.R = W.R;
    H.wqb = (k == 0) ? qb3[0] : (k == 1) ? qb3[1] : qb3[2];
    H.wqs = (k == 0) ? qs3[0] : (k == 1) ? qs3[1] : qs3[2];
    H.axis = k;
    H.inv = W.INV;
    H.mask = W.inter_mask;
    H.phantom = W.PHANTOM;
    wh_out[w] = H;
    // wall-free box: a point strictly inside is on the positive side of
    // every wall by far more than the rounding of calcW_PLANE; with a
    // curved wall the box is empty and every particle takes the search
    if (R.anycurved) {
      lo[0] = INFINITY;
    } else if (H.A == 0.0) {
      const bool safe = H.inv ? (H.R > 0.0) : (H.R < 0.0);  // calcW is the constant -R (or R)
      if (!safe) lo[0] = INFINITY;
    } else {
      const double b = H.wqb + H.R / H.A;
      const double m = 1e-9 * (1.0 + fabs(H.wqb) + fabs(H.R / H.A) + Lmax);
      if ((H.A > 0.0) == (H.inv != 0)) hi[k] = fmin(hi[k], b - m);
      else lo[k] = fmax(lo[k], b + m);
    }
  }
  for (int j = 0; j < 3; j++) {
    box_out[r * 6 + 2 * j] = lo[j];
    box_out[r * 6 + 2 * j + 1] = hi[j];
  }
}

__global__ void k_shift(Ctx C, Draws D, int do_shift, const double *wq_prev_b, double *shift_out, double *wqs_out,
                        double *wqb_out, WallHot *wh_out, double *box_out) {
  D.resolve();
  // after a bucket overflow the shifts of the last complete state are kept
  if (C.flags[0] & kFlagPoison) return;
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= C.nrep) return;
  shift_replica(C, D, r, do_shift, wq_prev_b, shift_out, wqs_out, wqb_out, wh_out, box_out);
}

// ---------------------------------------------------------------------------
// 2. histogram + composite key.  mode 0: initial binning (binin, mpc.c:924);
//    1: shifted bin; 2: materialise the pending unshifted bin
// ---------------------------------------------------------------------------
// hist[cell]++ for every active lane, returning each lane's old count: one
// atomic per distinct cell of the wavefront (sorted slots put most of a
// wavefront's particles in a few cells).  The counts a cell's lanes get are
// distinct, which is all the scatter needs: the keys are sorted in the cell.
__device__ __forceinline__ int hist_add(int32_t *hist, int cell) {
  const int lane = (int)__lane_id();
  uint64_t pending = __ballot(1);
  int loc = 0;
  while (pending) {
    const int leader = __ffsll((unsigned long long)pending) - 1;
    const int lead_cell = __shfl(cell, leader, 64);
    const uint64_t peers = __ballot(cell == lead_cell) & pending;
    int base = 0;
    if (lane == leader) base = atomicAdd(&hist[lead_cell], __popcll(peers));
    base = __shfl(base, leader, 64);
    if ((peers >> lane) & 1ull) loc = base + __popcll(peers & ((1ull << lane) - 1ull));
    pending &= ~peers;
  }
  return loc;
}

template <typename Real>
__global__ void k_count(Ctx C, State<Real> S, const BlkDev *blk, int mode, int pending, int cu_src) {
  const BlkDev B = blk[blockIdx.x];
  int s = B.beg + threadIdx.x;
  if (s >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  double q[3];
  load_q(S, s, q);
  int cnew, cu;
  uint32_t a1 = 0, a2 = 0;
  if (mode == 0) {
    cnew = cell_of(R, q, 0);
    if (cnew < 0) { atomicOr(&C.flags[0], kFlagOutside); cnew = R.cell_off; }
    cu = cnew;
  } else {
    int cs = S.cst[s];
    // the unshifted cell left by streaming: in cu[] or in the key the fused
    // kernel wrote for this slot (bits 32..62)
    cu = pending ? (cu_src ? (int)((C.key[s] >> 32) & 0x7FFFFFFFull) : C.cu[s]) : cs;
    a1 = (cu != cs) ? 1u : 0u;
    if (mode == 1) {
      shift_fwd(R, &C.shift[B.rep * 3], q);
      cnew = cell_of(R, q, 1);
      if (cnew < 0) cnew = cu;  // escaped: stays in its list (mpc.c:970-982)
    } else {
      cnew = cu;
    }
    a2 = (cnew != cu) ? 1u : 0u;
  }
  int l = hist_add(C.hist, cnew);
  C.cnew[s] = cnew;
  C.loc[s] = l;
  C.key[s] = ((uint64_t)a2 << 63) | ((uint64_t)(uint32_t)cu << 32) | ((uint64_t)a1 << 31) | (uint64_t)(uint32_t)s;
}

// ---------------------------------------------------------------------------
// 3. exclusive scan of the histogram (multi-level, 1024 cells per block)
// ---------------------------------------------------------------------------
__global__ void k_scan_block(const int32_t *in, int32_t *out, int32_t *bsum, int n) {
  __shared__ int32_t sh[1024];
  int base = blockIdx.x * 1024;
  for (int k = threadIdx.x; k < 1024; k += blockDim.x) sh[k] = (base + k < n) ? in[base + k] : 0;
  __syncthreads();
  // Hillis-Steele inclusive scan over 1024 values with 256 threads
  for (int off = 1; off < 1024; off <<= 1) {
    int32_t t[4];
    for (int j = 0; j < 4; j++) {
      int k = threadIdx.x + j * 256;
      t[j] = (k >= off) ? sh[k - off] : 0;
    }
    __syncthreads();
    for (int j = 0; j < 4; j++) sh[threadIdx.x + j * 256] += t[j];
    __syncthreads();
  }
  for (int k = threadIdx.x; k < 1024; k += blockDim.x)
    if (base + k < n) out[base + k] = (k == 0) ? 0 : sh[k - 1];
  if (threadIdx.x == 0 && bsum) bsum[blockIdx.x] = sh[1023];
}

__global__ void k_scan_add(int32_t *out, const int32_t *boff, int n) {
  int i = blockIdx.x * 1024;
  int32_t add = boff[blockIdx.x];
  for (int k = threadIdx.x; k < 1024; k += blockDim.x)
    if (i + k < n) out[i + k] += add;
}

__global__ void k_setu32(uint32_t *a, uint32_t v) { *a = v; }
__global__ void k_addu32(uint32_t *a, uint32_t v) { *a += v; }

__global__ void k_fill(int32_t *a, int32_t v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// 4. scatter the keys to their cells
// ---------------------------------------------------------------------------
__global__ void k_scatter(Ctx C) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= C.ntot) return;
  C.skey[C.start[C.cnew[s]] + C.loc[s]] = C.key[s];
}

// ---------------------------------------------------------------------------
// 5. sort every cell's keys (insertion sort; cells hold tens of particles)
//    and record the cell of every new slot
// ---------------------------------------------------------------------------
// One tile of consecutive cells per workgroup: the tile's keys (one
// contiguous range) are staged in LDS and every key's position in its cell
// is its rank there (the keys are distinct).  A tile above the capacity
// falls back to an insertion sort, one lane per cell.
__global__ __launch_bounds__(kLcThreads) void k_sortcells_t(Ctx C, int32_t *cst_new, const TileDev *tiles) {
  constexpr int NT = kLcThreads;
  __shared__ uint64_t s_k[kLcCap];
  __shared__ int16_t s_c[kLcCap];
  const TileDev Tl = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int b0 = C.start[Tl.c0], n = C.start[Tl.c1] - b0, ncl = Tl.c1 - Tl.c0;
  if (n > kLcCap) {
    for (int k = tid; k < ncl; k += NT) {
      const int c = Tl.c0 + k, b = C.start[c], e = C.start[c + 1];
      uint64_t *kk = C.skey;
      for (int i = b + 1; i < e; i++) {
        uint64_t v = kk[i];
        int j = i - 1;
        while (j >= b && kk[j] > v) {
          kk[j + 1] = kk[j];
          j--;
        }
        kk[j + 1] = v;
      }
      for (int i = b; i < e; i++) cst_new[i] = c;
    }
    return;
  }
  for (int i = tid; i < n; i += NT) s_k[i] = C.skey[b0 + i];
  for (int k = tid; k < ncl; k += NT)
    for (int i = C.start[Tl.c0 + k] - b0; i < C.start[Tl.c0 + k + 1] - b0; i++) s_c[i] = (int16_t)k;
  __syncthreads();
  for (int i = tid; i < n; i += NT) {
    const int k = s_c[i], c = Tl.c0 + k, ib = C.start[c] - b0, ie = C.start[c + 1] - b0;
    const uint64_t v = s_k[i];
    int r = 0;
    for (int j = ib; j < ie; j++) r += (s_k[j] < v) ? 1 : 0;
    C.skey[b0 + ib + r] = v;
    cst_new[b0 + i] = c;
  }
}

// ---------------------------------------------------------------------------
// 6. gather the state into the new order, applying the grid shift
// ---------------------------------------------------------------------------
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_gather(Ctx C, State<Real> S, State<Real> D, const BlkDev *blk, int shifted, int lc, Draws Dr) {
  Dr.resolve();
  const BlkDev B = blk[blockIdx.x];
  int t = B.beg + threadIdx.x;
  if (t >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  int src = (int)(C.skey[t] & 0x7FFFFFFFull);
  double q[3];
  load_q(S, src, q);
  double v[3] = {(double)S.vx[src], (double)S.vy[src], (double)S.vz[src]};
  int id = S.id[src], sp = S.sp[src];
  bool esc = false;
  double u[3] = {0.0, 0.0, 0.0};
  if (shifted) {
    shift_fwd(R, &C.shift[B.rep * 3], q);
    if (cell_of(R, q, 1) < 0) {
      // escape repair (mpc.c:970-982); the particle stays in its list
      atomicAdd(&C.events[0], 1ull);
      esc = true;
      if (!escape_repair(R, C.wh, 0, Dr, B.rep, (uint32_t)id, C.sp[sp].mass, lc != 0, q, v, u, C.wall, C.wq_s))
        atomicOr(&C.flags[0], kFlagEscape);
    }
  }
  D.x[t] = (Real)q[0];
  D.y[t] = (Real)q[1];
  D.z[t] = (Real)q[2];
  D.vx[t] = (Real)v[0];
  D.vy[t] = (Real)v[1];
  D.vz[t] = (Real)v[2];
  if (lc) {
    D.ux[t] = esc ? (Real)u[0] : S.ux[src];
    D.uy[t] = esc ? (Real)u[1] : S.uy[src];
    D.uz[t] = esc ? (Real)u[2] : S.uz[src];
  }
  D.id[t] = id;
  D.sp[t] = sp;
}

// the tangent's share of the gather: the same permutation; a particle the
// escape repair re-places gets fresh draws, which carry no derivative
template <typename Real>
__global__ void k_gather_tan(Ctx C, State<Real> S, TanDev Ts, TanDev Td, const BlkDev *blk, int shifted, int lc) {
  const BlkDev B = blk[blockIdx.x];
  int t = B.beg + threadIdx.x;
  if (t >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  int src = (int)(C.skey[t] & 0x7FFFFFFFull);
  bool fresh = false;
  if (shifted) {
    double q[3];
    load_q(S, src, q);
    shift_fwd(R, &C.shift[B.rep * 3], q);
    fresh = cell_of(R, q, 1) < 0;
  }
  for (int k = 0; k < 3; k++) {
    Td.q[k][t] = fresh ? 0.0 : Ts.q[k][src];
    Td.v[k][t] = fresh ? 0.0 : Ts.v[k][src];
    if (lc) Td.u[k][t] = Ts.u[k][src];
  }
}

// ---------------------------------------------------------------------------
// 7. cell properties: acc_all (mpc.c:697) then localPROP (mpc.c:64-216) with
//    localMomInertiaTensor (mpc.c:4209-4299) and the order tensor / director
//    (lc.c:1125-1148, mtools.c:1443)
// ---------------------------------------------------------------------------
template <typename Real, typename F>
__device__ void cellprops_cell(const Ctx &C, const State<Real> &S, int rep, int c, int do_acc, const TanDev &T) {
  using A = TA<F>;
  const RepDev &R = C.rep[rep];
  const int dim = R.dim;
  const bool ccm = (R.rtech == MPCD_RAT) || (R.lc != MPCD_ISOF);
  int b = C.start[c], e = C.start[c + 1];
  F vcm[3], cm[3];
  double mass = 0.0;
  for (int d = 0; d < 3; d++) {
    vcm[d] = A::ld(C.cell.vcm, T.vcm, c * 3 + d);
    cm[d] = A::ld(C.cell.cm, T.cm, c * 3 + d);
  }
  for (int d = 0; d < dim; d++) vcm[d] = 0.0;
  if (ccm) for (int d = 0; d < dim; d++) cm[d] = 0.0;
  F grav[3];
  for (int d = 0; d < 3; d++) grav[d] = A::par(R.grav[d], T.dgrav[d], th_slot(T, rep, kThGrav + d));
  // a tangent along g applies acc_all even where the force itself is zero
  const bool acc = do_acc && (R.gravflag || A::moves(grav[0]) || A::moves(grav[1]) || A::moves(grav[2]));
  for (int t = b; t < e; t++) {
    double m = C.sp[S.sp[t]].mass;
    F v[3] = {A::ld(S.vx, T.v[0], t), A::ld(S.vy, T.v[1], t), A::ld(S.vz, T.v[2], t)};
    if (acc) {
      for (int d = 0; d < dim; d++) v[d] = v[d] + R.dt * grav[d];
      A::st(S.vx, T.v[0], t, v[0]);
      A::st(S.vy, T.v[1], t, v[1]);
      if (dim == 3) A::st(S.vz, T.v[2], t, v[2]);
      v[0] = A::ld(S.vx, T.v[0], t);
      v[1] = A::ld(S.vy, T.v[1], t);
      v[2] = A::ld(S.vz, T.v[2], t);
    }
    mass += m;
    for (int d = 0; d < dim; d++) vcm[d] += v[d] * m;
    if (ccm) {
      F q[3];
      load_qt(S, T, t, q);
      for (int d = 0; d < dim; d++) cm[d] += q[d] * m;
    }
  }
  if (mass > 0.0)
    for (int d = 0; d < dim; d++) {
      vcm[d] /= mass;
      if (ccm) cm[d] /= mass;
    }
  C.cell.pop[c] = e - b;
  C.cell.mass[c] = mass;
  for (int d = 0; d < 3; d++) {
    A::st(C.cell.vcm, T.vcm, c * 3 + d, vcm[d]);
    A::st(C.cell.cm, T.cm, c * 3 + d, cm[d]);
  }
  if (ccm) {
    F I[6] = {0., 0., 0., 0., 0., 0.};  // I00 I11 I22 I01 I02 I12
    for (int t = b; t < e; t++) {
      double m = C.sp[S.sp[t]].mass;
      F q[3], r[3] = {0., 0., 0.};
      load_qt(S, T, t, q);
      for (int d = 0; d < dim; d++) r[d] = q[d] - cm[d];
      I[0] += m * (r[1] * r[1] + r[2] * r[2]);
      I[1] += m * (r[0] * r[0] + r[2] * r[2]);
      I[2] += m * (r[0] * r[0] + r[1] * r[1]);
      I[3] += m * r[0] * r[1];
      I[4] += m * r[0] * r[2];
      I[5] += m * r[1] * r[2];
    }
    I[3] *= -1.; I[4] *= -1.; I[5] *= -1.;
    for (int k = 0; k < 6; k++) A::st(C.cell.I, T.I, c * 6 + k, I[k]);
  }
  if (R.lc != MPCD_ISOF && e - b > 1) {
    F Sm[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    double fD = (double)dim, cc = 1. / (fD - 1.), POPinv = 1. / ((double)(e - b));
    for (int t = b; t < e; t++) {
      F u[3] = {A::ld(S.ux, T.u[0], t), A::ld(S.uy, T.u[1], t), A::ld(S.uz, T.u[2], t)};
      for (int i = 0; i < dim; i++)
        for (int j = i; j < dim; j++) Sm[i][j] += u[i] * u[j];
    }
    for (int i = 0; i < dim; i++)
      for (int j = i; j < dim; j++) Sm[i][j] *= cc * fD * POPinv;
    for (int i = 0; i < dim; i++) Sm[i][i] -= cc;
    for (int i = 0; i < dim; i++)
      for (int j = i + 1; j < dim; j++) Sm[j][i] = Sm[i][j];
    F Sv, dir[3] = {0., 0., 0.};
    int br = 0;
    if (dim == 3) {
      F ev[3];
      eig3_top(Sm, ev, dir, &br);
      Sv = -1. * (ev[1] + ev[2]);
    } else {
      F m2[2][2] = {{Sm[0][0], Sm[0][1]}, {Sm[1][0], Sm[1][1]}};
      Sv = eig2_top(m2, dir, &br);
    }
    const bool clamped = Sv > 1.0;
    if (Sv > 1.0) Sv = 1.0;
    if (C.cdec) C.cdec[c] = cell_decision_word(br, clamped);
    A::st(C.cell.S, T.S, c, Sv);
    for (int i = 0; i < dim; i++) A::st(C.cell.dir, T.dir, c * 3 + i, dir[i]);
    double avM = 0.0;
    for (int t = b; t < e; t++) avM += C.sp[S.sp[t]].mfpot;
    avM /= (float)(e - b);
    C.cell.avm[c] = avM;
  }
}

template <typename Real, typename F>
__global__ __launch_bounds__(kBlock) void k_cellprops(Ctx C, State<Real> S, const BlkDev *cblk, int do_acc, TanDev T) {
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c < B.end) cellprops_cell<Real, F>(C, S, B.rep, c, do_acc, T);
}

// order tensor, eigen-decomposition and the S <= 1 clamp of one cell
// (lc.c:1125-1148, mpc.c:189-209) from its sums
template <int DIM>
__device__ __forceinline__ void order_of_cell(Ctx &C, int c, double Sm[3][3], int n) {
  double fD = (double)DIM, cc = 1. / (fD - 1.), POPinv = 1. / ((double)n);
  for (int i = 0; i < DIM; i++)
    for (int j = i; j < DIM; j++) Sm[i][j] *= cc * fD * POPinv;
  for (int i = 0; i < DIM; i++) Sm[i][i] -= cc;
  for (int i = 0; i < DIM; i++)
    for (int j = i + 1; j < DIM; j++) Sm[j][i] = Sm[i][j];
  double Sv, dir[3] = {0., 0., 0.};
  int br = 0;
  if (DIM == 3) {
    double ev[3];
    eig3_top(Sm, ev, dir, &br);
    Sv = -1. * (ev[1] + ev[2]);
  } else {
    double m2[2][2] = {{Sm[0][0], Sm[0][1]}, {Sm[1][0], Sm[1][1]}};
    Sv = eig2_top(m2, dir, &br);
  }
  const bool clamped = Sv > 1.0;
  if (Sv > 1.0) Sv = 1.0;
  if (C.cdec) C.cdec[c] = cell_decision_word(br, clamped);
  C.cell.S[c] = Sv;
  for (int i = 0; i < DIM; i++) C.cell.dir[c * 3 + i] = dir[i];
}

// localPROP on a tile of consecutive cells (the nematic and unfused paths):
// one lane per slot applies the body force and stages v, r, u and the mass
// in LDS with coalesced loads; one lane per cell forms the sums in list
// order with the arithmetic of cellprops_cell, whose values it gives bit for
// bit.  A tile above the capacity falls back to cellprops_cell.
template <typename Real, int DIM>
__global__ __launch_bounds__(kLcThreads) void k_cellprops_t(Ctx C, State<Real> S, const TileDev *tiles, int do_acc) {
  constexpr int NT = kLcThreads;
  __shared__ double s_v[DIM][kLcCap], s_q[DIM][kLcCap], s_u[DIM][kLcCap], s_m[kLcCap], s_mf[kLcCap];
  const TileDev Tl = tiles[blockIdx.x];
  const RepDev &R = C.rep[Tl.rep];
  const int tid = threadIdx.x;
  const int b0 = C.start[Tl.c0], n = C.start[Tl.c1] - b0, ncl = Tl.c1 - Tl.c0;
  if (n > kLcCap) {
    const TanDev T0{};
    for (int k = tid; k < ncl; k += NT) cellprops_cell<Real, double>(C, S, Tl.rep, Tl.c0 + k, do_acc, T0);
    return;
  }
  const bool ccm = (R.rtech == MPCD_RAT) || (R.lc != MPCD_ISOF), lc = R.lc != MPCD_ISOF;
  const bool acc = do_acc && R.gravflag;
  for (int i = tid; i < n; i += NT) {
    const int t = b0 + i;
    const mpcd_gpu_species &sp = C.sp[S.sp[t]];
    s_m[i] = sp.mass;
    s_mf[i] = sp.mfpot;
    double v[3] = {(double)S.vx[t], (double)S.vy[t], (double)S.vz[t]};
    if (acc) {
      for (int d = 0; d < DIM; d++) v[d] = v[d] + R.dt * R.grav[d];
      S.vx[t] = (Real)v[0];
      S.vy[t] = (Real)v[1];
      if (DIM == 3) S.vz[t] = (Real)v[2];
      v[0] = (double)S.vx[t];
      v[1] = (double)S.vy[t];
      v[2] = (double)S.vz[t];
    }
    for (int d = 0; d < DIM; d++) s_v[d][i] = v[d];
    if (ccm) {
      double q[3];
      load_q(S, t, q);
      for (int d = 0; d < DIM; d++) s_q[d][i] = q[d];
    }
    if (lc) {
      const Real *u[3] = {S.ux, S.uy, S.uz};
      for (int d = 0; d < DIM; d++) s_u[d][i] = (double)u[d][t];
    }
  }
  __syncthreads();
  for (int k = tid; k < ncl; k += NT) {
    const int c = Tl.c0 + k, ib = C.start[c] - b0, ie = C.start[c + 1] - b0;
    double vcm[3], cm[3], mass = 0.0;
    for (int d = 0; d < 3; d++) { vcm[d] = C.cell.vcm[c * 3 + d]; cm[d] = C.cell.cm[c * 3 + d]; }
    for (int d = 0; d < DIM; d++) vcm[d] = 0.0;
    if (ccm) for (int d = 0; d < DIM; d++) cm[d] = 0.0;
    for (int i = ib; i < ie; i++) {
      const double m = s_m[i];
      mass += m;
      for (int d = 0; d < DIM; d++) vcm[d] += s_v[d][i] * m;
      if (ccm) for (int d = 0; d < DIM; d++) cm[d] += s_q[d][i] * m;
    }
    if (mass > 0.0)
      for (int d = 0; d < DIM; d++) {
        vcm[d] /= mass;
        if (ccm) cm[d] /= mass;
      }
    C.cell.pop[c] = ie - ib;
    C.cell.mass[c] = mass;
    for (int d = 0; d < 3; d++) { C.cell.vcm[c * 3 + d] = vcm[d]; C.cell.cm[c * 3 + d] = cm[d]; }
    if (ccm) {
      double I[6] = {0., 0., 0., 0., 0., 0.};
      for (int i = ib; i < ie; i++) {
        const double m = s_m[i];
        double r[3] = {0., 0., 0.};
        for (int d = 0; d < DIM; d++) r[d] = s_q[d][i] - cm[d];
        I[0] += m * (r[1] * r[1] + r[2] * r[2]);
        I[1] += m * (r[0] * r[0] + r[2] * r[2]);
        I[2] += m * (r[0] * r[0] + r[1] * r[1]);
        I[3] += m * r[0] * r[1];
        I[4] += m * r[0] * r[2];
        I[5] += m * r[1] * r[2];
      }
      I[3] *= -1.; I[4] *= -1.; I[5] *= -1.;
      for (int m = 0; m < 6; m++) C.cell.I[c * 6 + m] = I[m];
    }
    if (lc && ie - ib > 1) {
      double Sm[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
      for (int i = ib; i < ie; i++)
        for (int a = 0; a < DIM; a++)
          for (int b = a; b < DIM; b++) Sm[a][b] += s_u[a][i] * s_u[b][i];
      order_of_cell<DIM>(C, c, Sm, ie - ib);
      double avM = 0.0;
      for (int i = ib; i < ie; i++) avM += s_mf[i];
      avM /= (float)(ie - ib);
      C.cell.avm[c] = avM;
    }
  }
}

// ---------------------------------------------------------------------------
// 8. ghost particles and strong anchoring at phantom walls (ghostPart,
//    mpc.c:351-567; oriBC lc.c:1290-1393)
// ---------------------------------------------------------------------------
template <typename F>
__device__ __forceinline__ F wall_W(const mpcd_gpu_wall &W, const double *wq, const F *q, int dim) {
  F Wv = 0.;                                              // calcW_PLANE, bc.c:306
  for (int i = 0; i < dim; i++) Wv += W.A[i] * (q[i] - wq[i]);
  Wv -= W.R;
  if (W.INV) Wv *= -1.;
  return Wv;
}

__device__ __forceinline__ bool anchoring(const mpcd_gpu_wall &W) { return feq(W.MUN, 0.0) || feq(W.MUT, 0.0); }

template <typename F>
__device__ __forceinline__ void ori_bc(F *u, const mpcd_gpu_wall &W, const F *n, int dim) {
  F UN[3] = {0., 0., 0.}, UT[3] = {0., 0., 0.};
  normn(u, dim);
  if (dotn(u, n, dim) >= 0.0) { for (int i = 0; i < dim; i++) u[i] *= 1.; }
  else { for (int i = 0; i < dim; i++) u[i] *= -1.; }
  projn(u, n, UN, dim);
  for (int i = 0; i < dim; i++) UT[i] = u[i] - UN[i];
  if (feq(W.MUN, 0.0) && feq(W.MUT, 0.0)) {
    for (int i = 0; i < dim; i++) u[i] = W.DUxyz[i];
  } else {
    for (int i = 0; i < dim; i++) { UN[i] *= W.MUN; UT[i] *= W.MUT; }
    for (int i = 0; i < dim; i++) u[i] = UN[i] + UT[i];
    u[0] *= W.MUxyz[0];
    if (dim >= 2) u[1] *= W.MUxyz[1];
    if (dim >= 3) u[2] *= W.MUxyz[2];
  }
  normn(u, dim);
}

// CURVED: the general surfaces of mpcd_walls.h (FP64 only): corners tested
// with calcW, the anchoring normal at each particle (normal, bc.c:1516), and
// at a curved wall the cell's S and director from the order tensor of the
// anchored orientations (mpc.c:512-536)
template <typename Real, typename F, bool CURVED = false>
__device__ __forceinline__ void ghost_cell(const Ctx &C, const State<Real> &S, int rep, int c, const Draws &Dr,
                                           const TanDev &T) {
  using A = TA<F>;
  const RepDev &R = C.rep[rep];
  if (!R.anyphantom) return;
  int pop = C.cell.pop[c];
  if (pop <= 0) return;
  const int dim = R.dim, nco = (dim == 3) ? 8 : 4;
  int lc_ = c - R.cell_off;
  int a = lc_ / (R.ny * R.nz), bb = (lc_ / R.nz) % R.ny, cz = lc_ % R.nz;
  double tq[8][3];
  for (int j = 0; j < 8; j++) {
    tq[j][0] = (double)a + (double)(j & 1);
    tq[j][1] = (double)bb + (double)((j >> 1) & 1);
    tq[j][2] = (double)cz + (double)((j >> 2) & 1);
  }
  int numBC = 0, wallindex = -1, flagW = 0;
  for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
    const mpcd_gpu_wall &W = C.wall[w];
    if (!(W.PHANTOM && anchoring(W))) continue;
    int fw = 0;
    for (int j = 0; j < nco; j++) {
      const double Wc = CURVED ? calc_w(W, &C.wq_s[w * 3], tq[j], dim) : wall_W(W, &C.wq_s[w * 3], tq[j], dim);
      if (Wc < 0.0) { fw = 1; wallindex = w; }
    }
    if (fw) numBC++;
  }
  int b = C.start[c], e = C.start[c + 1];
  for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
    const mpcd_gpu_wall &W = C.wall[w];
    if (!W.PHANTOM) continue;
    for (int j = 0; j < nco; j++) {
      const double Wc = CURVED ? calc_w(W, &C.wq_s[w * 3], tq[j], dim) : wall_W(W, &C.wq_s[w * 3], tq[j], dim);
      if (Wc < 0.0) flagW = 1;
    }
    if (R.lc != MPCD_ISOF && flagW && numBC == 1 && wallindex == w && anchoring(W)) {
      for (int t = b; t < e; t++) {
        if (!((W.inter_mask >> S.sp[t]) & 1)) continue;
        F n[3] = {0., 0., 0.}, u[3] = {A::ld(S.ux, T.u[0], t), A::ld(S.uy, T.u[1], t), A::ld(S.uz, T.u[2], t)};
        if constexpr (CURVED) {
          double q[3];
          load_q(S, t, q);
          wall_normal(W, &C.wq_s[w * 3], q, dim, n);
        } else {
          for (int k = 0; k < dim; k++) n[k] = W.A[k];
        }
        normn(n, dim);
        ori_bc(u, W, n, dim);
        A::st(S.ux, T.u[0], t, u[0]);
        A::st(S.uy, T.u[1], t, u[1]);
        A::st(S.uz, T.u[2], t, u[2]);
      }
      const bool plane = feq(W.P[0], 1.0) && feq(W.P[1], 1.0) && feq(W.P[2], 1.0);
      if (!CURVED || plane || pop == 1) {
        // a plane, or a single particle: S = 1 along the first orientation
        A::st(C.cell.S, T.S, c, F(1.0));
        F d3[3] = {A::ld(S.ux, T.u[0], b), A::ld(S.uy, T.u[1], b), A::ld(S.uz, T.u[2], b)};
        normn(d3, dim);
        for (int k = 0; k < dim; k++) A::st(C.cell.dir, T.dir, c * 3 + k, d3[k]);
      } else if constexpr (CURVED) {
        // tensOrderParam + solveEigensystem of the anchored orientations;
        // the cell's recorded decisions stay those of localPROP
        double Sm[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
        for (int t = b; t < e; t++) {
          const double u[3] = {(double)S.ux[t], (double)S.uy[t], (double)S.uz[t]};
          for (int i = 0; i < dim; i++)
            for (int j = i; j < dim; j++) Sm[i][j] += u[i] * u[j];
        }
        Ctx Cn = C;
        Cn.cdec = nullptr;
        if (dim == 3) order_of_cell<3>(Cn, c, Sm, pop);
        else order_of_cell<2>(Cn, c, Sm, pop);
      }
    }
  }
  if (flagW && pop < R.gndnst && pop > 0) {
    double invN = 1.0 / (R.gndnst - (double)pop);
    for (int k = 0; k < dim; k++) {
      double sigma = sqrt(R.kbt / invN);
      double Rg = sigma * Dr.ghost_n(rep, c, k);
      F v = A::ld(C.cell.vcm, T.vcm, c * 3 + k);
      v += Rg;
      v /= R.gndnst;
      A::st(C.cell.vcm, T.vcm, c * 3 + k, v);
    }
  }
}

template <typename Real, typename F, bool CURVED = false>
__global__ __launch_bounds__(kBlock) void k_ghost(Ctx C, State<Real> S, const BlkDev *cblk, Draws Dr, TanDev T) {
  Dr.resolve();
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c < B.end) ghost_cell<Real, F, CURVED>(C, S, B.rep, c, Dr, T);
}

// ---------------------------------------------------------------------------
// 9. nematic LC: velocity gradient (velGradD2Q9 / velGradD3Q15,
//    lc.c:2070-2162), global order (avOrderParam lc.c:982), Maier-Saupe
//    orientation collision (LCcollision lc.c:427-544) and Jeffery torque
//    (lc.c:559-718)
// ---------------------------------------------------------------------------
__global__ void k_velgrad_bulk(Ctx C, const BlkDev *cblk) {
  const BlkDev B = cblk[blockIdx.x];
  int c = B.beg + threadIdx.x;
  if (c >= B.end) return;
  const RepDev &R = C.rep[B.rep];
  int lc_ = c - R.cell_off;
  int a = lc_ / (R.ny * R.nz), b = (lc_ / R.nz) % R.ny, cz = lc_ % R.nz;
  const double *V = C.cell.vcm;
  double *E = &C.cell.E[(size_t)c * 9];
  auto idx = [&](int aa, int bb, int cc) { return R.cell_off + (aa * R.ny + bb) * R.nz + cc; };
  if (R.dim == 2) {
    if (!(a >= 1 && a < R.L[0] && b >= 1 && b < R.L[1])) return;
    const int uv[8][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
    const double w1 = 1. / 9., w5 = 1. / 36.;
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) {
        double acc = 0.;
        for (int k = 0; k < 8; k++) {
          double w = (k < 4) ? w1 : w5;
          acc += w * uv[k][j] * V[idx(a + uv[k][0], b + uv[k][1], 0) * 3 + i];
        }
        acc *= 3.;
        E[i * 3 + j] = acc;
      }
  } else {
    if (!(a >= 1 && a < R.L[0] && b >= 1 && b < R.L[1] && cz >= 1 && cz < R.L[2])) return;
    const int uv[14][3] = {{1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0},  {0, 0, 1},
                           {0, 0, -1}, {1, 1, 1},  {-1, 1, 1}, {1, -1, 1},  {-1, -1, 1},
                           {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
    const double w1 = 1. / 9., w7 = 1. / 72.;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double acc = 0.;
        for (int k = 0; k < 14; k++) {
          double w = (k < 6) ? w1 : w7;
          acc += w * uv[k][j] * V[idx(a + uv[k][0], b + uv[k][1], cz + uv[k][2]) * 3 + i];
        }
        acc *= 3.;
        E[i * 3 + j] = acc;
      }
  }
}

// edge copies in the reference's order: x faces, then y faces, then z faces
// (corners read values not yet refreshed this call, as the reference does)
__global__ void k_velgrad_edges(Ctx C) {
  const RepDev &R = C.rep[blockIdx.x];
  const int X = R.L[0], Y = R.L[1], dim = R.dim;
  const int Z = (dim == 3) ? R.L[2] : 1;
  auto idx = [&](int aa, int bb, int cc) { return R.cell_off + (aa * R.ny + bb) * R.nz + cc; };
  double *E = C.cell.E;
  const int zr = (dim == 3) ? Z : 1;
  // x faces
  for (int k = threadIdx.x; k < Y * zr; k += blockDim.x) {
    int b = k / zr, c = k % zr;
    for (int ij = 0; ij < 9; ij++) {
      if ((ij / 3) >= dim || (ij % 3) >= dim) continue;
      E[(size_t)idx(0, b, c) * 9 + ij] = E[(size_t)idx(1, b, c) * 9 + ij];
      E[(size_t)idx(X, b, c) * 9 + ij] = E[(size_t)idx(X - 1, b, c) * 9 + ij];
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < X * zr; k += blockDim.x) {
    int a = k / zr, c = k % zr;
    for (int ij = 0; ij < 9; ij++) {
      if ((ij / 3) >= dim || (ij % 3) >= dim) continue;
      E[(size_t)idx(a, 0, c) * 9 + ij] = E[(size_t)idx(a, 1, c) * 9 + ij];
      E[(size_t)idx(a, Y, c) * 9 + ij] = E[(size_t)idx(a, Y - 1, c) * 9 + ij];
    }
  }
  if (dim == 3) {
    __syncthreads();
    for (int k = threadIdx.x; k < X * Y; k += blockDim.x) {
      int a = k / Y, b = k % Y;
      for (int ij = 0; ij < 9; ij++) {
        E[(size_t)idx(a, b, 0) * 9 + ij] = E[(size_t)idx(a, b, 1) * 9 + ij];
        E[(size_t)idx(a, b, Z) * 9 + ij] = E[(size_t)idx(a, b, Z - 1) * 9 + ij];
      }
    }
  }
}

// global order tensor partial sums (lc = 2): 6 unique entries per block
template <typename Real>
__global__ void k_avorder_part(Ctx C, State<Real> S, const BlkDev *blk, double *part) {
  const BlkDev B = blk[blockIdx.x];
  int t = B.beg + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (t < B.end) {
    double u[3] = {(double)S.ux[t], (double)S.uy[t], (double)S.uz[t]};
    v[0] = u[0] * u[0]; v[1] = u[0] * u[1]; v[2] = u[0] * u[2];
    v[3] = u[1] * u[1]; v[4] = u[1] * u[2]; v[5] = u[2] * u[2];
  }
  __shared__ double sh[6][kBlock];
  for (int k = 0; k < 6; k++) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 6; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; k++) part[blockIdx.x * 6 + k] = sh[k][0];
}

__global__ void k_avorder_final(Ctx C, const BlkDev *blk, int nblk, const double *part) {
  int r = threadIdx.x;
  if (r >= C.nrep) return;
  const RepDev &R = C.rep[r];
  double s6[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nblk; b++)
    if (blk[b].rep == r)
      for (int k = 0; k < 6; k++) s6[k] += part[b * 6 + k];
  const int dim = R.dim;
  double fD = (double)dim, cc = 1. / (fD - 1.), inv = 1. / ((double)R.npart);
  double Sm[3][3] = {{s6[0], s6[1], s6[2]}, {s6[1], s6[3], s6[4]}, {s6[2], s6[4], s6[5]}};
  for (int i = 0; i < dim; i++)
    for (int j = i; j < dim; j++) Sm[i][j] *= cc * fD * inv;
  for (int i = 0; i < dim; i++) Sm[i][i] -= cc;
  for (int i = 0; i < dim; i++)
    for (int j = i + 1; j < dim; j++) Sm[j][i] = Sm[i][j];
  double dir[3];
  if (dim == 3) {
    double ev[3];
    eig3_top(Sm, ev, dir);
    C.sg[r] = -1. * (ev[1] + ev[2]);
  } else {
    double m2[2][2] = {{Sm[0][0], Sm[0][1]}, {Sm[1][0], Sm[1][1]}};
    C.sg[r] = eig2_top(m2, dir);
  }
}

// Maier-Saupe samplers (lc.c:52-411); `slot` keys the draws
// `dec` receives the decision word of the draw (lc_decision_word)
template <typename F>
__device__ void ms_sample(const RepDev &R, const Draws &Dr, uint32_t rep, uint32_t slot, const F *DIR, F *rotAx,
                          const F &rotAngle, F *U, const F &KBT, const F &S, const F &effM, uint32_t &dec) {
  if (R.dim == 2) {
    F newU[3] = {U[0], U[1], 0.0};
    F ang0 = signed_angle(DIR, U, 2);
    const bool redraw = feq(ang0, 0.5 * kPI) || feq(ang0, -0.5 * kPI);
    if (redraw) ang0 = kPI * (0.5 - Dr.lc_u(rep, slot, 0));
    F uz0 = dotn(DIR, U, 2);
    int annealNum = (int)val(kMCINT + kMCSLOPE * effM * S / KBT);
    int cnt = 0;
    for (int i = 0; i < annealNum; i++) {
      int accept = 0;
      F w0 = effM * S * uz0 * uz0 / KBT;
      F ang1 = kPI * (0.5 - Dr.lc_u(rep, slot, 1 + 2 * i));
      F uz1 = cos(ang1);
      F w1 = effM * S * uz1 * uz1 / KBT;
      F dw = w1 - w0;
      if (dw >= 0.) accept = 1;
      else if (Dr.lc_u(rep, slot, 2 + 2 * i) <= exp(dw)) accept = 1;
      if (accept) { cnt += 1; ang0 = ang1; uz0 = uz1; }
    }
    newU[0] = uz0;
    newU[1] = sin(ang0);
    if (cnt > 0) {
      rodrigues(newU, rotAx, rotAngle);
      U[0] = newU[0];
      U[1] = newU[1];
    }
    dec = lc_decision_word(0, redraw, false, annealNum, cnt);
    return;
  }
  F BUS = effM * S / KBT;
  F newU[3] = {U[0], U[1], U[2]};
  if (BUS > kBUSMAX) {
    F sd = 1. / sqrt(2. * effM * S / KBT);
    F theta = sd * Dr.lc_n(rep, slot, 0) + 0.0;
    F uz = 1. - 0.5 * theta * theta;
    F pm = (Dr.lc_u(rep, slot, 1) <= 0.5) ? -1.0 : 1.0;
    newU[0] = pm * uz;
    F phi = Dr.lc_u(rep, slot, 2) * 2.0 * kPI;
    newU[1] = sqrt(1.0 - uz * uz) * cos(phi);
    newU[2] = sqrt(1.0 - uz * uz) * sin(phi);
    rodrigues(newU, rotAx, rotAngle);
    for (int d = 0; d < 3; d++) U[d] = newU[d];
    dec = lc_decision_word(1, false, val(pm) < 0.0, 0, 0);
  } else if (BUS < kBUSMIN) {
    F B2 = BUS;
    if (B2 < kTOL) B2 = kTOL;
    F Rr = Dr.lc_u(rep, slot, 0) * (1. + B2 / 3.);
    F c = sqrt(9. * B2 * B2 * B2 * B2 * Rr * Rr + 4. * B2 * B2 * B2);
    F num = pow(3. * B2 * B2 * Rr + c, 1. / 3.);
    F denom = pow(2., 1. / 3.);
    F uz = num / (denom * B2) - denom / num;
    F pm = (Dr.lc_u(rep, slot, 1) <= 0.5) ? -1.0 : 1.0;
    newU[0] = pm * uz;
    F phi = Dr.lc_u(rep, slot, 2) * 2.0 * kPI;
    newU[1] = sqrt(1.0 - uz * uz) * cos(phi);
    newU[2] = sqrt(1.0 - uz * uz) * sin(phi);
    rodrigues(newU, rotAx, rotAngle);
    for (int d = 0; d < 3; d++) U[d] = newU[d];
    dec = lc_decision_word(2, false, val(pm) < 0.0, 0, 0);
  } else {
    F uz0 = dotn(DIR, U, 3);
    int annealNum = (int)val(kMCINT + kMCSLOPE * effM * S / KBT);
    int cnt = 0;
    for (int i = 0; i < annealNum; i++) {
      int accept = 0;
      F w0 = effM * S * uz0 * uz0 / KBT;
      F uz1 = Dr.lc_u(rep, slot, 1 + 2 * i);
      F w1 = effM * S * uz1 * uz1 / KBT;
      F dw = w1 - w0;
      if (dw >= 0.) accept = 1;
      else if (Dr.lc_u(rep, slot, 2 + 2 * i) <= exp(dw)) accept = 1;
      if (accept) { cnt += 1; uz0 = uz1; }
    }
    F phi = Dr.lc_u(rep, slot, 1 + 2 * annealNum) * 2.0 * kPI;
    F pm = (Dr.lc_u(rep, slot, 2 + 2 * annealNum) <= 0.5) ? -1.0 : 1.0;
    newU[0] = pm * uz0;
    newU[1] = sqrt(1.0 - uz0 * uz0) * cos(phi);
    newU[2] = sqrt(1.0 - uz0 * uz0) * sin(phi);
    if (cnt > 0) {
      rodrigues(newU, rotAx, rotAngle);
      for (int d = 0; d < 3; d++) U[d] = newU[d];
    }
    dec = lc_decision_word(3, false, val(pm) < 0.0, annealNum, cnt);
  }
}

// LCcollision (lc.c:427-544) and jefferysTorque + larsonRotRate
// (lc.c:559-718) with one lane per particle: the cell's director, order,
// mean field and velocity gradient are recomputed by each lane with the
// per-cell arithmetic of the reference (the mean field summed over the cell
// in list order), and the Metropolis draws are keyed by slot, so the result
// does not depend on the mapping
template <typename Real, typename F>
__device__ __forceinline__ void lc_particle(const Ctx &C, const State<Real> &S, int rep, int t, const Draws &Dr,
                                            const TanDev &T) {
  using A = TA<F>;
  const RepDev &R = C.rep[rep];
  const int c = S.cst[t];
  const int b = C.start[c], e = C.start[c + 1];
  if (e - b <= 1) {
    if (C.pdec) C.pdec[R.part_off + S.id[t]] = 0u;
    return;
  }
  const int dim = R.dim;
  F DIR[3] = {0., 0., 0.}, rotAx[3], xaxis[3] = {1.0, 0.0, 0.0};
  F Sv = (R.lc == MPCD_LCL) ? A::ld(C.cell.S, T.S, c) : F(C.sg[rep]);
  for (int i = 0; i < dim; i++) DIR[i] = A::ld(C.cell.dir, T.dir, c * 3 + i);
  cross3(DIR, xaxis, rotAx);
  F rotAngle = -1. * abs_angle(DIR, xaxis, dim);
  normn(rotAx, 3);
  const double avM = C.cell.avm[c];  // summed in list order by localPROP
  double effM = avM * 0.5 * (double)dim;
  const mpcd_gpu_species &sp = C.sp[S.sp[t]];
  {
    F U[3] = {A::ld(S.ux, T.u[0], t), A::ld(S.uy, T.u[1], t), A::ld(S.uz, T.u[2], t)}, Tq[3], dU[3] = {0., 0., 0.};
    for (int i = 0; i < dim; i++) dU[i] = U[i];
    uint32_t dec = 0;
    ms_sample(R, Dr, rep, t, DIR, rotAx, rotAngle, U, F(R.kbt), Sv, F(effM), dec);
    if (C.pdec) C.pdec[R.part_off + S.id[t]] = dec;
    for (int i = 0; i < dim; i++) dU[i] = (U[i] - dU[i]) / R.dt;
    cross3(U, dU, Tq);
    for (int i = 0; i < 3; i++) Tq[i] *= sp.rfc;
    A::st(S.ux, T.u[0], t, U[0]);
    A::st(S.uy, T.u[1], t, U[1]);
    A::st(S.uz, T.u[2], t, U[2]);
    for (int i = 0; i < 3; i++) A::st(C.T, T.T, (size_t)t * 3 + i, Tq[i]);
  }
  F D[3][3], W[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      F Eij = A::ld(C.cell.E, T.E, (size_t)c * 9 + i * 3 + j), Eji = A::ld(C.cell.E, T.E, (size_t)c * 9 + j * 3 + i);
      D[j][i] = 0.5 * (Eij + Eji);
      W[j][i] = 0.5 * (Eij - Eji);
    }
  F u[3] = {A::ld(S.ux, T.u[0], t), A::ld(S.uy, T.u[1], t), A::ld(S.uz, T.u[2], t)}, dudt[3], w[3];
  F uuuD[3] = {0., 0., 0.}, uw[3] = {0., 0., 0.}, uD[3] = {0., 0., 0.};
  for (int j = 0; j < 3; j++) {
    for (int i = 0; i < 3; i++) {
      for (int k = 0; k < 3; k++) uuuD[j] += u[j] * (u[i] * D[i][k] * u[k]);
      uD[j] += u[i] * D[i][j];
      uw[j] += u[i] * W[i][j];
    }
    dudt[j] = uw[j] + sp.tumble * (uD[j] - uuuD[j]);
  }
  cross3(u, dudt, w);
  for (int i = 0; i < 3; i++) w[i] *= sp.chihi;
  for (int i = 0; i < 3; i++) w[i] *= R.dt;
  F theta = lengthn(w, 3);
  normn(w, 3);
  rodrigues(u, w, theta);
  A::st(S.ux, T.u[0], t, u[0]);
  A::st(S.uy, T.u[1], t, u[1]);
  A::st(S.uz, T.u[2], t, u[2]);
}

template <typename Real, typename F>
__global__ __launch_bounds__(kBlock) void k_lc(Ctx C, State<Real> S, const BlkDev *pblk, Draws Dr, TanDev T) {
  Dr.resolve();
  const BlkDev B = pblk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t < B.end) lc_particle<Real, F>(C, S, B.rep, t, Dr, T);
}

// ---------------------------------------------------------------------------
// 10. collision operators (MPCcollision switch, mpc.c:2777-2812), then the
//     kinetic-energy partial sums of TEMP (therm.c:209) and gridShift back
// ---------------------------------------------------------------------------
// dipole activity of the cell holding slots [b, e) (dipoleAndersenROT_LC,
// lc.c:1811-1841): ACT_i / m_i summed in list order over the particles that
// pass MINACTRATIO, normalised (AV, SIG) and cut off (SIG, SIG_SUM), times dt.
// spof(t) is the species of slot t; the species populations of the cell are
// counted only for a species with a MINACTRATIO
template <typename F, typename SpOf>
__device__ __forceinline__ F cell_activity(const mpcd_gpu_species *sps, SpOf spof, const RepDev &R, int rep, int b,
                                           int e, const TanDev &T) {
  const int n = e - b;
  F ACT = 0.;
  double sigW = 0., sigP = 0.;
  for (int t = b; t < e; t++) {
    const int k = spof(t);
    const mpcd_gpu_species &sp = sps[k];
    bool pass = (sp.minactratio == 0.0);
    if (!pass) {
      int pop = 0;
      for (int t2 = b; t2 < e; t2++) pop += (spof(t2) == k) ? 1 : 0;
      pass = sp.minactratio * R.gndnst < (double)pop;
    }
    if (pass) {
      ACT += TA<F>::par(sp.act, T.dact[k], th_slot(T, rep, kThAct + k)) / sp.mass;
      sigW += sp.sigwidth;
      sigP += sp.sigpos;
    }
  }
  if (R.rtech == MPCD_DIPOLE_DIR_AV || R.rtech == MPCD_DIPOLE_DIR_SIG) ACT *= R.gndnst / ((double)n);
  if (R.rtech == MPCD_DIPOLE_DIR_SIG || R.rtech == MPCD_DIPOLE_DIR_SIG_SUM) {
    sigW /= (double)n;
    sigP /= (double)n;
    double fall = (1 - tanh(((double)n - R.gndnst * (1 + sigP)) / (R.gndnst * sigW)));
    ACT *= fall / 2;
  }
  ACT *= R.dt;
  return ACT;
}

__device__ __forceinline__ int lc_mode(const RepDev &R) {
  return (R.lc == MPCD_ISOF) ? 0
         : (R.rtech == MPCD_DIPOLE_DIR_SUM || R.rtech == MPCD_DIPOLE_DIR_AV || R.rtech == MPCD_DIPOLE_DIR_SIG ||
            R.rtech == MPCD_DIPOLE_DIR_SIG_SUM)
             ? 2
             : 1;
}

// the collision of cell c of replica `rep` and its kinetic-energy sum (one
// lane per cell)
template <typename Real, typename F>
__device__ double collide_cell(const Ctx &C, const State<Real> &S, int rep, int c, const Draws &Dr, const TanDev &T) {
  using A = TA<F>;
  const RepDev &R = C.rep[rep];
  const int dim = R.dim;
  double ke = 0.0;
  {
    int b = C.start[c], e = C.start[c + 1], n = e - b;
    F vcm[3] = {A::ld(C.cell.vcm, T.vcm, c * 3), A::ld(C.cell.vcm, T.vcm, c * 3 + 1),
                A::ld(C.cell.vcm, T.vcm, c * 3 + 2)};
    if (n > 1) {
      if (R.lc == MPCD_ISOF && (R.rtech == MPCD_ARBAXIS || R.rtech == MPCD_ORTHAXIS)) {
        // stochrotMPC + rotate, mpc.c:1181-1333
        double RV[3] = {0., 0., 0.}, M[3][3];
        int CA = 0;
        if (R.rtech == MPCD_ARBAXIS) {
          if (dim == 3) Dr.cell_axis(rep, c, RV);
          else { double x = Dr.cell_u(rep, c, 0); RV[0] = x; RV[1] = sqrt(1. - x * x); }
        } else {
          CA = (int)((int)dim * Dr.cell_u(rep, c, 0));
        }
        double sgn = Dr.cell_u(rep, c, 3);
        double Cc = R.cosA, Sn;
        if (R.true_sign ? (sgn >= 0.5) : ((long)sgn > 0.5)) Sn = R.sinA;
        else Sn = -1. * R.sinA;
        if (dim == 3) {
          const double *L = RV;
          if (R.rtech == MPCD_ARBAXIS) {
            M[0][0] = L[0] * L[0] + (1. - L[0] * L[0]) * Cc;
            M[0][1] = L[0] * L[1] * (1. - Cc) - L[2] * Sn;
            M[0][2] = L[0] * L[2] * (1. - Cc) + L[1] * Sn;
            M[1][0] = L[0] * L[1] * (1. - Cc) + L[2] * Sn;
            M[1][1] = L[1] * L[1] + (1. - L[1] * L[1]) * Cc;
            M[1][2] = L[1] * L[2] * (1. - Cc) - L[0] * Sn;
            M[2][0] = L[0] * L[2] * (1. - Cc) - L[1] * Sn;
            M[2][1] = L[1] * L[2] * (1. - Cc) + L[0] * Sn;
            M[2][2] = L[2] * L[2] + (1. - L[2] * L[2]) * Cc;
          } else {
            // mpc.c:1206-1237, the x-axis entry -.1 S as written (mpc.c:1215)
            double M9[9];
            orth_matrix(CA, Cc, Sn, M9);
            for (int i = 0; i < 3; i++)
              for (int j = 0; j < 3; j++) M[i][j] = M9[3 * i + j];
          }
        } else {
          M[0][0] = Cc; M[0][1] = -1. * Sn;
          M[1][0] = Sn; M[1][1] = Cc;
        }
        for (int t = b; t < e; t++) {
          F V[3] = {A::ld(S.vx, T.v[0], t) - vcm[0], A::ld(S.vy, T.v[1], t) - vcm[1], A::ld(S.vz, T.v[2], t) - vcm[2]};
          F TM[3] = {0.0, 0.0, 0.0};
          for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) TM[i] += M[i][j] * V[j];
          A::st(S.vx, T.v[0], t, vcm[0] + TM[0]);
          A::st(S.vy, T.v[1], t, vcm[1] + TM[1]);
          if (dim == 3) A::st(S.vz, T.v[2], t, vcm[2] + TM[2]);
        }
      } else if (R.lc == MPCD_ISOF && R.rtech == MPCD_MPCAT) {
        // andersenMPC, mpc.c:1353-1444
        double RS[3] = {0, 0, 0};
        for (int t = b; t < e; t++) {
          double m = C.sp[S.sp[t]].mass, sig = sqrt(R.kbt / m);
          {
            double nrm[3];
            Dr.part_n_all(rep, t, dim, nrm);
            for (int j = 0; j < dim; j++) RS[j] += m * (sig * nrm[j]);
          }
        }
        for (int j = 0; j < dim; j++) RS[j] /= C.cell.mass[c];
        for (int t = b; t < e; t++) {
          const mpcd_gpu_species &sp = C.sp[S.sp[t]];
          double sig = sqrt(R.kbt / sp.mass);
          F v[3];
          double nrm[3];
          Dr.part_n_all(rep, t, dim, nrm);
          for (int j = 0; j < dim; j++) {
            double rv = sig * nrm[j];
            F dv = (sp.damp) * (vcm[j]) / ((double)n);
            v[j] = vcm[j] + rv - RS[j] - dv;
          }
          A::st(S.vx, T.v[0], t, v[0]);
          A::st(S.vy, T.v[1], t, v[1]);
          if (dim == 3) A::st(S.vz, T.v[2], t, v[2]);
        }
      } else {
        // andersenROT (mpc.c:1461-1657), andersenROT_LC (lc.c:1515-1728),
        // dipoleAndersenROT_LC (lc.c:1749-2043)
        const int mode = lc_mode(R);
        F cm[3] = {A::ld(C.cell.cm, T.cm, c * 3), A::ld(C.cell.cm, T.cm, c * 3 + 1), A::ld(C.cell.cm, T.cm, c * 3 + 2)};
        F ACT = 0., PA[3] = {0., 0., 0.};
        if (mode == 2) {
          ACT = cell_activity<F>(C.sp, [&](int t) { return (int)S.sp[t]; }, R, rep, b, e, T);
          for (int j = 0; j < dim; j++) PA[j] = A::ld(C.cell.dir, T.dir, c * 3 + j);
        }
        // a tangent along the activity keeps the active terms where the
        // activity itself is zero (they add exact zeros to the values)
        const bool act = (mode == 2) && (fneq(ACT, 0.0) || A::moves(ACT));
        double RS[3] = {0, 0, 0};
        F AS[3] = {0., 0., 0.};
        for (int t = b; t < e; t++) {
          const mpcd_gpu_species &sp = C.sp[S.sp[t]];
          double sig = sqrt(R.kbt / sp.mass);
          {
            double nrm[3];
            Dr.part_n_all(rep, t, dim, nrm);
            for (int j = 0; j < dim; j++) RS[j] += sp.mass * (sig * nrm[j]);
          }
          if (act) {
            F q[3], pW = 0.0;
            load_qt(S, T, t, q);
            for (int j = 0; j < dim; j++) pW += PA[j] * (q[j] - cm[j]);
            pW -= 0.0;
            double pm = (pW <= 0) ? -1. : 1.;
            if (C.pdec && pm > 0.0) C.pdec[R.part_off + S.id[t]] |= kDecSide;
            for (int j = 0; j < dim; j++) AS[j] += sp.mass * (pm * ACT * PA[j]);
          }
        }
        for (int j = 0; j < dim; j++) RS[j] /= C.cell.mass[c];
        if (mode == 2) for (int j = 0; j < dim; j++) AS[j] /= C.cell.mass[c];
        F II[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
        F I6[6];
        for (int k = 0; k < 6; k++) I6[k] = A::ld(C.cell.I, T.I, c * 6 + k);
        if (dim == 3) {
          if (n >= 3) {
            F Im[3][3] = {{I6[0], I6[3], I6[4]}, {I6[3], I6[1], I6[5]}, {I6[4], I6[5], I6[2]}};
            inv3(II, Im);
          }
        } else {
          II[2][2] = 1. / I6[2];
        }
        F L[3] = {0., 0., 0.}, Llm[3] = {0., 0., 0.}, Lnm[3] = {0., 0., 0.};
        for (int t = b; t < e; t++) {
          const mpcd_gpu_species &sp = C.sp[S.sp[t]];
          double sig = sqrt(R.kbt / sp.mass);
          F q[3], rq[3] = {0., 0., 0.}, dv[3] = {0., 0., 0.}, am[3];
          F v[3] = {A::ld(S.vx, T.v[0], t), A::ld(S.vy, T.v[1], t), A::ld(S.vz, T.v[2], t)};
          load_qt(S, T, t, q);
          for (int j = 0; j < dim; j++) rq[j] = q[j] - cm[j];
          double nrm[3];
          Dr.part_n_all(rep, t, dim, nrm);
          for (int j = 0; j < dim; j++) {
            double rv = sig * nrm[j];
            if (act) {
              F pW = 0.0;
              for (int jj = 0; jj < dim; jj++) pW += PA[jj] * (q[jj] - cm[jj]);
              pW -= 0.0;
              double pm = (pW <= 0) ? -1. : 1.;
              dv[j] = sp.mass * (v[j] - rv - pm * ACT * PA[j]);
            } else if (mode == 2) {
              dv[j] = sp.mass * (v[j] - rv - 0.0);
            } else {
              dv[j] = sp.mass * (v[j] - rv);
            }
          }
          cross3(rq, dv, am);
          if (mode == 0) for (int j = 0; j < 3; j++) L[j] += am[j];
          else {
            for (int j = 0; j < 3; j++) Llm[j] += am[j];
            for (int j = 0; j < 3; j++) Lnm[j] += A::ld(C.T, T.T, (size_t)t * 3 + j) * R.dt;
          }
        }
        if (mode != 0) for (int j = 0; j < 3; j++) L[j] = Llm[j] + Lnm[j];
        F Wv[3];
        for (int i = 0; i < 3; i++) {
          Wv[i] = 0.;
          for (int j = 0; j < 3; j++) Wv[i] += II[i][j] * L[j];
        }
        for (int t = b; t < e; t++) {
          const mpcd_gpu_species &sp = C.sp[S.sp[t]];
          double sig = sqrt(R.kbt / sp.mass);
          F q[3], rq[3] = {0., 0., 0.}, at[3], v[3];
          load_qt(S, T, t, q);
          for (int j = 0; j < dim; j++) rq[j] = q[j] - cm[j];
          cross3(Wv, rq, at);
          double pm = 0.0;
          if (act) {
            F pW = 0.0;
            for (int jj = 0; jj < dim; jj++) pW += PA[jj] * (q[jj] - cm[jj]);
            pW -= 0.0;
            pm = (pW <= 0) ? -1. : 1.;
          }
          double nrm[3];
          Dr.part_n_all(rep, t, dim, nrm);
          for (int j = 0; j < dim; j++) {
            double rv = sig * nrm[j];
            if (mode == 0) v[j] = vcm[j] + rv - RS[j] - (sp.damp) * (vcm[j]) + at[j];
            else if (mode == 1) v[j] = vcm[j] + rv - RS[j] + at[j];
            else v[j] = vcm[j] + rv - RS[j] + at[j] + (act ? pm * ACT * PA[j] : F(0.0)) - AS[j] - (sp.damp) * (vcm[j]);
          }
          A::st(S.vx, T.v[0], t, v[0]);
          A::st(S.vy, T.v[1], t, v[1]);
          if (dim == 3) A::st(S.vz, T.v[2], t, v[2]);
        }
      }
    }
    // TEMP partial sum (post collision) and gridShift back
    for (int t = b; t < e; t++) {
      double m = C.sp[S.sp[t]].mass;
      double v0 = (double)S.vx[t] - R.avv[0], v1 = (double)S.vy[t] - R.avv[1], v2 = (double)S.vz[t] - R.avv[2];
      double tmp = v0 * v0;
      tmp += v1 * v1;
      tmp += v2 * v2;
      ke += m * tmp;
      if (R.galinv) {
        F q[3];
        load_qt(S, T, t, q);
        shift_back(R, &C.shift[rep * 3], q);
        A::st(S.x, T.q[0], t, q[0]);
        A::st(S.y, T.q[1], t, q[1]);
        if (dim == 3) A::st(S.z, T.q[2], t, q[2]);
      }
    }
  }
  return ke;
}

template <typename Real, typename F>
__global__ __launch_bounds__(kBlock) void k_collide(Ctx C, State<Real> S, const BlkDev *cblk, Draws Dr, TanDev T) {
  Dr.resolve();
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  const double ke = (c < B.end) ? collide_cell<Real, F>(C, S, B.rep, c, Dr, T) : 0.0;
  // block reduction of the kinetic energy
  __shared__ double sh[kBlock];
  sh[threadIdx.x] = ke;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) C.tpart[blockIdx.x] = sh[0];
}

// The nematic collision operators (andersenROT_LC, lc.c:1515-1728, and
// dipoleAndersenROT_LC, lc.c:1749-2043) on a tile of consecutive cells whose
// slots fit LDS.  One lane per slot draws its thermal velocity once and
// computes its terms of the cell sums (momentum of the draws, active
// momentum, angular momentum, torque) with the expressions of collide_cell;
// one lane per cell adds the terms in list order and forms W = I^-1 L; one
// lane per slot applies the update.  The values are collide_cell's bit for
// bit; a tile above the capacity is left to k_collide_lc_over.

template <typename Real, int DIM>
__global__ __launch_bounds__(kLcThreads) void k_collide_lc(Ctx C, State<Real> S, const TileDev *tiles, Draws Dr,
                                                           int32_t *over) {
  constexpr int NT = kLcThreads, NTERM = (DIM == 3) ? 12 : 6;
  // terms per slot: 3D RS[3] AS[3] L[3] T dt[3]; 2D RS[2] AS[2] L_z (T dt)_z
  __shared__ double s_term[NTERM][kLcCap];
  __shared__ double s_act[kLcTileCells];
  __shared__ int s_actf[kLcTileCells];
  __shared__ double s_cell[kLcTileCells][9];  // RS, AS, W
  __shared__ double s_ke[NT / 64];
  __shared__ mpcd_gpu_species s_spc[MPCD_MAXSPEC];
  __shared__ int32_t s_sp[kLcCap];
  Dr.resolve();
  const TileDev Tl = tiles[blockIdx.x];
  const RepDev &R = C.rep[Tl.rep];
  const int tid = threadIdx.x;
  const int b0 = C.start[Tl.c0], n = C.start[Tl.c1] - b0, ncl = Tl.c1 - Tl.c0;
  const int mode = lc_mode(R);
  const TanDev T0{};
  double ke = 0.0;
  if (n > kLcCap) {
    // k_collide_lc_over takes the tile (and writes its kinetic energy)
    if (tid == 0) over[1 + atomicAdd(&over[0], 1)] = blockIdx.x;
  } else {
    // the species table and the tile's species ids in LDS
    for (int k = tid; k < MPCD_MAXSPEC; k += NT) s_spc[k] = C.sp[k];
    for (int i = tid; i < n; i += NT) s_sp[i] = S.sp[b0 + i];
    __syncthreads();
    // 0. one lane per cell: the activity
    for (int k = tid; k < ncl; k += NT) {
      const int c = Tl.c0 + k, b = C.start[c], e = C.start[c + 1];
      double ACT = 0.;
      if (mode == 2 && e - b > 1) ACT = cell_activity<double>(s_spc, [&](int t) { return s_sp[t - b0]; }, R, Tl.rep, b, e, T0);
      s_act[k] = ACT;
      s_actf[k] = (mode == 2) && fneq(ACT, 0.0);
    }
    __syncthreads();
    // 1. one lane per slot: draws and the terms of the cell sums
    double rvk[kLcPPT][3], rqk[kLcPPT][3], pmk[kLcPPT];
    int ck[kLcPPT];
#pragma unroll
    for (int j = 0; j < kLcPPT; j++) {
      const int i = tid + j * NT;
      ck[j] = -1;
      pmk[j] = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) rvk[j][d] = rqk[j][d] = 0.0;
      if (i >= n) continue;
      const int t = b0 + i, c = S.cst[t], k = c - Tl.c0;
      ck[j] = c;
      if (C.start[c + 1] - C.start[c] <= 1) continue;
      const mpcd_gpu_species &sp = s_spc[s_sp[i]];
      const double sig = sqrt(R.kbt / sp.mass);
      const double ACT = s_act[k];
      const bool act = s_actf[k] != 0;
      double cm[3] = {C.cell.cm[c * 3], C.cell.cm[c * 3 + 1], C.cell.cm[c * 3 + 2]}, PA[3] = {0., 0., 0.};
      if (mode == 2)
        for (int d = 0; d < DIM; d++) PA[d] = C.cell.dir[c * 3 + d];
      double q[3], rq[3] = {0., 0., 0.}, dv[3] = {0., 0., 0.}, am[3];
      const double v[3] = {(double)S.vx[t], (double)S.vy[t], (double)S.vz[t]};
      load_q(S, t, q);
      for (int d = 0; d < DIM; d++) rq[d] = q[d] - cm[d];
      double pm = 0.0;
      if (act) {
        double pW = 0.0;
        for (int jj = 0; jj < DIM; jj++) pW += PA[jj] * (q[jj] - cm[jj]);
        pW -= 0.0;
        pm = (pW <= 0) ? -1. : 1.;
        if (C.pdec && pm > 0.0) C.pdec[R.part_off + S.id[t]] |= kDecSide;
      }
      double nrm[3];
      Dr.part_n_all(Tl.rep, t, DIM, nrm);
#pragma unroll
      for (int d = 0; d < DIM; d++) {
        const double rv = sig * nrm[d];
        rvk[j][d] = rv;
        s_term[d][i] = sp.mass * rv;
        s_term[DIM + d][i] = act ? sp.mass * (pm * ACT * PA[d]) : 0.0;
        if (act) dv[d] = sp.mass * (v[d] - rv - pm * ACT * PA[d]);
        else if (mode == 2) dv[d] = sp.mass * (v[d] - rv - 0.0);
        else dv[d] = sp.mass * (v[d] - rv);
      }
      cross3(rq, dv, am);
      if (DIM == 3) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          s_term[6 + d][i] = am[d];
          s_term[9 + d][i] = C.T[(size_t)t * 3 + d] * R.dt;
        }
      } else {
        s_term[4][i] = am[2];
        s_term[5][i] = C.T[(size_t)t * 3 + 2] * R.dt;
      }
#pragma unroll
      for (int d = 0; d < 3; d++) rqk[j][d] = rq[d];
      pmk[j] = pm;
    }
    __syncthreads();
    // 2. one lane per cell: the sums in list order and W = I^-1 L
    for (int k = tid; k < ncl; k += NT) {
      const int c = Tl.c0 + k, ib = C.start[c] - b0, ie = C.start[c + 1] - b0, nn = ie - ib;
      if (nn <= 1) continue;
      const bool act = s_actf[k] != 0;
      double RS[3] = {0, 0, 0}, AS[3] = {0., 0., 0.};
      for (int i = ib; i < ie; i++) {
#pragma unroll
        for (int d = 0; d < DIM; d++) RS[d] += s_term[d][i];
        if (act)
#pragma unroll
          for (int d = 0; d < DIM; d++) AS[d] += s_term[DIM + d][i];
      }
      for (int d = 0; d < DIM; d++) RS[d] /= C.cell.mass[c];
      if (mode == 2)
        for (int d = 0; d < DIM; d++) AS[d] /= C.cell.mass[c];
      double II[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
      double I6[6];
      for (int m = 0; m < 6; m++) I6[m] = C.cell.I[c * 6 + m];
      if (DIM == 3) {
        if (nn >= 3) {
          double Im[3][3] = {{I6[0], I6[3], I6[4]}, {I6[3], I6[1], I6[5]}, {I6[4], I6[5], I6[2]}};
          inv3(II, Im);
        }
      } else {
        II[2][2] = 1. / I6[2];
      }
      double Llm[3] = {0., 0., 0.}, Lnm[3] = {0., 0., 0.}, L[3];
      for (int i = ib; i < ie; i++) {
        if (DIM == 3) {
#pragma unroll
          for (int d = 0; d < 3; d++) {
            Llm[d] += s_term[6 + d][i];
            Lnm[d] += s_term[9 + d][i];
          }
        } else {
          Llm[2] += s_term[4][i];
          Lnm[2] += s_term[5][i];
        }
      }
      for (int d = 0; d < 3; d++) L[d] = Llm[d] + Lnm[d];
      double Wv[3];
      for (int a = 0; a < 3; a++) {
        Wv[a] = 0.;
        for (int b = 0; b < 3; b++) Wv[a] += II[a][b] * L[b];
      }
      for (int d = 0; d < 3; d++) {
        s_cell[k][d] = RS[d];
        s_cell[k][3 + d] = AS[d];
        s_cell[k][6 + d] = Wv[d];
      }
    }
    __syncthreads();
    // 3. one lane per slot: the update, the kinetic energy and gridShift back
#pragma unroll
    for (int j = 0; j < kLcPPT; j++) {
      const int c = ck[j];
      if (c < 0) continue;
      const int t = b0 + tid + j * NT, k = c - Tl.c0;
      const mpcd_gpu_species &sp = s_spc[s_sp[tid + j * NT]];
      if (C.start[c + 1] - C.start[c] > 1) {
        const double vcm[3] = {C.cell.vcm[c * 3], C.cell.vcm[c * 3 + 1], C.cell.vcm[c * 3 + 2]};
        const double Wv[3] = {s_cell[k][6], s_cell[k][7], s_cell[k][8]};
        double at[3], v[3];
        cross3(Wv, rqk[j], at);
        const bool act = s_actf[k] != 0;
        const double ACT = s_act[k];
        for (int d = 0; d < DIM; d++) {
          const double rv = rvk[j][d], RSd = s_cell[k][d];
          if (mode == 1) {
            v[d] = vcm[d] + rv - RSd + at[d];
          } else {
            const double PAd = C.cell.dir[c * 3 + d];
            v[d] = vcm[d] + rv - RSd + at[d] + (act ? pmk[j] * ACT * PAd : 0.0) - s_cell[k][3 + d] -
                   (sp.damp) * (vcm[d]);
          }
        }
        S.vx[t] = (Real)v[0];
        S.vy[t] = (Real)v[1];
        if (DIM == 3) S.vz[t] = (Real)v[2];
      }
      const double m = sp.mass;
      double v0 = (double)S.vx[t] - R.avv[0], v1 = (double)S.vy[t] - R.avv[1], v2 = (double)S.vz[t] - R.avv[2];
      double tmp = v0 * v0;
      tmp += v1 * v1;
      tmp += v2 * v2;
      ke += m * tmp;
      if (R.galinv) {
        double q[3];
        load_q(S, t, q);
        shift_back(R, &C.shift[Tl.rep * 3], q);
        S.x[t] = (Real)q[0];
        S.y[t] = (Real)q[1];
        if (DIM == 3) S.z[t] = (Real)q[2];
      }
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ke += __shfl_xor(ke, o, 64);
  if ((tid & 63) == 0) s_ke[tid >> 6] = ke;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; w++) t += s_ke[w];
    C.tpart[blockIdx.x] = t;
  }
}

// the tiles k_collide_lc found above its capacity, one lane per cell (one
// workgroup, so it can reset the list for the next step when done)
template <typename Real>
__global__ __launch_bounds__(kLcThreads) void k_collide_lc_over(Ctx C, State<Real> S, const TileDev *tiles, Draws Dr,
                                                                int32_t *over) {
  constexpr int NT = kLcThreads;
  __shared__ double s_ke[NT / 64];
  Dr.resolve();
  const int tid = threadIdx.x, nover = over[0];
  const TanDev T0{};
  for (int j = 0; j < nover; j++) {
    const int tile = over[1 + j];
    const TileDev Tl = tiles[tile];
    double ke = 0.0;
    for (int k = tid; k < Tl.c1 - Tl.c0; k += NT) ke += collide_cell<Real, double>(C, S, Tl.rep, Tl.c0 + k, Dr, T0);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ke += __shfl_xor(ke, o, 64);
    if ((tid & 63) == 0) s_ke[tid >> 6] = ke;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < NT / 64; w++) t += s_ke[w];
      C.tpart[tile] = t;
    }
    __syncthreads();
  }
  if (tid == 0) over[0] = 0;
}

// ---------------------------------------------------------------------------
// 11. thermostat velocity scaling (scaleT, therm.c:52-143); the factor comes
//     from k_temp_reduce
// ---------------------------------------------------------------------------
// one particle's term of TEMP (therm.c:209-235): m |v - AVV|^2 over all three
// components, in the reference's order
__device__ __forceinline__ double temp_term(double m, const double v[3], const double *avv) {
  double tmp = (v[0] - avv[0]) * (v[0] - avv[0]);
  tmp += (v[1] - avv[1]) * (v[1] - avv[1]);
  tmp += (v[2] - avv[2]) * (v[2] - avv[2]);
  return m * tmp;
}

// the sum of x over a workgroup of kBlock threads, written by thread 0 to out
__device__ __forceinline__ void block_sum_to(double x, double *out) {
  __shared__ double s_w[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; w++) t += s_w[w];
    *out = t;
  }
}

// The thermostat's velocity scaling (scaleT, therm.c:52-143), with the
// factor from k_temp_reduce, and the block's term of the temperature the
// reference recomputes after it (mpc.c:4760, 4773) into C.tpart.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_scale(Ctx C, State<Real> S, const BlkDev *blk) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  const RepDev &R = C.rep[B.rep];
  double ke = 0.0;
  if (t < B.end) {
    const double g = C.gscale[B.rep];
    double v[3] = {(double)S.vx[t], (double)S.vy[t], (double)S.vz[t]};
    if (g != 1.0) {
      for (int d = 0; d < R.dim; d++) v[d] = (double)(Real)(v[d] * g);
      S.vx[t] = (Real)v[0];
      S.vy[t] = (Real)v[1];
      if (R.dim == 3) S.vz[t] = (Real)v[2];
    }
    ke = temp_term(C.sp[S.sp[t]].mass, v, R.avv);
  }
  block_sum_to(ke, &C.tpart[blockIdx.x]);
}

// KBTNOW after the velocity scaling: the sum of k_scale's (or
// k_stream_next's) block terms of each replica, blocks [rg.x, rg.y)
__global__ __launch_bounds__(kBlock) void k_temp_final(Ctx C, const int2 *ranges) {
  const int r = blockIdx.x;
  const RepDev &R = C.rep[r];
  const int2 rg = ranges[r];
  double acc = 0.0;
  for (int b = rg.x + threadIdx.x; b < rg.y; b += kBlock) acc += C.tpart[b];
  __shared__ double s_out;
  block_sum_to(acc, &s_out);
  if (threadIdx.x == 0) C.kbtnow[r] = s_out / ((double)(R.npart * R.dim));
}

// ---------------------------------------------------------------------------
// 11b. the periodic Galilean transformation to the rest frame (galileantrans,
//      mtools.c:673-718, with zeroExtraDims, mtools.c:727-744): the
//      mass-weighted mean velocity less AVV is taken off every particle
// ---------------------------------------------------------------------------
// per slot block: the sums of m and m v[d], d < 3, to part[4 * block]
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_netmom_part(Ctx C, State<Real> S, const BlkDev *blk, double *part) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  double x[4] = {0.0, 0.0, 0.0, 0.0};
  if (t < B.end) {
    const double m = C.sp[S.sp[t]].mass;
    x[0] = m;
    x[1] = m * (double)S.vx[t];
    x[2] = m * (double)S.vy[t];
    if (C.rep[B.rep].dim == 3) x[3] = m * (double)S.vz[t];
  }
  __shared__ double s_w[4][kBlock / 64];
#pragma unroll
  for (int k = 0; k < 4; k++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x[k] += __shfl_xor(x[k], o, 64);
    if ((threadIdx.x & 63) == 0) s_w[k][threadIdx.x >> 6] = x[k];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; w++) s += s_w[threadIdx.x][w];
    part[4 * blockIdx.x + threadIdx.x] = s;
  }
}

// per replica (one workgroup each): NET[d] = sum m v[d] / sum m - AVV[d]
// for d < DIM, to net[4 * r + d]
__global__ __launch_bounds__(kBlock) void k_netmom_final(Ctx C, const int2 *ranges, const double *part, double *net) {
  const int r = blockIdx.x;
  const RepDev &R = C.rep[r];
  const int2 rg = ranges[r];
  double x[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = rg.x + threadIdx.x; b < rg.y; b += kBlock)
    for (int k = 0; k < 4; k++) x[k] += part[4 * b + k];
  __shared__ double s_w[4][kBlock / 64];
#pragma unroll
  for (int k = 0; k < 4; k++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x[k] += __shfl_xor(x[k], o, 64);
    if ((threadIdx.x & 63) == 0) s_w[k][threadIdx.x >> 6] = x[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s[4];
    for (int k = 0; k < 4; k++) {
      s[k] = 0.0;
      for (int w = 0; w < kBlock / 64; w++) s[k] += s_w[k][w];
    }
    for (int d = 0; d < 3; d++) net[4 * r + d] = (d < R.dim) ? s[1 + d] / s[0] - R.avv[d] : 0.0;
  }
}

// v[d] -= NET[d] for d < DIM; in 2D the third components of Q and V are set
// to zero (zeroExtraDims)
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_netmom_apply(Ctx C, State<Real> S, const BlkDev *blk, const double *net) {
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t >= B.end) return;
  const double *n = net + 4 * B.rep;
  S.vx[t] = (Real)((double)S.vx[t] - n[0]);
  S.vy[t] = (Real)((double)S.vy[t] - n[1]);
  if (C.rep[B.rep].dim == 3) {
    S.vz[t] = (Real)((double)S.vz[t] - n[2]);
  } else {
    S.z[t] = (Real)0;
    S.vz[t] = (Real)0;
  }
}

// ---------------------------------------------------------------------------
// 12. streaming and planar walls (stream_all mpc.c:680; MPC_BCcollision
//     bc.c:1363-1496), then the unshifted cell of every particle
// ---------------------------------------------------------------------------
template <typename Real, typename F>
__device__ __forceinline__ void stream_particle(const Ctx &C, const State<Real> &S, int rep, int t, const Draws &Dr,
                                                float *wmul, const TanDev &T, const double *boxes) {
  using A = TA<F>;
  const RepDev &R = C.rep[rep];
  const int dim = R.dim;
  F Q[3], V[3] = {A::ld(S.vx, T.v[0], t), A::ld(S.vy, T.v[1], t), A::ld(S.vz, T.v[2], t)};
  F U[3] = {0., 0., 0.};
  load_qt(S, T, t, Q);
  const bool lc = R.lc != MPCD_ISOF;
  if (lc) { U[0] = A::ld(S.ux, T.u[0], t); U[1] = A::ld(S.uy, T.u[1], t); U[2] = A::ld(S.uz, T.u[2], t); }
  const int spid = S.sp[t];
  const mpcd_gpu_species &sp = C.sp[spid];
  for (int i = 0; i < dim; i++) Q[i] = Q[i] + R.dt * V[i];
  // a particle strictly inside the step's wall-free box crosses no wall: the
  // search below would end at once
  const double *box = &boxes[rep * 6];
  bool inside = true;
  for (int i = 0; i < dim; i++) inside = inside && (Q[i] > box[2 * i]) && (Q[i] < box[2 * i + 1]);
  // MPC_BCcollision
  F t_step = R.dt, t_left = t_step, t_coll = 0.0;
  int chosen = 0, cnt = 0, cnt2 = 0, flagCNT = 0, rk = 0;
  F n[3] = {0., 0., 0.};
  // adjoint tape: the product of the velocity factors of the bounces (MVN on
  // the wall's axis, MVT on the others); NaN when the velocity is redrawn
  float mul[3] = {1.f, 1.f, 1.f};
  while (!inside) {
    F tmin = t_left, chW = 1.0;
    for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
      const mpcd_gpu_wall &W = C.wall[w];
      if (!((W.inter_mask >> spid) & 1)) continue;
      const double *wq = &C.wq_b[w * 3];
      F tw = wall_W(W, wq, Q, dim);
      if (tw < 0.0) {
        F tc = W.R;
        for (int i = 0; i < dim; i++) tc += W.A[i] * (wq[i] - Q[i]);
        tc /= (W.A[0] * (-V[0]) + W.A[1] * (-V[1]) + W.A[2] * (-V[2]));
        F t1 = t_left - tc;
        if (t1 < tmin) { tmin = t1; chosen = w; chW = tw; }
      }
    }
    t_coll = tmin;
    if (chW > -kTOL) break;
    cnt++;
    const mpcd_gpu_wall &W = C.wall[chosen];
    const double *wq = &C.wq_b[chosen * 3];
    if (t_coll > -kTOL) {
      for (int i = 0; i < dim; i++) Q[i] = Q[i] - (t_left - t_coll) * V[i];
      for (int i = 0; i < dim; i++) n[i] = W.A[i];
      normn(n, dim);
      F Rr[3] = {0., 0., 0.}, VR[3], Vr[3] = {0., 0., 0.}, VN[3] = {0., 0., 0.}, VT[3] = {0., 0., 0.};
      F WL[3] = {W.L[0], W.L[1], W.L[2]};
      for (int i = 0; i < dim; i++) Rr[i] = Q[i] - wq[i];
      cross3(WL, Rr, VR);
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
      for (int i = 0; i < 3; i++) mul[i] *= (float)((W.A[i] != 0.0) ? W.MVN : W.MVT);
      F PN[3] = {0., 0., 0.}, PT[3] = {0., 0., 0.};
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
        mul[0] = mul[1] = mul[2] = __builtin_nanf("");
        if (Dr.reth_keyed() && rk + dim > kRethStride) {
          atomicOr(&C.flags[0], kFlagPushes);
          break;
        }
        const uint32_t pid = (uint32_t)S.id[t];
        double Vn[3] = {0.0, 0.0, 0.0};
        push_velocity(sp, R.kbt, dim, Dr, rep, (size_t)R.part_off + pid, pid, rk, Vn);
        for (int j = 0; j < dim; j++) V[j] = Vn[j];
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
  // the unshifted bin (mpc.c:4936): escaped particles are re-placed and stay
  // in the list they were in (mpc.c:970-982)
  int cu = cell_of(R, Q, 0);
  if (cu < 0) {
    atomicAdd(&C.events[0], 1ull);
    mul[0] = mul[1] = mul[2] = __builtin_nanf("");
    const uint32_t pid = (uint32_t)S.id[t];
    double q[3] = {val(Q[0]), val(Q[1]), val(Q[2])}, v[3], u[3] = {val(U[0]), val(U[1]), val(U[2])};
    if (!escape_repair(R, C.wh, 1, Dr, rep, pid, sp.mass, lc, q, v, u, C.wall, C.wq_b))
      atomicOr(&C.flags[0], kFlagEscape);
    for (int j = 0; j < 3; j++) {
      Q[j] = q[j];
      if (j < dim) V[j] = v[j];
      U[j] = u[j];
    }
    cu = S.cst[t];
  }
  A::st(S.x, T.q[0], t, Q[0]);
  A::st(S.y, T.q[1], t, Q[1]);
  A::st(S.z, T.q[2], t, Q[2]);
  A::st(S.vx, T.v[0], t, V[0]);
  A::st(S.vy, T.v[1], t, V[1]);
  A::st(S.vz, T.v[2], t, V[2]);
  if (lc) {
    A::st(S.ux, T.u[0], t, U[0]);
    A::st(S.uy, T.u[1], t, U[1]);
    A::st(S.uz, T.u[2], t, U[2]);
  }
  C.cu[t] = cu;
  if (wmul)
    for (int i = 0; i < 3; i++) wmul[3 * t + i] = mul[i];
}

// the general wall engine (mpcd_walls.h) on a drifted particle (Q after the
// drift; pid: its id, the key of its re-thermalisation draws):
// MPC_BCcollision (bc.c:1363-1496) with chooseBC over every wall the species
// meets (bc.c:1252-1328: planes by calcW_PLANE and the plane crossing time,
// wavy planes and curved surfaces by calcW, crosstimeReverse and chooseT, the
// curved ones at their periodic image nearest the particle), the bounce at
// the surface normal of the crossing point, and the NBOUNCE rewinds and
// re-thermalisations of the planar path.  FP64 arithmetic; the derivative
// modes are refused for contexts with such walls.
__device__ __forceinline__ void general_wall_loop(const Ctx &C, const RepDev &R, int rep, int spid, uint32_t pid,
                                                  const Draws &Dr, bool lc, double Q[3], double V[3], double U[3]) {
  const int dim = R.dim;
  const mpcd_gpu_species &sp = C.sp[spid];
  const double t_step = R.dt;
  double t_left = t_step, t_coll = 0.0;
  int chosen = 0, cnt = 0, cnt2 = 0, flagCNT = 0, rk = 0;
  for (;;) {
    // chooseBC: the earliest crossing among the walls the particle is inside
    double tmin = t_left, chW = 1.0;
    for (int w = R.wall_off; w < R.wall_off + R.nwall; w++) {
      const mpcd_gpu_wall &W = C.wall[w];
      if (!((W.inter_mask >> spid) & 1)) continue;
      const double *wq = &C.wq_b[w * 3];
      double wqi[3], tw, t1, t2, tc;
      if (wall_is_plane(W)) {
        tw = calc_w_plane(W, wq, Q, dim);
        if (tw < 0.0) {
          crosstime_reverse(W, true, wq, Q, V, dim, t_left, &t1, &t2);
          tc = t_left - t1;
          if (tc < tmin) { tmin = tc; chosen = w; chW = tw; }
        }
        continue;
      }
      const bool wavy_plane = !feq(W.B[0], 0.0) && feq(W.P[0], 1.0) && feq(W.P[1], 1.0) && feq(W.P[2], 1.0);
      if (wavy_plane) for (int k = 0; k < 3; k++) wqi[k] = wq[k];
      else wall_image(W, R.L, dim, wq, Q, wqi);
      tw = calc_w(W, wqi, Q, dim);
      if (tw < 0.0) {
        crosstime_reverse(W, false, wqi, Q, V, dim, t_left, &t1, &t2);
        tc = choose_t(t_left, t_left - t1, t_left - t2);
        if (tc < tmin) { tmin = tc; chosen = w; chW = tw; }
      }
    }
    t_coll = tmin;
    if (chW > -kTOL) break;
    cnt++;
    const mpcd_gpu_wall &W = C.wall[chosen];
    double wq[3];
    wall_image(W, R.L, dim, &C.wq_b[chosen * 3], Q, wq);
    if (t_coll > -kTOL) {
      double n[3] = {0., 0., 0.};
      for (int i = 0; i < dim; i++) Q[i] = Q[i] - (t_left - t_coll) * V[i];  // rewind_P
      wall_normal(W, wq, Q, dim, n);
      normn(n, dim);
      // velBC, rule-based collisions (bc.c:598-672)
      double Rr[3] = {0., 0., 0.}, VR[3], Vr[3] = {0., 0., 0.}, VN[3] = {0., 0., 0.}, VT[3] = {0., 0., 0.};
      const double WL[3] = {W.L[0], W.L[1], W.L[2]};
      for (int i = 0; i < dim; i++) Rr[i] = Q[i] - wq[i];
      cross3(WL, Rr, VR);
      for (int i = 0; i < dim; i++) Vr[i] = V[i] - W.V[i] - VR[i];
      projn(Vr, n, VN, dim);
      for (int i = 0; i < dim; i++) VT[i] = Vr[i] - VN[i];
      const double IMpart = 1.0 / sp.mass;
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
      // posBC (bc.c:760-774)
      double PN[3] = {0., 0., 0.}, PT[3] = {0., 0., 0.};
      projn(Q, n, PN, dim);
      for (int i = 0; i < dim; i++) PT[i] = Q[i] - PN[i];
      for (int i = 0; i < dim; i++) {
        PN[i] += W.DN * n[i] * (1. - 2. * kTOL);
        PT[i] += W.DT;
      }
      for (int i = 0; i < dim; i++) Q[i] = PN[i] + PT[i];
      if (lc) ori_bc(U, W, n, dim);
      t_left -= t_coll;
      for (int i = 0; i < dim; i++) Q[i] = Q[i] + t_left * V[i];  // stream_P
      if (calc_w(W, wq, Q, dim) < -kTOL)
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
        if (Dr.reth_keyed() && rk + dim > kRethStride) {
          atomicOr(&C.flags[0], kFlagPushes);
          break;
        }
        push_velocity(sp, R.kbt, dim, Dr, rep, (size_t)R.part_off + pid, pid, rk, V);
        rk += 3;
        t_left = t_step;
        t_coll = 0.0;
        cnt2++;
        cnt = 0;
      }
      if (cnt2 > 1000) {  // walls that leave the particle no room: the reference would not return
        atomicOr(&C.flags[0], kFlagStuck);
        break;
      }
    }
  }
  if (flagCNT) atomicAdd(&C.events[1], 1ull);
}

// streaming with the general wall engine, the unshifted bin and its escape
// repair, for one slot of the split path
template <typename Real>
__device__ __forceinline__ void stream_particle_walls(const Ctx &C, const State<Real> &S, int rep, int t,
                                                      const Draws &Dr) {
  const RepDev &R = C.rep[rep];
  const int dim = R.dim;
  double Q[3], V[3] = {(double)S.vx[t], (double)S.vy[t], (double)S.vz[t]}, U[3] = {0., 0., 0.};
  load_q(S, t, Q);
  const bool lc = R.lc != MPCD_ISOF;
  if (lc) { U[0] = (double)S.ux[t]; U[1] = (double)S.uy[t]; U[2] = (double)S.uz[t]; }
  const int spid = S.sp[t];
  const mpcd_gpu_species &sp = C.sp[spid];
  for (int i = 0; i < dim; i++) Q[i] = Q[i] + R.dt * V[i];
  general_wall_loop(C, R, rep, spid, (uint32_t)S.id[t], Dr, lc, Q, V, U);
  // the unshifted bin (mpc.c:4936) with bin()'s escape repair (mpc.c:970-982)
  int cu = cell_of(R, Q, 0);
  if (cu < 0) {
    atomicAdd(&C.events[0], 1ull);
    const uint32_t pid = (uint32_t)S.id[t];
    if (!escape_repair(R, C.wh, 1, Dr, rep, pid, sp.mass, lc, Q, V, U, C.wall, C.wq_b))
      atomicOr(&C.flags[0], kFlagEscape);
    cu = S.cst[t];
  }
  S.x[t] = (Real)Q[0];
  S.y[t] = (Real)Q[1];
  S.z[t] = (Real)Q[2];
  S.vx[t] = (Real)V[0];
  S.vy[t] = (Real)V[1];
  S.vz[t] = (Real)V[2];
  if (lc) {
    S.ux[t] = (Real)U[0];
    S.uy[t] = (Real)U[1];
    S.uz[t] = (Real)U[2];
  }
  C.cu[t] = cu;
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_stream_walls(Ctx C, State<Real> S, const BlkDev *blk, Draws Dr) {
  Dr.resolve();
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t < B.end) stream_particle_walls<Real>(C, S, B.rep, t, Dr);
}

template <typename Real, typename F>
__global__ __launch_bounds__(kBlock) void k_stream(Ctx C, State<Real> S, const BlkDev *blk, Draws Dr, float *wmul, TanDev T,
                         const double *boxes) {
  Dr.resolve();
  const BlkDev B = blk[blockIdx.x];
  const int t = B.beg + threadIdx.x;
  if (t < B.end) stream_particle<Real, F>(C, S, B.rep, t, Dr, wmul, T, boxes);
}

// ---------------------------------------------------------------------------
// output helpers
// ---------------------------------------------------------------------------
__global__ void k_order_ids(const uint64_t *skey, const int32_t *id, int32_t *out, int beg, int n) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  out[t] = id[(int)(skey[beg + t] & 0x7FFFFFFFull)];
}

template <typename Real>
__global__ void k_unsort(State<Real> S, int beg, int n, double *Q, double *V, double *U, int lc) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int s = beg + t;
  int i = S.id[s];
  Q[i * 3 + 0] = (double)S.x[s];
  Q[i * 3 + 1] = (double)S.y[s];
  Q[i * 3 + 2] = (double)S.z[s];
  V[i * 3 + 0] = (double)S.vx[s];
  V[i * 3 + 1] = (double)S.vy[s];
  V[i * 3 + 2] = (double)S.vz[s];
  if (lc && U) {
    U[i * 3 + 0] = (double)S.ux[s];
    U[i * 3 + 1] = (double)S.uy[s];
    U[i * 3 + 2] = (double)S.uz[s];
  }
}

template <typename Real>
__global__ void k_stage(State<Real> S, int beg, int n, const double *Q, const double *V, const double *U,
                        const int32_t *spid) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  int s = beg + t;
  S.x[s] = (Real)Q[t * 3 + 0];
  S.y[s] = (Real)Q[t * 3 + 1];
  S.z[s] = (Real)Q[t * 3 + 2];
  S.vx[s] = (Real)V[t * 3 + 0];
  S.vy[s] = (Real)V[t * 3 + 1];
  S.vz[s] = (Real)V[t * 3 + 2];
  if (U && S.ux) {
    S.ux[s] = (Real)U[t * 3 + 0];
    S.uy[s] = (Real)U[t * 3 + 1];
    S.uz[s] = (Real)U[t * 3 + 2];
  }
  S.id[s] = t;
  S.sp[s] = spid ? spid[t] : 0;
  S.cst[s] = -1;
}

}  // namespace mpcd
