// mpcd_exact.h -- the exact-sum FP32 form of the isotropic step (precision
// MPCD_FP32_EXACT): FP32 particle state, stepped without the reference's
// linked-list order.
//
// Every per-cell sum is built from quantised terms, integer multiples of
// 2^-28 held in FP64.  While a sum stays below 2^25 in magnitude, each FP64
// addition of such terms is exact, so a cell sum is the same number in
// whatever order its particles are added.  A step is therefore one function
// of the particle state, whatever the layout:
//   * the tiles of the bucketed layout add their particles into per-cell LDS
//     accumulators with atomics: no in-cell rank, no 64-bit list keys, and a
//     particle record of 28 bytes (x y z vx | vy vz | id);
//   * the dense form of the same step (the fallback after a bucket overflow,
//     single steps, runs with the bucketed layout switched off) adds them
//     with global atomics and gives the same bits.
// What the step computes, per replica:
//   * cell sums: particle count, per-species counts (mass = sum of count x
//     mass in species order), and m v (MPC-AT: also m v_rand; MPC-AT+a also
//     m r, m r r and r x m (v - v_rand), with r relative to the cell's
//     corner) quantised to 2^-28; from them, in FP64, the centre-of-mass
//     velocity, the ghost particles of phantom walls (ghostPart,
//     mpc.c:351-567), the SRD rotation matrix (rotate, mpc.c:1181-1257) or
//     the MPC-AT(+a) terms (andersenMPC/andersenROT, mpc.c:1353-1657: the
//     random-velocity sum, the inertia about the centre of mass and the
//     angular velocity I^-1 L), rounded to FP32 for the particle phase;
//   * particle arithmetic in FP32: the grid shift (gridShift_all,
//     mpc.c:720-766), the body force (acc_all, mpc.c:697), the collision,
//     the shift back, the drift (stream_all, mpc.c:680) and the next shifted
//     cell.  A drifted particle strictly inside the step's wall-free box
//     crosses no wall.  Outside it, a replica whose walls are exactly the
//     periodic faces of its box wraps the position by the box length; any
//     other replica runs the reference's FP64 bounce loop (MPC_BCcollision,
//     bc.c:1363-1496) from the pre-drift state and rounds the result;
//   * draws: the FP64 path's generator -- the SRD cell axis keyed by cell,
//     the MPC-AT thermal normals keyed by particle id (the FP64 path keys
//     them by list slot), escapes keyed by id;
//   * kinetic energy m |v - AVV|^2 per particle, quantised to 2^-28 and
//     summed exactly in int64 per replica (TEMP, therm.c:209-235).
// Escapes of the shifted bin (bin(shifted=1), mpc.c:970-982) are re-placed
// with the reference's escape repair and then binned in the cell of their
// new position (the reference keeps them in their old list for that step).
// Thermostats, nematic modes and injected draws are outside this form.
//
// A bucketed step is one launch of k_xbucket, plus k_xwalls when a replica
// has walls other than its box's periodic faces.  The step kernel also
// clears the bucket counts of the step after next (three count arrays turn
// over) and, in workgroup 0, writes the next step's grid shift and wall
// tables; every workgroup draws the next shift of its replica itself.
#pragma once

#include "mpcd_bucket.h"

namespace mpcd {

constexpr float kXQ = 268435456.0f;  // 2^28: the quantum of the cell sums is 2^-28
constexpr double kXQd = 268435456.0;
constexpr double kXQinv = 1.0 / 268435456.0;

// quantised terms (integer-valued FP64)
__device__ __forceinline__ double xq(float x) { return (double)rintf(x * kXQ); }
__device__ __forceinline__ double xqd(double x) { return rint(x * kXQd); }

// number of quantised sums per cell: m v [3]; MPC-AT adds m v_rand [3];
// MPC-AT+a adds m r [3], m r r [6] (xx yy zz xy xz yz) and r x m (v - v_rand) [3]
__host__ __device__ constexpr int xnsum(int op) { return op == OP_SRD ? 3 : op == OP_AT ? 6 : 18; }

// The bucketed records of the exact form: x y z vx | vy vz | id (28 B), plus
// the species when there are several.  Bit 31 of the id marks a record that
// the preparation kernel re-placed in this step's shifted frame.
struct XState {
  float4 *a;
  float2 *b;
  int32_t *id;
  int32_t *sp;
  int32_t *cnt;
};

// records waiting for the wall kernel: the pre-drift state of particles
// that leave the wall-free box (m.w = 0), and streamed particles whose next
// shifted bin they escape (m.w = 1)
struct XList {
  float4 *a;
  float2 *b;
  int4 *m;      // id, species, replica, kind
  int32_t *cnt;  // [1]
  int cap;
};

// what the step kernel writes for the next step: its grid shift and wall
// tables (shift_replica, from this step's shifted-back walls), and the
// cleared count of the next step's wall list
struct XNext {
  Draws D1;  // the next step's draws
  const double *wq_prev_b;
  double *shift, *wqs, *wqb, *box;
  WallHot *wh;
  int32_t *wl_cnt;
};

__device__ __forceinline__ void xrec_load(const XState &B, size_t o, bool onesp, float q[3], float v[3], int &id,
                                          int &sp) {
  const float4 a = B.a[o];
  const float2 b = B.b[o];
  q[0] = a.x;
  q[1] = a.y;
  q[2] = a.z;
  v[0] = a.w;
  v[1] = b.x;
  v[2] = b.y;
  id = B.id[o];
  sp = onesp ? 0 : B.sp[o];
}

__device__ __forceinline__ void xrec_store(const XState &B, size_t o, bool onesp, const float q[3], const float v[3],
                                           int id, int sp) {
  B.a[o] = make_float4(q[0], q[1], q[2], v[0]);
  B.b[o] = make_float2(v[1], v[2]);
  B.id[o] = id;
  if (!onesp) B.sp[o] = sp;
}

// appends a record to a list; false when the list is full
__device__ __forceinline__ bool xlist_push(const XList &L, const float q[3], const float v[3], int id, int sp,
                                           int rep, int kind) {
  const int e = atomicAdd(L.cnt, 1);
  if (e >= L.cap) return false;
  L.a[e] = make_float4(q[0], q[1], q[2], v[0]);
  L.b[e] = make_float2(v[1], v[2]);
  L.m[e] = make_int4(id, sp, rep, kind);
  return true;
}

// FP32 forms of the grid shift and the cell index (gridShift_all and bin,
// mpc.c:720-766, 948-968)
template <int DIM>
__device__ __forceinline__ void xshift_fwd(const float L[3], const float s[3], float q[3]) {
#pragma unroll
  for (int j = 0; j < DIM; j++) {
    const float t = q[j] + s[j];
    q[j] = ((t > L[j]) & (t <= L[j] + s[j])) ? t - L[j] : t;
  }
}

template <int DIM>
__device__ __forceinline__ void xshift_back(const float L[3], const float s[3], float q[3]) {
#pragma unroll
  for (int j = 0; j < DIM; j++) {
    const float sg = -s[j];
    const float t = q[j] + sg;
    q[j] = ((t < 0.0f) & (t >= sg)) ? t + L[j] : t;
  }
}

template <int DIM>
__device__ __forceinline__ int xcell(const int Li[3], const int pbc[3], int cell_off, int ny, int nz, const float q[3],
                                     bool shifted) {
  int a = (int)q[0], b = (int)q[1], c = (DIM == 3) ? (int)q[2] : 0;
  const int Lz = (DIM == 3) ? Li[2] : 1;
  if (shifted) {
    a = (pbc[0] & (a == Li[0])) ? 0 : a;
    b = (pbc[1] & (b == Li[1])) ? 0 : b;
    if (DIM == 3) c = (pbc[2] & (c == Lz)) ? 0 : c;
  }
  // (int) truncates toward zero, as the reference's cast does
  bool bad = (a > Li[0]) | (b > Li[1]) | (a < 0) | (b < 0);
  if (DIM == 3) bad = bad | (c > Lz) | (c < 0);
  return bad ? -1 : cell_off + (a * ny + b) * nz + c;
}

// the escape repair of a record whose coming shifted bin it escapes (bin,
// mpc.c:970-982; replacePos_WithCheck, init.c:1462) with the draws and wall
// positions of that coming step (Cn, Dn): the re-placed position is in that
// step's shifted frame.  Returns its shifted cell.
__device__ __forceinline__ int xrepair(const Ctx &Cn, const RepDev &R, const Draws &Dn, int rep, int pid, double mass,
                                       float q[3], float v[3]) {
  double qd[3] = {(double)q[0], (double)q[1], (double)q[2]}, vd[3] = {(double)v[0], (double)v[1], (double)v[2]};
  double u0[3] = {0.0, 0.0, 0.0};
  atomicAdd(&Cn.events[0], 1ull);
  if (!escape_repair(R, Cn.wh, 0, Dn, rep, (uint32_t)pid, mass, false, qd, vd, u0, Cn.wall, Cn.wq_s))
    atomicOr(&Cn.flags[0], kFlagEscape);
  for (int d = 0; d < 3; d++) {
    q[d] = (float)qd[d];
    v[d] = (float)vd[d];
  }
  int cn = (R.dim == 3) ? xcell<3>(R.L, R.pbc, R.cell_off, R.ny, R.nz, q, true)
                        : xcell<2>(R.L, R.pbc, R.cell_off, R.ny, R.nz, q, true);
  if (cn < 0) {
    // the FP32 rounding of a re-placed position can reach the box's edge
    atomicOr(&Cn.flags[0], kFlagEscape);
    cn = R.cell_off;
  }
  return cn;
}

// the wall-free box in FP32, rounded inwards, and cut to the positions
// whose unshifted cell exists (a truncated coordinate in [0, L], so the
// position in (-1, L + 1)): a drifted particle outside it takes the FP64
// path, which bounces it off the walls and re-places it if it left the box
// (the escape repair of bin(shifted=0), mpc.c:970-982)
__device__ __forceinline__ void xbox(const double *box, const float L[3], float out[6]) {
#pragma unroll
  for (int k = 0; k < 6; k++) {
    float f = (float)box[k];
    if (k & 1) {
      if ((double)f > box[k]) f = nextafterf(f, -INFINITY);
      f = fminf(f, L[k >> 1] + 1.0f);
    } else {
      if ((double)f < box[k]) f = nextafterf(f, INFINITY);
      f = fmaxf(f, -1.0f);
    }
    out[k] = f;
  }
}

// the FP64 path of a drift that left the FP32 box: the reference's bounce
// loop from the pre-drift state (planar walls; the general engine when a
// wall of the replica is curved or wavy), then the unshifted bin's escape
// repair
template <int DIM>
__device__ __forceinline__ void xwalls_fp64(const Ctx &C, const RepDev &R, const mpcd_gpu_species *__restrict__ sps,
                                            const mpcd_gpu_wall *__restrict__ walls,
                                            const WallHot *__restrict__ wh, const double *__restrict__ wq_b,
                                            const double *__restrict__ box, const Draws &Dr, int rep, int spid,
                                            int pid, double q[3], double v[3]) {
  double U0[3] = {0, 0, 0};
  if (!R.anycurved) {
    (void)stream_walls<DIM>(C, R, sps, walls, wh, wq_b, box, Dr, rep, spid, pid, 0, q, v, U0, false);
    return;
  }
  for (int d = 0; d < DIM; d++) q[d] = q[d] + R.dt * v[d];
  general_wall_loop(C, R, rep, spid, (uint32_t)pid, Dr, false, q, v, U0);
  if (cell_of(R, q, 0) < 0) {
    atomicAdd(&C.events[0], 1ull);
    if (!escape_repair(R, C.wh, 1, Dr, rep, (uint32_t)pid, sps[spid].mass, false, q, v, U0, C.wall, C.wq_b))
      atomicOr(&C.flags[0], kFlagEscape);
  }
}

// per replica: does its wall set consist of exactly the periodic faces of
// the box (planes at 0 and L with DN = L and the identity rule)?
inline bool walls_are_box_faces(const mpcd_gpu_params &p) {
  if (p.nwall != 2 * p.dim) return false;
  int seen[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  for (int w = 0; w < p.nwall; w++) {
    const mpcd_gpu_wall &W = p.wall[w];
    int ax = -1, nz = 0;
    for (int i = 0; i < 3; i++)
      if (W.A[i] != 0.0) {
        ax = i;
        nz++;
      }
    if (nz != 1 || ax >= p.dim) return false;
    const bool plane = (W.P[0] == 0.0 && W.P[1] == 0.0 && W.P[2] == 0.0 && W.P[3] == 0.0) ||
                       (W.P[0] == 1.0 && W.P[1] == 1.0 && W.P[2] == 1.0);
    if (!plane || W.ABS || W.B[0] != 0.0 || W.INV || W.PHANTOM || W.COLL_TYPE != 1) return false;
    if (W.MVN != 1.0 || W.MVT != 1.0 || W.DT != 0.0 || W.DVN != 0.0 || W.DVT != 0.0) return false;
    for (int i = 0; i < 3; i++)
      if (W.DVxyz[i] != 0.0 || W.V[i] != 0.0 || W.L[i] != 0.0) return false;
    if (W.DN != (double)p.L[ax] || W.Q[0] != 0.0 || W.Q[1] != 0.0 || W.Q[2] != 0.0) return false;
    const double L = (double)p.L[ax];
    if (W.A[ax] == 1.0 && W.R == 0.0) seen[ax][0]++;
    else if (W.A[ax] == -1.0 && W.R == -L) seen[ax][1]++;
    else return false;
  }
  for (int i = 0; i < p.dim; i++)
    if (seen[i][0] != 1 || seen[i][1] != 1) return false;
  return true;
}

// drift of a collided particle (q unshifted): true when the result is final
// (inside the wall-free box, or wrapped on a periodic box); false when the
// FP64 wall loop must take it from the pre-drift state
template <int DIM>
__device__ __forceinline__ bool xdrift(const float boxf[6], const float L[3], bool wrap, float dt, const float v[3],
                                       float q[3]) {
  float Q[3] = {q[0], q[1], q[2]};
  bool inside = true;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    Q[d] = Q[d] + dt * v[d];
    inside = inside & (Q[d] > boxf[2 * d]) & (Q[d] < boxf[2 * d + 1]);
  }
  if (!inside && !wrap) return false;
  if (!inside) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      const float t = Q[d];
      Q[d] = (t < 0.0f) ? t + L[d] : ((t >= L[d]) ? t - L[d] : t);
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; d++) q[d] = Q[d];
  return true;
}

// thermal velocity of MPC-AT(+a), keyed by particle id: the FP64 normals of
// the generator (sqrt(kT/m) N(0,1)), rounded to FP32
template <int DIM>
__device__ __forceinline__ void xthermal(const Draws &Dr, int rep, int pid, double sig, float rv[3]) {
  double n0[2], n1[2] = {0.0, 0.0};
  Dr.rng.normal_pair_turn(rep, P_NORMAL, (uint32_t)pid, 0, n0);
  if (DIM == 3) Dr.rng.normal_pair_turn(rep, P_NORMAL, (uint32_t)pid, 1, n1);
  rv[0] = (float)(sig * n0[0]);
  rv[1] = (float)(sig * n0[1]);
  rv[2] = (DIM == 3) ? (float)(sig * n1[0]) : 0.0f;
}

// the quantised terms of one particle (r relative to its cell's corner)
template <int OP, int DIM>
__device__ __forceinline__ void xterms(double m, const float q[3], const float v[3], const float rv[3], int a,
                                       int b, int c, double t[18]) {
  if (m == 1.0) {
#pragma unroll
    for (int d = 0; d < 3; d++) t[d] = xq(v[d]);
  } else {
#pragma unroll
    for (int d = 0; d < 3; d++) t[d] = xqd(m * (double)v[d]);
  }
  if (OP == OP_SRD) return;
#pragma unroll
  for (int d = 0; d < 3; d++) t[3 + d] = xqd(m * (double)rv[d]);
  if (OP != OP_RAT) return;
  // exact in FP32: the cell corner and the position have the same binade
  const float r[3] = {q[0] - (float)a, q[1] - (float)b, (DIM == 3) ? q[2] - (float)c : 0.0f};
  const double rd[3] = {(double)r[0], (double)r[1], (double)r[2]};
#pragma unroll
  for (int d = 0; d < 3; d++) t[6 + d] = xqd(m * rd[d]);
  t[9] = xqd(m * (rd[0] * rd[0]));
  t[10] = xqd(m * (rd[1] * rd[1]));
  t[11] = xqd(m * (rd[2] * rd[2]));
  t[12] = xqd(m * (rd[0] * rd[1]));
  t[13] = xqd(m * (rd[0] * rd[2]));
  t[14] = xqd(m * (rd[1] * rd[2]));
  const double dv[3] = {m * ((double)v[0] - (double)rv[0]), m * ((double)v[1] - (double)rv[1]),
                        m * ((double)v[2] - (double)rv[2])};
  t[15] = xqd(rd[1] * dv[2] - rd[2] * dv[1]);
  t[16] = xqd(rd[2] * dv[0] - rd[0] * dv[2]);
  t[17] = xqd(rd[0] * dv[1] - rd[1] * dv[0]);
}

// the collision data of one cell from its exact sums, rounded to FP32:
//   cd[0..2] centre-of-mass velocity after the body force's kick g dt (with
//   ghost particles at phantom walls); for SRD instead the offset
//   b = vcm - M (vcm - g dt) of v' = M v + b.  The particles' velocities enter
//   unkicked: the kick is added once per cell in FP64, not per particle in
//   FP32, whose rounding of a fixed small increment would be biased by the
//   binade of v;
//   SRD: cd[3..11] the rotation matrix;
//   MPC-AT: cd[3..5] the random-velocity sum RS = sum m v_rand / M;
//   MPC-AT+a: also cd[6..8] the angular velocity W and cd[9..11] the centre
//   of mass relative to the cell's corner.
// np_ particles of total mass `mass`; s the quantised sums (xnsum(OP)).
template <int OP, int DIM, bool GHOST>
__device__ __forceinline__ void xcell_terms(const RepDev &R, const WallHot *__restrict__ wh, const Draws &Dr, int rep,
                                            int c, int np_, double mass, const double *s, float cd[12]) {
  // one division per cell: the sums' quantum over the cell mass
  const double im = (mass > 0.0) ? kXQinv / mass : 0.0;
  double vcm[3] = {0.0, 0.0, 0.0}, dtg[3] = {0.0, 0.0, 0.0};
  if (R.gravflag)
    for (int d = 0; d < DIM; d++) dtg[d] = R.dt * R.grav[d];
  for (int d = 0; d < DIM; d++) vcm[d] = s[d] * im + dtg[d];
  if (GHOST && R.anyphantom && np_ > 0) {
    // ghostPart (mpc.c:351-567): a cell cut by a phantom wall and holding
    // fewer than GnDNST particles gains the ghosts' momentum
    const int lc_ = c - R.cell_off;
    const int a = lc_ / (R.ny * R.nz), bb = (lc_ / R.nz) % R.ny, cz = lc_ % R.nz;
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
      const double invN = 1.0 / (R.gndnst - (double)np_);
      for (int d = 0; d < DIM; d++) {
        const double sigma = sqrt(R.kbt / invN);
        vcm[d] += sigma * Dr.ghost_n(rep, c, d);
        vcm[d] /= R.gndnst;
      }
    }
  }
  for (int d = 0; d < 3; d++) cd[d] = (float)vcm[d];
  for (int d = 3; d < 12; d++) cd[d] = 0.0f;
  if (np_ <= 1) return;
  if (OP == OP_SRD) {
    double RV[3] = {0.0, 0.0, 0.0};
    int CA = 0;
    if (R.rtech == MPCD_ARBAXIS) {
      if (DIM == 3) Dr.rng.unitvec_turn(rep, P_CELL, c, 0, RV);
    } else {
      CA = (int)((int)DIM * Dr.rng.uniform(rep, P_CELL, c, 0));
    }
    // the sign of the angle (mpc.c:1291): a uniform stored in a long is 0,
    // so the reference always rotates by -RA
    const bool pos = R.true_sign ? (Dr.rng.uniform(rep, P_CELL, c, 3) >= 0.5) : false;
    const double Cc = R.cosA, Sn = pos ? R.sinA : -1. * R.sinA;
    double M[9];
    if (DIM == 3) {
      const double *L = RV;
      if (R.rtech == MPCD_ARBAXIS) {
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
    for (int k = 0; k < 9; k++) cd[3 + k] = (float)M[k];
    // v' = vcm + M (v + g dt - vcm) = M v + b with the unkicked v
    for (int d = 0; d < 3; d++) {
      double b = vcm[d];
      for (int k = 0; k < 3; k++) b -= M[3 * d + k] * (vcm[k] - dtg[k]);
      cd[d] = (float)b;
    }
    return;
  }
  double RS[3];
  for (int d = 0; d < 3; d++) RS[d] = (d < DIM) ? s[3 + d] * im : 0.0;
  for (int d = 0; d < 3; d++) cd[3 + d] = (float)RS[d];
  if (OP != OP_RAT) return;
  // centre of mass, inertia about it (mpc.c:4209-4299, off-diagonals negated)
  // and angular momentum about it (mpc.c:1554-1618), from the moments about
  // the cell's corner
  double cm[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < DIM; d++) cm[d] = s[6 + d] * im;
  const double Sxx = s[9] * kXQinv - mass * cm[0] * cm[0], Syy = s[10] * kXQinv - mass * cm[1] * cm[1];
  const double Szz = s[11] * kXQinv - mass * cm[2] * cm[2];
  const double Sxy = s[12] * kXQinv - mass * cm[0] * cm[1], Sxz = s[13] * kXQinv - mass * cm[0] * cm[2];
  const double Syz = s[14] * kXQinv - mass * cm[1] * cm[2];
  double I[6] = {Syy + Szz, Sxx + Szz, Sxx + Syy, -Sxy, -Sxz, -Syz};
  // sum m (v - v_rand) = sum m v - sum m v_rand
  const double P[3] = {(s[0] - s[3]) * kXQinv, (s[1] - s[4]) * kXQinv, (s[2] - s[5]) * kXQinv};
  double Lc[3] = {s[15] * kXQinv, s[16] * kXQinv, s[17] * kXQinv};
  Lc[0] -= cm[1] * P[2] - cm[2] * P[1];
  Lc[1] -= cm[2] * P[0] - cm[0] * P[2];
  Lc[2] -= cm[0] * P[1] - cm[1] * P[0];
  double II[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (DIM == 3) {
    if (np_ >= 3) {
      double Im[3][3] = {{I[0], I[3], I[4]}, {I[3], I[1], I[5]}, {I[4], I[5], I[2]}};
      inv3(II, Im);
    }
  } else {
    II[2][2] = 1. / I[2];
  }
  for (int i = 0; i < 3; i++) {
    double w = 0.;
    for (int j = 0; j < 3; j++) w += II[i][j] * Lc[j];
    cd[6 + i] = (float)w;
    cd[9 + i] = (float)cm[i];
  }
}

// the collision of one particle of a cell of np_ > 1 particles, in FP32;
// r: position relative to the cell's corner (MPC-AT+a)
template <int OP, int DIM>
__device__ __forceinline__ void xcollide(const float cd[12], int np_, float damp, const float rv[3], const float r[3],
                                         float v[3]) {
  const float vc[3] = {cd[0], cd[1], cd[2]};
  if (OP == OP_SRD) {
    // v' = M v + b (b in cd[0..2])
    const float *M = cd + 3;
    const float d0 = v[0], d1 = v[1], d2 = v[2];
    if (DIM == 3) {
      float t0 = M[0] * d0;
      t0 += M[1] * d1;
      t0 += M[2] * d2;
      float t1 = M[3] * d0;
      t1 += M[4] * d1;
      t1 += M[5] * d2;
      float t2 = M[6] * d0;
      t2 += M[7] * d1;
      t2 += M[8] * d2;
      v[0] = vc[0] + t0;
      v[1] = vc[1] + t1;
      v[2] = vc[2] + t2;
    } else {
      float t0 = M[0] * d0;
      t0 += M[1] * d1;
      float t1 = M[3] * d0;
      t1 += M[4] * d1;
      v[0] = vc[0] + t0;
      v[1] = vc[1] + t1;
    }
  } else if (OP == OP_AT) {
    const float inv = 1.0f / (float)np_;
#pragma unroll
    for (int d = 0; d < DIM; d++) v[d] = vc[d] + rv[d] - cd[3 + d] - damp * vc[d] * inv;
  } else {
    const float rr[3] = {r[0] - cd[9], r[1] - cd[10], (DIM == 3) ? r[2] - cd[11] : 0.0f};
    const float W[3] = {cd[6], cd[7], cd[8]};
    float at[3];
    cross3(W, rr, at);
#pragma unroll
    for (int d = 0; d < DIM; d++) v[d] = vc[d] + rv[d] - cd[3 + d] - damp * vc[d] + at[d];
  }
}

// kinetic-energy term m |v - AVV|^2 of TEMP (therm.c:209-235), quantised
__device__ __forceinline__ double xke(double m, const float avv[3], const float v[3]) {
  const float a = v[0] - avv[0], b = v[1] - avv[1], c = v[2] - avv[2];
  float e = a * a;
  e += b * b;
  e += c * c;
  return rint((double)e * m * kXQd);
}

// per-replica values of a tile kernel in FP32
struct XRep {
  float L[3], sc[3], box[6], dt, dtg[3], avv[3];
  int Li[3], pbc[3], cell_off, ny, nz;
  bool wrap;
};

__device__ __forceinline__ XRep xrep(const RepDev &R, const double *shc, const double *box) {
  XRep X;
  for (int j = 0; j < 3; j++) {
    X.L[j] = (float)R.L[j];
    X.sc[j] = (float)shc[j];
    X.dtg[j] = (float)(R.dt * R.grav[j]);
    X.avv[j] = (float)R.avv[j];
    X.Li[j] = R.L[j];
    X.pbc[j] = R.pbc[j];
  }
  xbox(box, X.L, X.box);
  X.dt = (float)R.dt;
  X.cell_off = R.cell_off;
  X.ny = R.ny;
  X.nz = R.nz;
  X.wrap = R.xwrap != 0;
  return X;
}

// ---------------------------------------------------------------------------
// the bucketed step
// ---------------------------------------------------------------------------
constexpr int kXMaxCells = kBktMaxCells;

// the bucketed step: one workgroup per bucket.  It also clears bucket count
// cnt_zero[blockIdx.x] of the step after next, and workgroup 0 writes the
// next step's tables (Nx) once its tile is done.  The SRD kernel asks for 5
// waves per SIMD (at most 96 VGPRs) and spreads its cell phase over the four
// waves (cell k on wave k mod 4); the MPC-AT kernels ask for 3 (at most 168
// VGPRs, which the two-species ghost variant would exceed) and keep cell k on
// lane k, since their heavier cell phase gains nothing from the spread.
template <int OP, bool ONESP, int DIM, bool GHOST>
__global__ __launch_bounds__(kBktThreads) __attribute__((amdgpu_waves_per_eu(OP == OP_SRD ? 5 : 3))) void k_xbucket(
    Ctx C, XState I, XState O, int32_t *__restrict__ cnt_zero, const TileDev *__restrict__ bkts,
    const BucketRep *__restrict__ brep, const RepDev *__restrict__ reps, const mpcd_gpu_species *__restrict__ sps,
    const WallHot *__restrict__ wh, const double *__restrict__ boxes, Draws Dr, const double *__restrict__ shift_cur,
    XNext Nx, int cap_lim, XList WL, int nspec) {
  constexpr int NT = kBktThreads, PPT = bkt_ppt(OP), CAP = NT * PPT, NS = xnsum(OP);
  constexpr int NSP = ONESP ? 1 : MPCD_MAXSPEC;
  const int tid = threadIdx.x;
  __shared__ int s_flag;
  __shared__ int32_t s_n[kXMaxCells];
  __shared__ int32_t s_nsp[ONESP ? 1 : kXMaxCells][NSP];
  __shared__ double s_sum[kXMaxCells][NS];
  __shared__ float s_cd[kXMaxCells][12];
  __shared__ int32_t s_hkey[kBktHash], s_hcnt[kBktHash], s_hbase[kBktHash];
  __shared__ double s_ke[NT / 64];
  __shared__ double s_shn[3];
  __shared__ int32_t s_id[CAP];  // particle ids, out of the registers during the cell phase

  // every record load of the thread first (a slot past the count reads the
  // last filled slot again; its values are not used)
  const size_t off = (size_t)blockIdx.x * CAP;
  const int n = min(I.cnt[blockIdx.x], CAP);
  const int last = max(n - 1, 0);
  float rq[PPT][3], rv[PPT][3];
  int rid[PPT], rsp[PPT], rk[PPT], ra[PPT];
#pragma unroll
  for (int j = 0; j < PPT; j++) xrec_load(I, off + min(tid + j * NT, last), ONESP, rq[j], rv[j], rid[j], rsp[j]);
  if (tid == 0) s_flag = C.flags[0];
  const TileDev T = bkts[blockIdx.x];
  const int ncl = T.c1 - T.c0;
  for (int k = tid; k < ncl; k += NT) {
    s_n[k] = 0;
#pragma unroll
    for (int j = 0; j < NS; j++) s_sum[k][j] = 0.0;
    if (!ONESP)
      for (int j = 0; j < NSP; j++) s_nsp[k][j] = 0;
  }
  for (int k = tid; k < kBktHash; k += NT) {
    s_hkey[k] = -1;
    s_hcnt[k] = 0;
  }
  __syncthreads();
  if (s_flag & kFlagPoison) return;
  if (tid == 0) cnt_zero[blockIdx.x] = 0;
  Dr.resolve();
  const RepDev &R = reps[T.rep];
  const BucketRep BR = brep[T.rep];
  const XRep X = xrep(R, &shift_cur[T.rep * 3], &boxes[T.rep * 6]);
  const double sig1 = (OP == OP_SRD) ? 0.0 : sqrt(R.kbt / sps[0].mass);
  float rt[PPT][OP == OP_SRD ? 1 : 3];  // thermal velocities

  // 1. grid shift, body force, cell, quantised sums
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    rk[j] = -1;
    if (tid + j * NT < n) {
      const bool repaired = rid[j] < 0;
      rid[j] &= 0x7FFFFFFF;
      s_id[tid + j * NT] = rid[j];
      if (!repaired) xshift_fwd<DIM>(X.L, X.sc, rq[j]);
      const int cn = xcell<DIM>(X.Li, X.pbc, X.cell_off, X.ny, X.nz, rq[j], true);
      const int k = cn - T.c0;
      if ((unsigned)k < (unsigned)ncl) {
        rk[j] = k;
        const int spid = ONESP ? 0 : rsp[j];
        const double m = ONESP ? sps[0].mass : sps[spid].mass;
        float rvv[3] = {0.0f, 0.0f, 0.0f};
        if (OP != OP_SRD) {
          xthermal<DIM>(Dr, T.rep, rid[j], ONESP ? sig1 : sqrt(R.kbt / m), rvv);
#pragma unroll
          for (int d = 0; d < 3; d++) rt[j][OP == OP_SRD ? 0 : d] = rvv[d];
        }
        const int lc_ = cn - X.cell_off;
        const int ca = lc_ / (X.ny * X.nz), cb = (lc_ / X.nz) % X.ny, cc = lc_ % X.nz;
        double t[18];
        xterms<OP, DIM>(ONESP ? 1.0 : m, rq[j], rv[j], rvv, ca, cb, cc, t);
        atomicAdd(&s_n[k], 1);
        if (!ONESP) atomicAdd(&s_nsp[k][spid], 1);
#pragma unroll
        for (int d = 0; d < NS; d++)
          if (DIM == 3 || (d % 3) != 2 || OP == OP_RAT) atomicAdd(&s_sum[k][d], t[d]);
      } else {
        atomicOr(&C.flags[0], kFlagBadCell);
      }
    }
  }
  __syncthreads();
  // 2. one lane per cell: the collision data; the last three lanes, which
  //    hold no cell in either cell order, draw the next step's grid shift
  //    of this replica as shift_replica draws it
  static_assert(kXMaxCells <= NT - 3, "the last lanes hold no cell");
  if (tid >= NT - 3) {
    const int j = NT - 1 - tid;
    Nx.D1.resolve();
    s_shn[j] = (j < DIM && R.galinv) ? Nx.D1.rng.uniform(T.rep, P_SHIFT, 0, j) : 0.0;
  }
  for (int k = (OP == OP_SRD) ? (((tid & 63) << 2) | (tid >> 6)) : tid; k < ncl; k += NT) {
    const int np_ = s_n[k];
    if (np_ == 0) continue;
    double mass;
    if (ONESP) {
      mass = (double)np_ * sps[0].mass;
    } else {
      mass = 0.0;
      for (int s = 0; s < nspec; s++) mass += (double)s_nsp[k][s] * sps[s].mass;
    }
    double sm[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) sm[j] = s_sum[k][j];
    if (ONESP && OP != OP_SRD) {
      // the terms were quantised without the common mass
#pragma unroll
      for (int j = 0; j < NS; j++) sm[j] *= sps[0].mass;
    } else if (ONESP) {
#pragma unroll
      for (int j = 0; j < 3; j++) sm[j] *= sps[0].mass;
    }
    xcell_terms<OP, DIM, GHOST>(R, wh, Dr, T.rep, T.c0 + k, np_, mass, sm, s_cd[k]);
  }
  __syncthreads();
  // 3. one lane per particle: collision, kinetic energy, shift back, drift,
  //    next shifted cell and bucket
  double ke = 0.0;
  int wl = 0;  // bit j: particle j takes the wall list; bit 8 + j: the escape list
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    if (rk[j] < 0) continue;
    const int k = rk[j];
    const int np_ = s_n[k];
    const int spid = ONESP ? 0 : rsp[j];
    const double m = ONESP ? sps[0].mass : sps[spid].mass;
    float *v = rv[j], *q = rq[j];
    if (np_ > 1) {
      const float4 c0 = *(const float4 *)&s_cd[k][0], c1 = *(const float4 *)&s_cd[k][4],
                   c2 = *(const float4 *)&s_cd[k][8];
      const float cd[12] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y, c2.z, c2.w};
      float rvv[3] = {0.0f, 0.0f, 0.0f}, r[3] = {0.0f, 0.0f, 0.0f};
      if (OP != OP_SRD)
#pragma unroll
        for (int d = 0; d < 3; d++) rvv[d] = rt[j][OP == OP_SRD ? 0 : d];
      if (OP == OP_RAT) {
        const int lc_ = T.c0 + k - X.cell_off;
        r[0] = q[0] - (float)(lc_ / (X.ny * X.nz));
        r[1] = q[1] - (float)((lc_ / X.nz) % X.ny);
        r[2] = (DIM == 3) ? q[2] - (float)(lc_ % X.nz) : 0.0f;
      }
      xcollide<OP, DIM>(cd, np_, (float)sps[spid].damp, rvv, r, v);
    } else if (R.gravflag) {
      // a particle alone in its cell takes the kick itself
#pragma unroll
      for (int d = 0; d < DIM; d++) v[d] = v[d] + X.dtg[d];
    }
    ke += xke(m, X.avv, v);
    xshift_back<DIM>(X.L, X.sc, q);
    const float q0[3] = {q[0], q[1], q[2]};
    if (!xdrift<DIM>(X.box, X.L, X.wrap, X.dt, v, q)) {
      // pre-drift state to the wall list
#pragma unroll
      for (int d = 0; d < 3; d++) q[d] = q0[d];
      wl |= 1 << j;
      rk[j] = -2;
      continue;
    }
    float qn[3] = {q[0], q[1], q[2]};
    const float sn[3] = {(float)s_shn[0], (float)s_shn[1], (float)s_shn[2]};
    xshift_fwd<DIM>(X.L, sn, qn);
    const int cn = xcell<DIM>(X.Li, X.pbc, X.cell_off, X.ny, X.nz, qn, true);
    if (cn < 0) {
      wl |= 1 << (8 + j);
      rk[j] = -3;
      continue;
    }
    const int b2 = bucket_of(BR, R, cn);
    rk[j] = b2;
    int h = (int)(((uint32_t)b2 * 2654435761u) >> 26);
    int hs = (s_hkey[h] == b2) ? h : -1;
    for (int t = 0; hs < 0 && t < kBktHash; t++) {
      int old = s_hkey[h];
      if (old == -1) old = atomicCAS(&s_hkey[h], -1, b2);
      if (old == -1 || old == b2) hs = h;
      else h = (h + 1) & (kBktHash - 1);
    }
    ra[j] = (hs >= 0) ? ((hs << 16) | atomicAdd(&s_hcnt[hs], 1)) : -1;
  }
  __syncthreads();
  // 4. one global atomic per destination bucket, then the writes
  for (int h = tid; h < kBktHash; h += NT)
    if (s_hcnt[h] > 0) s_hbase[h] = atomicAdd(&O.cnt[s_hkey[h]], s_hcnt[h]);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    const int b2 = rk[j];
    if (b2 == -2 || b2 == -3) {
      const int spid = ONESP ? 0 : rsp[j];
      if (!xlist_push(WL, rq[j], rv[j], s_id[tid + j * NT], spid, T.rep, b2 == -2 ? 0 : 1)) {
        atomicOr(&C.flags[0], kFlagPoison);
        atomicMin(&C.flags[1], (int)Dr.rng.step);
      }
      continue;
    }
    if (b2 < 0) continue;
    const int p2 = (ra[j] >= 0) ? s_hbase[ra[j] >> 16] + (ra[j] & 0xFFFF) : atomicAdd(&O.cnt[b2], 1);
    if (p2 >= cap_lim) {
      atomicOr(&C.flags[0], kFlagPoison);
      atomicMin(&C.flags[1], (int)Dr.rng.step);
      continue;
    }
    xrec_store(O, (size_t)b2 * CAP + p2, ONESP, rq[j], rv[j], s_id[tid + j * NT], ONESP ? 0 : rsp[j]);
  }
  (void)wl;
  // exact kinetic-energy sum of the tile (integer-valued FP64)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ke += __shfl_xor(ke, o, 64);
  if ((tid & 63) == 0) s_ke[tid >> 6] = ke;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; w++) t += s_ke[w];
    C.tpart[blockIdx.x] = t;
  }
  // workgroup 0, with the tile done (its registers free): the next step's
  // shift and wall tables, in the other parity slot
  if (blockIdx.x == 0) {
    Nx.D1.resolve();
    for (int r = tid; r < C.nrep; r += NT)
      shift_replica(C, Nx.D1, r, 1, Nx.wq_prev_b, Nx.shift, Nx.wqs, Nx.wqb, Nx.wh, Nx.box);
    if (tid == 0) *Nx.wl_cnt = 0;
  }
}

// the wall list of a bucketed step: the reference's FP64 bounce loop from
// the pre-drift state (kind 0), then the next shifted cell; a record whose
// next shifted bin it escapes (kind 1, or after the bounce loop) is re-placed
// with the next step's draws and wall positions (Cn, Dn) and appended
// already in that step's shifted frame
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_xwalls(Ctx C, XList WL, XState O, const BucketRep *__restrict__ brep,
                                                   const RepDev *__restrict__ reps,
                                                   const mpcd_gpu_species *__restrict__ sps,
                                                   const mpcd_gpu_wall *__restrict__ walls,
                                                   const WallHot *__restrict__ wh, const double *__restrict__ boxes,
                                                   Draws Dr, const double *__restrict__ wq_b, Ctx Cn, Draws Dn,
                                                   int cap, int cap_lim, int onesp) {
  if (C.flags[0] & kFlagPoison) return;
  const int n = min(*WL.cnt, WL.cap);
  Dr.resolve();
  Dn.resolve();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 a = WL.a[i];
    const float2 b = WL.b[i];
    const int4 m = WL.m[i];
    const int pid = m.x, spid = m.y, rep = m.z;
    const RepDev &R = reps[rep];
    float rq[3] = {a.x, a.y, a.z}, rv[3] = {a.w, b.x, b.y};
    if (m.w == 0) {
      double q[3] = {(double)a.x, (double)a.y, (double)a.z}, v[3] = {(double)a.w, (double)b.x, (double)b.y};
      xwalls_fp64<DIM>(C, R, sps, walls, wh, wq_b, &boxes[rep * 6], Dr, rep, spid, pid, q, v);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        rq[d] = (float)q[d];
        rv[d] = (float)v[d];
      }
    }
    float L[3] = {(float)R.L[0], (float)R.L[1], (float)R.L[2]}, sn[3];
    for (int d = 0; d < 3; d++) sn[d] = (float)Cn.shift[rep * 3 + d];
    float qn[3] = {rq[0], rq[1], rq[2]};
    xshift_fwd<DIM>(L, sn, qn);
    int cn = xcell<DIM>(R.L, R.pbc, R.cell_off, R.ny, R.nz, qn, true);
    int id = pid;
    if (cn < 0) {
      cn = xrepair(Cn, R, Dn, rep, pid, sps[spid].mass, rq, rv);
      id = pid | kIdRepaired;
    }
    const int b2 = bucket_of(brep[rep], R, cn);
    const int p2 = atomicAdd(&O.cnt[b2], 1);
    if (p2 >= cap_lim) {
      atomicOr(&C.flags[0], kFlagPoison);
      atomicMin(&C.flags[1], (int)Dr.rng.step);
      continue;
    }
    xrec_store(O, (size_t)b2 * cap + p2, onesp != 0, rq, rv, id, spid);
  }
}

// dense -> bucketed (step s, whose shift and tables are drawn): every
// particle goes to the bucket of its shifted cell; an escape of that bin is
// re-placed first (with step s's draws) and marked as already shifted
__global__ void k_xbucketize(Ctx C, State<float> S, XState O, Draws D, const BucketRep *__restrict__ brep, int cap,
                             int cap_lim, int onesp, int32_t *shifted) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= C.ntot) return;
  int r = 0;
  while (r + 1 < C.nrep && C.rep[r + 1].part_off <= s) r++;
  const RepDev &R = C.rep[r];
  const float q[3] = {S.x[s], S.y[s], S.z[s]}, v[3] = {S.vx[s], S.vy[s], S.vz[s]};
  const int id = S.id[s], sp = onesp ? 0 : S.sp[s];
  const float L[3] = {(float)R.L[0], (float)R.L[1], (float)R.L[2]};
  const float sc[3] = {(float)C.shift[r * 3], (float)C.shift[r * 3 + 1], (float)C.shift[r * 3 + 2]};
  float qs[3] = {q[0], q[1], q[2]};
  int cn;
  // a particle re-placed by the step that left it: already in this step's
  // frame.  The flag stays: after an overflow of this step the dense state,
  // flags included, is where the step starts again.
  const bool pre = shifted[R.part_off + id] != 0;
  if (R.dim == 3) {
    if (!pre) xshift_fwd<3>(L, sc, qs);
    cn = xcell<3>(R.L, R.pbc, R.cell_off, R.ny, R.nz, qs, true);
  } else {
    if (!pre) xshift_fwd<2>(L, sc, qs);
    cn = xcell<2>(R.L, R.pbc, R.cell_off, R.ny, R.nz, qs, true);
  }
  float rq[3] = {q[0], q[1], q[2]}, rv[3] = {v[0], v[1], v[2]};
  int rid = pre ? (id | kIdRepaired) : id;
  if (cn < 0) {
    D.resolve();
    cn = xrepair(C, R, D, r, id, C.sp[sp].mass, rq, rv);
    rid = id | kIdRepaired;
  }
  const int b = bucket_of(brep[r], R, cn);
  const int p = atomicAdd(&O.cnt[b], 1);
  if (p >= cap_lim) {
    atomicOr(&C.flags[0], kFlagPoison);
    atomicMin(&C.flags[1], -1);
    return;
  }
  xrec_store(O, (size_t)b * cap + p, onesp != 0, rq, rv, rid, sp);
}

// bucketed -> dense: every record to the slot of its particle (replica offset
// + id), so the dense state is in particle order.  A re-placed record (an
// escape of the coming step's shifted bin, already in that step's shifted
// frame) keeps its position and raises its particle's flag (by particle
// index, so it survives any reordering of the dense state), which tells the
// next step not to shift it; every other record clears its particle's flag.
__global__ void k_xdensify(Ctx C, XState I, State<float> S, const TileDev *__restrict__ bkts, int cap,
                           int cap_lim, int onesp, int32_t *shifted) {
  const TileDev T = bkts[blockIdx.x];
  const RepDev &R = C.rep[T.rep];
  const int n = min(I.cnt[blockIdx.x], min(cap, cap_lim));
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const size_t o = (size_t)blockIdx.x * cap + i;
    float q[3], v[3];
    int id, sp;
    xrec_load(I, o, onesp != 0, q, v, id, sp);
    const int pre = id < 0 ? 1 : 0;
    id &= 0x7FFFFFFF;
    const int s = R.part_off + id;
    shifted[s] = pre;
    S.x[s] = q[0];
    S.y[s] = q[1];
    S.z[s] = q[2];
    S.vx[s] = v[0];
    S.vy[s] = v[1];
    S.vz[s] = v[2];
    S.id[s] = id;
    S.sp[s] = onesp ? 0 : sp;
    S.cst[s] = -1;
  }
}

// ---------------------------------------------------------------------------
// the dense form of the same step (any particle order): global atomics for
// the sums, one lane per cell, one lane per particle
// ---------------------------------------------------------------------------
struct XDense {
  int32_t *n;     // [nctot]
  int32_t *nsp;   // [nctot * MPCD_MAXSPEC] (several species)
  double *sum;    // [nctot * NS]
  float *cd;      // [nctot * 12]
  int32_t *cell;  // [ntot] shifted cell of the slot, -1 none
  float *rt;      // [3 * ntot] thermal velocities (MPC-AT)
};

template <int OP, bool ONESP, int DIM>
__global__ __launch_bounds__(kBlock) void k_xd_sums(Ctx C, State<float> S, XDense X, const BlkDev *__restrict__ blk,
                                                    const mpcd_gpu_species *__restrict__ sps, Draws Dr,
                                                    int32_t *shifted) {
  constexpr int NS = xnsum(OP);
  const BlkDev B = blk[blockIdx.x];
  const int s = B.beg + threadIdx.x;
  if (s >= B.end) return;
  Dr.resolve();
  const RepDev &R = C.rep[B.rep];
  const float L[3] = {(float)R.L[0], (float)R.L[1], (float)R.L[2]};
  const float sc[3] = {(float)C.shift[B.rep * 3], (float)C.shift[B.rep * 3 + 1], (float)C.shift[B.rep * 3 + 2]};
  float q[3] = {S.x[s], S.y[s], S.z[s]}, v[3] = {S.vx[s], S.vy[s], S.vz[s]};
  const int pid = S.id[s], spid = ONESP ? 0 : S.sp[s];
  const double m = sps[spid].mass;
  // a particle re-placed by the step that left it is already in this frame
  if (shifted[R.part_off + pid]) shifted[R.part_off + pid] = 0;
  else xshift_fwd<DIM>(L, sc, q);
  int cn = xcell<DIM>(R.L, R.pbc, R.cell_off, R.ny, R.nz, q, true);
  if (cn < 0) {
    // escape repair (mpc.c:970-982), then binned where it was placed
    double qd[3] = {(double)S.x[s], (double)S.y[s], (double)S.z[s]}, vd[3] = {v[0], v[1], v[2]}, u0[3] = {0, 0, 0};
    atomicAdd(&C.events[0], 1ull);
    if (!escape_repair(R, C.wh, 0, Dr, B.rep, (uint32_t)pid, m, false, qd, vd, u0, C.wall, C.wq_s))
      atomicOr(&C.flags[0], kFlagEscape);
    for (int d = 0; d < 3; d++) {
      q[d] = (float)qd[d];
      v[d] = (float)vd[d];
    }
    cn = xcell<DIM>(R.L, R.pbc, R.cell_off, R.ny, R.nz, q, true);
    if (cn < 0) {
      atomicOr(&C.flags[0], kFlagEscape);
      cn = R.cell_off;
    }
  }
  float rvv[3] = {0.0f, 0.0f, 0.0f};
  if (OP != OP_SRD) {
    xthermal<DIM>(Dr, B.rep, pid, sqrt(R.kbt / m), rvv);
    for (int d = 0; d < 3; d++) X.rt[3 * (size_t)s + d] = rvv[d];
  }
  const int lc_ = cn - R.cell_off;
  double t[18];
  xterms<OP, DIM>(ONESP ? 1.0 : m, q, v, rvv, lc_ / (R.ny * R.nz), (lc_ / R.nz) % R.ny, lc_ % R.nz, t);
  atomicAdd(&X.n[cn], 1);
  if (!ONESP) atomicAdd(&X.nsp[(size_t)cn * MPCD_MAXSPEC + spid], 1);
  for (int d = 0; d < NS; d++)
    if (DIM == 3 || (d % 3) != 2 || OP == OP_RAT) atomicAdd(&X.sum[(size_t)cn * NS + d], t[d]);
  S.x[s] = q[0];
  S.y[s] = q[1];
  S.z[s] = q[2];
  S.vx[s] = v[0];
  S.vy[s] = v[1];
  S.vz[s] = v[2];
  X.cell[s] = cn;
}

template <int OP, bool ONESP, int DIM, bool GHOST>
__global__ __launch_bounds__(kBlock) void k_xd_cells(Ctx C, XDense X, const BlkDev *__restrict__ cblk,
                                                     const mpcd_gpu_species *__restrict__ sps, Draws Dr, int nspec) {
  constexpr int NS = xnsum(OP);
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c >= B.end) return;
  Dr.resolve();
  const RepDev &R = C.rep[B.rep];
  const int np_ = X.n[c];
  if (np_ == 0) return;
  double mass = 0.0;
  if (ONESP) mass = (double)np_ * sps[0].mass;
  else
    for (int s = 0; s < nspec; s++) mass += (double)X.nsp[(size_t)c * MPCD_MAXSPEC + s] * sps[s].mass;
  double sm[NS];
  for (int j = 0; j < NS; j++) sm[j] = X.sum[(size_t)c * NS + j];
  if (ONESP) {
    const int nm = (OP == OP_SRD) ? 3 : NS;
    for (int j = 0; j < nm; j++) sm[j] *= sps[0].mass;
  }
  xcell_terms<OP, DIM, GHOST>(R, C.wh, Dr, B.rep, c, np_, mass, sm, &X.cd[(size_t)c * 12]);
}

// the particle phase of a dense step; a particle whose next shifted bin it
// escapes is re-placed here with the next step's draws and wall positions
// (Cn, Dn) and flagged as already in that step's frame, as the bucketed
// layout does
template <int OP, bool ONESP, int DIM>
__global__ __launch_bounds__(kBlock) void k_xd_part(Ctx C, State<float> S, XDense X, const BlkDev *__restrict__ blk,
                                                    const mpcd_gpu_species *__restrict__ sps,
                                                    const mpcd_gpu_wall *__restrict__ walls,
                                                    const double *__restrict__ boxes, Draws Dr, Ctx Cn, Draws Dn,
                                                    int32_t *shifted) {
  const BlkDev B = blk[blockIdx.x];
  const int s = B.beg + threadIdx.x;
  double ke = 0.0;
  if (s < B.end) {
    Dr.resolve();
    const RepDev &R = C.rep[B.rep];
    const XRep XR = xrep(R, &C.shift[B.rep * 3], &boxes[B.rep * 6]);
    float q[3] = {S.x[s], S.y[s], S.z[s]}, v[3] = {S.vx[s], S.vy[s], S.vz[s]};
    const int pid = S.id[s], spid = ONESP ? 0 : S.sp[s];
    const double m = sps[spid].mass;
    const int c = X.cell[s];
    const int np_ = X.n[c];
    if (np_ > 1) {
      float cd[12];
      for (int k = 0; k < 12; k++) cd[k] = X.cd[(size_t)c * 12 + k];
      float rvv[3] = {0.0f, 0.0f, 0.0f}, r[3] = {0.0f, 0.0f, 0.0f};
      if (OP != OP_SRD)
        for (int d = 0; d < 3; d++) rvv[d] = X.rt[3 * (size_t)s + d];
      if (OP == OP_RAT) {
        const int lc_ = c - R.cell_off;
        r[0] = q[0] - (float)(lc_ / (R.ny * R.nz));
        r[1] = q[1] - (float)((lc_ / R.nz) % R.ny);
        r[2] = (DIM == 3) ? q[2] - (float)(lc_ % R.nz) : 0.0f;
      }
      xcollide<OP, DIM>(cd, np_, (float)sps[spid].damp, rvv, r, v);
    } else if (R.gravflag) {
      for (int d = 0; d < DIM; d++) v[d] = v[d] + XR.dtg[d];
    }
    ke = xke(m, XR.avv, v);
    xshift_back<DIM>(XR.L, XR.sc, q);
    if (!xdrift<DIM>(XR.box, XR.L, XR.wrap, XR.dt, v, q)) {
      double qd[3] = {q[0], q[1], q[2]}, vd[3] = {v[0], v[1], v[2]};
      xwalls_fp64<DIM>(C, R, sps, walls, C.wh, C.wq_b, &boxes[B.rep * 6], Dr, B.rep, spid, pid, qd, vd);
      for (int d = 0; d < 3; d++) {
        q[d] = (float)qd[d];
        v[d] = (float)vd[d];
      }
    }
    {
      float sn[3] = {(float)Cn.shift[B.rep * 3], (float)Cn.shift[B.rep * 3 + 1], (float)Cn.shift[B.rep * 3 + 2]};
      float qn[3] = {q[0], q[1], q[2]};
      xshift_fwd<DIM>(XR.L, sn, qn);
      if (xcell<DIM>(R.L, R.pbc, R.cell_off, R.ny, R.nz, qn, true) < 0) {
        Dn.resolve();
        (void)xrepair(Cn, R, Dn, B.rep, pid, m, q, v);
        shifted[R.part_off + pid] = 1;
      }
    }
    S.x[s] = q[0];
    S.y[s] = q[1];
    S.z[s] = q[2];
    S.vx[s] = v[0];
    S.vy[s] = v[1];
    S.vz[s] = v[2];
  }
  // exact per-block sum of the kinetic-energy terms
  __shared__ double s_ke[kBlock / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) ke += __shfl_xor(ke, o, 64);
  if ((threadIdx.x & 63) == 0) s_ke[threadIdx.x >> 6] = ke;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; w++) t += s_ke[w];
    C.tpart[blockIdx.x] = t;
  }
}

// TEMP of the exact form: the integer-valued partials of a replica summed in
// int64, so the temperature does not depend on how they were grouped
__global__ __launch_bounds__(kTempThreads) void k_xtemp_reduce(Ctx C, const int2 *ranges) {
  const int r = blockIdx.x;
  const RepDev &R = C.rep[r];
  const int2 rg = ranges[r];
  __shared__ unsigned long long sh;
  if (threadIdx.x == 0) sh = 0ull;
  __syncthreads();
  long long acc = 0;
  for (int b = rg.x + threadIdx.x; b < rg.y; b += kTempThreads) acc += (long long)C.tpart[b];
  atomicAdd(&sh, (unsigned long long)acc);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const double K = ((double)(long long)sh * kXQinv) / ((double)(R.npart * R.dim));
  C.gscale[r] = 1.0;
  C.kbtnow[r] = K;
}

}  // namespace mpcd
