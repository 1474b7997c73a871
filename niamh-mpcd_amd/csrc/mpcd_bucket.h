// mpcd_bucket.h -- the bucketed particle layout of multi-step runs and its
// step kernel (isotropic collision operators, no thermostat).
//
// The cells of every replica are cut into tiles of consecutive cells (the
// x-major order of the reference's cell array), and tile b owns the bucket
// [b * cap, (b + 1) * cap) of the bucketed state.  Between steps a bucket
// holds, in no particular order, the particles whose shifted cell of the
// coming step lies in its tile, each with the composite list-order key of
// the two binnings (bin_key).  One workgroup per tile
//   1. reads its bucket (contiguous), applies the grid shift with the escape
//      repair of bin(shifted=1) (mpc.c:970-982) and the body force, and
//      finds each particle's cell in the tile,
//   2. counting-sorts the keys by cell in LDS and ranks them inside every
//      cell: the key order is the reference's linked-list order, and the
//      bucket-count prefix plus the rank in the tile is the particle's slot
//      in the cell-sorted order -- the slot the dense layout gives it, so
//      keyed draws and keys are the same in both layouts,
//   3. computes the cell terms (one lane per cell, FP64 sums in list order)
//      and, one lane per particle, the collision, shift back, streaming and
//      walls, and the next step's shifted cell and key,
//   4. appends every particle to the bucket of its next tile.  The counts
//      per destination bucket are gathered in an LDS hash table first, so a
//      single global atomic reserves the room of all the workgroup's
//      particles for one destination.
// A bucket that would overflow sets the poison flag: that step and all later
// ones (and the shift kernels) do nothing, and the host resumes from the last
// complete state on the cell-sorted path.
#pragma once

#include "mpcd_fused.h"

namespace mpcd {

constexpr int kBktThreads = 256;
constexpr int kBktMaxCells = 128;  // cells per tile at most
constexpr int kBktHash = 64;       // destination buckets gathered per workgroup

// particles per thread, and the bucket capacity (one tile's particles must
// fit LDS): the MPC-AT kernels also stage three FP64 draws per particle
__host__ __device__ constexpr int bkt_ppt(int op) { return op == OP_SRD ? 4 : 2; }
__host__ __device__ constexpr int bkt_cap(int op) { return kBktThreads * bkt_ppt(op); }

// first bucket of a replica, the cells of its tiles and their reciprocal
struct BucketRep {
  int b0, tc;
  double inv_tc;
};

// The bucketed state: `cap` slots per bucket (of which the first cap_lim,
// at most cap, may be filled), `cnt` particles in each.  A slot is one
// record packed into vector words, so a particle costs three or four memory
// instructions each way:
//   FP32: a = (x, y, z, vx), b = (vy, vz, id, key bits 0..31),
//         c = key bits 32..63 (36 B), plus sp when there are several species;
//   FP64: a = (x, y), b = (z, vx), c = (vy, vz), d = (id, key bits 0..31,
//         key bits 32..63, species) (64 B).
template <typename Real>
struct BState;
template <>
struct BState<float> {
  float4 *a, *b;
  uint32_t *c;
  int32_t *sp;
  int32_t *cnt;
};
template <>
struct BState<double> {
  double2 *a, *b, *c;
  int4 *d;
  int32_t *cnt;
};

__device__ __forceinline__ void rec_load(const BState<float> &B, size_t o, bool onesp, float q[3], float v[3], int &id,
                                         int &sp, uint64_t &key) {
  const float4 a = B.a[o], b = B.b[o];
  const uint32_t hi = B.c[o];
  q[0] = a.x;
  q[1] = a.y;
  q[2] = a.z;
  v[0] = a.w;
  v[1] = b.x;
  v[2] = b.y;
  id = __float_as_int(b.z);
  key = ((uint64_t)hi << 32) | (uint64_t)(uint32_t)__float_as_int(b.w);
  sp = onesp ? 0 : B.sp[o];
}

__device__ __forceinline__ void rec_store(const BState<float> &B, size_t o, bool onesp, const float q[3],
                                          const float v[3], int id, int sp, uint64_t key) {
  B.a[o] = make_float4(q[0], q[1], q[2], v[0]);
  B.b[o] = make_float4(v[1], v[2], __int_as_float(id), __int_as_float((int)(uint32_t)key));
  B.c[o] = (uint32_t)(key >> 32);
  if (!onesp) B.sp[o] = sp;
}

__device__ __forceinline__ void rec_load(const BState<double> &B, size_t o, bool, double q[3], double v[3], int &id,
                                         int &sp, uint64_t &key) {
  const double2 a = B.a[o], b = B.b[o], c = B.c[o];
  const int4 d = B.d[o];
  q[0] = a.x;
  q[1] = a.y;
  q[2] = b.x;
  v[0] = b.y;
  v[1] = c.x;
  v[2] = c.y;
  id = d.x;
  key = ((uint64_t)(uint32_t)d.z << 32) | (uint64_t)(uint32_t)d.y;
  sp = d.w;
}

__device__ __forceinline__ void rec_store(const BState<double> &B, size_t o, bool, const double q[3], const double v[3],
                                          int id, int sp, uint64_t key) {
  B.a[o] = make_double2(q[0], q[1]);
  B.b[o] = make_double2(q[2], v[0]);
  B.c[o] = make_double2(v[1], v[2]);
  B.d[o] = make_int4(id, (int)(uint32_t)key, (int)(uint32_t)(key >> 32), sp);
}

// (cell - cell_off) / tc: an FP64 estimate (off by at most one for offsets
// below 2^31) and one correction step
// particles that leave the step's wall-free box, or its cells, while
// streaming: their state before the drift (exactly representable in the
// state precision), finished by k_bucket_walls
template <typename Real>
struct WallList {
  Real *x, *y, *z, *vx, *vy, *vz;
  int4 *meta;    // particle id, shifted cell of the step, slot, replica << 8 | species
  int32_t *cnt;  // [1]
};

// Slots (bucket * cap + position) of the records whose particle escapes in
// the coming step's shifted bin (bin(shifted=1), mpc.c:970-982).  The step
// that writes a record knows the next grid shift, so it lists such records;
// the next step's preparation kernel re-places them (escape_repair) before
// k_bucket runs and marks them (sign bit of the id: already in the shifted
// frame).  The step kernel itself never runs the repair, which would raise
// its register count for every particle.
struct EscList {
  int32_t *idx;
  int32_t *cnt;  // [1]
  int cap;
};

__device__ __forceinline__ void esc_push(const Ctx &C, const EscList &E, size_t o) {
  const int e = atomicAdd(E.cnt, 1);
  if (e < E.cap) E.idx[e] = (int32_t)o;
  else atomicOr(&C.flags[0], kFlagPoison);  // the cell-sorted path finishes the run
}

constexpr int32_t kIdRepaired = (int32_t)0x80000000;

__device__ __forceinline__ int bucket_of(const BucketRep &B, const RepDev &R, int cell) {
  const int x = cell - R.cell_off;
  int q = (int)((double)x * B.inv_tc);
  const int r = x - q * B.tc;
  if (r < 0) q--;
  else if (r >= B.tc) q++;
  return B.b0 + q;
}

// Branch-free forms of shift_fwd_d, shift_back_d and cell_of_d for the hot
// path (same arithmetic: selects instead of branches, so the work of several
// particles can overlap).  L: box lengths as doubles; Li: as integers.
template <int DIM>
__device__ __forceinline__ void shift_fwd_bf(const double L[3], const double s[3], double q[3]) {
#pragma unroll
  for (int j = 0; j < DIM; j++) {
    const double t = q[j] + s[j];
    q[j] = ((t > L[j]) & (t <= L[j] + s[j])) ? t - L[j] : t;
  }
}

template <int DIM>
__device__ __forceinline__ void shift_back_bf(const double L[3], const double s[3], double q[3]) {
#pragma unroll
  for (int j = 0; j < DIM; j++) {
    const double sg = -1.0 * s[j];
    const double t = q[j] + sg;
    q[j] = ((t < 0.0) & (t >= sg)) ? t + L[j] : t;
  }
}

template <int DIM>
__device__ __forceinline__ int cell_bf(const int Li[3], const int pbc[3], int cell_off, int ny, int nz,
                                       const double q[3], bool shifted) {
  int a = (int)q[0], b = (int)q[1], c = (int)q[2];
  const int Lz = (DIM == 3) ? Li[2] : 1;
  if (shifted) {
    a = (pbc[0] & (a == Li[0])) ? 0 : a;
    b = (pbc[1] & (b == Li[1])) ? 0 : b;
    c = (pbc[2] & (c == Lz)) ? 0 : c;
  }
  const bool bad = (a > Li[0]) | (b > Li[1]) | (c > Lz) | (a < 0) | (b < 0) | (c < 0);
  if (DIM < 3) c = 0;
  return bad ? -1 : cell_off + (a * ny + b) * nz + c;
}

template <typename Real, int OP, bool ONESP, int DIM, int PPT, bool GHOST>
__global__ __launch_bounds__(kBktThreads) void k_bucket(
    Ctx C, BState<Real> I, BState<Real> O, const TileDev *__restrict__ bkts, const BucketRep *__restrict__ brep,
    const RepDev *__restrict__ reps, const mpcd_gpu_species *__restrict__ sps, const mpcd_gpu_wall *__restrict__ walls,
    const WallHot *__restrict__ wh, const double *__restrict__ boxes, Draws Dr,
    const double *__restrict__ shift_cur, const double *__restrict__ wq_b, const double *__restrict__ shift_next,
    const int32_t *__restrict__ bsum, int32_t *start_out, int cap_lim, WallList<Real> WL, EscList EN) {
  constexpr int NT = kBktThreads, CAP = NT * PPT;
  const int tid = threadIdx.x;
  __shared__ int s_flag;
  __shared__ int32_t s_cnt[kBktMaxCells + 1];  // per-cell counts, then starts (s_cnt[ncl] = total)
  __shared__ uint64_t s_key[CAP];
  __shared__ Real s_v[3][CAP];
  __shared__ Real s_q[OP == OP_RAT ? 3 : 1][OP == OP_RAT ? CAP : 1];
  __shared__ double s_rv[OP == OP_SRD ? 1 : 3][OP == OP_SRD ? 1 : CAP];
  __shared__ uint8_t s_sp[ONESP ? 1 : CAP];
  __shared__ double s_cd[kBktMaxCells][12];
  __shared__ int32_t s_hkey[kBktHash], s_hcnt[kBktHash], s_hbase[kBktHash];
  __shared__ int32_t s_nwl, s_wlbase;  // wall-list entries of the workgroup, their first index
  __shared__ double s_ke[NT / 64];

  // The loads of a tile are issued before any of their values is used, so
  // a tile waits for the bucket's count and then for one memory latency,
  // not one per particle:
  //  - every record of the thread.  These loads are unconditional: a slot at
  //    or past the bucket's count reads the last filled slot again (a cached
  //    line, no HBM traffic), and its values are never used (rk = -1);
  //  - the terms of the bucket's slot base: the counts of the earlier
  //    buckets of its block of kBlock (one a thread) plus the block sums of
  //    the preparation kernel;
  //  - the poison flag, read once for the whole workgroup.
  static_assert(NT == kBlock, "one count of the bucket's block a thread");
  const size_t off = (size_t)blockIdx.x * CAP;
  const int n_raw = I.cnt[blockIdx.x];
  const int last = max(min(n_raw, CAP) - 1, 0);
  Real rq[PPT][3], rv[PPT][3];
  int rid[PPT], rsp[PPT], rk[PPT], ra[PPT];
  uint64_t rkey[PPT];
#pragma unroll
  for (int j = 0; j < PPT; j++)
    rec_load(I, off + min(tid + j * NT, last), ONESP, rq[j], rv[j], rid[j], rsp[j], rkey[j]);
  __shared__ int s_base[NT / 64];
  {
    const int bb = blockIdx.x / kBlock, k1 = bb * kBlock + tid;
    const int c1 = I.cnt[min(k1, (int)blockIdx.x)];
    int flag = 0;
    if (tid == 0) flag = C.flags[0];
    int part = 0;
    for (int k = tid; k < bb; k += NT) part += bsum[k];
    part += (k1 < (int)blockIdx.x) ? c1 : 0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    if ((tid & 63) == 0) s_base[tid >> 6] = part;
    if (tid == 0) s_flag = flag;
  }
  const TileDev T = bkts[blockIdx.x];
  const int ncl = T.c1 - T.c0;
  for (int k = tid; k <= ncl; k += NT) s_cnt[k] = 0;
  for (int k = tid; k < kBktHash; k += NT) {
    s_hkey[k] = -1;
    s_hcnt[k] = 0;
  }
  if (tid == 0) s_nwl = 0;
  __syncthreads();
  if (s_flag & kFlagPoison) return;
  int gbase = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) gbase += s_base[w];
  Dr.resolve();
  const RepDev &R = reps[T.rep];
  const BucketRep BR = brep[T.rep];
  const int n = min(n_raw, CAP);
  // the replica values of the hot path, loaded once
  const double Ld[3] = {(double)R.L[0], (double)R.L[1], (double)R.L[2]};
  const int Li[3] = {R.L[0], R.L[1], R.L[2]}, pbc[3] = {R.pbc[0], R.pbc[1], R.pbc[2]};
  const int cell_off = R.cell_off, ny = R.ny, nz = R.nz;
  const double shc[3] = {shift_cur[T.rep * 3], shift_cur[T.rep * 3 + 1], shift_cur[T.rep * 3 + 2]};
  const double shn[3] = {shift_next[T.rep * 3], shift_next[T.rep * 3 + 1], shift_next[T.rep * 3 + 2]};
  double box[6];
#pragma unroll
  for (int k = 0; k < 6; k++) box[k] = boxes[T.rep * 6 + k];
  const double dt = R.dt;

  // 1. the bucket: grid shift, escape repair, body force, cell in the tile
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    const int i = tid + j * NT;
    rk[j] = -1;
    ra[j] = 0;
    if (i < n) {
      const bool repaired = rid[j] < 0;
      rid[j] &= 0x7FFFFFFF;
      const int spid = ONESP ? 0 : rsp[j];
      const uint64_t key = rkey[j];
      double q[3] = {(double)rq[j][0], (double)rq[j][1], (double)rq[j][2]};
      double v[3] = {(double)rv[j][0], (double)rv[j][1], (double)rv[j][2]};
      if (R.galinv && !repaired) shift_fwd_bf<DIM>(Ld, shc, q);
      int cn = cell_bf<DIM>(Li, pbc, cell_off, ny, nz, q, true);
      // an escape of this shifted bin was re-placed by the preparation
      // kernel and stays in the list it was in (its unshifted cell, bits
      // 32..62 of the key); an unlisted escape would be an internal error
      if (repaired) cn = (int)((key >> 32) & 0x7FFFFFFFull);
      Real vv[3] = {(Real)v[0], (Real)v[1], (Real)v[2]};
      if (R.gravflag) {
        // acc_all (mpc.c:697), in the state precision
        for (int d = 0; d < DIM; d++) vv[d] = (Real)((double)vv[d] + R.dt * R.grav[d]);
      }
#pragma unroll
      for (int d = 0; d < 3; d++) {
        rq[j][d] = (Real)q[d];
        rv[j][d] = vv[d];
      }
      rsp[j] = spid;
      const int k = cn - T.c0;
      if ((unsigned)k < (unsigned)ncl) {
        rk[j] = k;
        ra[j] = atomicAdd(&s_cnt[k], 1);
      } else {
        atomicOr(&C.flags[0], kFlagBadCell);
      }
    }
  }
  __syncthreads();
  // 2. exclusive scan of the cell counts (one wavefront, two cells a lane)
  if (tid < 64) {
    const int k0 = 2 * tid;
    const int a = (k0 < ncl) ? s_cnt[k0] : 0, b = (k0 + 1 < ncl) ? s_cnt[k0 + 1] : 0;
    int incl = a + b;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (tid >= o) incl += t;
    }
    const int excl = incl - a - b;
    if (k0 < ncl) s_cnt[k0] = excl;
    if (k0 + 1 < ncl) s_cnt[k0 + 1] = excl + a;
    if (tid == 63) s_cnt[ncl] = incl;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PPT; j++)
    if (rk[j] >= 0) s_key[s_cnt[rk[j]] + ra[j]] = rkey[j];
  __syncthreads();
  // rank inside the cell (list order) and stage the cell sums' inputs
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    if (rk[j] < 0) continue;
    const int b = s_cnt[rk[j]], e = s_cnt[rk[j] + 1];
    int r = 0;
#pragma unroll 4
    for (int f = b; f < e; f++) r += (s_key[f] < rkey[j]) ? 1 : 0;
    const int pos = b + r;
    ra[j] = pos;
#pragma unroll
    for (int d = 0; d < 3; d++) s_v[d][pos] = rv[j][d];
    if (OP == OP_RAT)
#pragma unroll
      for (int d = 0; d < 3; d++) s_q[d][pos] = rq[j][d];
    if (!ONESP) s_sp[pos] = (uint8_t)rsp[j];
    if (OP != OP_SRD) {
      // thermal velocities of the collision (mpc.c:1382, 1506), keyed by slot
      const double sig = sqrt(R.kbt / sps[rsp[j]].mass);
      double nrm[3];
      Dr.part_n_all(T.rep, gbase + pos, DIM, nrm);
      for (int d = 0; d < DIM; d++) s_rv[d][pos] = sig * nrm[d];
    }
  }
  __syncthreads();
  // 3. one lane per cell: sums in list order and the collision's cell terms;
  //    the cell starts in the dense slot numbering
  for (int k = tid; k < ncl; k += NT) {
    cell_terms<Real, OP, ONESP, DIM, GHOST>(R, sps, wh, Dr, T.rep, T.c0 + k, s_cnt[k], s_cnt[k + 1], &s_v[0][0],
                                     &s_q[0][0], OP == OP_SRD ? nullptr : &s_rv[0][0], ONESP ? nullptr : s_sp, CAP,
                                     s_cd[k]);
    start_out[T.c0 + k] = gbase + s_cnt[k];
  }
  if (tid == 0) start_out[T.c1] = gbase + s_cnt[ncl];
  __syncthreads();
  // one lane per particle: collision, kinetic energy, shift back, stream,
  // next cell and key, destination bucket.  A particle that leaves the
  // wall-free box (or its cells) goes to the wall list instead.
  double ke = 0.0;
  uint32_t escm = 0;  // bit j: particle j escapes in the next shifted bin
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    if (rk[j] < 0) continue;
    const int k = rk[j], c = T.c0 + k, pos = ra[j];
    const int np_ = s_cnt[k + 1] - s_cnt[k];
    const int spid = ONESP ? 0 : rsp[j];
    const mpcd_gpu_species &sp = sps[spid];
    double v[3] = {(double)rv[j][0], (double)rv[j][1], (double)rv[j][2]};
    double q[3] = {(double)rq[j][0], (double)rq[j][1], (double)rq[j][2]};
    {
      // collided velocity, kept for cells of more than one particle
      double vc[3] = {v[0], v[1], v[2]}, dr[3] = {0., 0., 0.};
      if (OP != OP_SRD)
#pragma unroll
        for (int d = 0; d < DIM; d++) dr[d] = s_rv[d][pos];
      collide_one<Real, OP, DIM>(sp, s_cd[k], np_, dr, q, vc);
#pragma unroll
      for (int d = 0; d < 3; d++) v[d] = (np_ > 1) ? vc[d] : v[d];
    }
    {
      double v0 = v[0] - R.avv[0], v1 = v[1] - R.avv[1], v2 = v[2] - R.avv[2];
      double tmp = v0 * v0;
      tmp += v1 * v1;
      tmp += v2 * v2;
      ke += sp.mass * tmp;
    }
    if (R.galinv) {
      shift_back_bf<DIM>(Ld, shc, q);
#pragma unroll
      for (int d = 0; d < DIM; d++) q[d] = (double)(Real)q[d];
    }
    // stream_all (mpc.c:680) and the wall-free box of stream_walls
    double Q[3] = {q[0], q[1], q[2]};
    bool inside = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      Q[d] = Q[d] + dt * v[d];
      inside = inside & (Q[d] > box[2 * d]) & (Q[d] < box[2 * d + 1]);
    }
    const int cu0 = cell_bf<DIM>(Li, pbc, cell_off, ny, nz, Q, false);
    const int cu = inside ? cu0 : -1;
    if (cu < 0) {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        rq[j][d] = (Real)q[d];
        rv[j][d] = (Real)v[d];
      }
      rk[j] = -2 - atomicAdd(&s_nwl, 1);
      ra[j] = pos | (k << 16);
      continue;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
      rq[j][d] = (Real)Q[d];
      rv[j][d] = (Real)v[d];
      q[d] = (double)rq[j][d];
    }
    // next step's shifted cell (next_cell); an escape stays in its list
    double qn[3] = {q[0], q[1], q[2]};
    shift_fwd_bf<DIM>(Ld, shn, qn);
    const int cn0 = cell_bf<DIM>(Li, pbc, cell_off, ny, nz, qn, true);
    const int cn = cn0 < 0 ? cu : cn0;
    if (cn0 < 0) escm |= 1u << j;
    rkey[j] = bin_key(cn, cu, c, gbase + pos);
    const int b2 = bucket_of(BR, R, cn);
    rk[j] = b2;
    // LDS hash of the destination buckets (open addressing)
    int h = (int)(((uint32_t)b2 * 2654435761u) >> 26);
    // a plain read first: after the first few particles the slot is taken
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
  if (tid == NT - 1 && s_nwl > 0) s_wlbase = atomicAdd(WL.cnt, s_nwl);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < PPT; j++) {
    const int b2 = rk[j];
    if (b2 <= -2) {
      const int w = s_wlbase + (-2 - b2);
      WL.x[w] = rq[j][0];
      WL.y[w] = rq[j][1];
      WL.z[w] = rq[j][2];
      WL.vx[w] = rv[j][0];
      WL.vy[w] = rv[j][1];
      WL.vz[w] = rv[j][2];
      WL.meta[w] = make_int4(rid[j], T.c0 + (ra[j] >> 16), gbase + (ra[j] & 0xFFFF), (T.rep << 8) | rsp[j]);
      continue;
    }
    if (b2 < 0) continue;
    const int p2 = (ra[j] >= 0) ? s_hbase[ra[j] >> 16] + (ra[j] & 0xFFFF) : atomicAdd(&O.cnt[b2], 1);
    if (p2 >= cap_lim) {
      atomicOr(&C.flags[0], kFlagPoison);
      atomicMin(&C.flags[1], (int)Dr.rng.step);
      continue;
    }
    const size_t o = (size_t)b2 * CAP + p2;
    const Real q3[3] = {rq[j][0], rq[j][1], rq[j][2]}, v3[3] = {rv[j][0], rv[j][1], rv[j][2]};
    rec_store(O, o, ONESP, q3, v3, rid[j], rsp[j], rkey[j]);
    if ((escm >> j) & 1u) esc_push(C, EN, o);
  }
  // kinetic-energy partial of the tile
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

// The preparation of a bucketed step, one launch of kBlock-thread blocks:
//   blocks [0, nbs): the sums of this step's bucket counts per block of
//     kBlock buckets (the bucket kernel adds them up to its slot base), and
//     the zeroing of the output counts and of the wall-list count;
//   block nbs: the next step's grid shift and wall positions (shift_replica);
//   blocks above: the SRD cell draws of this step (as k_celldraws).
// Nothing happens after a bucket overflow.
template <typename Real>
__global__ __launch_bounds__(kBlock) void k_bucket_prep(
    Ctx C, Draws D1, const double *wq_prev_b, double *shift_out, double *wqs_out, double *wqb_out, WallHot *wh_out,
    double *box_out, const int32_t *__restrict__ cnt_in, int32_t *bsum, int32_t *cnt_zero, int32_t *wl_cnt, int nbkt,
    int nbs, Draws D, const BlkDev *__restrict__ cblk, double *celldraw, int ncb, BState<Real> I, EscList EC,
    EscList EN, const TileDev *__restrict__ bkts, int cap, int onesp, const WallHot *__restrict__ wh_cur, int nesc) {
  if (C.flags[0] & kFlagPoison) return;  // the flag does not change during this kernel
  const int tid = threadIdx.x, blk = blockIdx.x;
  if (blk >= nbs + 1 + ncb) {
    // the escapes of this step's shifted bin, listed by the step that wrote
    // their records: grid shift, escape repair, record marked as shifted
    D.resolve();
    const int n = min(*EC.cnt, EC.cap);
    for (int e = (blk - nbs - 1 - ncb) * kBlock + tid; e < n; e += nesc * kBlock) {
      const size_t o = (size_t)EC.idx[e];
      const TileDev T = bkts[o / (size_t)cap];
      const RepDev &R = C.rep[T.rep];
      Real q_[3], v_[3];
      int id, sp;
      uint64_t key;
      rec_load(I, o, onesp != 0, q_, v_, id, sp, key);
      double q[3] = {(double)q_[0], (double)q_[1], (double)q_[2]}, v[3] = {(double)v_[0], (double)v_[1], (double)v_[2]};
      if (R.galinv) shift_fwd(R, &C.shift[T.rep * 3], q);
      if (cell_of(R, q, 1) >= 0) continue;  // listed only when it escapes
      atomicAdd(&C.events[0], 1ull);
      double u0[3] = {0.0, 0.0, 0.0};
      if (!escape_repair(R, wh_cur, 0, D, T.rep, (uint32_t)id, C.sp[sp].mass, false, q, v, u0))
        atomicOr(&C.flags[0], kFlagEscape);
      const Real q3[3] = {(Real)q[0], (Real)q[1], (Real)q[2]}, v3[3] = {(Real)v[0], (Real)v[1], (Real)v[2]};
      rec_store(I, o, onesp != 0, q3, v3, id | kIdRepaired, sp, key);
    }
    return;
  }
  if (blk < nbs) {
    const int b = blk * kBlock + tid;
    int v = (b < nbkt) ? cnt_in[b] : 0;
    if (b <= nbkt) cnt_zero[b] = 0;
    if (b == 0) {
      *wl_cnt = 0;
      *EN.cnt = 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __shared__ int s_w[kBlock / 64];
    if ((tid & 63) == 0) s_w[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
      int t = 0;
      for (int w = 0; w < kBlock / 64; w++) t += s_w[w];
      bsum[blk] = t;
    }
  } else if (blk == nbs) {
    D1.resolve();
    for (int r = tid; r < C.nrep; r += kBlock)
      shift_replica(C, D1, r, 1, wq_prev_b, shift_out, wqs_out, wqb_out, wh_out, box_out);
  } else if (celldraw) {
    D.resolve();
    const BlkDev B = cblk[blk - nbs - 1];
    const int c = B.beg + tid;
    if (c >= B.end) return;
    const RepDev &R = C.rep[B.rep];
    double *o = &celldraw[(size_t)c * 4];
    if (R.rtech == MPCD_ARBAXIS && R.dim == 3) {
      double u[3];
      D.rng.unitvec(B.rep, P_CELL, c, 0, u);
      o[0] = u[0];
      o[1] = u[1];
      o[2] = u[2];
    } else {
      o[0] = D.rng.uniform(B.rep, P_CELL, c, 0);
    }
    o[3] = D.rng.uniform(B.rep, P_CELL, c, 3);
  }
}

// the wall list of a bucketed step: stream with the planar walls
// (MPC_BCcollision, bc.c:1363-1496), unshifted cell with escape repair, next
// cell and key, one global atomic per particle on its destination bucket
// walls staged in LDS by k_bucket_walls (the wall search reads every wall
// of the replica on each pass)
constexpr int kWallStage = 32;

template <typename Real, int DIM>
__global__ __launch_bounds__(kBlock) void k_bucket_walls(
    Ctx C, WallList<Real> WL, BState<Real> O, const BucketRep *__restrict__ brep, const RepDev *__restrict__ reps,
    const mpcd_gpu_species *__restrict__ sps, const mpcd_gpu_wall *__restrict__ walls,
    const WallHot *__restrict__ wh, const double *__restrict__ boxes, Draws Dr, const double *__restrict__ wq_b,
    const double *__restrict__ shift_next, int cap, int cap_lim, int nwall, EscList EN) {
  if (C.flags[0] & kFlagPoison) return;
  const int n = *WL.cnt;
  if ((int)(blockIdx.x * blockDim.x) >= n) return;  // the whole block has no entry
  __shared__ WallHot s_wh[kWallStage];
  __shared__ mpcd_gpu_wall s_wall[kWallStage];
  __shared__ double s_wqb[3 * kWallStage];
  if (nwall <= kWallStage) {
    for (int w = threadIdx.x; w < nwall; w += blockDim.x) {
      s_wh[w] = wh[w];
      s_wall[w] = walls[w];
      s_wqb[3 * w] = wq_b[3 * w];
      s_wqb[3 * w + 1] = wq_b[3 * w + 1];
      s_wqb[3 * w + 2] = wq_b[3 * w + 2];
    }
    __syncthreads();
    wh = s_wh;
    walls = s_wall;
    wq_b = s_wqb;
  }
  Dr.resolve();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int4 m = WL.meta[i];
    const int pid = m.x, c = m.y, slot = m.z, rep = m.w >> 8, spid = m.w & 255;
    const RepDev &R = reps[rep];
    double q[3] = {(double)WL.x[i], (double)WL.y[i], (double)WL.z[i]};
    double v[3] = {(double)WL.vx[i], (double)WL.vy[i], (double)WL.vz[i]};
    double U0[3] = {0, 0, 0};
    const int cu = stream_walls<DIM>(C, R, sps, walls, wh, wq_b, &boxes[rep * 6], Dr, rep, spid, pid, c, q, v, U0,
                                     false);
    Real rq[3], rv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      rq[d] = (Real)q[d];
      rv[d] = (Real)v[d];
      q[d] = (double)rq[d];
    }
    double qn[3] = {q[0], q[1], q[2]};
    shift_fwd_d<DIM>(R, &shift_next[rep * 3], qn);
    const int cn0 = cell_of_d<DIM>(R, qn, 1);
    const int cn = cn0 < 0 ? cu : cn0;  // next_cell: an escape stays in its list
    const int b2 = bucket_of(brep[rep], R, cn);
    const int p2 = atomicAdd(&O.cnt[b2], 1);
    if (p2 >= cap_lim) {
      atomicOr(&C.flags[0], kFlagPoison);
      atomicMin(&C.flags[1], (int)Dr.rng.step);
      continue;
    }
    const size_t o = (size_t)b2 * cap + p2;
    rec_store(O, o, false, rq, rv, pid, spid, bin_key(cn, cu, c, slot));
    if (cn0 < 0) esc_push(C, EN, o);
  }
}

// dense -> bucketed: every slot of the cell-sorted state, with the pending
// shifted bin of the coming step (cnew, key), goes to the bucket of its cell
template <typename Real>
__global__ void k_bucketize(Ctx C, State<Real> S, BState<Real> O, const BucketRep *__restrict__ brep, int cap,
                            int cap_lim, EscList EN) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= C.ntot) return;
  const int cn = C.cnew[s];
  // the replica of a cell: the cell offsets are non-decreasing
  int r = 0;
  while (r + 1 < C.nrep && C.rep[r + 1].cell_off <= cn) r++;
  const int b = bucket_of(brep[r], C.rep[r], cn);
  const int p = hist_add(O.cnt, b);  // consecutive slots share a few buckets
  if (p >= cap_lim) {
    atomicOr(&C.flags[0], kFlagPoison);
    atomicMin(&C.flags[1], -1);
    return;
  }
  const size_t o = (size_t)b * cap + p;
  const Real q[3] = {S.x[s], S.y[s], S.z[s]}, v[3] = {S.vx[s], S.vy[s], S.vz[s]};
  rec_store(O, o, false, q, v, S.id[s], S.sp[s], C.key[s]);
  // an escape of the coming shifted bin (C.shift) goes to the escape list
  const RepDev &R = C.rep[r];
  double qs[3] = {(double)q[0], (double)q[1], (double)q[2]};
  if (R.galinv) shift_fwd(R, &C.shift[r * 3], qs);
  if (cell_of(R, qs, 1) < 0) esc_push(C, EN, o);
}

// bucketed -> dense: each particle goes back to its slot of the last step's
// cell-sorted order (bits 0..30 of its key) with the pending shifted bin of
// the coming step (cell, key, histogram) and the cell of the slot, found in
// the last step's cell starts
template <typename Real, int DIM>
__global__ void k_densify(Ctx C, BState<Real> I, State<Real> S, const TileDev *__restrict__ bkts, int cap, int onesp,
                          const double *__restrict__ shift, const int32_t *__restrict__ start_prev) {
  const TileDev T = bkts[blockIdx.x];
  const RepDev &R = C.rep[T.rep];
  const int n = min(I.cnt[blockIdx.x], cap);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const size_t o = (size_t)blockIdx.x * cap + i;
    Real q_[3], v_[3];
    int id, sp;
    uint64_t key;
    rec_load(I, o, onesp != 0, q_, v_, id, sp, key);
    const int slot = (int)(key & 0x7FFFFFFFull);
    const int cu = (int)((key >> 32) & 0x7FFFFFFFull);
    if (id < 0) {
      // re-placed by a step that did not complete (bucket overflow): the
      // cell-sorted step repeats the repair with the same draws; any position
      // that escapes its shifted bin stands for the one it left
      id &= 0x7FFFFFFF;
      q_[0] = (Real)(-2.0 * (double)R.L[0] - 2.0);
    }
    const double q[3] = {(double)q_[0], (double)q_[1], (double)q_[2]};
    S.x[slot] = q_[0];
    S.y[slot] = q_[1];
    S.z[slot] = q_[2];
    S.vx[slot] = v_[0];
    S.vy[slot] = v_[1];
    S.vz[slot] = v_[2];
    S.id[slot] = id;
    S.sp[slot] = onesp ? 0 : sp;
    const int cn = next_cell<DIM>(R, &shift[T.rep * 3], cu, q);
    C.cnew[slot] = cn;
    C.loc[slot] = hist_add(C.hist, cn);
    C.key[slot] = key;
    // last cell of the replica whose start is <= slot
    int lo = R.cell_off, hi = R.cell_off + R.ncell;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (start_prev[mid] <= slot) lo = mid;
      else hi = mid;
    }
    S.cst[slot] = lo;
  }
}

}  // namespace mpcd
