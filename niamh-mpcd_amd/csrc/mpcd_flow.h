// mpcd_flow.h -- the running flow field of the reference's flow output.
//
// With flow output on, the end of every timestep (mpc.c:4946-4951) computes
// the centre-of-mass velocity of each cell of the unshifted binning
// (localFLOW, mpc.c:226-233, through localMPCVCM, mpc.c:3889-3910:
// sum m v / sum m, divided by 1 in an empty cell) and adds it to the cell's
// running sum FLOW (sumFLOW, mpc.c:241-246); flowout writes FLOW divided by
// the output interval and clears it (pout.c:1681-1700).
//
// Cell-sorted layout: one launch over the particles gathers the sums of m v
// and m of every unshifted cell into an FP64 accumulator with hardware
// atomics (one per value and run of equal cells in a wavefront), and one launch over the cells adds the quotient to FLOW and clears
// the accumulator with plain stores (so the pair can sit in a captured
// graph).  The atomics add in no fixed order: a cell's sum equals the
// reference's list-order sum to round-off, not bit for bit.
//
// Bucketed layout, without global atomics: after a step, bucket b holds the
// particles whose next shifted cell cn lies in tile b, and a particle's
// unshifted cell cu differs from cn by 0 or +1 (periodic wrap) along each
// axis, since the grid shift lies in [0, 1).  One workgroup per bucket sums
// m v and m in LDS per (cell cn of the tile, offset class delta in
// {0, 1}^DIM) and writes every partial of its tile; the fold of cell cu then
// adds the 2^DIM partials that can hold its particles, in a fixed order.  A
// particle outside both candidates of an axis (none is expected) goes to the
// atomic accumulator instead, so nothing is lost.
#pragma once

#include "mpcd_bucket.h"

namespace mpcd {

// accumulator of cell c: acc[4c .. 4c+2] = sum m v, acc[4c+3] = sum m
template <int DIM>
__device__ __forceinline__ void flow_add(double *acc, int c, double m, double vx, double vy, double vz) {
  double *a = acc + (size_t)c * 4;
  unsafeAtomicAdd(a, vx * m);
  unsafeAtomicAdd(a + 1, vy * m);
  if (DIM == 3) unsafeAtomicAdd(a + 2, vz * m);
  unsafeAtomicAdd(a + 3, m);
}

// the cell-sorted state after a step: the unshifted cell that streaming
// left for slot s is in cu[] or in bits 32..62 of the slot's key (cu_src)
// (slots are in shifted-cell order, so neighbouring lanes often share an
// unshifted cell: each run of equal cells in a wavefront is summed with
// shuffles first, and only the run's last lane issues the atomics)
template <typename Real, int DIM>
__global__ __launch_bounds__(kBlock) void k_flow_dense(Ctx C, State<Real> S, const BlkDev *__restrict__ blk,
                                                        int cu_src, double *acc) {
  const BlkDev B = blk[blockIdx.x];
  const int s = B.beg + threadIdx.x;
  const bool live = s < B.end;
  int cu = -1;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  if (live) {
    cu = cu_src ? (int)((C.key[s] >> 32) & 0x7FFFFFFFull) : C.cu[s];
    const double m = C.sp[S.sp[s]].mass;
    a[0] = (double)S.vx[s] * m;
    a[1] = (double)S.vy[s] * m;
    a[2] = DIM == 3 ? (double)S.vz[s] * m : 0.0;
    a[3] = m;
  }
  // segmented inclusive scan over runs of equal cu in lane order; a cell can
  // occur in several runs of a wavefront, so a run starts wherever cu changes
  const int lane = threadIdx.x & 63;
  const int cu_prev = __shfl_up(cu, 1, 64);
  int head = (lane == 0 || cu_prev != cu) ? lane : 0;  // first lane of this lane's run: a max-scan of the starts
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int hu = __shfl_up(head, o, 64);
    if (lane >= o && hu > head) head = hu;
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    double t[4];
#pragma unroll
    for (int k = 0; k < 4; k++) t[k] = __shfl_up(a[k], o, 64);
    if (lane - o >= head)
#pragma unroll
      for (int k = 0; k < 4; k++) a[k] += t[k];
  }
  const int cu_next = __shfl_down(cu, 1, 64);
  const bool last = lane == 63 || cu_next != cu;
  if (live && last) {
    double *p = acc + (size_t)cu * 4;
    unsafeAtomicAdd(p, a[0]);
    unsafeAtomicAdd(p + 1, a[1]);
    if (DIM == 3) unsafeAtomicAdd(p + 2, a[2]);
    unsafeAtomicAdd(p + 3, a[3]);
  }
}

// the next cell along an axis under the grid shift: u + 1, or 0 past the
// last cell of a periodic axis (the wrap of bin(shifted=1), mpc.c:964-968);
// -1 beyond the cell array
__device__ __forceinline__ int flow_next(const RepDev &R, int ax, int u) {
  const int n = u + 1;
  if (R.pbc[ax] && n == R.L[ax]) return 0;
  return (n <= R.L[ax]) ? n : -1;
}

// the coordinates of the cell of q as cell_bf finds them (truncation toward
// zero, the periodic wrap of a shifted position at L); false where cell_bf
// gives -1
template <int DIM>
__device__ __forceinline__ bool cell_coords_bf(const RepDev &R, const double q[3], bool shifted, int w[3]) {
  w[0] = (int)q[0];
  w[1] = (int)q[1];
  w[2] = (DIM == 3) ? (int)q[2] : 0;
  const int Lz = (DIM == 3) ? R.L[2] : 1;
  if (shifted) {
    w[0] = (R.pbc[0] & (w[0] == R.L[0])) ? 0 : w[0];
    w[1] = (R.pbc[1] & (w[1] == R.L[1])) ? 0 : w[1];
    w[2] = (R.pbc[2] & (w[2] == Lz)) ? 0 : w[2];
  }
  const bool bad = (w[0] > R.L[0]) | (w[1] > R.L[1]) | (w[2] > Lz) | (w[0] < 0) | (w[1] < 0) | (w[2] < 0);
  if (DIM < 3) w[2] = 0;
  return !bad;
}

__device__ __forceinline__ int cell_index(const RepDev &R, const int w[3]) {
  return R.cell_off + (w[0] * R.ny + w[1]) * R.nz + w[2];
}

__device__ __forceinline__ void cell_decompose(const RepDev &R, int c, int w[3]) {
  const int l = c - R.cell_off;
  w[0] = l / (R.ny * R.nz);
  w[1] = (l / R.nz) % R.ny;
  w[2] = l % R.nz;
}

// partials of bucketed runs: part[(cls * nctot + cell) * 4 + j], offset class
// cls < 2^DIM, j: sum m v [3], sum m (class-major, so that the fold's reads
// of one class by consecutive cells are contiguous)
constexpr int kFlowCls = 8;

// the partials of the tile of bucket b from the records just written;
// shift_next: the grid shift of the coming step (cn is found as the
// bucketed step finds it, escapes staying in their list)
template <typename Real, int DIM>
__global__ __launch_bounds__(kBlock) void k_flow_tiles(const int32_t *__restrict__ flags, BState<Real> B,
                                                        const TileDev *__restrict__ bkts,
                                                        const RepDev *__restrict__ reps,
                                                        const mpcd_gpu_species *__restrict__ sps,
                                                        const double *__restrict__ shift_next, int cap, int onesp,
                                                        int nctot, double *part, double *acc) {
  constexpr int NCLS = 1 << DIM;
  __shared__ double s_part[kBktMaxCells * kFlowCls * 4];  // [cell of the tile][class][4]
  if (flags[0] & kFlagPoison) return;
  const TileDev T = bkts[blockIdx.x];
  const int ncl = T.c1 - T.c0;
  for (int k = threadIdx.x; k < ncl * kFlowCls * 4; k += kBlock) s_part[k] = 0.0;
  __syncthreads();
  const RepDev &R = reps[T.rep];
  const double Ld[3] = {(double)R.L[0], (double)R.L[1], (double)R.L[2]};
  const double sh[3] = {shift_next[T.rep * 3], shift_next[T.rep * 3 + 1], shift_next[T.rep * 3 + 2]};
  const int n = min(B.cnt[blockIdx.x], cap);
  for (int i = threadIdx.x; i < n; i += kBlock) {
    Real q_[3], v[3];
    int id, sp;
    uint64_t key;
    rec_load(B, (size_t)blockIdx.x * cap + i, onesp != 0, q_, v, id, sp, key);
    const int cu = (int)((key >> 32) & 0x7FFFFFFFull);
    double q[3] = {(double)q_[0], (double)q_[1], (double)q_[2]};
    // the unshifted cell's coordinates from the position; the key's cell
    // decides where the two differ (FP32 rounding at a cell face)
    int u[3], w[3];
    if (!cell_coords_bf<DIM>(R, q, false, u) || cell_index(R, u) != cu) cell_decompose(R, cu, u);
    if (R.galinv) shift_fwd_bf<DIM>(Ld, sh, q);
    int cn;
    if (cell_coords_bf<DIM>(R, q, true, w)) {
      cn = cell_index(R, w);
    } else {
      cn = cu;  // escaped: stays in its list
      w[0] = u[0];
      w[1] = u[1];
      w[2] = u[2];
    }
    int cls = 0;
    bool ok = (unsigned)(cn - T.c0) < (unsigned)ncl;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      if (w[d] == u[d]) continue;
      cls |= 1 << d;
      ok = ok && (w[d] == flow_next(R, d, u[d]));
    }
    const double m = sps[sp].mass;
    if (ok) {
      double *a = &s_part[((cn - T.c0) * kFlowCls + cls) * 4];
      unsafeAtomicAdd(a, (double)v[0] * m);
      unsafeAtomicAdd(a + 1, (double)v[1] * m);
      if (DIM == 3) unsafeAtomicAdd(a + 2, (double)v[2] * m);
      unsafeAtomicAdd(a + 3, m);
    } else {
      flow_add<DIM>(acc, cu, m, (double)v[0], (double)v[1], (double)v[2]);
    }
  }
  __syncthreads();
  // class-major write: for each class, the tile's cells are contiguous
  const int per = ncl * 4;
  for (int e = threadIdx.x; e < NCLS * per; e += kBlock) {
    const int cls = e / per, r = e - cls * per, k = r >> 2, j = r & 3;
    part[((size_t)cls * nctot + T.c0 + k) * 4 + j] = s_part[(k * kFlowCls + cls) * 4 + j];
  }
}

// FLOW += sum m v / sum m of every cell from the tile partials that can hold
// its particles (offset classes in a fixed order) and the atomic
// accumulator (touched only where a particle fell back to it); one block per
// cell block of a replica, and block 0 counts the step for every replica
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_flow_fold_tiles(const int32_t *__restrict__ flags,
                                                             const RepDev *__restrict__ reps,
                                                             const BlkDev *__restrict__ cblk, int nrep, int nctot,
                                                             const double *__restrict__ part, double *acc,
                                                             double *flow, int64_t *nsum) {
  if (flags[0] & kFlagPoison) return;
  if (blockIdx.x == 0)
    for (int r = threadIdx.x; r < nrep; r += kBlock) nsum[r] += 1;
  const BlkDev B = cblk[blockIdx.x];
  const int c = B.beg + threadIdx.x;
  if (c >= B.end) return;
  const RepDev &R = reps[B.rep];
  int u[3];
  cell_decompose(R, c, u);
  double sum[4] = {0.0, 0.0, 0.0, 0.0};
  for (int cls = 0; cls < (1 << DIM); cls++) {
    int w[3] = {u[0], u[1], u[2]};
    bool ok = true;
#pragma unroll
    for (int d = 0; d < DIM; d++)
      if (cls & (1 << d)) {
        w[d] = flow_next(R, d, u[d]);
        ok = ok && w[d] >= 0;
      }
    if (!ok) continue;
    const double *p = part + ((size_t)cls * nctot + cell_index(R, w)) * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) sum[k] += p[k];
  }
  double *a = acc + (size_t)c * 4;
  if (a[3] != 0.0) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      sum[k] += a[k];
      a[k] = 0.0;
    }
  }
  const double den = (sum[3] != 0.0) ? sum[3] : 1.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) flow[(size_t)c * 3 + d] += sum[d] / den;
}

// FLOW += sum m v / (sum m, or 1 when the cell is empty), components < DIM
// (mpc.c:3909, 244); the accumulator is cleared for the next step, and
// thread r < nrep counts the step for replica r
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_flow_fold(const int32_t *__restrict__ flags, int nctot, int nrep,
                                                       double *acc, double *flow, int64_t *nsum) {
  if (flags[0] & kFlagPoison) return;
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c < nrep) nsum[c] += 1;
  if (c >= nctot) return;
  double *a = acc + (size_t)c * 4;
  const double m = a[3];
  const double den = (m != 0.0) ? m : 1.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) flow[(size_t)c * 3 + d] += a[d] / den;
#pragma unroll
  for (int k = 0; k < 4; k++) a[k] = 0.0;
}

}  // namespace mpcd
