// mpcd_noalloc.hip -- mpcd_gpu_bin_noalloc: the binning of mpcd_gpu_bin for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  It takes no context: the
// particle positions and the outputs are the caller's device arrays, and the
// working memory (histogram, cell starts, in-cell ranks, sort keys, scan
// partials) is the file-scope device array `scratch`, of a fixed size that
// each GPU the library runs on holds until the process ends.
//
// The binning is the reference's (bin(), mpc.c:948-992; binin, mpc.c:924):
//  - without a shift: binin, every cell's list in particle-index order;
//  - with a shift: gridShift_all(+shift) (mpc.c:728-736, the periodic wrap of
//    bin(shifted=1), mpc.c:964-968) and the relinking of the lists of a
//    previous binning (its order and cells given): in every cell the
//    particles that stayed keep their order and are followed by the
//    arrivals in the previous order.  An escaped particle stays in the list
//    it was in (its re-placement, mpc.c:970-982, is left to the caller);
//    without a previous binning it is listed after every cell with cell -1.
// It launches kernels only (no memset, copy or synchronisation), on the
// caller's stream.  Calls that overlap on one GPU share `scratch` and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mpcd_params.h"

using namespace mpcd;

namespace {

constexpr int kNaMaxPart = MPCD_NOALLOC_MAX_PARTICLES;
constexpr int kNaMaxCell = MPCD_NOALLOC_MAX_CELLS;  // (cells + 2) / 1024 scan blocks stay within one block of sums
constexpr int kNaThreads = 256;
constexpr int kNaScanBlock = 1024;   // values per block of the histogram scan
// scratch layout in int32 words: histogram and starts (ncell + 2 each: the
// extra bucket holds the escapes of a binin), in-cell ranks, inverse of the
// previous order, scan block sums (two levels), then the 64-bit sort keys
constexpr size_t kNaHist = 0;
constexpr size_t kNaStart = kNaHist + kNaMaxCell + 2;
constexpr size_t kNaLoc = kNaStart + kNaMaxCell + 2;
constexpr size_t kNaInv = kNaLoc + kNaMaxPart;
constexpr size_t kNaBsum = kNaInv + kNaMaxPart;
constexpr size_t kNaBoff = kNaBsum + 2048;
constexpr size_t kNaKeys = ((kNaBoff + 2048 + 63) / 64) * 64;  // 256-byte aligned
constexpr size_t kNaWords = kNaKeys + 2 * (size_t)kNaMaxPart;   // about 41.6 MB
static_assert(kNaMaxCell + 2 <= kNaScanBlock * kNaScanBlock, "the histogram scan has two levels");

static __device__ __align__(256) int32_t scratch[kNaWords];

thread_local std::string g_err;

std::mutex g_mu;
int32_t *g_scratch_dev[64] = {};  // hipGetSymbolAddress(scratch) per device

mpcd_status na_fail(mpcd_status st, const char *m) {
  g_err = m;
  return st;
}

struct NaBox {
  int dim, L[3], ny, nz, pbc[3], ncell;
  double s[3];
};

__global__ void k_na_fill(int32_t *a, int32_t v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// position in the previous list order of every particle
__global__ void k_na_inverse(const int32_t *prev_perm, int32_t *inv, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) inv[prev_perm[t]] = t;
}

// cell of every particle, its in-cell sort key and its count.  Key: with a
// previous binning (moved << 31 | previous position), else the particle index
__global__ void k_na_count(NaBox B, const double *__restrict__ q, const int32_t *__restrict__ prev_cell,
                           const int32_t *__restrict__ inv, int shifted, int32_t *hist, int32_t *loc,
                           uint32_t *key, int32_t *cell_out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
  if (shifted)
    for (int j = 0; j < B.dim; j++) {
      x[j] += B.s[j];
      if (x[j] > B.L[j] && x[j] <= B.L[j] + B.s[j]) x[j] -= B.L[j];
    }
  int a = (int)x[0], b = (int)x[1], c = (int)x[2];
  const int Lz = (B.dim == 3) ? B.L[2] : 1;
  if (shifted) {
    if (B.pbc[0] && a == B.L[0]) a = 0;
    if (B.pbc[1] && b == B.L[1]) b = 0;
    if (B.pbc[2] && c == Lz) c = 0;
  }
  const bool esc = a > B.L[0] || b > B.L[1] || c > Lz || a < 0 || b < 0 || c < 0;
  if (B.dim < 3) c = 0;
  int cell = esc ? -1 : (a * B.ny + b) * B.nz + c;
  uint32_t k = (uint32_t)i;
  if (prev_cell) {
    const int pc = prev_cell[i];
    if (esc) cell = pc;  // stays in its list
    k = ((cell != pc) ? 0x80000000u : 0u) | (uint32_t)inv[i];
  }
  cell_out[i] = cell;
  const int h = (cell < 0) ? B.ncell : cell;
  loc[i] = atomicAdd(&hist[h], 1);
  key[i] = k;
}

// exclusive scan of n values, kNaScanBlock per block; block totals to bsum
__global__ void k_na_scan(const int32_t *in, int32_t *out, int32_t *bsum, int n) {
  __shared__ int32_t sh[kNaScanBlock];
  const int base = blockIdx.x * kNaScanBlock;
  for (int k = threadIdx.x; k < kNaScanBlock; k += blockDim.x) sh[k] = (base + k < n) ? in[base + k] : 0;
  __syncthreads();
  for (int off = 1; off < kNaScanBlock; off <<= 1) {
    int32_t t[kNaScanBlock / kNaThreads];
    for (int j = 0; j < kNaScanBlock / kNaThreads; j++) {
      const int k = threadIdx.x + j * kNaThreads;
      t[j] = (k >= off) ? sh[k - off] : 0;
    }
    __syncthreads();
    for (int j = 0; j < kNaScanBlock / kNaThreads; j++) sh[threadIdx.x + j * kNaThreads] += t[j];
    __syncthreads();
  }
  for (int k = threadIdx.x; k < kNaScanBlock; k += blockDim.x)
    if (base + k < n) out[base + k] = (k == 0) ? 0 : sh[k - 1];
  if (threadIdx.x == 0 && bsum) bsum[blockIdx.x] = sh[kNaScanBlock - 1];
}

__global__ void k_na_scan_add(int32_t *out, const int32_t *boff, int n) {
  const int base = blockIdx.x * kNaScanBlock;
  const int32_t add = boff[blockIdx.x];
  for (int k = threadIdx.x; k < kNaScanBlock; k += blockDim.x)
    if (base + k < n) out[base + k] += add;
}

__global__ void k_na_scatter(const int32_t *__restrict__ cell, const int32_t *__restrict__ loc,
                             const uint32_t *__restrict__ key, const int32_t *__restrict__ start, uint64_t *sorted,
                             int ncell, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int h = (cell[i] < 0) ? ncell : cell[i];
  sorted[start[h] + loc[i]] = ((uint64_t)key[i] << 32) | (uint32_t)i;
}

// every list (and the escapes' bucket) sorted by key; the permutation and the
// populations out
__global__ void k_na_sort(const int32_t *__restrict__ start, uint64_t *sorted, int32_t *perm, int32_t *count,
                          int ncell) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncell) return;
  const int b = start[c], e = start[c + 1];
  for (int i = b + 1; i < e; i++) {
    const uint64_t v = sorted[i];
    int j = i - 1;
    while (j >= b && sorted[j] > v) {
      sorted[j + 1] = sorted[j];
      j--;
    }
    sorted[j + 1] = v;
  }
  for (int i = b; i < e; i++) perm[i] = (int32_t)(uint32_t)sorted[i];
  if (c < ncell && count) count[c] = e - b;
}

// the scratch address on the current device, looked up once per device
mpcd_status scratch_ptr(int32_t **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return na_fail(MPCD_EDEVICE, "hipGetDevice");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return na_fail(MPCD_EDEVICE, "hipGetSymbolAddress(scratch) failed");
    g_scratch_dev[dev] = (int32_t *)ptr;
  }
  *out = g_scratch_dev[dev];
  return MPCD_OK;
}

}  // namespace

namespace mpcd {
const char *noalloc_error() { return g_err.c_str(); }
}  // namespace mpcd

extern "C" mpcd_status mpcd_gpu_bin_noalloc(const mpcd_gpu_params *p, int64_t n, const double *q_dev,
                                            const double *shift, const int32_t *prev_perm_dev,
                                            const int32_t *prev_cell_dev, int32_t *cell_dev, int32_t *count_dev,
                                            int32_t *perm_dev, void *stream) {
  if (!p) return na_fail(MPCD_EINVAL, "bin_noalloc: null parameters");
  if (p->dim != 2 && p->dim != 3) return na_fail(MPCD_EINVAL, "bin_noalloc: dim must be 2 or 3");
  if (n <= 0 || n > kNaMaxPart)
    return na_fail(MPCD_EINVAL, "bin_noalloc: n must lie in [1, MPCD_NOALLOC_MAX_PARTICLES = 2097152]");
  if ((prev_perm_dev == nullptr) != (prev_cell_dev == nullptr))
    return na_fail(MPCD_EINVAL, "bin_noalloc: give both the previous order and the previous cells, or neither");
  if (prev_perm_dev && !shift) return na_fail(MPCD_EINVAL, "bin_noalloc: a previous binning needs a shift");
  NaBox B{};
  B.dim = p->dim;
  int64_t ncell = 1;
  for (int i = 0; i < 3; i++) {
    B.L[i] = (i < p->dim) ? p->L[i] : 1;
    if (i < p->dim && B.L[i] < 2) return na_fail(MPCD_EINVAL, "bin_noalloc: box lengths must be at least 2");
    B.s[i] = (shift && i < p->dim) ? shift[i] : 0.0;
  }
  B.ny = B.L[1] + 1;
  B.nz = (p->dim == 3) ? B.L[2] + 1 : 1;
  ncell = (int64_t)(B.L[0] + 1) * B.ny * B.nz;
  if (ncell > kNaMaxCell) return na_fail(MPCD_EINVAL, "bin_noalloc: more than MPCD_NOALLOC_MAX_CELLS = 1000000 cells");
  B.ncell = (int)ncell;
  if (!q_dev || !cell_dev || !perm_dev) return na_fail(MPCD_EINVAL, "bin_noalloc: null device array");
  set_pbc(*p, B.pbc);
  int32_t *sc = nullptr;
  mpcd_status st = scratch_ptr(&sc);
  if (st) return st;
  int32_t *hist = sc + kNaHist, *start = sc + kNaStart, *loc = sc + kNaLoc, *inv = sc + kNaInv;
  int32_t *bsum = sc + kNaBsum, *boff = sc + kNaBoff;
  uint32_t *key = (uint32_t *)(sc + kNaInv);  // the keys reuse the inverse's words once it is read
  uint64_t *sorted = (uint64_t *)(sc + kNaKeys);
  hipStream_t s = (hipStream_t)stream;
  const int N = (int)n, nb = (N + kNaThreads - 1) / kNaThreads;
  const int nh = B.ncell + 1;  // buckets: the cells and the escapes
  // the histogram starts at zero: the last call's counts are still there
  k_na_fill<<<(nh + 1 + kNaThreads - 1) / kNaThreads, kNaThreads, 0, s>>>(hist, 0, nh + 1);
  if (prev_perm_dev) k_na_inverse<<<nb, kNaThreads, 0, s>>>(prev_perm_dev, inv, N);
  // the key of particle i is written after its inverse entry was read (same thread, same word)
  k_na_count<<<nb, kNaThreads, 0, s>>>(B, q_dev, prev_cell_dev, prev_perm_dev ? inv : nullptr, shift ? 1 : 0, hist,
                                       loc, key, cell_dev, N);
  // exclusive scan of nh + 1 values (the last gives the total)
  const int nsb = (nh + 1 + kNaScanBlock - 1) / kNaScanBlock;  // at most 1025
  k_na_scan<<<nsb, kNaThreads, 0, s>>>(hist, start, bsum, nh + 1);
  if (nsb > 1) {
    // nsb <= kNaScanBlock (the static_assert above): one block scans the sums
    k_na_scan<<<1, kNaThreads, 0, s>>>(bsum, boff, nullptr, nsb);
    k_na_scan_add<<<nsb, kNaThreads, 0, s>>>(start, boff, nh + 1);
  }
  k_na_scatter<<<nb, kNaThreads, 0, s>>>(cell_dev, loc, key, start, sorted, B.ncell, N);
  k_na_sort<<<(nh + kNaThreads - 1) / kNaThreads, kNaThreads, 0, s>>>(start, sorted, perm_dev, count_dev, B.ncell);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return na_fail(MPCD_EDEVICE, hipGetErrorString(e));
  return MPCD_OK;
}
