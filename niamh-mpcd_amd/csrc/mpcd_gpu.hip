// mpcd_gpu.hip -- C-ABI of the MPCD solvent library for gfx950 (see
// include/mpcd_gpu.h for the entry points and the reference calls they
// replace).  Host orchestration: device buffers, block and tile tables, the
// per-phase kernel sequences, the fused whole-step sequence and the hipGraph
// used by mpcd_gpu_run.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "mpcd_adjoint.h"
#include "mpcd_bucket.h"
#include "mpcd_exact.h"
#include "mpcd_flow.h"
#include "mpcd_fused.h"
#include "mpcd_kernels.h"

using namespace mpcd;

struct mpcd_gpu_ctx {
  int nrep = 0, dim = 0, lc = 0, rtech = 0, precision = 0, nspec = 1;
  int ntot = 0, nctot = 0, nwtot = 0;
  uint64_t seed = 0;
  uint32_t step = 0;
  std::string err;
  std::vector<mpcd_gpu_params> params;
  std::vector<RepDev> reps;
  std::vector<BlkDev> pblk, cblk;  // slot blocks, cell blocks
  std::vector<TileDev> tiles;      // cell tiles of the fused kernel
  std::vector<TileDev> lct;        // cell tiles of the nematic collision kernel
  std::vector<int2> lct_rng;
  TileDev *d_lct = nullptr;
  int2 *d_lct_rng = nullptr;
  int32_t *d_lct_over = nullptr;  // [1 + tiles]: count, then the tiles above k_collide_lc's capacity
  std::vector<int2> cblk_rng, tile_rng, pblk_rng;
  std::vector<int> uploaded;
  bool need_init = true;
  // binning state: the logical list order is the storage order, or (pending)
  // the unshifted bin implied by the unshifted cells in cu[] (cu_src 0) or in
  // the keys (cu_src 1); hist_ready: the next shifted bin's histogram and
  // keys are already built (by the fused kernel)
  bool pending = false, hist_ready = false;
  int cu_src = 0;
  int cur = 0;
  bool any_therm = false, any_phantom = false, any_lcg = false, fusable = false;
  bool any_curved = false;  // a wall that is not an axis-aligned plane: the split path with the general walls
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // mpcd_gpu_run graphs: two fused steps captured from an even step with the
  // ping-pong buffer `graph_cur` current; the step number is read from d_step
  bool capturing = false;
  uint32_t cap_base = 0;
  int graph_cur = -1;
  uint32_t *d_step = nullptr;
  double tile_fill = 0.45;  // mean tile population as a fraction of the LDS capacity (autotuned)
  // the bucketed layout of mpcd_gpu_run (isotropic operators, no thermostat):
  // tiles, their buckets of bkt_cap slots, the two bucketed states and the
  // per-bucket counts; bcur is the buffer holding the coming step
  bool bkt_on = true, bkt_dirty = true;
  bool in_bucket = false;  // the state at rest is in the bucketed layout
  double bkt_fill = 0.6;  // mean bucket population as a fraction of its capacity (autotuned)
  int nbkt = 0, bcur = 0, bkt_caplim = 0;  // bkt_caplim: usable slots per bucket (0: all)
  int64_t bkt_steps = 0, bkt_overflows = 0;
  size_t bkt_slots = 0;
  uint32_t bkt_step0 = 0;
  std::vector<TileDev> bkts;
  std::vector<BucketRep> brep;
  std::vector<int2> bkt_rng;
  TileDev *d_bkts = nullptr;
  BucketRep *d_brep = nullptr;
  int2 *d_bkt_rng = nullptr;
  BState<double> bd[2]{};
  BState<float> bf[2]{};
  WallList<double> wld{};
  WallList<float> wlf{};
  int32_t *bcnt[2] = {nullptr, nullptr}, *bbase = nullptr, *bstart[2] = {nullptr, nullptr};
  // records whose particle escapes in the next shifted bin, by step parity
  int32_t *esc_idx[2] = {nullptr, nullptr}, *esc_cnt = nullptr;  // esc_cnt [2]
  int esc_cap = 0;
  // the exact-sum FP32 form (mpcd_exact.h): its bucketed records, wall list,
  // escape lists by step parity, and the cell arrays of its dense steps
  bool exact = false;
  XState bx[2]{};
  XList xwl[2]{};                       // wall lists by step parity
  int32_t *xcnt[3] = {nullptr, nullptr, nullptr};  // bucket counts, turning over
  int xci0 = 0;                          // count array of the first bucketed step (bkt_step0)
  int32_t *xshifted = nullptr;           // per particle: re-placed, already in the coming step's frame
  uint32_t x_tables = 0xFFFFFFFFu;      // the step whose shift and wall tables are in its parity slot
  XDense xd{};
  hipGraph_t bgraph = nullptr;
  hipGraphExec_t bgexec = nullptr;
  int bgraph_cur = -1;
  // adjoint: a tape of recorded steps (split path), lambda = dJ/dv per particle
  bool recording = false;
  int tape_cap = 0, tape_len = 0;
  int32_t *tp_id = nullptr, *tp_sp = nullptr, *tp_start = nullptr;
  double *tp_v1 = nullptr, *tp_vcm = nullptr, *tp_a = nullptr, *tp_mass = nullptr, *tp_M = nullptr, *tp_dM = nullptr;
  float *tp_wmul = nullptr;
  double *lam = nullptr, *lam_s = nullptr, *lam1_s = nullptr, *adj_dev = nullptr, *dtheta = nullptr;
  // forward-mode tangent (mpcd_gpu_tangent): 9 slot arrays (q, v, u) per
  // ping-pong buffer, 25 doubles per cell, the torque tangent per slot, and
  // the direction in parameter space
  bool tan_on = false;
  double *tan_part[2] = {nullptr, nullptr};
  double *tan_cell = nullptr, *tan_T = nullptr;
  double tan_dgrav[3] = {0.0, 0.0, 0.0};
  double tan_dact[MPCD_MAXSPEC] = {};
  // device checkpoints: (destination, source, bytes) copies plus the host step state
  struct Checkpoint {
    std::vector<void *> bufs;
    uint32_t step;
    int cur;
    bool pending, hist_ready, need_init;
    int cu_src;
  };
  std::vector<Checkpoint> ckpts;
  // reverse mode of the whole step (mpcd_gpu_step_backward_state): the start
  // state of every recorded step, the context's own state around a backward
  // call, the adjoints in slot order (9 per slot, 25 per cell, the torque 3
  // per slot), in particle order (9 per particle), of the parameters
  // (kNTheta per replica), and the tape of one reverse launch
  std::vector<Checkpoint> rev_steps;
  Checkpoint rev_here;
  int rev_len = 0, rev_cap = 0;
  bool rev_replay = false, rev_bad = false, lin_ok = false;
  double *ra_part = nullptr, *ra_cell = nullptr, *ra_T = nullptr, *ra_lam = nullptr, *ra_in = nullptr;
  double *ra_th = nullptr, *ra_obs = nullptr;
  int2 *rt_par = nullptr;
  double2 *rt_dp = nullptr;
  double *rt_adj = nullptr;
  double **rt_optr = nullptr;
  int *rt_ovar = nullptr, *rt_need = nullptr;
  size_t rt_entries = 0, rt_outs = 0;  // capacities (entries x lanes, outputs x lanes)

  std::vector<void *> allocs;
  RepDev *d_rep = nullptr;
  BlkDev *d_pblk = nullptr, *d_cblk = nullptr;
  TileDev *d_tiles = nullptr;
  int2 *d_cblk_rng = nullptr, *d_tile_rng = nullptr, *d_pblk_rng = nullptr;
  mpcd_gpu_species *d_sp = nullptr;
  mpcd_gpu_wall *d_wall = nullptr;
  Ctx C{};
  double *shift2 = nullptr, *wqs2 = nullptr, *wqb2 = nullptr;  // per step parity
  WallHot *wh2 = nullptr;                                         // per step parity
  double *box2 = nullptr;                                         // per step parity: wall-free boxes
  double *d_celldraw = nullptr;  // SRD cell draws of the current step (fused path, internal generator)
  // running flow field (mpcd_gpu_flow): sums per cell [3 * nctot], the
  // step's accumulator [4 * nctot], steps summed per replica [nrep]
  bool flow_on = false;
  double *d_flow = nullptr, *d_flow_acc = nullptr;
  double *d_flow_part = nullptr;  // bucketed runs: partials per offset class and cell [kFlowCls * 4 * nctot]
  int64_t *d_flow_n = nullptr;
  int32_t *hist2[2] = {nullptr, nullptr};
  State<double> sd[2]{};
  State<float> sf[2]{};
  double *d_avpart = nullptr;
  double *d_netmom = nullptr;  // mpcd_gpu_zero_net_momentum: 4 sums per slot block, then 4 per replica
  double *d_dshift = nullptr, *d_dcell = nullptr, *d_dghost = nullptr, *d_dpart = nullptr, *d_dlc = nullptr;
  // injected escape / re-thermalisation draws of the current step (null
  // arrays: the generator), and whether the step's streaming uses them
  double *d_desc = nullptr, *d_dreth = nullptr;
  bool esc_inj = false, reth_inj = false, stream_inj = false;
  int dlc_stride = 0;
  size_t dlc_cap = 0;
  int lc_stride = 0;
  double *d_q = nullptr, *d_v = nullptr, *d_u = nullptr;
  int32_t *d_spid = nullptr, *d_i32 = nullptr;
  size_t stage_cap = 0;
  std::vector<int32_t *> scan_lv;
  // kernel timing with HIP events on the library stream (mpcd_gpu_profile)
  bool prof = false;
  struct ProfRec {
    int fam;
    hipEvent_t a, b;
  };
  std::vector<ProfRec> prof_recs;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
};

// mean particles per cell above which the split path is used (cells must
// fit the fused kernel's LDS capacity; populations are Poisson distributed)
constexpr double kFusedMaxPPC = 40.0;

// kernel families reported by mpcd_gpu_timing
enum { F_COUNT = 0, F_SCAN, F_SCATTER, F_SORT, F_GATHER, F_PROPS, F_GHOST, F_LC, F_COLLIDE, F_TEMP, F_STREAM, F_FUSED,
       F_BUCKET, F_FLOW, F_N };

static hipEvent_t prof_event(mpcd_gpu_ctx *c) {
  if (c->ev_used == c->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[c->ev_used++];
}

// records an event pair on the library stream around the launches in scope
struct ProfScope {
  mpcd_gpu_ctx *c;
  int fam;
  hipEvent_t a = nullptr;
  ProfScope(mpcd_gpu_ctx *c_, int f) : c(c_), fam(f) {
    if (c->prof && !c->capturing && (a = prof_event(c))) (void)hipEventRecord(a, c->stream);
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = prof_event(c);
    if (!b) return;
    (void)hipEventRecord(b, c->stream);
    c->prof_recs.push_back({fam, a, b});
  }
};

static mpcd_status fail(mpcd_gpu_ctx *c, mpcd_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return st;
}

#define HIPCHK(ctx, call)                                                                           \
  do {                                                                                              \
    hipError_t e_ = (call);                                                                         \
    if (e_ != hipSuccess) return fail(ctx, MPCD_EDEVICE, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// the cell-sorted state for the entry points that read or change it (the
// state at rest may be in the bucketed layout after mpcd_gpu_run)
static mpcd_status ensure_dense(mpcd_gpu_ctx *c);
static mpcd_status zero_tangent(mpcd_gpu_ctx *c);

template <typename T>
static T *dalloc(mpcd_gpu_ctx *c, size_t n) {
  void *p = nullptr;
  if (n == 0) n = 1;
  if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
  c->allocs.push_back(p);
  // zeroed on the library stream: the stream is non-blocking, so a memset on
  // the null stream would not be ordered before the library's own kernels
  if (hipMemsetAsync(p, 0, n * sizeof(T), c->stream) != hipSuccess) return nullptr;
  return (T *)p;
}

static int anneal_stride(const mpcd_gpu_params &p) {
  double effm = 0.0;
  for (int i = 0; i < p.nspec; i++)
    if (p.sp[i].mfpot > effm) effm = p.sp[i].mfpot;
  effm = effm * 0.5 * (double)p.dim;
  int amax = (int)(kMCINT + kMCSLOPE * effm * 1.0 / p.kbt) + 1;
  return 3 + 2 * amax;
}

template <typename Real>
static State<Real> alloc_state(mpcd_gpu_ctx *c, int n, bool lc) {
  State<Real> s{};
  s.x = dalloc<Real>(c, n);
  s.y = dalloc<Real>(c, n);
  s.z = dalloc<Real>(c, n);
  s.vx = dalloc<Real>(c, n);
  s.vy = dalloc<Real>(c, n);
  s.vz = dalloc<Real>(c, n);
  if (lc) {
    s.ux = dalloc<Real>(c, n);
    s.uy = dalloc<Real>(c, n);
    s.uz = dalloc<Real>(c, n);
  }
  s.id = dalloc<int32_t>(c, n);
  s.sp = dalloc<int32_t>(c, n);
  s.cst = dalloc<int32_t>(c, n);
  return s;
}

template <typename Real>
static bool state_ok(const State<Real> &s, bool lc) {
  return s.x && s.y && s.z && s.vx && s.vy && s.vz && s.id && s.sp && s.cst && (!lc || (s.ux && s.uy && s.uz));
}

static double mean_ppc(const RepDev &R) {
  double vol = (double)R.L[0] * R.L[1] * (R.dim == 3 ? R.L[2] : 1);
  return R.npart / vol;
}

static int fused_op(int rtech) { return (rtech == MPCD_MPCAT) ? OP_AT : (rtech == MPCD_RAT) ? OP_RAT : OP_SRD; }

// cells per fused tile for a replica: the mean tile population is `fill`
// times the LDS capacity (0.45 by default), so a chunk rarely needs splitting
static int tile_cells(const RepDev &R, int cap, double fill) {
  int tc = (int)lround(fill * cap / std::max(mean_ppc(R), 1.0));
  return std::max(1, std::min(kTileCells, tc));
}

// the fused kernel's tile table for a given fill fraction (host side)
static void build_tiles(mpcd_gpu_ctx *c, double fill) {
  c->tiles.clear();
  c->tile_rng.clear();
  const int cap = tile_cap(c->precision == MPCD_FP64, fused_op(c->rtech));
  for (int r = 0; r < c->nrep; r++) {
    const RepDev &R = c->reps[r];
    const int t0 = (int)c->tiles.size(), tc = tile_cells(R, cap, fill);
    for (int b = 0; b < R.ncell; b += tc)
      c->tiles.push_back(TileDev{r, R.cell_off + b, R.cell_off + std::min(b + tc, R.ncell), 0});
    c->tile_rng.push_back(make_int2(t0, (int)c->tiles.size()));
  }
  c->tile_fill = fill;
}

extern "C" mpcd_status mpcd_gpu_create(const mpcd_gpu_params *params, int nrep, const int64_t *n_particles,
                                       uint64_t seed, mpcd_gpu_ctx **out) {
  if (!params || !n_particles || !out || nrep <= 0) return MPCD_EINVAL;
  *out = nullptr;
  mpcd_gpu_ctx *c = new mpcd_gpu_ctx();
  c->nrep = nrep;
  c->seed = seed;
  c->dim = params[0].dim;
  c->lc = params[0].lc;
  c->rtech = params[0].rtech;
  c->precision = params[0].precision;
  c->exact = c->precision == MPCD_FP32_EXACT;
  c->nspec = params[0].nspec;
  c->params.assign(params, params + nrep);
  auto bad = [&](mpcd_status st, const char *m) {
    fprintf(stderr, "mpcd_gpu_create: %s\n", m);
    delete c;
    return st;
  };
  int64_t ntot = 0, nctot = 0;
  for (int r = 0; r < nrep; r++) {
    const mpcd_gpu_params &p = params[r];
    if (p.dim != c->dim || p.lc != c->lc || p.rtech != c->rtech || p.precision != c->precision)
      return bad(MPCD_EINVAL, "dim, lc, rtech and precision must agree across replicas");
    if (p.dim != 2 && p.dim != 3) return bad(MPCD_EINVAL, "dim must be 2 or 3");
    if (p.precision != MPCD_FP64 && p.precision != MPCD_FP32 && p.precision != MPCD_FP32_EXACT)
      return bad(MPCD_EINVAL, "precision");
    bool curved_phantom = false;
    for (int w = 0; w < p.nwall; w++) {
      mpcd_gpu_wall W = p.wall[w];
      if (W.P[0] == 0.0 && W.P[1] == 0.0 && W.P[2] == 0.0 && W.P[3] == 0.0)
        for (int i = 0; i < 4; i++) W.P[i] = 1.0;  // a P left zero is a plane
      if (W.PHANTOM && wall_is_general(W)) curved_phantom = true;
    }
    if (p.precision == MPCD_FP32_EXACT && curved_phantom)
      return bad(MPCD_EUNSUPPORTED, "the exact-sum FP32 form takes phantom walls that are axis-aligned planes");
    if (p.precision == MPCD_FP32_EXACT &&
        (p.lc != MPCD_ISOF || p.tstech != MPCD_NOTHERM ||
         !(p.rtech == MPCD_ARBAXIS || p.rtech == MPCD_ORTHAXIS || p.rtech == MPCD_MPCAT || p.rtech == MPCD_RAT)))
      return bad(MPCD_EUNSUPPORTED,
                 "the exact-sum FP32 form covers SRD, MPC-AT and MPC-AT+a without a thermostat or nematic mode");
    if (p.nspec < 1 || p.nspec > MPCD_MAXSPEC || p.nwall < 0 || p.nwall > MPCD_MAXWALL)
      return bad(MPCD_EINVAL, "species or wall count out of range");
    if (memcmp(p.sp, params[0].sp, sizeof(mpcd_gpu_species) * MPCD_MAXSPEC) != 0 || p.nspec != params[0].nspec)
      return bad(MPCD_EINVAL, "the species table must agree across replicas");
    for (int i = 0; i < p.nspec; i++) {
      if (!(p.sp[i].mass > 0.0)) return bad(MPCD_EINVAL, "species mass must be positive");
      // push() (init.c:1261-1293): READ (1) needs a file, anything else exits
      if (!(p.sp[i].vdist == 0 || p.sp[i].vdist == 2 || p.sp[i].vdist == 3 || p.sp[i].vdist == 4))
        return bad(MPCD_EUNSUPPORTED, "species velocity distribution (vDist) outside this path");
    }
    if (p.lc == MPCD_ISOF) {
      if (!(p.rtech == MPCD_ARBAXIS || p.rtech == MPCD_ORTHAXIS || p.rtech == MPCD_MPCAT || p.rtech == MPCD_RAT))
        return bad(MPCD_EUNSUPPORTED, "collision operator outside this path");
    } else if (p.lc != MPCD_LCL && p.lc != MPCD_LCG) {
      return bad(MPCD_EUNSUPPORTED, "liquid-crystal mode outside this path");
    }
    if (!(p.tstech == MPCD_NOTHERM || p.tstech == MPCD_VSC || p.tstech == MPCD_BEREND))
      return bad(MPCD_EUNSUPPORTED, "thermostat outside this path");
    for (int w = 0; w < p.nwall; w++) {
      const mpcd_gpu_wall &W = p.wall[w];
      int nz = 0;
      for (int i = 0; i < 3; i++) nz += fneq(W.A[i], 0.0) ? 1 : 0;
      if (nz == 0) return bad(MPCD_EINVAL, "a wall needs a nonzero A");
      if (W.COLL_TYPE != 1) return bad(MPCD_EUNSUPPORTED, "only rule-based wall collisions are on this path");
    }
    for (int i = 0; i < p.dim; i++)
      if (p.L[i] < 2) return bad(MPCD_EINVAL, "box lengths must be at least 2 cells");
    if (n_particles[r] <= 0) return bad(MPCD_EINVAL, "replica without particles");
    ntot += n_particles[r];
    int nz = (p.dim == 3) ? p.L[2] + 1 : 1;
    nctot += (int64_t)(p.L[0] + 1) * (p.L[1] + 1) * nz;
  }
  if (ntot >= (1ll << 31) - 1 || nctot >= (1ll << 31) - 1) return bad(MPCD_EINVAL, "too many particles or cells");
  c->ntot = (int)ntot;
  c->nctot = (int)nctot;

  int poff = 0, coff = 0, woff = 0;
  std::vector<mpcd_gpu_wall> walls;
  for (int r = 0; r < nrep; r++) {
    const mpcd_gpu_params &p = params[r];
    RepDev R{};
    R.dim = p.dim;
    for (int i = 0; i < 3; i++) R.L[i] = (i < p.dim) ? p.L[i] : 1;
    R.nx = R.L[0] + 1;
    R.ny = R.L[1] + 1;
    R.nz = (p.dim == 3) ? R.L[2] + 1 : 1;
    set_pbc(p, R.pbc);
    R.cell_off = coff;
    R.ncell = R.nx * R.ny * R.nz;
    R.part_off = poff;
    R.npart = (int)n_particles[r];
    R.wall_off = woff;
    R.nwall = p.nwall;
    R.rtech = p.rtech;
    R.lc = p.lc;
    R.tstech = p.tstech;
    R.galinv = p.galinv;
    R.true_sign = p.true_random_sign;
    R.dt = p.dt;
    R.kbt = p.kbt;
    R.cosA = cos(p.rotang);  // in->C, in->S (init.c:1194-1195)
    R.sinA = sin(p.rotang);
    R.tau = p.tau;
    for (int i = 0; i < 3; i++) {
      R.grav[i] = p.grav[i];
      R.avv[i] = p.avv[i];
    }
    R.gravflag = (fneq(p.grav[0], 0.0) || fneq(p.grav[1], 0.0) || fneq(p.grav[2], 0.0)) ? 1 : 0;
    R.gndnst = p.gndnst;
    R.anyphantom = 0;
    R.anycurved = 0;
    R.xwrap = (p.precision == MPCD_FP32_EXACT && walls_are_box_faces(p)) ? 1 : 0;
    for (int w = 0; w < p.nwall; w++) {
      mpcd_gpu_wall W = p.wall[w];
      if (W.P[0] == 0.0 && W.P[1] == 0.0 && W.P[2] == 0.0 && W.P[3] == 0.0)
        for (int i = 0; i < 4; i++) W.P[i] = 1.0;
      walls.push_back(W);
      if (W.PHANTOM) R.anyphantom = 1;
      if (wall_is_general(W)) R.anycurved = 1;
    }
    if (R.anyphantom) c->any_phantom = true;
    if (R.anycurved) c->any_curved = true;
    if (p.tstech != MPCD_NOTHERM || (p.rtech >= 7 && p.rtech <= 17)) c->any_therm = true;
    if (p.lc == MPCD_LCG) c->any_lcg = true;
    c->reps.push_back(R);
    int cb0 = (int)c->cblk.size(), pb0 = (int)c->pblk.size();
    for (int b = 0; b < R.npart; b += kBlock)
      c->pblk.push_back(BlkDev{r, poff + b, poff + std::min(b + kBlock, R.npart), 0});
    c->pblk_rng.push_back(make_int2(pb0, (int)c->pblk.size()));
    for (int b = 0; b < R.ncell; b += kBlock)
      c->cblk.push_back(BlkDev{r, coff + b, coff + std::min(b + kBlock, R.ncell), 0});
    c->cblk_rng.push_back(make_int2(cb0, (int)c->cblk.size()));
    poff += R.npart;
    coff += R.ncell;
    woff += p.nwall;
  }
  c->nwtot = woff;
  build_tiles(c, c->tile_fill);
  // nematic collision tiles: about half the kernel's LDS capacity on average
  for (int r = 0; r < nrep; r++) {
    const RepDev &R = c->reps[r];
    const int t0 = (int)c->lct.size();
    const int tc = std::max(1, std::min(kLcTileCells, (int)lround(0.5 * kLcCap / std::max(mean_ppc(R), 1.0))));
    for (int b = 0; b < R.ncell; b += tc)
      c->lct.push_back(TileDev{r, R.cell_off + b, R.cell_off + std::min(b + tc, R.ncell), 0});
    c->lct_rng.push_back(make_int2(t0, (int)c->lct.size()));
  }
  c->uploaded.assign(nrep, 0);
  c->lc_stride = anneal_stride(params[0]);
  // the fused kernel stages a cell in LDS: dense systems take the split path
  c->fusable = (c->lc == MPCD_ISOF) && !c->any_curved;
  for (const RepDev &R : c->reps)
    if (mean_ppc(R) > kFusedMaxPPC) c->fusable = false;

  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bad(MPCD_EDEVICE, "stream");
  bool ok = true;
  const int nw3 = 3 * std::max(1, woff);
  c->d_rep = dalloc<RepDev>(c, nrep);
  c->d_pblk = dalloc<BlkDev>(c, c->pblk.size());
  c->d_cblk = dalloc<BlkDev>(c, c->cblk.size());
  c->d_tiles = dalloc<TileDev>(c, c->nctot);  // room for the finest tiling
  c->d_cblk_rng = dalloc<int2>(c, nrep);
  c->d_pblk_rng = dalloc<int2>(c, nrep);
  c->d_tile_rng = dalloc<int2>(c, nrep);
  c->d_lct = dalloc<TileDev>(c, c->lct.size());
  c->d_lct_rng = dalloc<int2>(c, nrep);
  c->d_lct_over = dalloc<int32_t>(c, c->lct.size() + 1);
  c->d_sp = dalloc<mpcd_gpu_species>(c, MPCD_MAXSPEC);
  c->d_wall = dalloc<mpcd_gpu_wall>(c, std::max<size_t>(1, walls.size()));
  c->shift2 = dalloc<double>(c, 2 * 3 * nrep);
  c->wqs2 = dalloc<double>(c, 2 * nw3);
  c->wqb2 = dalloc<double>(c, 2 * nw3);
  c->wh2 = dalloc<WallHot>(c, 2 * std::max(1, woff));
  c->box2 = dalloc<double>(c, 2 * 6 * nrep);
  c->d_step = dalloc<uint32_t>(c, 1);
  ok = ok && c->d_step;
  if (c->fusable && fused_op(c->rtech) == OP_SRD) {
    c->d_celldraw = dalloc<double>(c, 4 * (size_t)c->nctot);
    ok = ok && c->d_celldraw;
  }
  c->hist2[0] = dalloc<int32_t>(c, c->nctot + 1);
  c->hist2[1] = dalloc<int32_t>(c, c->nctot + 1);
  if (c->exact) {
    c->xshifted = dalloc<int32_t>(c, c->ntot);
    ok = ok && c->xshifted;
  }
  ok = ok && c->d_rep && c->d_pblk && c->d_cblk && c->d_tiles && c->d_cblk_rng && c->d_pblk_rng && c->d_tile_rng &&
       c->d_sp &&
       c->d_lct && c->d_lct_rng && c->d_lct_over &&
       c->d_wall && c->shift2 && c->wqs2 && c->wqb2 && c->wh2 && c->box2 && c->hist2[0] && c->hist2[1];
  Ctx &C = c->C;
  C.rep = c->d_rep;
  C.sp = c->d_sp;
  C.wall = c->d_wall;
  C.start = dalloc<int32_t>(c, c->nctot + 1 + 1024);
  C.cnew = dalloc<int32_t>(c, c->ntot);
  C.loc = dalloc<int32_t>(c, c->ntot);
  C.cu = dalloc<int32_t>(c, c->ntot);
  C.key = dalloc<uint64_t>(c, c->ntot);
  C.skey = dalloc<uint64_t>(c, c->ntot);
  C.cell.pop = dalloc<int32_t>(c, c->nctot);
  C.cell.mass = dalloc<double>(c, c->nctot);
  C.cell.vcm = dalloc<double>(c, 3 * (size_t)c->nctot);
  C.cell.cm = dalloc<double>(c, 3 * (size_t)c->nctot);
  C.cell.I = dalloc<double>(c, 6 * (size_t)c->nctot);
  C.cell.S = dalloc<double>(c, c->nctot);
  C.cell.dir = dalloc<double>(c, 3 * (size_t)c->nctot);
  C.cell.E = (c->lc != MPCD_ISOF) ? dalloc<double>(c, 9 * (size_t)c->nctot) : nullptr;
  C.cell.avm = (c->lc != MPCD_ISOF) ? dalloc<double>(c, (size_t)c->nctot) : nullptr;
  C.T = (c->lc != MPCD_ISOF) ? dalloc<double>(c, 3 * (size_t)c->ntot) : nullptr;
  C.tpart = dalloc<double>(c, std::max<size_t>(std::max<size_t>(c->cblk.size(), c->nctot), c->pblk.size()));
  C.kbtnow = dalloc<double>(c, nrep);
  C.gscale = dalloc<double>(c, nrep);
  C.sg = dalloc<double>(c, nrep);
  C.events = dalloc<unsigned long long>(c, 4);
  C.flags = dalloc<int>(c, 4);
  C.ntot = c->ntot;
  C.nctot = c->nctot;
  C.nrep = nrep;
  c->d_avpart = (c->lc == MPCD_LCG) ? dalloc<double>(c, 6 * c->pblk.size()) : nullptr;
  ok = ok && C.start && C.cnew && C.loc && C.cu && C.key && C.skey && C.cell.pop && C.cell.mass && C.cell.vcm &&
       C.cell.cm && C.cell.I && C.cell.S && C.cell.dir && C.tpart && C.kbtnow && C.gscale && C.sg && C.events &&
       C.flags;
  bool lcm = c->lc != MPCD_ISOF;
  if (lcm) ok = ok && C.cell.E && C.cell.avm && C.T;
  if (c->any_lcg) ok = ok && c->d_avpart;
  if (c->precision == MPCD_FP64) {
    c->sd[0] = alloc_state<double>(c, c->ntot, lcm);
    c->sd[1] = alloc_state<double>(c, c->ntot, lcm);
    ok = ok && state_ok(c->sd[0], lcm) && state_ok(c->sd[1], lcm);
  } else {
    c->sf[0] = alloc_state<float>(c, c->ntot, lcm);
    c->sf[1] = alloc_state<float>(c, c->ntot, lcm);
    ok = ok && state_ok(c->sf[0], lcm) && state_ok(c->sf[1], lcm);
  }
  if (!ok) {
    mpcd_gpu_destroy(c);
    fprintf(stderr, "mpcd_gpu_create: device allocation failed\n");
    return MPCD_ENOMEM;
  }
  std::vector<double> wq(2 * nw3, 0.0);
  for (int w = 0; w < woff; w++)
    for (int j = 0; j < 3; j++) wq[w * 3 + j] = wq[nw3 + w * 3 + j] = walls[w].Q[j];
  bool cp = true;
  auto up = [&](void *dst, const void *src, size_t bytes) {
    cp = cp && hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream) == hipSuccess;
  };
  up(c->d_rep, c->reps.data(), sizeof(RepDev) * nrep);
  up(c->d_pblk, c->pblk.data(), sizeof(BlkDev) * c->pblk.size());
  up(c->d_cblk, c->cblk.data(), sizeof(BlkDev) * c->cblk.size());
  up(c->d_tiles, c->tiles.data(), sizeof(TileDev) * c->tiles.size());
  up(c->d_cblk_rng, c->cblk_rng.data(), sizeof(int2) * nrep);
  up(c->d_pblk_rng, c->pblk_rng.data(), sizeof(int2) * nrep);
  up(c->d_tile_rng, c->tile_rng.data(), sizeof(int2) * nrep);
  up(c->d_lct, c->lct.data(), sizeof(TileDev) * c->lct.size());
  up(c->d_lct_rng, c->lct_rng.data(), sizeof(int2) * nrep);
  up(c->d_sp, params[0].sp, sizeof(mpcd_gpu_species) * MPCD_MAXSPEC);
  if (woff) up(c->d_wall, walls.data(), sizeof(mpcd_gpu_wall) * woff);
  up(c->wqs2, wq.data(), sizeof(double) * wq.size());
  up(c->wqb2, wq.data(), sizeof(double) * wq.size());
  // host tables are borrowed: wait for the copies (and the zeroing) here
  if (!cp || hipStreamSynchronize(c->stream) != hipSuccess) {
    mpcd_gpu_destroy(c);
    return MPCD_EDEVICE;
  }
  *out = c;
  return MPCD_OK;
}

extern "C" void mpcd_gpu_destroy(mpcd_gpu_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  if (c->bgexec) (void)hipGraphExecDestroy(c->bgexec);
  if (c->bgraph) (void)hipGraphDestroy(c->bgraph);
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

extern "C" const char *mpcd_gpu_last_error(const mpcd_gpu_ctx *c) { return c ? c->err.c_str() : noalloc_error(); }

extern "C" int mpcd_gpu_ncells(const mpcd_gpu_ctx *c, int r) {
  if (!c || r < 0 || r >= c->nrep) return -1;
  return c->reps[r].ncell;
}

extern "C" int mpcd_gpu_lc_stride(const mpcd_gpu_ctx *c) { return c ? c->lc_stride : -1; }

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static int nblocks(int n, int b) { return (n + b - 1) / b; }

static Draws make_draws(mpcd_gpu_ctx *c, int injected, uint32_t step) {
  Draws D{};
  D.rng.seed = c->seed;
  D.rng.step = step;
  D.injected = injected;
  D.shift = c->d_dshift;
  D.cell = c->d_dcell;
  D.ghost = c->d_dghost;
  D.part = c->d_dpart;
  D.lc = c->d_dlc;
  D.lc_stride = c->dlc_stride;
  D.esc = (injected && c->esc_inj) ? c->d_desc : nullptr;
  D.reth = (injected && c->reth_inj) ? c->d_dreth : nullptr;
  D.ntot = c->ntot;
  if (c->capturing) {
    D.step_dev = c->d_step;
    D.step_off = step - c->cap_base;
  }
  return D;
}

static TapeDev tape_at(mpcd_gpu_ctx *c, int k) {
  TapeDev T;
  const size_t n = c->ntot, nc = c->nctot;
  T.id = c->tp_id + k * n;
  T.sp = c->tp_sp + k * n;
  T.start = c->tp_start + k * (nc + 1);
  T.v1 = c->tp_v1 + k * 3 * n;
  T.vcm = c->tp_vcm + k * 3 * nc;
  T.a = c->tp_a + k * nc;
  T.mass = c->tp_mass + k * nc;
  T.M = c->tp_M + k * 9 * nc;
  T.dM = c->tp_dM + k * 9 * nc;
  T.wmul = c->tp_wmul + k * 3 * n;
  return T;
}

// the kernels' view of the tangent arrays of ping-pong buffer `buf`
static TanDev tan_view(mpcd_gpu_ctx *c, int buf) {
  TanDev T{};
  if (!c->tan_part[0]) return T;
  const size_t n = c->ntot, nc = c->nctot;
  for (int k = 0; k < 3; k++) {
    T.q[k] = c->tan_part[buf] + k * n;
    T.v[k] = c->tan_part[buf] + (3 + k) * n;
    T.u[k] = c->tan_part[buf] + (6 + k) * n;
  }
  T.vcm = c->tan_cell;
  T.cm = c->tan_cell + 3 * nc;
  T.I = c->tan_cell + 6 * nc;
  T.S = c->tan_cell + 12 * nc;
  T.dir = c->tan_cell + 13 * nc;
  T.E = c->tan_cell + 16 * nc;
  T.T = c->tan_T;
  for (int k = 0; k < 3; k++) T.dgrav[k] = c->tan_dgrav[k];
  for (int k = 0; k < MPCD_MAXSPEC; k++) T.dact[k] = c->tan_dact[k];
  return T;
}

// the kernels' view of the context for the step with parity `par`
static Ctx view(mpcd_gpu_ctx *c, int par) {
  Ctx C = c->C;
  const int nw3 = 3 * std::max(1, c->nwtot);
  C.shift = c->shift2 + par * 3 * c->nrep;
  C.wq_s = c->wqs2 + par * nw3;
  C.wq_b = c->wqb2 + par * nw3;
  C.wh = c->wh2 + par * std::max(1, c->nwtot);
  C.hist = c->hist2[par];
  return C;
}

static mpcd_status stage_draws(mpcd_gpu_ctx *c, const mpcd_draws *d) {
  if (!d) return MPCD_OK;
  if (!c->d_dshift) {
    c->d_dshift = dalloc<double>(c, 3 * c->nrep);
    c->d_dcell = dalloc<double>(c, 4 * (size_t)c->nctot);
    c->d_dghost = dalloc<double>(c, 3 * (size_t)c->nctot);
    c->d_dpart = dalloc<double>(c, 3 * (size_t)c->ntot);
    if (!c->d_dshift || !c->d_dcell || !c->d_dghost || !c->d_dpart) return fail(c, MPCD_ENOMEM, "draw staging");
  }
  if (d->shift) HIPCHK(c, hipMemcpyAsync(c->d_dshift, d->shift, sizeof(double) * 3 * c->nrep, hipMemcpyHostToDevice, c->stream));
  if (d->cell) HIPCHK(c, hipMemcpyAsync(c->d_dcell, d->cell, sizeof(double) * 4 * (size_t)c->nctot, hipMemcpyHostToDevice, c->stream));
  if (d->ghost) HIPCHK(c, hipMemcpyAsync(c->d_dghost, d->ghost, sizeof(double) * 3 * (size_t)c->nctot, hipMemcpyHostToDevice, c->stream));
  if (d->part) HIPCHK(c, hipMemcpyAsync(c->d_dpart, d->part, sizeof(double) * 3 * (size_t)c->ntot, hipMemcpyHostToDevice, c->stream));
  if (d->lc) {
    if (d->lc_stride <= 0) return fail(c, MPCD_EINVAL, "lc_stride");
    size_t need = (size_t)d->lc_stride * c->ntot;
    if (need > c->dlc_cap) {
      c->d_dlc = dalloc<double>(c, need);
      if (!c->d_dlc) return fail(c, MPCD_ENOMEM, "lc draw staging");
      c->dlc_cap = need;
    }
    c->dlc_stride = d->lc_stride;
    HIPCHK(c, hipMemcpyAsync(c->d_dlc, d->lc, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
  }
  c->esc_inj = d->escape != nullptr;
  if (d->escape) {
    const size_t need = 2 * (size_t)c->ntot * kEscStride;
    if (!c->d_desc && !(c->d_desc = dalloc<double>(c, need))) return fail(c, MPCD_ENOMEM, "escape draw staging");
    HIPCHK(c, hipMemcpyAsync(c->d_desc, d->escape, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
  }
  c->reth_inj = d->retherm != nullptr;
  if (d->retherm) {
    const size_t need = (size_t)c->ntot * kRethStride;
    if (!c->d_dreth && !(c->d_dreth = dalloc<double>(c, need))) return fail(c, MPCD_ENOMEM, "retherm draw staging");
    HIPCHK(c, hipMemcpyAsync(c->d_dreth, d->retherm, sizeof(double) * need, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));  // host arrays are borrowed only for the call
  return MPCD_OK;
}

static void scan_fixed(mpcd_gpu_ctx *c, int32_t *in, int32_t *out, int n, size_t level) {
  int nb = nblocks(n + 1, 1024);
  if (nb == 1) {
    k_scan_block<<<1, 256, 0, c->stream>>>(in, out, nullptr, n + 1);
    return;
  }
  int32_t *bsum = c->scan_lv[2 * level], *boff = c->scan_lv[2 * level + 1];
  k_scan_block<<<nb, 256, 0, c->stream>>>(in, out, bsum, n + 1);
  scan_fixed(c, bsum, boff, nb - 1, level + 1);
  k_scan_add<<<nb, 256, 0, c->stream>>>(out, boff, n + 1);
}

static bool alloc_scan(mpcd_gpu_ctx *c) {
  if (!c->scan_lv.empty() || c->nctot + 1 <= 1024) return true;
  int n = c->nctot;
  while (true) {
    int nb = nblocks(n + 1, 1024);
    if (nb == 1) break;
    int32_t *a = dalloc<int32_t>(c, nb + 1), *b = dalloc<int32_t>(c, nb + 1);
    if (!a || !b) return false;
    c->scan_lv.push_back(a);
    c->scan_lv.push_back(b);
    n = nb - 1;
  }
  return true;
}

template <typename Real>
static State<Real> *states(mpcd_gpu_ctx *c);
template <>
State<double> *states<double>(mpcd_gpu_ctx *c) { return c->sd; }
template <>
State<float> *states<float>(mpcd_gpu_ctx *c) { return c->sf; }

// grid shift draw of step `step` and the wall positions it implies; reads
// the previous step's shifted-back walls, writes this step's parity slot
static void launch_shift(mpcd_gpu_ctx *c, uint32_t step, int injected) {
  const int par = step & 1, nw3 = 3 * std::max(1, c->nwtot);
  Draws D = make_draws(c, injected, step);
  k_shift<<<nblocks(c->nrep, 64), 64, 0, c->stream>>>(c->C, D, 1, c->wqb2 + (par ^ 1) * nw3,
                                                        c->shift2 + par * 3 * c->nrep, c->wqs2 + par * nw3,
                                                        c->wqb2 + par * nw3, c->wh2 + par * std::max(1, c->nwtot),
                                                        c->box2 + par * 6 * c->nrep);
}

// binning: mode 0 initial, 1 shifted (this step), 2 materialise the pending
// unshifted bin.  gather = false only builds the sorted keys.
static void rev_save_step(mpcd_gpu_ctx *c, int injected);

template <typename Real>
static void run_bin(mpcd_gpu_ctx *c, int mode, bool gather, int injected) {
  // a recorded step starts with its shifted bin: keep its start state
  if (mode == 1 && c->recording && !c->rev_replay) rev_save_step(c, injected);
  State<Real> *S = states<Real>(c);
  const int par = c->step & 1;
  Ctx C = view(c, par);
  Draws D = make_draws(c, injected, c->step);
  {
    ProfScope ps(c, F_COUNT);
    k_fill<<<nblocks(c->nctot + 1, kBlock), kBlock, 0, c->stream>>>(C.hist, 0, c->nctot + 1);
    if (mode == 1) launch_shift(c, c->step, injected);
    k_count<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, mode, c->pending ? 1 : 0,
                                                             c->cu_src);
  }
  c->hist_ready = false;
  {
    ProfScope ps(c, F_SCAN);
    scan_fixed(c, C.hist, C.start, c->nctot, 0);
  }
  {
    ProfScope ps(c, F_SCATTER);
    k_scatter<<<nblocks(c->ntot, kBlock), kBlock, 0, c->stream>>>(C);
  }
  {
    ProfScope ps(c, F_SORT);
    k_sortcells_t<<<c->lct.size(), kLcThreads, 0, c->stream>>>(C, S[1 - c->cur].cst, c->d_lct);
  }
  if (gather) {
    ProfScope ps(c, F_GATHER);
    k_gather<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], S[1 - c->cur], c->d_pblk, mode == 1 ? 1 : 0,
                                                              c->lc != MPCD_ISOF, D);
    if (c->tan_on)
      k_gather_tan<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], tan_view(c, c->cur),
                                                                    tan_view(c, 1 - c->cur), c->d_pblk,
                                                                    mode == 1 ? 1 : 0, c->lc != MPCD_ISOF);
    c->cur = 1 - c->cur;
    c->pending = false;
  }
}

template <typename Real>
static void run_props(mpcd_gpu_ctx *c, int do_acc) {
  State<Real> *S = states<Real>(c);
  Ctx C = view(c, c->step & 1);
  // the cell decisions are those of the collision's localPROP
  if (!do_acc) C.cdec = nullptr;
  else if (C.cdec) k_fill<<<nblocks(c->nctot, kBlock), kBlock, 0, c->stream>>>((int32_t *)C.cdec, 0, c->nctot);
  ProfScope ps(c, F_PROPS);
  const bool tan = std::is_same<Real, double>::value && c->tan_on;
  const TanDev T = tan ? tan_view(c, c->cur) : TanDev{};
  if constexpr (std::is_same<Real, double>::value)
    if (tan) {
      k_cellprops<Real, Dual><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_cblk, do_acc, T);
    }
  if (!tan) {
    if (c->dim == 3)
      k_cellprops_t<Real, 3><<<c->lct.size(), kLcThreads, 0, c->stream>>>(C, S[c->cur], c->d_lct, do_acc);
    else
      k_cellprops_t<Real, 2><<<c->lct.size(), kLcThreads, 0, c->stream>>>(C, S[c->cur], c->d_lct, do_acc);
  }
  if (c->lc != MPCD_ISOF) {
    k_velgrad_bulk<<<c->cblk.size(), kBlock, 0, c->stream>>>(C, c->d_cblk);
    k_velgrad_edges<<<c->nrep, 256, 0, c->stream>>>(C);
    if (tan) {
      // the velocity gradient is linear in the cell velocities: the same
      // stencil maps their tangents to its tangent
      Ctx Ct = C;
      Ct.cell.vcm = T.vcm;
      Ct.cell.E = T.E;
      k_velgrad_bulk<<<c->cblk.size(), kBlock, 0, c->stream>>>(Ct, c->d_cblk);
      k_velgrad_edges<<<c->nrep, 256, 0, c->stream>>>(Ct);
    }
  }
}

template <typename Real>
static void run_collide(mpcd_gpu_ctx *c, int injected) {
  State<Real> *S = states<Real>(c);
  Ctx C = view(c, c->step & 1);
  Draws D = make_draws(c, injected, c->step);
  run_props<Real>(c, 1);
  const bool tan = std::is_same<Real, double>::value && c->tan_on;
  const TanDev T = tan ? tan_view(c, c->cur) : TanDev{};
  // the nematic operators run on cell tiles staged in LDS (the tangent keeps
  // the lane-per-cell kernel, whose values are the same)
  const bool lc_tiled = c->lc != MPCD_ISOF && !tan;
  if (c->any_phantom) {
    ProfScope ps(c, F_GHOST);
    if constexpr (std::is_same<Real, double>::value)
      if (tan) {
        k_ghost<Real, Dual><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_cblk, D, T);
      }
    if (!tan && c->any_curved)
      k_ghost<Real, double, true><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_cblk, D, T);
    else if (!tan)
      k_ghost<Real, double><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_cblk, D, T);
  }
  if (c->recording && c->tape_len < c->tape_cap)
    k_tape_cells<Real><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_cblk, D, tape_at(c, c->tape_len));
  if (c->lc != MPCD_ISOF) {
    ProfScope ps(c, F_LC);
    if (c->any_lcg) {
      k_avorder_part<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, c->d_avpart);
      k_avorder_final<<<1, 64, 0, c->stream>>>(C, c->d_pblk, (int)c->pblk.size(), c->d_avpart);
    }
    if constexpr (std::is_same<Real, double>::value)
      if (tan) {
        k_lc<Real, Dual><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, D, T);
      }
    // the orientation step computes in the state's precision
    using FL = typename std::conditional<std::is_same<Real, float>::value, float, double>::type;
    if (!tan) k_lc<Real, FL><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, D, T);
  }
  {
    ProfScope ps(c, F_COLLIDE);
    if constexpr (std::is_same<Real, double>::value)
      if (tan) {
        k_collide<Real, Dual><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_cblk, D, T);
      }
    if (!tan) {
      if (lc_tiled) {
        if (c->dim == 3)
          k_collide_lc<Real, 3><<<c->lct.size(), kLcThreads, 0, c->stream>>>(C, S[c->cur], c->d_lct, D,
                                                                             c->d_lct_over);
        else
          k_collide_lc<Real, 2><<<c->lct.size(), kLcThreads, 0, c->stream>>>(C, S[c->cur], c->d_lct, D,
                                                                             c->d_lct_over);
        k_collide_lc_over<Real><<<1, kLcThreads, 0, c->stream>>>(C, S[c->cur], c->d_lct, D, c->d_lct_over);
      } else
        k_collide<Real, double><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_cblk, D, T);
    }
  }
  ProfScope ps(c, F_TEMP);
  k_temp_reduce<<<c->nrep, kTempThreads, 0, c->stream>>>(C, lc_tiled ? c->d_lct_rng : c->d_cblk_rng);
  if (c->any_therm) {
    k_scale<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk);
    k_temp_final<<<c->nrep, kBlock, 0, c->stream>>>(C, c->d_pblk_rng);
  }
}

// the running flow field after a completed step (mpc.c:4946-4951): the
// sums of the unshifted cells, then FLOW += their quotient
static void flow_fold(mpcd_gpu_ctx *c) {
  const int nb = nblocks(std::max(c->nctot, c->nrep), kBlock);
  if (c->dim == 3)
    k_flow_fold<3><<<nb, kBlock, 0, c->stream>>>(c->C.flags, c->nctot, c->nrep, c->d_flow_acc, c->d_flow, c->d_flow_n);
  else
    k_flow_fold<2><<<nb, kBlock, 0, c->stream>>>(c->C.flags, c->nctot, c->nrep, c->d_flow_acc, c->d_flow, c->d_flow_n);
}

template <typename Real>
static void flow_after_dense(mpcd_gpu_ctx *c) {
  if (!c->flow_on) return;
  ProfScope ps(c, F_FLOW);
  State<Real> *S = states<Real>(c);
  if (c->dim == 3)
    k_flow_dense<Real, 3><<<c->pblk.size(), kBlock, 0, c->stream>>>(c->C, S[c->cur], c->d_pblk, c->cu_src,
                                                                   c->d_flow_acc);
  else
    k_flow_dense<Real, 2><<<c->pblk.size(), kBlock, 0, c->stream>>>(c->C, S[c->cur], c->d_pblk, c->cu_src,
                                                                   c->d_flow_acc);
  flow_fold(c);
}

template <typename Real>
static void run_stream(mpcd_gpu_ctx *c, int injected) {
  State<Real> *S = states<Real>(c);
  Ctx C = view(c, c->step & 1);
  Draws D = make_draws(c, injected, c->step);
  ProfScope ps(c, F_STREAM);
  const bool tape = c->recording && c->tape_len < c->tape_cap;
  float *wmul = tape ? tape_at(c, c->tape_len).wmul : nullptr;
  const double *box = c->box2 + (c->step & 1) * 6 * c->nrep;  // this step's wall-free boxes
  const bool tan = std::is_same<Real, double>::value && c->tan_on;
  if constexpr (std::is_same<Real, double>::value)
    if (tan) {
      k_stream<Real, Dual><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, D, wmul,
                                                                     tan_view(c, c->cur), box);
    }
  if (!tan && c->any_curved)
    k_stream_walls<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, D);
  else if (!tan)
    k_stream<Real, double><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, D, wmul, TanDev{}, box);
  if (tape) c->tape_len++;
  c->pending = true;
  c->cu_src = 0;
  c->hist_ready = false;
  flow_after_dense<Real>(c);
  c->step++;
}

template <typename Real, int OP, bool ONESP, int DIM>
static void launch_fused(mpcd_gpu_ctx *c, const Ctx &C, State<Real> *S, const Draws &D, int par, int do_next,
                         int do_stream) {
  const int nw3 = 3 * std::max(1, c->nwtot);
  k_fused<Real, OP, ONESP, DIM><<<c->tiles.size(), kTileThreads, 0, c->stream>>>(
      C, S[c->cur], S[1 - c->cur], c->d_tiles, c->d_rep, c->d_sp, c->d_wall, c->wh2 + par * std::max(1, c->nwtot),
      c->box2 + par * 6 * c->nrep, D, c->shift2 + par * 3 * c->nrep, c->wqb2 + par * nw3,
      c->shift2 + (par ^ 1) * 3 * c->nrep, c->hist2[par ^ 1], do_next, do_stream);
}

template <typename Real, int OP, bool ONESP>
static void launch_fused_dim(mpcd_gpu_ctx *c, const Ctx &C, State<Real> *S, const Draws &D, int par, int do_next,
                             int do_stream) {
  if (c->dim == 3) launch_fused<Real, OP, ONESP, 3>(c, C, S, D, par, do_next, do_stream);
  else launch_fused<Real, OP, ONESP, 2>(c, C, S, D, par, do_next, do_stream);
}

template <typename Real>
static void dispatch_fused(mpcd_gpu_ctx *c, const Ctx &C, State<Real> *S, const Draws &D, int par, int do_next,
                           int do_stream) {
  const bool one = c->nspec == 1;
  const int op = fused_op(c->rtech);
  if (op == OP_SRD) {
    if (one) launch_fused_dim<Real, OP_SRD, true>(c, C, S, D, par, do_next, do_stream);
    else launch_fused_dim<Real, OP_SRD, false>(c, C, S, D, par, do_next, do_stream);
  } else if (op == OP_AT) {
    if (one) launch_fused_dim<Real, OP_AT, true>(c, C, S, D, par, do_next, do_stream);
    else launch_fused_dim<Real, OP_AT, false>(c, C, S, D, par, do_next, do_stream);
  } else {
    if (one) launch_fused_dim<Real, OP_RAT, true>(c, C, S, D, par, do_next, do_stream);
    else launch_fused_dim<Real, OP_RAT, false>(c, C, S, D, par, do_next, do_stream);
  }
}

// one whole step through the fused kernel (isotropic operators).  With the
// internal generator the next step's shift is drawn first, so the kernel
// also builds the next step's histogram and keys.
template <typename Real>
static void run_step_fused(mpcd_gpu_ctx *c, int injected) {
  State<Real> *S = states<Real>(c);
  const int par = c->step & 1;
  Ctx C = view(c, par);
  Draws D = make_draws(c, injected, c->step);
  // a prebuilt histogram used the internal shift: injected draws recount
  if (!c->hist_ready || injected) {
    ProfScope ps(c, F_COUNT);
    k_fill<<<nblocks(c->nctot + 1, kBlock), kBlock, 0, c->stream>>>(C.hist, 0, c->nctot + 1);
    launch_shift(c, c->step, injected);
    k_count<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S[c->cur], c->d_pblk, 1, c->pending ? 1 : 0,
                                                             c->cu_src);
  }
  const int do_next = injected ? 0 : 1;
  const int do_stream = c->any_therm ? 0 : 1;
  {
    ProfScope ps(c, F_SCAN);
    if (do_next) launch_shift(c, c->step + 1, 0);
    scan_fixed(c, C.hist, C.start, c->nctot, 0);
  }
  if (c->d_celldraw && !injected) {
    // the fused kernel reads the SRD cell draws from the keyed array
    ProfScope ps(c, F_COLLIDE);
    k_celldraws<<<c->cblk.size(), kBlock, 0, c->stream>>>(C, c->d_cblk, D, c->d_celldraw);
    D.cell = c->d_celldraw;
  }
  {
    ProfScope ps(c, F_SCATTER);
    k_scatter<<<nblocks(c->ntot, kBlock), kBlock, 0, c->stream>>>(C);
    if (do_next) k_fill<<<nblocks(c->nctot + 1, kBlock), kBlock, 0, c->stream>>>(c->hist2[par ^ 1], 0, c->nctot + 1);
  }
  {
    ProfScope ps(c, F_FUSED);
    dispatch_fused<Real>(c, C, S, D, par, do_next, do_stream);
  }
  c->cur = 1 - c->cur;
  {
    ProfScope ps(c, F_TEMP);
    k_temp_reduce<<<c->nrep, kTempThreads, 0, c->stream>>>(C, c->d_tile_rng);
  }
  if (!do_stream) {
    ProfScope ps(c, F_STREAM);
    const int nw3 = 3 * std::max(1, c->nwtot);
    const int nwm = std::max(1, c->nwtot);
    if (c->dim == 3)
      k_stream_next<Real, 3><<<c->pblk.size(), kBlock, 0, c->stream>>>(
          C, S[c->cur], c->d_pblk, c->d_rep, c->d_sp, c->d_wall, c->wh2 + par * nwm, c->box2 + par * 6 * c->nrep, D,
          c->wqb2 + par * nw3, c->shift2 + (par ^ 1) * 3 * c->nrep, c->hist2[par ^ 1], do_next);
    else
      k_stream_next<Real, 2><<<c->pblk.size(), kBlock, 0, c->stream>>>(
          C, S[c->cur], c->d_pblk, c->d_rep, c->d_sp, c->d_wall, c->wh2 + par * nwm, c->box2 + par * 6 * c->nrep, D,
          c->wqb2 + par * nw3, c->shift2 + (par ^ 1) * 3 * c->nrep, c->hist2[par ^ 1], do_next);
    k_temp_final<<<c->nrep, kBlock, 0, c->stream>>>(C, c->d_pblk_rng);
  }
  c->pending = true;
  c->cu_src = do_next ? 1 : 0;
  c->hist_ready = do_next != 0;
  flow_after_dense<Real>(c);
  c->step++;
}

static void run_step_xdense(mpcd_gpu_ctx *c);

template <typename Real>
static void run_step(mpcd_gpu_ctx *c, int injected) {
  if (c->exact) {
    run_step_xdense(c);
    return;
  }
  if (c->fusable && !c->recording && !c->tan_on) {
    run_step_fused<Real>(c, injected);
    return;
  }
  run_bin<Real>(c, 1, true, injected);
  run_collide<Real>(c, injected);
  run_stream<Real>(c, injected);
}

static mpcd_status check_launch(mpcd_gpu_ctx *c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, MPCD_EDEVICE, "kernel launch: %s", hipGetErrorString(e));
  return MPCD_OK;
}

static mpcd_status check_flags(mpcd_gpu_ctx *c) {
  int flags[4];
  HIPCHK(c, hipMemcpyAsync(flags, c->C.flags, sizeof flags, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (flags[0] & (kFlagEscape | kFlagPushes | kFlagStuck)) {
    // reported once: the bits are cleared
    const int keep = flags[0] & ~(kFlagEscape | kFlagPushes | kFlagStuck);
    k_fill<<<1, 1, 0, c->stream>>>(c->C.flags, keep, 1);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (flags[0] & kFlagEscape)
      return fail(c, MPCD_EINVAL,
                  "escape repair found no position clear of the walls (%d tries with injected draws, %d with the "
                  "generator)", (kEscStride - kEscPos0) / c->dim, kEscMaxTries);
    if (flags[0] & kFlagPushes)
      return fail(c, MPCD_EINVAL, "a particle needed more than %d wall re-thermalisations in one step (injected draws)",
                  kRethStride / 3);
    return fail(c, MPCD_EINVAL, "a particle was re-thermalised 1000 times in one step: the walls leave it no room");
  }
  if (flags[0] & 2) return fail(c, MPCD_EINVAL, "initial positions outside the box");
  if (flags[0] & 1)
    return fail(c, MPCD_ECAPACITY, "a cell holds more than %d particles (fused-kernel LDS capacity)",
                tile_cap(c->precision == MPCD_FP64, fused_op(c->rtech)));
  return MPCD_OK;
}

static mpcd_status ensure_init(mpcd_gpu_ctx *c) {
  if (!c->need_init) return MPCD_OK;
  for (int r = 0; r < c->nrep; r++)
    if (!c->uploaded[r]) return fail(c, MPCD_EINVAL, "replica %d has no state (mpcd_gpu_upload)", r);
  if (!alloc_scan(c)) return fail(c, MPCD_ENOMEM, "scan scratch");
  if (c->precision == MPCD_FP64) {
    run_bin<double>(c, 0, true, 0);
    run_props<double>(c, 0);
  } else {
    run_bin<float>(c, 0, true, 0);
    run_props<float>(c, 0);
  }
  mpcd_status st = check_launch(c);
  if (st) return st;
  st = check_flags(c);
  if (st) return st;
  c->need_init = false;
  return MPCD_OK;
}

static mpcd_status ensure_stage(mpcd_gpu_ctx *c, size_t n) {
  if (n <= c->stage_cap) return MPCD_OK;
  c->d_q = dalloc<double>(c, 3 * n);
  c->d_v = dalloc<double>(c, 3 * n);
  c->d_u = dalloc<double>(c, 3 * n);
  c->d_spid = dalloc<int32_t>(c, n);
  c->d_i32 = dalloc<int32_t>(c, n + 1);
  if (!c->d_q || !c->d_v || !c->d_u || !c->d_spid || !c->d_i32) return fail(c, MPCD_ENOMEM, "staging");
  c->stage_cap = n;
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_upload(mpcd_gpu_ctx *c, int r, const double *Q, const double *V, const double *U,
                                       const int32_t *spid) {
  if (!c || r < 0 || r >= c->nrep || !Q || !V) return c ? fail(c, MPCD_EINVAL, "upload: bad argument") : MPCD_EINVAL;
  {
    mpcd_status st0 = ensure_dense(c);
    if (st0) return st0;
  }
  const RepDev &R = c->reps[r];
  mpcd_status st = ensure_stage(c, R.npart);
  if (st) return st;
  size_t n = R.npart;
  if (spid)
    for (size_t i = 0; i < n; i++)
      if (spid[i] < 0 || spid[i] >= c->params[r].nspec) return fail(c, MPCD_EINVAL, "species id out of range");
  for (size_t i = 0; i < 3 * n; i++)
    if (!isfinite(Q[i]) || !isfinite(V[i])) return fail(c, MPCD_EINVAL, "non-finite position or velocity");
  HIPCHK(c, hipMemcpyAsync(c->d_q, Q, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_v, V, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  if (U) HIPCHK(c, hipMemcpyAsync(c->d_u, U, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  if (spid) HIPCHK(c, hipMemcpyAsync(c->d_spid, spid, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
  if (c->precision == MPCD_FP64)
    k_stage<double><<<nblocks(R.npart, kBlock), kBlock, 0, c->stream>>>(c->sd[c->cur], R.part_off, R.npart, c->d_q,
                                                                      c->d_v, U ? c->d_u : nullptr,
                                                                      spid ? c->d_spid : nullptr);
  else
    k_stage<float><<<nblocks(R.npart, kBlock), kBlock, 0, c->stream>>>(c->sf[c->cur], R.part_off, R.npart, c->d_q,
                                                                     c->d_v, U ? c->d_u : nullptr,
                                                                     spid ? c->d_spid : nullptr);
  {
    mpcd_status zs = zero_tangent(c);  // a new state has no derivative yet
    if (zs) return zs;
  }
  if (c->xshifted)  // a new state has no re-placed particle
    k_fill<<<nblocks(R.npart, kBlock), kBlock, 0, c->stream>>>(c->xshifted + R.part_off, 0, R.npart);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->uploaded[r] = 1;
  c->need_init = true;
  c->pending = false;
  c->hist_ready = false;
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_download(mpcd_gpu_ctx *c, int r, double *Q, double *V, double *U) {
  if (!c || r < 0 || r >= c->nrep || !Q || !V) return c ? fail(c, MPCD_EINVAL, "download: bad argument") : MPCD_EINVAL;
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  const RepDev &R = c->reps[r];
  st = ensure_stage(c, R.npart);
  if (st) return st;
  size_t n = R.npart;
  int lc = c->lc != MPCD_ISOF;
  if (c->precision == MPCD_FP64)
    k_unsort<double><<<nblocks(R.npart, kBlock), kBlock, 0, c->stream>>>(c->sd[c->cur], R.part_off, R.npart, c->d_q,
                                                                       c->d_v, c->d_u, lc);
  else
    k_unsort<float><<<nblocks(R.npart, kBlock), kBlock, 0, c->stream>>>(c->sf[c->cur], R.part_off, R.npart, c->d_q,
                                                                      c->d_v, c->d_u, lc);
  HIPCHK(c, hipMemcpyAsync(Q, c->d_q, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(V, c->d_v, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  if (U) {
    if (lc) HIPCHK(c, hipMemcpyAsync(U, c->d_u, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    else memset(U, 0, sizeof(double) * 3 * n);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

// the records of an array of structures (the reference's particleMPC): base
// + i * stride + offset holds particle i's field
static bool aos_fields_ok(int64_t stride, int64_t off, int64_t bytes) { return off < 0 || off + bytes <= stride; }

extern "C" mpcd_status mpcd_gpu_upload_aos(mpcd_gpu_ctx *c, int r, const void *base, int64_t stride, int64_t off_q,
                                           int64_t off_v, int64_t off_u, int64_t off_spid) {
  if (!c || r < 0 || r >= c->nrep || !base || stride <= 0 || off_q < 0 || off_v < 0 ||
      !aos_fields_ok(stride, off_q, 24) || !aos_fields_ok(stride, off_v, 24) || !aos_fields_ok(stride, off_u, 24) ||
      !aos_fields_ok(stride, off_spid, 4))
    return c ? fail(c, MPCD_EINVAL, "upload_aos: bad argument") : MPCD_EINVAL;
  const size_t n = c->reps[r].npart;
  std::vector<double> Q(3 * n), V(3 * n), U(off_u >= 0 ? 3 * n : 0);
  std::vector<int32_t> sp(off_spid >= 0 ? n : 0);
  const char *b = (const char *)base;
  for (size_t i = 0; i < n; i++) {
    const char *rec = b + i * (size_t)stride;
    memcpy(&Q[3 * i], rec + off_q, 24);
    memcpy(&V[3 * i], rec + off_v, 24);
    if (off_u >= 0) memcpy(&U[3 * i], rec + off_u, 24);
    if (off_spid >= 0) memcpy(&sp[i], rec + off_spid, 4);
  }
  return mpcd_gpu_upload(c, r, Q.data(), V.data(), off_u >= 0 ? U.data() : nullptr,
                         off_spid >= 0 ? sp.data() : nullptr);
}

extern "C" mpcd_status mpcd_gpu_download_aos(mpcd_gpu_ctx *c, int r, void *base, int64_t stride, int64_t off_q,
                                             int64_t off_v, int64_t off_u) {
  if (!c || r < 0 || r >= c->nrep || !base || stride <= 0 || off_q < 0 || off_v < 0 ||
      !aos_fields_ok(stride, off_q, 24) || !aos_fields_ok(stride, off_v, 24) || !aos_fields_ok(stride, off_u, 24))
    return c ? fail(c, MPCD_EINVAL, "download_aos: bad argument") : MPCD_EINVAL;
  const size_t n = c->reps[r].npart;
  std::vector<double> Q(3 * n), V(3 * n), U(3 * n);
  mpcd_status st = mpcd_gpu_download(c, r, Q.data(), V.data(), U.data());
  if (st) return st;
  char *b = (char *)base;
  for (size_t i = 0; i < n; i++) {
    char *rec = b + i * (size_t)stride;
    memcpy(rec + off_q, &Q[3 * i], 24);
    memcpy(rec + off_v, &V[3 * i], 24);
    if (off_u >= 0) memcpy(rec + off_u, &U[3 * i], 24);
  }
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_bin(mpcd_gpu_ctx *c, const mpcd_draws *draws, int shifted, int32_t *cell_out,
                                    int32_t *count_out, int32_t *perm_out) {
  if (!c) return MPCD_EINVAL;
  if (c->exact) return fail(c, MPCD_EUNSUPPORTED, "the exact-sum FP32 form steps whole steps only");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  st = stage_draws(c, draws);
  if (st) return st;
  if (shifted) c->stream_inj = draws != nullptr;
  if (shifted == 0 && !c->pending) {
    // already materialised
  } else if (c->precision == MPCD_FP64) {
    run_bin<double>(c, shifted ? 1 : 2, true, draws != nullptr);
  } else {
    run_bin<float>(c, shifted ? 1 : 2, true, draws != nullptr);
  }
  st = check_launch(c);
  if (st) return st;
  if (cell_out || count_out || perm_out) {
    const int32_t *id = (c->precision == MPCD_FP64) ? c->sd[c->cur].id : c->sf[c->cur].id;
    const int32_t *cst = (c->precision == MPCD_FP64) ? c->sd[c->cur].cst : c->sf[c->cur].cst;
    std::vector<int32_t> hid(c->ntot), hcst(c->ntot), hstart(c->nctot + 1);
    HIPCHK(c, hipMemcpyAsync(hid.data(), id, sizeof(int32_t) * c->ntot, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hcst.data(), cst, sizeof(int32_t) * c->ntot, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hstart.data(), c->C.start, sizeof(int32_t) * (c->nctot + 1), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (count_out)
      for (int i = 0; i < c->nctot; i++) count_out[i] = hstart[i + 1] - hstart[i];
    for (int r = 0; r < c->nrep; r++) {
      const RepDev &R = c->reps[r];
      for (int t = R.part_off; t < R.part_off + R.npart; t++) {
        if (perm_out) perm_out[t] = hid[t];
        if (cell_out) cell_out[R.part_off + hid[t]] = hcst[t] - R.cell_off;
      }
    }
  }
  return check_flags(c);
}

static mpcd_status read_kbt(mpcd_gpu_ctx *c, double *kbt_now_out) {
  if (!kbt_now_out) return MPCD_OK;
  HIPCHK(c, hipMemcpyAsync(kbt_now_out, c->C.kbtnow, sizeof(double) * c->nrep, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_collide(mpcd_gpu_ctx *c, const mpcd_draws *draws, double *kbt_now_out) {
  if (!c) return MPCD_EINVAL;
  if (c->exact) return fail(c, MPCD_EUNSUPPORTED, "the exact-sum FP32 form steps whole steps only");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  st = stage_draws(c, draws);
  if (st) return st;
  if (draws) c->stream_inj = true;
  if (c->precision == MPCD_FP64) run_collide<double>(c, draws != nullptr);
  else run_collide<float>(c, draws != nullptr);
  st = check_launch(c);
  if (st) return st;
  return read_kbt(c, kbt_now_out);
}

// the periodic rest-frame transformation of the output step
// (galileantrans mtools.c:673-718 + zeroExtraDims mtools.c:727-744, called
// by outputResults pout.c:2557-2562 when runtime % zeroNetMom == 0)
template <typename Real>
static void run_netmom(mpcd_gpu_ctx *c) {
  if (c->pending) {
    // outputResults bins and runs localPROP before the transformation
    // (pout.c:2540-2545): the cell values stay those of the step's velocities
    run_bin<Real>(c, 2, true, 0);
    run_props<Real>(c, 0);
  }
  State<Real> S = states<Real>(c)[c->cur];
  double *part = c->d_netmom, *net = part + 4 * c->pblk.size();
  k_netmom_part<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(c->C, S, c->d_pblk, part);
  k_netmom_final<<<c->nrep, kBlock, 0, c->stream>>>(c->C, c->d_pblk_rng, part, net);
  k_netmom_apply<Real><<<c->pblk.size(), kBlock, 0, c->stream>>>(c->C, S, c->d_pblk, net);
}

extern "C" mpcd_status mpcd_gpu_zero_net_momentum(mpcd_gpu_ctx *c) {
  if (!c) return MPCD_EINVAL;
  if (c->recording || c->tan_on)
    return fail(c, MPCD_EUNSUPPORTED, "zero net momentum: not while an adjoint tape or a tangent is running");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  if (!c->d_netmom) {
    c->d_netmom = dalloc<double>(c, 4 * (c->pblk.size() + c->nrep));
    if (!c->d_netmom) return fail(c, MPCD_ENOMEM, "net momentum sums");
  }
  if (c->precision == MPCD_FP64) run_netmom<double>(c);
  else run_netmom<float>(c);
  st = check_launch(c);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MPCD_OK;
}

static mpcd_status read_events(mpcd_gpu_ctx *c, int64_t *out) {
  unsigned long long ev[4];
  HIPCHK(c, hipMemcpyAsync(ev, c->C.events, sizeof ev, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out)
    for (int i = 0; i < 4; i++) out[i] = (int64_t)ev[i];
  return ev[0] ? MPCD_EESCAPE : MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_stream(mpcd_gpu_ctx *c, int64_t *bc_event_counts_out) {
  if (!c) return MPCD_EINVAL;
  if (c->exact) return fail(c, MPCD_EUNSUPPORTED, "the exact-sum FP32 form steps whole steps only");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  const int inj = c->stream_inj ? 1 : 0;
  c->stream_inj = false;
  if (c->precision == MPCD_FP64) run_stream<double>(c, inj);
  else run_stream<float>(c, inj);
  st = check_launch(c);
  if (st) return st;
  st = check_flags(c);
  if (st) return st;
  if (bc_event_counts_out) return read_events(c, bc_event_counts_out);
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_step(mpcd_gpu_ctx *c, const mpcd_draws *draws, double *kbt_now_out) {
  if (!c) return MPCD_EINVAL;
  if (c->exact && draws)
    return fail(c, MPCD_EUNSUPPORTED, "the exact-sum FP32 form draws from its generator (no injected draws)");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  st = stage_draws(c, draws);
  if (st) return st;
  if (c->precision == MPCD_FP64) run_step<double>(c, draws != nullptr);
  else run_step<float>(c, draws != nullptr);
  st = check_launch(c);
  if (st) return st;
  st = read_kbt(c, kbt_now_out);
  if (st) return st;
  return check_flags(c);
}

// records two fused steps (from an even step) and the advance of the device
// step counter into c->gexec; nothing runs while recording
static mpcd_status capture_pair(mpcd_gpu_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  const uint32_t step0 = c->step;
  const int cur0 = c->cur;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  c->cap_base = step0;
  for (int k = 0; k < 2; k++) {
    if (c->precision == MPCD_FP64) run_step_fused<double>(c, 0);
    else run_step_fused<float>(c, 0);
  }
  k_addu32<<<1, 1, 0, c->stream>>>(c->d_step, 2u);
  c->capturing = false;
  hipError_t e = hipStreamEndCapture(c->stream, &c->graph);
  // the recorded steps have not run: restore the host-side step state
  c->step = step0;
  c->cur = cur0;
  if (e != hipSuccess) return fail(c, MPCD_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  HIPCHK(c, hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
  c->graph_cur = cur0;
  return MPCD_OK;
}

// ---------------------------------------------------------------------------
// the bucketed layout (mpcd_bucket.h)
// ---------------------------------------------------------------------------
template <typename Real>
static BState<Real> *bstates(mpcd_gpu_ctx *c);
template <>
BState<double> *bstates<double>(mpcd_gpu_ctx *c) { return c->bd; }
template <>
BState<float> *bstates<float>(mpcd_gpu_ctx *c) { return c->bf; }
template <typename Real>
static WallList<Real> &wall_list(mpcd_gpu_ctx *c);
template <>
WallList<double> &wall_list<double>(mpcd_gpu_ctx *c) { return c->wld; }
template <>
WallList<float> &wall_list<float>(mpcd_gpu_ctx *c) { return c->wlf; }

template <typename Real>
static bool alloc_wall_list(mpcd_gpu_ctx *c, WallList<Real> &w) {
  const size_t n = c->ntot;
  w.x = dalloc<Real>(c, n);
  w.y = dalloc<Real>(c, n);
  w.z = dalloc<Real>(c, n);
  w.vx = dalloc<Real>(c, n);
  w.vy = dalloc<Real>(c, n);
  w.vz = dalloc<Real>(c, n);
  w.meta = dalloc<int4>(c, n);
  w.cnt = dalloc<int32_t>(c, 1);
  return w.x && w.y && w.z && w.vx && w.vy && w.vz && w.meta && w.cnt;
}

static bool bucket_usable(const mpcd_gpu_ctx *c) {
  if (c->exact) return c->bkt_on;  // no thermostat, tape or tangent in this form
  return c->bkt_on && c->fusable && !c->recording && !c->tan_on && !c->any_therm;
}

static void drop_bucket_graph(mpcd_gpu_ctx *c) {
  if (c->bgexec) (void)hipGraphExecDestroy(c->bgexec);
  if (c->bgraph) (void)hipGraphDestroy(c->bgraph);
  c->bgexec = nullptr;
  c->bgraph = nullptr;
}

// tiles of consecutive cells per replica whose mean population is `fill`
// times the bucket capacity
static void build_buckets(mpcd_gpu_ctx *c, double fill) {
  c->bkts.clear();
  c->brep.clear();
  c->bkt_rng.clear();
  const int cap = bkt_cap(fused_op(c->rtech));
  for (int r = 0; r < c->nrep; r++) {
    const RepDev &R = c->reps[r];
    int tc = (int)floor(fill * cap / std::max(mean_ppc(R), 1e-9));
    tc = std::max(1, std::min(kBktMaxCells, tc));
    const int b0 = (int)c->bkts.size();
    for (int b = 0; b < R.ncell; b += tc)
      c->bkts.push_back(TileDev{r, R.cell_off + b, R.cell_off + std::min(b + tc, R.ncell), 0});
    c->brep.push_back(BucketRep{b0, tc, 1.0 / tc});
    c->bkt_rng.push_back(make_int2(b0, (int)c->bkts.size()));
  }
  c->nbkt = (int)c->bkts.size();
  c->bkt_fill = fill;
  c->bkt_dirty = true;
  drop_bucket_graph(c);
}

static void dfree(mpcd_gpu_ctx *c, void *p) {
  if (!p) return;
  auto it = std::find(c->allocs.begin(), c->allocs.end(), p);
  if (it != c->allocs.end()) c->allocs.erase(it);
  (void)hipFree(p);
}

static bool alloc_xstate(mpcd_gpu_ctx *c, XState &b, size_t n) {
  void *old[] = {b.a, b.b, b.id, b.sp};
  for (void *p : old) dfree(c, p);
  b.a = dalloc<float4>(c, n);
  b.b = dalloc<float2>(c, n);
  b.id = dalloc<int32_t>(c, n);
  b.sp = (c->nspec > 1) ? dalloc<int32_t>(c, n) : nullptr;
  return b.a && b.b && b.id && (c->nspec == 1 || b.sp);
}

static bool alloc_xlist(mpcd_gpu_ctx *c, XList &L, int cap) {
  L.a = dalloc<float4>(c, cap);
  L.b = dalloc<float2>(c, cap);
  L.m = dalloc<int4>(c, cap);
  L.cnt = dalloc<int32_t>(c, 1);
  L.cap = cap;
  return L.a && L.b && L.m && L.cnt;
}

static bool alloc_bstate(mpcd_gpu_ctx *c, BState<float> &b, size_t n) {
  void *old[] = {b.a, b.b, b.c, b.sp};
  for (void *p : old) dfree(c, p);
  b.a = dalloc<float4>(c, n);
  b.b = dalloc<float4>(c, n);
  b.c = dalloc<uint32_t>(c, n);
  b.sp = dalloc<int32_t>(c, n);
  return b.a && b.b && b.c && b.sp;
}

static bool alloc_bstate(mpcd_gpu_ctx *c, BState<double> &b, size_t n) {
  void *old[] = {b.a, b.b, b.c, b.d};
  for (void *p : old) dfree(c, p);
  b.a = dalloc<double2>(c, n);
  b.b = dalloc<double2>(c, n);
  b.c = dalloc<double2>(c, n);
  b.d = dalloc<int4>(c, n);
  return b.a && b.b && b.c && b.d;
}

// device tables and buffers of the current bucket layout
static mpcd_status ensure_buckets(mpcd_gpu_ctx *c) {
  if (c->bkts.empty()) build_buckets(c, c->bkt_fill);
  if (!c->d_bkts) {
    const size_t nc1 = (size_t)c->nctot + 1;
    c->d_bkts = dalloc<TileDev>(c, c->nctot);
    c->d_brep = dalloc<BucketRep>(c, c->nrep);
    c->d_bkt_rng = dalloc<int2>(c, c->nrep);
    c->bcnt[0] = dalloc<int32_t>(c, nc1);
    c->bcnt[1] = dalloc<int32_t>(c, nc1);
    c->bbase = dalloc<int32_t>(c, nc1);
    c->bstart[0] = dalloc<int32_t>(c, nc1);
    c->bstart[1] = dalloc<int32_t>(c, nc1);
    if (!c->d_bkts || !c->d_brep || !c->d_bkt_rng || !c->bcnt[0] || !c->bcnt[1] || !c->bbase || !c->bstart[0] ||
        !c->bstart[1])
      return fail(c, MPCD_ENOMEM, "bucket tables");
    if (!alloc_scan(c)) return fail(c, MPCD_ENOMEM, "scan scratch");
    // escape lists: a step with more escapes finishes on the cell-sorted path
    c->esc_cap = std::max(4096, c->ntot / 8);
    if (c->exact) {
      // a wall list holds at most every particle once per step
      for (int k = 0; k < 3; k++) c->xcnt[k] = dalloc<int32_t>(c, nc1);
      if (!alloc_xlist(c, c->xwl[0], c->ntot) || !alloc_xlist(c, c->xwl[1], c->ntot) || !c->xcnt[0] ||
          !c->xcnt[1] || !c->xcnt[2])
        return fail(c, MPCD_ENOMEM, "wall lists and bucket counts");
    } else {
      if (!(c->precision == MPCD_FP64 ? alloc_wall_list<double>(c, c->wld) : alloc_wall_list<float>(c, c->wlf)))
        return fail(c, MPCD_ENOMEM, "wall list");
      c->esc_idx[0] = dalloc<int32_t>(c, c->esc_cap);
      c->esc_idx[1] = dalloc<int32_t>(c, c->esc_cap);
      c->esc_cnt = dalloc<int32_t>(c, 2);
      if (!c->esc_idx[0] || !c->esc_idx[1] || !c->esc_cnt) return fail(c, MPCD_ENOMEM, "escape lists");
    }
  }
  const int cap = bkt_cap(fused_op(c->rtech));
  const size_t slots = (size_t)c->nbkt * cap;
  if (slots > c->bkt_slots) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    bool ok = true;
    for (int k = 0; k < 2; k++)
      ok = ok && (c->exact                       ? alloc_xstate(c, c->bx[k], slots)
                  : c->precision == MPCD_FP64 ? alloc_bstate(c, c->bd[k], slots)
                                              : alloc_bstate(c, c->bf[k], slots));
    if (!ok) {
      c->bkt_slots = 0;
      return fail(c, MPCD_ENOMEM, "bucketed state (%zu slots)", slots);
    }
    c->bkt_slots = slots;
    drop_bucket_graph(c);
  }
  if (c->bkt_dirty) {
    HIPCHK(c, hipMemcpyAsync(c->d_bkts, c->bkts.data(), sizeof(TileDev) * c->nbkt, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_brep, c->brep.data(), sizeof(BucketRep) * c->nrep, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_bkt_rng, c->bkt_rng.data(), sizeof(int2) * c->nrep, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->bkt_dirty = false;
  }
  for (int k = 0; k < 2; k++) {
    c->bd[k].cnt = c->bcnt[k];
    c->bf[k].cnt = c->bcnt[k];
  }
  return MPCD_OK;
}

// the escape list of the step with parity par
static EscList esc_list(mpcd_gpu_ctx *c, int par) { return EscList{c->esc_idx[par], c->esc_cnt + par, c->esc_cap}; }

static int bkt_limit(const mpcd_gpu_ctx *c) {
  const int cap = bkt_cap(fused_op(c->rtech));
  return (c->bkt_caplim > 0 && c->bkt_caplim < cap) ? c->bkt_caplim : cap;
}

template <typename Real, int OP, bool ONESP, int DIM, bool GHOST>
static void launch_bucket_g(mpcd_gpu_ctx *c, const Ctx &C, const Draws &D, int par) {
  BState<Real> *B = bstates<Real>(c);
  const int nw3 = 3 * std::max(1, c->nwtot), nwm = std::max(1, c->nwtot);
  k_bucket<Real, OP, ONESP, DIM, bkt_ppt(OP), GHOST><<<c->nbkt, kBktThreads, 0, c->stream>>>(
      C, B[c->bcur], B[1 - c->bcur], c->d_bkts, c->d_brep, c->d_rep, c->d_sp, c->d_wall, c->wh2 + par * nwm,
      c->box2 + par * 6 * c->nrep, D, c->shift2 + par * 3 * c->nrep, c->wqb2 + par * nw3,
      c->shift2 + (par ^ 1) * 3 * c->nrep, c->bbase, c->bstart[par], bkt_limit(c), wall_list<Real>(c),
      esc_list(c, par ^ 1));
}

// the ghost-particle code only where a replica has phantom walls (it sets
// the kernel's register count)
template <typename Real, int OP, bool ONESP, int DIM>
static void launch_bucket(mpcd_gpu_ctx *c, const Ctx &C, const Draws &D, int par) {
  if (c->any_phantom) launch_bucket_g<Real, OP, ONESP, DIM, true>(c, C, D, par);
  else launch_bucket_g<Real, OP, ONESP, DIM, false>(c, C, D, par);
}

template <typename Real, int OP, bool ONESP>
static void launch_bucket_dim(mpcd_gpu_ctx *c, const Ctx &C, const Draws &D, int par) {
  if (c->dim == 3) launch_bucket<Real, OP, ONESP, 3>(c, C, D, par);
  else launch_bucket<Real, OP, ONESP, 2>(c, C, D, par);
}

template <typename Real>
static void dispatch_bucket(mpcd_gpu_ctx *c, const Ctx &C, const Draws &D, int par) {
  const bool one = c->nspec == 1;
  const int op = fused_op(c->rtech);
  if (op == OP_SRD) {
    if (one) launch_bucket_dim<Real, OP_SRD, true>(c, C, D, par);
    else launch_bucket_dim<Real, OP_SRD, false>(c, C, D, par);
  } else if (op == OP_AT) {
    if (one) launch_bucket_dim<Real, OP_AT, true>(c, C, D, par);
    else launch_bucket_dim<Real, OP_AT, false>(c, C, D, par);
  } else {
    if (one) launch_bucket_dim<Real, OP_RAT, true>(c, C, D, par);
    else launch_bucket_dim<Real, OP_RAT, false>(c, C, D, par);
  }
}

// ---------------------------------------------------------------------------
// the exact-sum FP32 form (mpcd_exact.h): bucketed steps and dense steps
// ---------------------------------------------------------------------------
struct XLaunch {
  Ctx C;
  XState I, O;
  int32_t *zero;
  XNext Nx;
  Draws D;
  int par;
};

template <int OP, bool ONESP, int DIM, bool GHOST>
static void launch_xbucket_g(mpcd_gpu_ctx *c, const XLaunch &a) {
  const int nwm = std::max(1, c->nwtot);
  k_xbucket<OP, ONESP, DIM, GHOST><<<c->nbkt, kBktThreads, 0, c->stream>>>(
      a.C, a.I, a.O, a.zero, c->d_bkts, c->d_brep, c->d_rep, c->d_sp, c->wh2 + a.par * nwm,
      c->box2 + a.par * 6 * c->nrep, a.D, c->shift2 + a.par * 3 * c->nrep, a.Nx, bkt_limit(c), c->xwl[a.par],
      c->nspec);
}

template <int OP, bool ONESP, int DIM>
static void launch_xbucket(mpcd_gpu_ctx *c, const XLaunch &a) {
  if (c->any_phantom) launch_xbucket_g<OP, ONESP, DIM, true>(c, a);
  else launch_xbucket_g<OP, ONESP, DIM, false>(c, a);
}

template <int OP, bool ONESP>
static void launch_xbucket_dim(mpcd_gpu_ctx *c, const XLaunch &a) {
  if (c->dim == 3) launch_xbucket<OP, ONESP, 3>(c, a);
  else launch_xbucket<OP, ONESP, 2>(c, a);
}

static void dispatch_xbucket(mpcd_gpu_ctx *c, const Ctx &C, const XState &I, const XState &O, int32_t *zero,
                             const XNext &Nx, const Draws &D, int par) {
  const XLaunch a{C, I, O, zero, Nx, D, par};
  const bool one = c->nspec == 1;
  const int op = fused_op(c->rtech);
  if (op == OP_SRD) {
    if (one) launch_xbucket_dim<OP_SRD, true>(c, a);
    else launch_xbucket_dim<OP_SRD, false>(c, a);
  } else if (op == OP_AT) {
    if (one) launch_xbucket_dim<OP_AT, true>(c, a);
    else launch_xbucket_dim<OP_AT, false>(c, a);
  } else {
    if (one) launch_xbucket_dim<OP_RAT, true>(c, a);
    else launch_xbucket_dim<OP_RAT, false>(c, a);
  }
}

// does any particle of the context ever take the wall list?  Not when every
// replica's walls are its box's periodic faces (they wrap in the step kernel)
static bool x_needs_walls(const mpcd_gpu_ctx *c) {
  for (const RepDev &R : c->reps)
    if (!R.xwrap) return true;
  return false;
}

// the count array of the input of bucketed step s (three turn over)
static int x_ci(const mpcd_gpu_ctx *c, uint32_t s) { return (int)((c->xci0 + (s - c->bkt_step0)) % 3u); }

static void run_step_xbucket(mpcd_gpu_ctx *c) {
  const int par = c->step & 1, p1 = par ^ 1;
  Ctx C = view(c, par);
  Draws D = make_draws(c, 0, c->step);
  const int cap = bkt_cap(fused_op(c->rtech)), one = c->nspec == 1 ? 1 : 0;
  const int nw3 = 3 * std::max(1, c->nwtot), nwm = std::max(1, c->nwtot);
  const int ci = x_ci(c, c->step);
  XState I = c->bx[c->bcur], O = c->bx[1 - c->bcur];
  I.cnt = c->xcnt[ci];
  O.cnt = c->xcnt[(ci + 1) % 3];
  XNext Nx{};
  Nx.D1 = make_draws(c, 0, c->step + 1);
  Nx.wq_prev_b = c->wqb2 + par * nw3;
  Nx.shift = c->shift2 + p1 * 3 * c->nrep;
  Nx.wqs = c->wqs2 + p1 * nw3;
  Nx.wqb = c->wqb2 + p1 * nw3;
  Nx.box = c->box2 + p1 * 6 * c->nrep;
  Nx.wh = c->wh2 + p1 * nwm;
  Nx.wl_cnt = c->xwl[p1].cnt;
  {
    ProfScope ps(c, F_BUCKET);
    dispatch_xbucket(c, C, I, O, c->xcnt[(ci + 2) % 3], Nx, D, par);
  }
  if (x_needs_walls(c)) {
    ProfScope ps(c, F_STREAM);
    const int nb = 512;
    Ctx Cn = view(c, p1);
    if (c->dim == 3)
      k_xwalls<3><<<nb, kBlock, 0, c->stream>>>(C, c->xwl[par], O, c->d_brep, c->d_rep, c->d_sp, c->d_wall,
                                                c->wh2 + par * nwm, c->box2 + par * 6 * c->nrep, D,
                                                c->wqb2 + par * nw3, Cn, Nx.D1, cap, bkt_limit(c), one);
    else
      k_xwalls<2><<<nb, kBlock, 0, c->stream>>>(C, c->xwl[par], O, c->d_brep, c->d_rep, c->d_sp, c->d_wall,
                                                c->wh2 + par * nwm, c->box2 + par * 6 * c->nrep, D,
                                                c->wqb2 + par * nw3, Cn, Nx.D1, cap, bkt_limit(c), one);
  }
  c->bcur ^= 1;
  c->step++;
  c->x_tables = c->step;  // the step kernel wrote them
}

// the shift and wall tables of step s in its parity slot (from step s - 1's)
static void x_ensure_tables(mpcd_gpu_ctx *c, uint32_t s) {
  if (c->x_tables == s) return;
  launch_shift(c, s, 0);
  c->x_tables = s;
}

// dense (any particle order) state -> bucketed, at the current step
static mpcd_status xbucket_enter(mpcd_gpu_ctx *c) {
  mpcd_status st = ensure_buckets(c);
  if (st) return st;
  const int par = c->step & 1;
  x_ensure_tables(c, c->step);
  Ctx C = view(c, par);
  Draws D = make_draws(c, 0, c->step);
  c->bkt_step0 = c->step;
  c->xci0 = 0;
  for (int k = 0; k < 3; k++) k_fill<<<nblocks(c->nbkt + 1, kBlock), kBlock, 0, c->stream>>>(c->xcnt[k], 0, c->nbkt + 1);
  k_fill<<<1, 1, 0, c->stream>>>(c->C.flags + 1, 0x7FFFFFFF, 1);
  k_fill<<<1, 1, 0, c->stream>>>(c->xwl[par].cnt, 0, 1);
  XState O = c->bx[0];
  O.cnt = c->xcnt[0];
  k_xbucketize<<<nblocks(c->ntot, kBlock), kBlock, 0, c->stream>>>(C, c->sf[c->cur], O, D, c->d_brep,
                                                                    bkt_cap(fused_op(c->rtech)), bkt_limit(c),
                                                                    c->nspec == 1 ? 1 : 0, c->xshifted);
  c->bcur = 0;
  return check_launch(c);
}

// bucketed -> dense (particle order) at step s from buffer bsrc; the binned
// structures of the cell-sorted layout are rebuilt when something reads them
static void xbucket_densify(mpcd_gpu_ctx *c, uint32_t s, int bsrc) {
  Ctx C = view(c, s & 1);
  XState I = c->bx[bsrc];
  I.cnt = c->xcnt[x_ci(c, s)];
  k_xdensify<<<c->nbkt, kBlock, 0, c->stream>>>(C, I, c->sf[c->cur], c->d_bkts, bkt_cap(fused_op(c->rtech)),
                                                bkt_limit(c), c->nspec == 1 ? 1 : 0, c->xshifted);
  c->pending = false;
  c->hist_ready = false;
  c->need_init = true;
}

static mpcd_status ensure_xdense(mpcd_gpu_ctx *c) {
  if (c->xd.n) return MPCD_OK;
  const size_t nc = c->nctot, ns = xnsum(fused_op(c->rtech));
  c->xd.n = dalloc<int32_t>(c, nc);
  c->xd.nsp = (c->nspec > 1) ? dalloc<int32_t>(c, nc * MPCD_MAXSPEC) : nullptr;
  c->xd.sum = dalloc<double>(c, nc * ns);
  c->xd.cd = dalloc<float>(c, nc * 12);
  c->xd.cell = dalloc<int32_t>(c, c->ntot);
  c->xd.rt = (fused_op(c->rtech) != OP_SRD) ? dalloc<float>(c, 3 * (size_t)c->ntot) : nullptr;
  if (!c->xd.n || !c->xd.sum || !c->xd.cd || !c->xd.cell || (c->nspec > 1 && !c->xd.nsp) ||
      (fused_op(c->rtech) != OP_SRD && !c->xd.rt)) {
    c->xd.n = nullptr;
    return fail(c, MPCD_ENOMEM, "cell sums of the exact form");
  }
  return MPCD_OK;
}

template <int OP, bool ONESP, int DIM>
static void launch_xdense(mpcd_gpu_ctx *c, const Ctx &C, const Draws &D, int par) {
  State<float> S = c->sf[c->cur];
  const int npb = (int)c->pblk.size(), ncb = (int)c->cblk.size();
  k_xd_sums<OP, ONESP, DIM><<<npb, kBlock, 0, c->stream>>>(C, S, c->xd, c->d_pblk, c->d_sp, D, c->xshifted);
  if (c->any_phantom)
    k_xd_cells<OP, ONESP, DIM, true><<<ncb, kBlock, 0, c->stream>>>(C, c->xd, c->d_cblk, c->d_sp, D, c->nspec);
  else
    k_xd_cells<OP, ONESP, DIM, false><<<ncb, kBlock, 0, c->stream>>>(C, c->xd, c->d_cblk, c->d_sp, D, c->nspec);
  k_xd_part<OP, ONESP, DIM><<<npb, kBlock, 0, c->stream>>>(C, S, c->xd, c->d_pblk, c->d_sp, c->d_wall,
                                                           c->box2 + par * 6 * c->nrep, D, view(c, par ^ 1),
                                                           make_draws(c, 0, c->step + 1), c->xshifted);
}

template <int OP, bool ONESP>
static void launch_xdense_dim(mpcd_gpu_ctx *c, const Ctx &C, const Draws &D, int par) {
  if (c->dim == 3) launch_xdense<OP, ONESP, 3>(c, C, D, par);
  else launch_xdense<OP, ONESP, 2>(c, C, D, par);
}

// one step of the exact form on the dense state (any particle order)
static void run_step_xdense(mpcd_gpu_ctx *c) {
  if (ensure_xdense(c)) return;
  const int par = c->step & 1;
  x_ensure_tables(c, c->step);
  Ctx C = view(c, par);
  Draws D = make_draws(c, 0, c->step);
  const size_t nc = c->nctot, ns = xnsum(fused_op(c->rtech));
  k_fill<<<nblocks((int)nc, kBlock), kBlock, 0, c->stream>>>(c->xd.n, 0, (int)nc);
  if (c->xd.nsp)
    k_fill<<<nblocks((int)(nc * MPCD_MAXSPEC), kBlock), kBlock, 0, c->stream>>>(c->xd.nsp, 0,
                                                                                 (int)(nc * MPCD_MAXSPEC));
  // zero doubles: all-zero words
  k_fill<<<nblocks((int)(2 * nc * ns), kBlock), kBlock, 0, c->stream>>>((int32_t *)c->xd.sum, 0, (int)(2 * nc * ns));
  // the next step's tables: the particle phase re-places the particles that
  // would escape its shifted bin
  launch_shift(c, c->step + 1, 0);
  const bool one = c->nspec == 1;
  const int op = fused_op(c->rtech);
  if (op == OP_SRD) {
    if (one) launch_xdense_dim<OP_SRD, true>(c, C, D, par);
    else launch_xdense_dim<OP_SRD, false>(c, C, D, par);
  } else if (op == OP_AT) {
    if (one) launch_xdense_dim<OP_AT, true>(c, C, D, par);
    else launch_xdense_dim<OP_AT, false>(c, C, D, par);
  } else {
    if (one) launch_xdense_dim<OP_RAT, true>(c, C, D, par);
    else launch_xdense_dim<OP_RAT, false>(c, C, D, par);
  }
  k_xtemp_reduce<<<c->nrep, kTempThreads, 0, c->stream>>>(C, c->d_pblk_rng);
  c->pending = false;
  c->hist_ready = false;
  c->need_init = true;
  c->step++;
  c->x_tables = c->step;
}

// one step in the bucketed layout (internal generator)
template <typename Real>
static void run_step_bucket(mpcd_gpu_ctx *c) {
  if (c->exact) {
    run_step_xbucket(c);
    return;
  }
  const int par = c->step & 1;
  Ctx C = view(c, par);
  Draws D = make_draws(c, 0, c->step);
  {
    // next step's shift, bucket-count block sums, zeroed output counts and
    // this step's SRD cell draws in one launch
    ProfScope ps(c, F_SCAN);
    const int p1 = (c->step + 1) & 1, nw3 = 3 * std::max(1, c->nwtot);
    const int nbs = nblocks(c->nbkt + 1, kBlock);
    const int ncb = c->d_celldraw ? (int)c->cblk.size() : 0;
    Draws D1 = make_draws(c, 0, c->step + 1);
    const int nesc = 4;  // blocks re-placing this step's listed escapes
    k_bucket_prep<Real><<<nbs + 1 + ncb + nesc, kBlock, 0, c->stream>>>(
        C, D1, c->wqb2 + (p1 ^ 1) * nw3, c->shift2 + p1 * 3 * c->nrep, c->wqs2 + p1 * nw3, c->wqb2 + p1 * nw3,
        c->wh2 + p1 * std::max(1, c->nwtot), c->box2 + p1 * 6 * c->nrep, c->bcnt[c->bcur], c->bbase,
        c->bcnt[1 - c->bcur], wall_list<Real>(c).cnt, c->nbkt, nbs, D, c->d_cblk, c->d_celldraw, ncb,
        bstates<Real>(c)[c->bcur], esc_list(c, par), esc_list(c, p1), c->d_bkts, bkt_cap(fused_op(c->rtech)),
        c->nspec == 1 ? 1 : 0, c->wh2 + par * std::max(1, c->nwtot), nesc);
    if (c->d_celldraw) D.cell = c->d_celldraw;
  }
  {
    ProfScope ps(c, F_BUCKET);
    dispatch_bucket<Real>(c, C, D, par);
  }
  {
    ProfScope ps(c, F_STREAM);
    BState<Real> *B = bstates<Real>(c);
    const int nw3 = 3 * std::max(1, c->nwtot), nwm = std::max(1, c->nwtot);
    const int cap = bkt_cap(fused_op(c->rtech)), nb = std::min(2048, nblocks(c->ntot, kBlock));
    if (c->dim == 3)
      k_bucket_walls<Real, 3><<<nb, kBlock, 0, c->stream>>>(
          C, wall_list<Real>(c), B[1 - c->bcur], c->d_brep, c->d_rep, c->d_sp, c->d_wall, c->wh2 + par * nwm,
          c->box2 + par * 6 * c->nrep, D, c->wqb2 + par * nw3, c->shift2 + (par ^ 1) * 3 * c->nrep, cap, bkt_limit(c),
          c->nwtot, esc_list(c, par ^ 1));
    else
      k_bucket_walls<Real, 2><<<nb, kBlock, 0, c->stream>>>(
          C, wall_list<Real>(c), B[1 - c->bcur], c->d_brep, c->d_rep, c->d_sp, c->d_wall, c->wh2 + par * nwm,
          c->box2 + par * 6 * c->nrep, D, c->wqb2 + par * nw3, c->shift2 + (par ^ 1) * 3 * c->nrep, cap, bkt_limit(c),
          c->nwtot, esc_list(c, par ^ 1));
  }
  c->bcur ^= 1;
  if (c->flow_on) {
    // the running flow field from the records just written: partials per
    // tile cell and offset class, then the fold per cell
    ProfScope ps(c, F_FLOW);
    BState<Real> *B = bstates<Real>(c);
    const int cap = bkt_cap(fused_op(c->rtech)), one = c->nspec == 1 ? 1 : 0;
    const double *shn = c->shift2 + ((c->step & 1) ^ 1) * 3 * c->nrep;
    const int ncb = (int)c->cblk.size();
    if (c->dim == 3) {
      k_flow_tiles<Real, 3><<<c->nbkt, kBlock, 0, c->stream>>>(c->C.flags, B[c->bcur], c->d_bkts, c->d_rep, c->d_sp,
                                                                shn, cap, one, c->nctot, c->d_flow_part,
                                                                c->d_flow_acc);
      k_flow_fold_tiles<3><<<ncb, kBlock, 0, c->stream>>>(c->C.flags, c->d_rep, c->d_cblk, c->nrep, c->nctot,
                                                           c->d_flow_part, c->d_flow_acc, c->d_flow, c->d_flow_n);
    } else {
      k_flow_tiles<Real, 2><<<c->nbkt, kBlock, 0, c->stream>>>(c->C.flags, B[c->bcur], c->d_bkts, c->d_rep, c->d_sp,
                                                                shn, cap, one, c->nctot, c->d_flow_part,
                                                                c->d_flow_acc);
      k_flow_fold_tiles<2><<<ncb, kBlock, 0, c->stream>>>(c->C.flags, c->d_rep, c->d_cblk, c->nrep, c->nctot,
                                                           c->d_flow_part, c->d_flow_acc, c->d_flow, c->d_flow_n);
    }
  }
  c->step++;
}

// dense -> bucketed, from a dense state whose next shifted bin is built
template <typename Real>
static mpcd_status bucket_enter(mpcd_gpu_ctx *c) {
  if (c->exact) return xbucket_enter(c);
  mpcd_status st = ensure_buckets(c);
  if (st) return st;
  State<Real> *S = states<Real>(c);
  BState<Real> *B = bstates<Real>(c);
  Ctx C = view(c, c->step & 1);
  k_fill<<<nblocks(c->nbkt + 1, kBlock), kBlock, 0, c->stream>>>(c->bcnt[0], 0, c->nbkt + 1);
  k_fill<<<1, 1, 0, c->stream>>>(c->C.flags + 1, 0x7FFFFFFF, 1);
  k_fill<<<1, 1, 0, c->stream>>>(c->esc_cnt + (c->step & 1), 0, 1);
  k_bucketize<Real><<<nblocks(c->ntot, kBlock), kBlock, 0, c->stream>>>(C, S[c->cur], B[0], c->d_brep,
                                                                          bkt_cap(fused_op(c->rtech)), bkt_limit(c),
                                                                          esc_list(c, c->step & 1));
  c->bcur = 0;
  c->bkt_step0 = c->step;
  return check_launch(c);
}

// bucketed -> dense at step s from bucket buffer bsrc (the buffer holding
// step s's particles), with the pending shifted bin of step s
template <typename Real>
static void bucket_densify(mpcd_gpu_ctx *c, uint32_t s, int bsrc) {
  if (c->exact) {
    xbucket_densify(c, s, bsrc);
    return;
  }
  Ctx C = view(c, s & 1);
  State<Real> *S = states<Real>(c);
  BState<Real> *B = bstates<Real>(c);
  const int cap = bkt_cap(fused_op(c->rtech)), one = c->nspec == 1 ? 1 : 0;
  const double *sh = c->shift2 + (s & 1) * 3 * c->nrep;
  k_fill<<<nblocks(c->nctot + 1, kBlock), kBlock, 0, c->stream>>>(C.hist, 0, c->nctot + 1);
  if (c->dim == 3)
    k_densify<Real, 3><<<c->nbkt, kBlock, 0, c->stream>>>(C, B[bsrc], S[c->cur], c->d_bkts, cap, one, sh,
                                                           c->bstart[(s - 1) & 1]);
  else
    k_densify<Real, 2><<<c->nbkt, kBlock, 0, c->stream>>>(C, B[bsrc], S[c->cur], c->d_bkts, cap, one, sh,
                                                           c->bstart[(s - 1) & 1]);
  c->pending = true;
  c->cu_src = 1;
  c->hist_ready = true;
}

// the end of a bucketed run: the temperature of its last step and the
// overflow check.  After a bucket overflow the state is the one before the
// first overflowing step -- the input buffer of that step, untouched since
// every later bucketed kernel was a no-op -- in the cell-sorted layout, and
// *done counts the steps completed since run_step0.  Without an overflow the
// state stays bucketed.
template <typename Real>
static mpcd_status bucket_finish(mpcd_gpu_ctx *c, uint32_t run_step0, int *done) {
  const uint32_t s0 = c->bkt_step0;
  if (c->step != run_step0) {
    ProfScope ps(c, F_TEMP);
    if (c->exact) k_xtemp_reduce<<<c->nrep, kTempThreads, 0, c->stream>>>(view(c, (c->step - 1) & 1), c->d_bkt_rng);
    else k_temp_reduce<<<c->nrep, kTempThreads, 0, c->stream>>>(view(c, (c->step - 1) & 1), c->d_bkt_rng);
  }
  int fl[2];
  HIPCHK(c, hipMemcpyAsync(fl, c->C.flags, sizeof fl, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (fl[0] & kFlagBadCell) return fail(c, MPCD_EDEVICE, "bucketed step: a particle outside its bucket's tile");
  *done = (int)(c->step - run_step0);
  if (!(fl[0] & kFlagPoison)) {
    c->bkt_steps += *done;
    return MPCD_OK;
  }
  const uint32_t s = (fl[1] < (int)s0) ? s0 : (uint32_t)fl[1];
  c->bkt_overflows++;
  int clr[2] = {fl[0] & ~kFlagPoison, 0x7FFFFFFF};
  HIPCHK(c, hipMemcpyAsync(c->C.flags, clr, sizeof clr, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  // s == s0: the cell-sorted state the bucketed layout started from is intact
  if (s != s0) bucket_densify<Real>(c, s, (int)((s - s0) & 1u));
  c->in_bucket = false;
  if (c->exact) c->x_tables = s;  // written by step s - 1 (or on entry)
  *done = (int)(s - run_step0);
  c->bkt_steps += *done;
  c->step = s;
  return check_launch(c);
}

// records two bucketed steps (from an even step) and the advance of the
// device step counter into c->bgexec; nothing runs while recording
// the steps of one recorded group of the bucketed layout: after them the
// buffers (two), the parity slots (two) and, in the exact-sum form, the
// bucket count arrays (three) are where they were
static int bucket_group(const mpcd_gpu_ctx *c) { return c->exact ? 6 : 2; }

// what a recorded group depends on besides the buffers: the exact-sum form's
// count arrays turn over with the step since the layout was entered
static int bucket_graph_key(const mpcd_gpu_ctx *c) {
  if (!c->exact) return c->bcur;
  return (int)((c->step - c->bkt_step0) % 6u) + 6 * (int)(c->bkt_step0 & 1u);
}

template <typename Real>
static mpcd_status capture_bucket_pair(mpcd_gpu_ctx *c) {
  drop_bucket_graph(c);
  const uint32_t step0 = c->step, tables0 = c->x_tables;
  const int bcur0 = c->bcur, key0 = bucket_graph_key(c), G = bucket_group(c);
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  c->cap_base = step0;
  for (int k = 0; k < G; k++) run_step_bucket<Real>(c);
  k_addu32<<<1, 1, 0, c->stream>>>(c->d_step, (uint32_t)G);
  c->capturing = false;
  hipError_t e = hipStreamEndCapture(c->stream, &c->bgraph);
  c->step = step0;
  c->bcur = bcur0;
  c->x_tables = tables0;
  if (e != hipSuccess) return fail(c, MPCD_EDEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
  HIPCHK(c, hipGraphInstantiate(&c->bgexec, c->bgraph, nullptr, nullptr, 0));
  c->bgraph_cur = key0;
  return MPCD_OK;
}

// nsteps steps in the bucketed layout; *done < nsteps after a bucket overflow
template <typename Real>
static mpcd_status bucket_run(mpcd_gpu_ctx *c, int nsteps, int use_graph, int *done) {
  *done = 0;
  mpcd_status st;
  if (!c->in_bucket) {
    st = bucket_enter<Real>(c);
    if (st) return st;
    c->in_bucket = true;
  }
  const uint32_t run_step0 = c->step;
  int left = nsteps;
  if (use_graph) {
    // eager steps up to the start of a group, so that one recorded group
    // serves every later call
    const int G = bucket_group(c);
    while (left > 0 && (c->exact ? (c->step - c->bkt_step0) % (uint32_t)G != 0u : (c->step & 1u) != 0u)) {
      run_step_bucket<Real>(c);
      left--;
    }
    const int groups = left / G;
    if (groups > 0) {
      if (!c->bgexec || c->bgraph_cur != bucket_graph_key(c)) {
        st = check_launch(c);
        if (st) return st;
        st = capture_bucket_pair<Real>(c);
        if (st) return st;
      }
      k_setu32<<<1, 1, 0, c->stream>>>(c->d_step, c->step);
      for (int k = 0; k < groups; k++) HIPCHK(c, hipGraphLaunch(c->bgexec, c->stream));
      c->step += (uint32_t)(G * groups);
      left -= G * groups;
      if (c->exact) c->x_tables = c->step;
    }
  }
  for (; left > 0; left--) run_step_bucket<Real>(c);
  st = check_launch(c);
  if (st) return st;
  return bucket_finish<Real>(c, run_step0, done);
}

static mpcd_status ensure_dense(mpcd_gpu_ctx *c) {
  if (!c->in_bucket) return MPCD_OK;
  const uint32_t s = c->step;
  const int bsrc = (int)((s - c->bkt_step0) & 1u);
  if (s != c->bkt_step0) {
    if (c->precision == MPCD_FP64) bucket_densify<double>(c, s, bsrc);
    else bucket_densify<float>(c, s, bsrc);
  }
  c->in_bucket = false;
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_set_layout(mpcd_gpu_ctx *c, int bucketed, double fill, int capacity) {
  if (!c) return MPCD_EINVAL;
  mpcd_status st = ensure_dense(c);
  if (st) return st;
  if (fill > 0.0) {
    if (fill > 4.0) return fail(c, MPCD_EINVAL, "set_layout: fill %g above 4", fill);
    build_buckets(c, fill);
  }
  if (capacity < 0) return fail(c, MPCD_EINVAL, "set_layout: capacity %d", capacity);
  c->bkt_caplim = capacity;
  c->bkt_on = bucketed != 0;
  drop_bucket_graph(c);
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_layout_stats(const mpcd_gpu_ctx *c, int64_t *out) {
  if (!c || !out) return MPCD_EINVAL;
  out[0] = c->bkt_steps;
  out[1] = c->bkt_overflows;
  out[2] = bkt_limit(c);
  out[3] = c->nbkt;
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_run(mpcd_gpu_ctx *c, int nsteps, int use_graph) {
  if (!c || nsteps < 0) return MPCD_EINVAL;
  mpcd_status st = ensure_init(c);
  if (st) return st;
  auto eager = [&]() {
    if (c->precision == MPCD_FP64) run_step<double>(c, 0);
    else run_step<float>(c, 0);
  };
  if (bucket_usable(c) && nsteps > 0) {
    // the bucketed layout starts from a cell-sorted state whose next bin is
    // built (the exact form bins the dense state itself)
    while (!c->exact && !c->in_bucket && nsteps > 0 && !c->hist_ready) {
      eager();
      nsteps--;
    }
    // after a bucket overflow the cell-sorted path takes the overflowed
    // step; the run goes back to buckets when the stretch before held at
    // least 16 steps (or it is the run's first overflow), else the
    // cell-sorted path finishes it
    bool first = true;
    while (nsteps > 0) {
      int done = 0;
      st = c->precision == MPCD_FP64 ? bucket_run<double>(c, nsteps, use_graph, &done)
                                     : bucket_run<float>(c, nsteps, use_graph, &done);
      if (st) return st;
      nsteps -= done;
      if (nsteps == 0 || !(first || done >= 16)) break;
      first = false;
      st = ensure_dense(c);
      if (st) return st;
      eager();
      nsteps--;
      while (!c->exact && nsteps > 0 && !c->hist_ready) {
        eager();
        nsteps--;
      }
    }
  }
  if (nsteps > 0) {
    st = ensure_dense(c);
    if (st) return st;
  }
  if (use_graph && c->fusable && !c->exact && !c->recording && !c->tan_on) {
    // graphs start from an even step whose histogram the previous step built
    while (nsteps > 0 && (!c->hist_ready || (c->step & 1))) {
      eager();
      nsteps--;
    }
    const int pairs = nsteps / 2;
    if (pairs > 0) {
      if (!c->gexec || c->graph_cur != c->cur) {
        st = check_launch(c);
        if (st) return st;
        st = capture_pair(c);
        if (st) return st;
      }
      k_setu32<<<1, 1, 0, c->stream>>>(c->d_step, c->step);
      for (int k = 0; k < pairs; k++) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
      c->step += 2u * pairs;
      nsteps -= 2 * pairs;
    }
  }
  for (int k = 0; k < nsteps; k++) eager();
  return check_launch(c);
}

// whether the buckets as built leave every replica's mean bucket population
// mu at least 6 sqrt(mu) slots below the limit: a bucket's count fluctuates
// like a Poisson count, and one overflow in a run costs a step on the
// cell-sorted path and a re-bucketing
static bool buckets_have_headroom(const mpcd_gpu_ctx *c) {
  const int lim = bkt_limit(c);
  for (int r = 0; r < c->nrep; r++) {
    const double mu = c->brep[r].tc * mean_ppc(c->reps[r]);
    if (lim - mu < 6.0 * sqrt(mu)) return false;
  }
  return true;
}

// the bucket fill of the bucketed layout: candidates in increasing order, so
// the first allocates the most slots; a trial whose buckets overflow, or
// (after the first) whose fill leaves too little headroom, loses.  Every
// candidate runs its trial steps, so the number of steps taken is fixed.
template <typename Real>
static mpcd_status autotune_buckets(mpcd_gpu_ctx *c, int steps, double *ms_out) {
  const double fills[] = {0.4, 0.5, 0.6, 0.7, 0.75, 0.8};
  double best = 1e300, best_fill = c->bkt_fill;
  hipEvent_t a, b;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  mpcd_status st = MPCD_OK;
  for (double f : fills) {
    // 1 + steps steps per candidate: the untimed one is a cell-sorted step
    // when the next bin is not built yet
    build_buckets(c, f);
    const bool eligible = f == fills[0] || buckets_have_headroom(c);
    int untimed = 0;
    if (!c->exact && !c->hist_ready) run_step<Real>(c, 0);
    else untimed = 1;
    st = bucket_enter<Real>(c);
    if (st) break;
    c->in_bucket = true;
    const uint32_t s0 = c->step;
    if (untimed) run_step_bucket<Real>(c);
    HIPCHK(c, hipEventRecord(a, c->stream));
    for (int k = 0; k < steps; k++) run_step_bucket<Real>(c);
    HIPCHK(c, hipEventRecord(b, c->stream));
    int done = 0;
    st = bucket_finish<Real>(c, s0, &done);
    if (st) break;
    st = ensure_dense(c);  // the next candidate has another layout
    if (st) break;
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, a, b));
    if (eligible && done == steps + untimed && ms / steps < best) {
      best = ms / steps;
      best_fill = f;
    }
    for (int k = done; k < steps + untimed; k++) run_step<Real>(c, 0);  // after an overflow
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (st) return st;
  build_buckets(c, best_fill);
  if (ms_out) *ms_out = best;
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_autotune(mpcd_gpu_ctx *c, int steps_per_trial, double *ms_per_step_out) {
  if (!c || steps_per_trial < 1) return MPCD_EINVAL;
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  if (!c->fusable && !c->exact) {
    if (ms_per_step_out) *ms_per_step_out = 0.0;
    return MPCD_OK;  // the split path has no tile parameter
  }
  if (c->exact && !bucket_usable(c)) {
    if (ms_per_step_out) *ms_per_step_out = 0.0;
    return MPCD_OK;  // dense steps of the exact form have no tile parameter
  }
  if (bucket_usable(c)) {
    st = c->precision == MPCD_FP64 ? autotune_buckets<double>(c, steps_per_trial, ms_per_step_out)
                                   : autotune_buckets<float>(c, steps_per_trial, ms_per_step_out);
    if (st) return st;
    st = check_launch(c);
    if (st) return st;
    return check_flags(c);
  }
  const double fills[] = {0.25, 0.35, 0.45, 0.6, 0.8};
  double best = 1e300, best_fill = c->tile_fill;
  hipEvent_t a, b;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  for (double f : fills) {
    build_tiles(c, f);
    HIPCHK(c, hipMemcpyAsync(c->d_tiles, c->tiles.data(), sizeof(TileDev) * c->tiles.size(), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_tile_rng, c->tile_rng.data(), sizeof(int2) * c->nrep, hipMemcpyHostToDevice,
                             c->stream));
    if (c->precision == MPCD_FP64) run_step<double>(c, 0);  // untimed: settles the histogram state
    else run_step<float>(c, 0);
    HIPCHK(c, hipEventRecord(a, c->stream));
    for (int k = 0; k < steps_per_trial; k++) {
      if (c->precision == MPCD_FP64) run_step<double>(c, 0);
      else run_step<float>(c, 0);
    }
    HIPCHK(c, hipEventRecord(b, c->stream));
    HIPCHK(c, hipEventSynchronize(b));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, a, b));
    if (ms / steps_per_trial < best) {
      best = ms / steps_per_trial;
      best_fill = f;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  build_tiles(c, best_fill);
  HIPCHK(c, hipMemcpyAsync(c->d_tiles, c->tiles.data(), sizeof(TileDev) * c->tiles.size(), hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tile_rng, c->tile_rng.data(), sizeof(int2) * c->nrep, hipMemcpyHostToDevice,
                           c->stream));
  // the tile count is part of a recorded graph
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  if (ms_per_step_out) *ms_per_step_out = best;
  st = check_launch(c);
  if (st) return st;
  return check_flags(c);
}

extern "C" double mpcd_gpu_tile_fill(const mpcd_gpu_ctx *c) {
  if (!c) return 0.0;
  return bucket_usable(c) ? c->bkt_fill : c->tile_fill;
}

// ---------------------------------------------------------------------------
// checkpoints and the adjoint
// ---------------------------------------------------------------------------
// every device buffer whose contents carry over from one step to the next
static void state_buffers(mpcd_gpu_ctx *c, std::vector<std::pair<void *, size_t>> &v) {
  const size_t n = c->ntot, nc = c->nctot, nw3 = 3 * std::max(1, c->nwtot);
  const size_t rs = (c->precision == MPCD_FP64) ? sizeof(double) : sizeof(float);
  const bool lc = c->lc != MPCD_ISOF;
  auto add_state = [&](auto &S) {
    v.push_back({S.x, n * rs}); v.push_back({S.y, n * rs}); v.push_back({S.z, n * rs});
    v.push_back({S.vx, n * rs}); v.push_back({S.vy, n * rs}); v.push_back({S.vz, n * rs});
    if (lc) { v.push_back({S.ux, n * rs}); v.push_back({S.uy, n * rs}); v.push_back({S.uz, n * rs}); }
    v.push_back({S.id, n * 4}); v.push_back({S.sp, n * 4}); v.push_back({S.cst, n * 4});
  };
  if (c->precision == MPCD_FP64) add_state(c->sd[c->cur]);
  else add_state(c->sf[c->cur]);
  v.push_back({c->C.key, n * 8}); v.push_back({c->C.cnew, n * 4}); v.push_back({c->C.loc, n * 4});
  v.push_back({c->C.cu, n * 4});
  v.push_back({c->hist2[0], (nc + 1) * 4}); v.push_back({c->hist2[1], (nc + 1) * 4});
  v.push_back({c->shift2, 2 * 3 * c->nrep * sizeof(double)});
  v.push_back({c->wqs2, 2 * nw3 * sizeof(double)}); v.push_back({c->wqb2, 2 * nw3 * sizeof(double)});
  v.push_back({c->wh2, 2 * std::max(1, c->nwtot) * sizeof(WallHot)});
  v.push_back({c->box2, 2 * 6 * c->nrep * sizeof(double)});
  v.push_back({c->C.kbtnow, c->nrep * sizeof(double)}); v.push_back({c->C.gscale, c->nrep * sizeof(double)});
  v.push_back({c->C.cell.S, nc * sizeof(double)}); v.push_back({c->C.cell.dir, 3 * nc * sizeof(double)});
  if (c->xshifted) v.push_back({c->xshifted, n * sizeof(int32_t)});
}

// copies of the cell-sorted state (state_buffers) and the host step state
static mpcd_status save_state(mpcd_gpu_ctx *c, mpcd_gpu_ctx::Checkpoint &k) {
  std::vector<std::pair<void *, size_t>> v;
  state_buffers(c, v);
  if (k.bufs.empty()) {
    for (auto &b : v) {
      void *p = dalloc<uint8_t>(c, b.second);
      if (!p) return fail(c, MPCD_ENOMEM, "checkpoint");
      k.bufs.push_back(p);
    }
  }
  for (size_t i = 0; i < v.size(); i++)
    HIPCHK(c, hipMemcpyAsync(k.bufs[i], v[i].first, v[i].second, hipMemcpyDeviceToDevice, c->stream));
  k.step = c->step;
  k.cur = c->cur;
  k.pending = c->pending;
  k.hist_ready = c->hist_ready;
  k.need_init = c->need_init;
  k.cu_src = c->cu_src;
  return MPCD_OK;
}

static mpcd_status load_state(mpcd_gpu_ctx *c, const mpcd_gpu_ctx::Checkpoint &k) {
  c->in_bucket = false;  // the restored state is cell-sorted
  c->cur = k.cur;        // the saved buffer set is the one that was current
  std::vector<std::pair<void *, size_t>> v;
  state_buffers(c, v);
  for (size_t i = 0; i < v.size(); i++)
    HIPCHK(c, hipMemcpyAsync(v[i].first, k.bufs[i], v[i].second, hipMemcpyDeviceToDevice, c->stream));
  c->step = k.step;
  c->pending = k.pending;
  c->hist_ready = k.hist_ready;
  c->need_init = k.need_init;
  c->cu_src = k.cu_src;
  c->x_tables = 0xFFFFFFFFu;  // redrawn from the restored wall positions
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_checkpoint(mpcd_gpu_ctx *c, int slot) {
  if (!c || slot < 0) return MPCD_EINVAL;
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  if ((int)c->ckpts.size() <= slot) c->ckpts.resize(slot + 1);
  st = save_state(c, c->ckpts[slot]);
  if (st) return st;
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_restore(mpcd_gpu_ctx *c, int slot) {
  if (!c || slot < 0 || slot >= (int)c->ckpts.size() || c->ckpts[slot].bufs.empty())
    return c ? fail(c, MPCD_EINVAL, "restore: no checkpoint in slot %d", slot) : MPCD_EINVAL;
  mpcd_status st = load_state(c, c->ckpts[slot]);
  if (st) return st;
  st = zero_tangent(c);  // checkpoints hold no tangent: it restarts from zero
  if (st) return st;
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_record(mpcd_gpu_ctx *c, int max_steps) {
  if (!c || max_steps < 0) return MPCD_EINVAL;
  {
    mpcd_status st0 = ensure_dense(c);
    if (st0) return st0;
  }
  if (max_steps == 0) {
    c->recording = false;
    return MPCD_OK;
  }
  if (c->tan_on) return fail(c, MPCD_EINVAL, "the tangent and the adjoint tape cannot run together");
  if (c->any_curved) return fail(c, MPCD_EUNSUPPORTED, "the adjoint covers axis-aligned planar walls only");
  // the reverse mode of the whole step covers what the tangent covers; the
  // velocity-only tape of mpcd_gpu_step_backward, SRD and MPC-AT between
  // resting walls
  c->lin_ok = true;
  for (const mpcd_gpu_params &p : c->params) {
    if (p.tstech != MPCD_NOTHERM) return fail(c, MPCD_EUNSUPPORTED, "the adjoint needs tsTech = 0");
    if (p.lc == MPCD_LCG) return fail(c, MPCD_EUNSUPPORTED, "the adjoint needs the local order parameter (lc 1)");
    if (p.lc != MPCD_ISOF || !(p.rtech == MPCD_ARBAXIS || p.rtech == MPCD_ORTHAXIS || p.rtech == MPCD_MPCAT))
      c->lin_ok = false;
    for (int w = 0; w < p.nwall; w++)
      for (int j = 0; j < 3; j++)
        if (p.wall[w].L[j] != 0.0 || p.wall[w].V[j] != 0.0) c->lin_ok = false;
  }
  if (c->precision != MPCD_FP64) return fail(c, MPCD_EUNSUPPORTED, "the adjoint needs the FP64 state");
  c->rev_len = 0;
  c->rev_cap = max_steps;
  c->rev_bad = false;
  if (c->lin_ok && max_steps > c->tape_cap) {
    const size_t n = c->ntot, nc = c->nctot, k = max_steps;
    c->tp_id = dalloc<int32_t>(c, k * n);
    c->tp_sp = dalloc<int32_t>(c, k * n);
    c->tp_start = dalloc<int32_t>(c, k * (nc + 1));
    c->tp_v1 = dalloc<double>(c, k * 3 * n);
    c->tp_vcm = dalloc<double>(c, k * 3 * nc);
    c->tp_a = dalloc<double>(c, k * nc);
    c->tp_mass = dalloc<double>(c, k * nc);
    c->tp_M = dalloc<double>(c, k * 9 * nc);
    c->tp_dM = dalloc<double>(c, k * 9 * nc);
    c->tp_wmul = dalloc<float>(c, k * 3 * n);
    if (!c->lam) {
      c->lam = dalloc<double>(c, 3 * n);
      c->lam_s = dalloc<double>(c, 3 * n);
      c->lam1_s = dalloc<double>(c, 3 * n);
      c->adj_dev = dalloc<double>(c, 3 * n);
      c->dtheta = dalloc<double>(c, 4 * c->nrep);
    }
    if (!c->tp_id || !c->tp_sp || !c->tp_start || !c->tp_v1 || !c->tp_vcm || !c->tp_a || !c->tp_mass || !c->tp_M ||
        !c->tp_dM || !c->tp_wmul || !c->lam || !c->lam_s || !c->lam1_s || !c->adj_dev || !c->dtheta)
      return fail(c, MPCD_ENOMEM, "adjoint tape");
    c->tape_cap = max_steps;
  }
  c->tape_len = 0;
  c->recording = true;
  return MPCD_OK;
}

static mpcd_status rev_alloc(mpcd_gpu_ctx *c);

extern "C" mpcd_status mpcd_gpu_adjoint_reset(mpcd_gpu_ctx *c) {
  if (!c) return MPCD_EINVAL;
  if (!c->lam && c->rev_cap == 0) return fail(c, MPCD_EINVAL, "no adjoint state (mpcd_gpu_record first)");
  if (c->lam) {
    HIPCHK(c, hipMemsetAsync(c->lam, 0, sizeof(double) * 3 * c->ntot, c->stream));
    HIPCHK(c, hipMemsetAsync(c->dtheta, 0, sizeof(double) * 4 * c->nrep, c->stream));
  }
  mpcd_status st = rev_alloc(c);
  if (st) return st;
  HIPCHK(c, hipMemsetAsync(c->ra_lam, 0, sizeof(double) * 9 * c->ntot, c->stream));
  HIPCHK(c, hipMemsetAsync(c->ra_th, 0, sizeof(double) * kNTheta * c->nrep, c->stream));
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_step_backward(mpcd_gpu_ctx *c, const double *adj_in, double *adj_out,
                                              double *dJ_dtheta) {
  if (!c) return MPCD_EINVAL;
  if (!c->lin_ok)
    return fail(c, MPCD_EUNSUPPORTED,
                "step_backward: the velocity-only adjoint covers SRD and MPC-AT between resting walls "
                "(mpcd_gpu_step_backward_state covers every operator)");
  if (c->tape_len <= 0 || !c->lam) return fail(c, MPCD_EINVAL, "step_backward: the tape is empty");
  const TapeDev T = tape_at(c, --c->tape_len);
  const size_t n = c->ntot;
  if (adj_in) HIPCHK(c, hipMemcpyAsync(c->adj_dev, adj_in, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  k_adj_gather<<<nblocks(c->ntot, kBlock), kBlock, 0, c->stream>>>(T, c->lam, adj_in ? c->adj_dev : nullptr, c->lam_s,
                                                                    c->ntot);
  k_adj_cells<<<c->cblk.size(), kBlock, 0, c->stream>>>(c->C, T, c->d_cblk, c->lam_s, c->lam1_s, c->dtheta);
  k_adj_scatter<<<nblocks(c->ntot, kBlock), kBlock, 0, c->stream>>>(T, c->lam1_s, c->lam, c->ntot);
  if (adj_out) HIPCHK(c, hipMemcpyAsync(adj_out, c->lam, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  if (dJ_dtheta)
    HIPCHK(c, hipMemcpyAsync(dJ_dtheta, c->dtheta, sizeof(double) * 4 * c->nrep, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

// ---------------------------------------------------------------------------
// reverse mode of the whole step: the step is replayed from its recorded
// start state up to each kernel in turn, and that kernel's reverse launch
// (mpcd_adjoint.h) turns the adjoints of its outputs into those of its inputs
// ---------------------------------------------------------------------------
enum { RS_PROPS = 0, RS_VELGRAD, RS_GHOST, RS_LC, RS_COLLIDE, RS_STREAM };

static void rev_save_step(mpcd_gpu_ctx *c, int injected) {
  if (c->rev_len >= c->rev_cap) return;
  if (injected) c->rev_bad = true;  // the replay draws from the generator
  if ((int)c->rev_steps.size() <= c->rev_len) c->rev_steps.resize(c->rev_len + 1);
  if (save_state(c, c->rev_steps[c->rev_len]) != MPCD_OK) c->rev_bad = true;
  c->rev_len++;
}

static mpcd_status rev_alloc(mpcd_gpu_ctx *c) {
  if (c->ra_part) return MPCD_OK;
  const size_t n = c->ntot, nc = c->nctot;
  c->ra_part = dalloc<double>(c, 9 * n);
  c->ra_cell = dalloc<double>(c, 25 * nc);
  c->ra_T = dalloc<double>(c, 3 * n);
  c->ra_lam = dalloc<double>(c, 9 * n);
  c->ra_in = dalloc<double>(c, 9 * n);
  c->ra_th = dalloc<double>(c, (size_t)kNTheta * c->nrep);
  c->ra_obs = dalloc<double>(c, 2 * (size_t)c->nrep);
  c->rt_need = dalloc<int>(c, 4);
  if (!c->ra_part || !c->ra_cell || !c->ra_T || !c->ra_lam || !c->ra_in || !c->ra_th || !c->ra_obs || !c->rt_need) {
    c->ra_part = nullptr;
    return fail(c, MPCD_ENOMEM, "adjoint arrays");
  }
  HIPCHK(c, hipMemsetAsync(c->ra_lam, 0, sizeof(double) * 9 * n, c->stream));
  HIPCHK(c, hipMemsetAsync(c->ra_th, 0, sizeof(double) * kNTheta * c->nrep, c->stream));
  return MPCD_OK;
}

// the kernels' view of the adjoint arrays (the tangent's layout)
static TanDev rev_view(mpcd_gpu_ctx *c, bool with_theta) {
  TanDev T{};
  const size_t n = c->ntot, nc = c->nctot;
  for (int k = 0; k < 3; k++) {
    T.q[k] = c->ra_part + k * n;
    T.v[k] = c->ra_part + (3 + k) * n;
    T.u[k] = c->ra_part + (6 + k) * n;
  }
  T.vcm = c->ra_cell;
  T.cm = c->ra_cell + 3 * nc;
  T.I = c->ra_cell + 6 * nc;
  T.S = c->ra_cell + 12 * nc;
  T.dir = c->ra_cell + 13 * nc;
  T.E = c->ra_cell + 16 * nc;
  T.T = c->ra_T;
  T.dth = with_theta ? c->ra_th : nullptr;
  return T;
}

// one reverse launch over nblk workgroups: a counting launch sizes the tape,
// then the recording launch sweeps it.  A reverse launch stores the forward
// values as it goes, so prep() puts the kernel's inputs back in between.
template <typename L, typename P>
static mpcd_status rev_launch(mpcd_gpu_ctx *c, int nblk, L launch, P prep) {
  if (nblk <= 0) return MPCD_OK;
  RevTape t{};
  t.stride = nblk * kBlock;
  t.need = c->rt_need;
  t.counting = 1;
  k_fill<<<1, 4, 0, c->stream>>>(c->rt_need, 0, 4);
  launch(t);
  int need[4];
  HIPCHK(c, hipMemcpyAsync(need, c->rt_need, sizeof need, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  mpcd_status st = check_launch(c);
  if (st) return st;
  const size_t ent = (size_t)std::max(need[0], 1), outs = (size_t)std::max(need[1], 1);
  if (ent * t.stride > c->rt_entries) {
    dfree(c, c->rt_par);
    dfree(c, c->rt_dp);
    dfree(c, c->rt_adj);
    c->rt_entries = 0;
    const size_t m = ent * t.stride;
    c->rt_par = dalloc<int2>(c, m);
    c->rt_dp = dalloc<double2>(c, m);
    c->rt_adj = dalloc<double>(c, m);
    if (!c->rt_par || !c->rt_dp || !c->rt_adj) return fail(c, MPCD_ENOMEM, "adjoint tape (%zu entries)", m);
    c->rt_entries = m;
  }
  if (outs * t.stride > c->rt_outs) {
    dfree(c, c->rt_optr);
    dfree(c, c->rt_ovar);
    c->rt_outs = 0;
    const size_t m = outs * t.stride;
    c->rt_optr = dalloc<double *>(c, m);
    c->rt_ovar = dalloc<int>(c, m);
    if (!c->rt_optr || !c->rt_ovar) return fail(c, MPCD_ENOMEM, "adjoint outputs (%zu)", m);
    c->rt_outs = m;
  }
  prep();
  t.counting = 0;
  t.cap = (int)ent;
  t.ocap = (int)outs;
  t.par = c->rt_par;
  t.dp = c->rt_dp;
  t.adj = c->rt_adj;
  t.optr = c->rt_optr;
  t.ovar = c->rt_ovar;
  launch(t);
  HIPCHK(c, hipMemcpyAsync(need, c->rt_need, sizeof need, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  st = check_launch(c);
  if (st) return st;
  if (need[2]) return fail(c, MPCD_EDEVICE, "adjoint tape: a lane recorded more than its counting launch");
  return MPCD_OK;
}

// step k of the tape replayed (FP64, generator draws) from its start state
// up to, not including, the kernel of `stage`
static void rev_replay(mpcd_gpu_ctx *c, int k, int stage) {
  (void)load_state(c, c->rev_steps[k]);
  run_bin<double>(c, 1, true, 0);
  if (stage <= RS_VELGRAD) return;
  run_props<double>(c, 1);
  Ctx C = view(c, c->step & 1);
  Draws D = make_draws(c, 0, c->step);
  State<double> &S = c->sd[c->cur];
  if (stage <= RS_GHOST) return;
  if (c->any_phantom) k_ghost<double, double><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S, c->d_cblk, D, TanDev{});
  if (stage <= RS_LC) return;
  if (c->lc != MPCD_ISOF) k_lc<double, double><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, S, c->d_pblk, D, TanDev{});
  if (stage <= RS_COLLIDE) return;
  k_collide<double, double><<<c->cblk.size(), kBlock, 0, c->stream>>>(C, S, c->d_cblk, D, TanDev{});
}

static mpcd_status rev_stage(mpcd_gpu_ctx *c, int k, int stage) {
  Ctx C = view(c, c->step & 1);
  C.cdec = nullptr;
  C.pdec = nullptr;
  Draws D = make_draws(c, 0, c->step);
  State<double> S = c->sd[c->cur];
  const TanDev A = rev_view(c, true);
  const int npb = (int)c->pblk.size(), ncb = (int)c->cblk.size();
  auto prep = [&]() { rev_replay(c, k, stage); };
  switch (stage) {
    case RS_STREAM: {
      const double *box = c->box2 + (c->step & 1) * 6 * c->nrep;
      return rev_launch(c, npb, [&](const RevTape &t) {
        k_rev_stream<double><<<npb, kBlock, 0, c->stream>>>(C, S, c->d_pblk, D, A, box, t);
      }, prep);
    }
    case RS_COLLIDE:
      return rev_launch(c, ncb, [&](const RevTape &t) {
        k_rev_collide<double><<<ncb, kBlock, 0, c->stream>>>(C, S, c->d_cblk, D, A, t);
      }, prep);
    case RS_LC:
      return rev_launch(c, npb, [&](const RevTape &t) {
        k_rev_lc<double><<<npb, kBlock, 0, c->stream>>>(C, S, c->d_pblk, D, A, t);
      }, prep);
    case RS_GHOST:
      return rev_launch(c, ncb, [&](const RevTape &t) {
        k_rev_ghost<double><<<ncb, kBlock, 0, c->stream>>>(C, S, c->d_cblk, D, A, t);
      }, prep);
    case RS_VELGRAD:
      k_velgrad_edges_adj<<<c->nrep, 256, 0, c->stream>>>(C, A.E);
      k_velgrad_bulk_adj<<<ncb, kBlock, 0, c->stream>>>(C, c->d_cblk, A.vcm, A.E);
      return check_launch(c);
    default:
      return rev_launch(c, ncb, [&](const RevTape &t) {
        k_rev_cellprops<double><<<ncb, kBlock, 0, c->stream>>>(C, S, c->d_cblk, 1, A, t);
      }, prep);
  }
}

extern "C" mpcd_status mpcd_gpu_step_backward_state(mpcd_gpu_ctx *c, const double *adj_in, double *adj_out,
                                                    double *dJ_dtheta) {
  if (!c) return MPCD_EINVAL;
  if (c->rev_len <= 0) return fail(c, MPCD_EINVAL, "step_backward_state: the tape is empty");
  if (c->rev_bad)
    return fail(c, MPCD_EUNSUPPORTED,
                "step_backward_state: a recorded step used injected draws (the replay draws from the generator)");
  mpcd_status st = rev_alloc(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  const size_t n = c->ntot, nc = c->nctot;
  // the context's own state and event counts come back at the end
  st = save_state(c, c->rev_here);
  if (st) return st;
  unsigned long long events[4];
  HIPCHK(c, hipMemcpyAsync(events, c->C.events, sizeof events, hipMemcpyDeviceToHost, c->stream));
  const bool rec = c->recording;
  c->recording = false;
  c->rev_replay = true;
  const int k = --c->rev_len;
  if (adj_in) {
    HIPCHK(c, hipMemcpyAsync(c->ra_in, adj_in, sizeof(double) * 9 * n, hipMemcpyHostToDevice, c->stream));
    k_add_f64<<<nblocks((int)(9 * n), kBlock), kBlock, 0, c->stream>>>(c->ra_lam, c->ra_in, 9 * n);
  }
  // the slot order of the step: its shifted bin
  rev_replay(c, k, RS_PROPS);
  {
    const TanDev A = rev_view(c, true);
    k_rev_to_slots<double><<<c->pblk.size(), kBlock, 0, c->stream>>>(c->sd[c->cur], c->d_pblk, c->d_rep, c->ra_lam,
                                                                     A);
    k_fill_f64<<<nblocks((int)(25 * nc), kBlock), kBlock, 0, c->stream>>>(c->ra_cell, 0.0, 25 * nc);
    k_fill_f64<<<nblocks((int)(3 * n), kBlock), kBlock, 0, c->stream>>>(c->ra_T, 0.0, 3 * n);
  }
  for (int stage = RS_STREAM; stage >= RS_PROPS && !st; stage--) {
    if ((stage == RS_LC || stage == RS_VELGRAD) && c->lc == MPCD_ISOF) continue;
    if (stage == RS_GHOST && !c->any_phantom) continue;
    if (stage != RS_VELGRAD) rev_replay(c, k, stage);
    st = rev_stage(c, k, stage);
  }
  if (!st) {
    // the cell-sorted state before the bin is the replay's other buffer
    const TanDev A = rev_view(c, true);
    k_rev_from_slots<double><<<c->pblk.size(), kBlock, 0, c->stream>>>(view(c, c->step & 1), c->sd[1 - c->cur],
                                                                       c->d_pblk, A, c->ra_lam);
  }
  c->rev_replay = false;
  c->recording = rec;
  {
    mpcd_status s2 = load_state(c, c->rev_here);
    if (!st) st = s2;
  }
  HIPCHK(c, hipMemcpyAsync(c->C.events, events, sizeof events, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (st) return st;
  if (adj_out) HIPCHK(c, hipMemcpyAsync(adj_out, c->ra_lam, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->stream));
  if (dJ_dtheta)
    HIPCHK(c, hipMemcpyAsync(dJ_dtheta, c->ra_th, sizeof(double) * kNTheta * c->nrep, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_observable(mpcd_gpu_ctx *c, int kind, const double *arg, double *value,
                                           double *grad) {
  if (!c || !value) return c ? fail(c, MPCD_EINVAL, "observable: null argument") : MPCD_EINVAL;
  if (kind != MPCD_OBS_BAND_VELOCITY && kind != MPCD_OBS_MEAN_S)
    return fail(c, MPCD_EINVAL, "observable: unknown kind %d", kind);
  if (kind == MPCD_OBS_MEAN_S && c->lc != MPCD_LCL)
    return fail(c, MPCD_EINVAL, "observable: the mean order parameter needs the local nematic (lc 1)");
  int comp = 0, axis = 1, lo = 0, hi = 0;
  if (kind == MPCD_OBS_BAND_VELOCITY) {
    if (!arg) return fail(c, MPCD_EINVAL, "observable: the band needs its arguments");
    comp = (int)arg[0];
    axis = (int)arg[1];
    lo = (int)arg[2];
    hi = (int)arg[3];
    if (comp < 0 || comp > 2 || axis < 0 || axis > 2) return fail(c, MPCD_EINVAL, "observable: bad component or axis");
  }
  if (grad && c->precision != MPCD_FP64) return fail(c, MPCD_EUNSUPPORTED, "observable: the gradient needs FP64");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  st = rev_alloc(c);
  if (st) return st;
  const bool fp64 = c->precision == MPCD_FP64;
  if (c->pending) {
    // the cells of the re-bin that ends the step (mpc.c:4936-4942)
    if (fp64) {
      run_bin<double>(c, 2, true, 0);
      run_props<double>(c, 0);
    } else {
      run_bin<float>(c, 2, true, 0);
      run_props<float>(c, 0);
    }
  }
  Ctx C = view(c, c->step & 1);
  const size_t n = c->ntot;
  k_fill_f64<<<1, 64, 0, c->stream>>>(c->ra_obs, 0.0, 2 * (size_t)c->nrep);
  if (kind == MPCD_OBS_BAND_VELOCITY) {
    if (fp64)
      k_obs_band<double><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, c->sd[c->cur], c->d_pblk, comp, axis, lo, hi,
                                                                   c->ra_obs);
    else
      k_obs_band<float><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, c->sf[c->cur], c->d_pblk, comp, axis, lo, hi,
                                                                  c->ra_obs);
  } else {
    k_obs_meanS<<<c->cblk.size(), kBlock, 0, c->stream>>>(C, c->d_cblk, c->ra_obs);
  }
  std::vector<double> sums(2 * c->nrep);
  HIPCHK(c, hipMemcpyAsync(sums.data(), c->ra_obs, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int r = 0; r < c->nrep; r++) value[r] = sums[2 * r] > 0.0 ? sums[2 * r + 1] / sums[2 * r] : 0.0;
  if (!grad) return check_launch(c);
  if (kind == MPCD_OBS_BAND_VELOCITY) {
    k_obs_band_adj<double><<<c->pblk.size(), kBlock, 0, c->stream>>>(C, c->sd[c->cur], c->d_pblk, comp, axis, lo, hi,
                                                                      c->ra_obs, c->ra_in);
  } else {
    // dS/d(orientations) of the cells' localPROP: its reverse launch with the
    // adjoint 1 / count at every counted cell's S
    const size_t nc = c->nctot;
    k_fill_f64<<<nblocks((int)(25 * nc), kBlock), kBlock, 0, c->stream>>>(c->ra_cell, 0.0, 25 * nc);
    k_fill_f64<<<nblocks((int)(9 * n), kBlock), kBlock, 0, c->stream>>>(c->ra_part, 0.0, 9 * n);
    k_fill_f64<<<nblocks((int)(3 * n), kBlock), kBlock, 0, c->stream>>>(c->ra_T, 0.0, 3 * n);
    const TanDev A = rev_view(c, false);
    k_obs_meanS_seed<<<c->cblk.size(), kBlock, 0, c->stream>>>(C, c->d_cblk, c->ra_obs, A.S);
    Ctx Cr = C;
    Cr.cdec = nullptr;
    State<double> S = c->sd[c->cur];
    const int ncb = (int)c->cblk.size();
    // without the body force the cell properties leave the particles as
    // they are: the recording launch sees the counting launch's inputs
    st = rev_launch(c, ncb, [&](const RevTape &t) {
      k_rev_cellprops<double><<<ncb, kBlock, 0, c->stream>>>(Cr, S, c->d_cblk, 0, A, t);
    }, []() {});
    if (st) return st;
    k_rev_unslot<double><<<c->pblk.size(), kBlock, 0, c->stream>>>(S, c->d_pblk, c->d_rep, A, c->ra_in);
  }
  HIPCHK(c, hipMemcpyAsync(grad, c->ra_in, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

static mpcd_status zero_tangent(mpcd_gpu_ctx *c) {
  if (!c->tan_part[0]) return MPCD_OK;
  const size_t n = c->ntot, nc = c->nctot;
  for (int b = 0; b < 2; b++) HIPCHK(c, hipMemsetAsync(c->tan_part[b], 0, sizeof(double) * 9 * n, c->stream));
  HIPCHK(c, hipMemsetAsync(c->tan_cell, 0, sizeof(double) * 25 * nc, c->stream));
  HIPCHK(c, hipMemsetAsync(c->tan_T, 0, sizeof(double) * 3 * n, c->stream));
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_tangent(mpcd_gpu_ctx *c, int on, const double *dgrav, const double *dact) {
  if (!c) return MPCD_EINVAL;
  if (!on) {
    c->tan_on = false;
    return MPCD_OK;
  }
  if (c->precision != MPCD_FP64) return fail(c, MPCD_EUNSUPPORTED, "the tangent needs the FP64 state");
  if (c->recording) return fail(c, MPCD_EINVAL, "the tangent and the adjoint tape cannot run together");
  if (c->any_curved) return fail(c, MPCD_EUNSUPPORTED, "the tangent covers axis-aligned planar walls only");
  for (const mpcd_gpu_params &p : c->params) {
    if (p.tstech != MPCD_NOTHERM) return fail(c, MPCD_EUNSUPPORTED, "the tangent needs tsTech = 0");
    if (p.lc == MPCD_LCG) return fail(c, MPCD_EUNSUPPORTED, "the tangent needs the local order parameter (lc 1)");
  }
  mpcd_status st = ensure_dense(c);
  if (st) return st;
  if (!c->tan_part[0]) {
    const size_t n = c->ntot, nc = c->nctot;
    c->tan_part[0] = dalloc<double>(c, 9 * n);
    c->tan_part[1] = dalloc<double>(c, 9 * n);
    c->tan_cell = dalloc<double>(c, 25 * nc);
    c->tan_T = dalloc<double>(c, 3 * n);
    if (!c->tan_part[0] || !c->tan_part[1] || !c->tan_cell || !c->tan_T) {
      c->tan_part[0] = c->tan_part[1] = nullptr;
      return fail(c, MPCD_ENOMEM, "tangent arrays");
    }
  } else {
    st = zero_tangent(c);
    if (st) return st;
  }
  for (int k = 0; k < 3; k++) c->tan_dgrav[k] = dgrav ? dgrav[k] : 0.0;
  for (int k = 0; k < MPCD_MAXSPEC; k++) c->tan_dact[k] = (dact && k < c->params[0].nspec) ? dact[k] : 0.0;
  c->tan_on = true;
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_tangent_download(mpcd_gpu_ctx *c, int r, double *dQ, double *dV, double *dU) {
  if (!c || r < 0 || r >= c->nrep || !dQ || !dV)
    return c ? fail(c, MPCD_EINVAL, "tangent_download: bad argument") : MPCD_EINVAL;
  if (!c->tan_part[0]) return fail(c, MPCD_EINVAL, "no tangent (mpcd_gpu_tangent first)");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  const RepDev &R = c->reps[r];
  st = ensure_stage(c, R.npart);
  if (st) return st;
  // the tangent arrays in the layout of a state, with the slots' particle ids
  const TanDev T = tan_view(c, c->cur);
  State<double> V{};
  V.x = T.q[0]; V.y = T.q[1]; V.z = T.q[2];
  V.vx = T.v[0]; V.vy = T.v[1]; V.vz = T.v[2];
  V.ux = T.u[0]; V.uy = T.u[1]; V.uz = T.u[2];
  V.id = c->sd[c->cur].id;
  const int lc = c->lc != MPCD_ISOF;
  const size_t n = R.npart;
  k_unsort<double><<<nblocks(R.npart, kBlock), kBlock, 0, c->stream>>>(V, R.part_off, R.npart, c->d_q, c->d_v, c->d_u,
                                                                     lc);
  HIPCHK(c, hipMemcpyAsync(dQ, c->d_q, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(dV, c->d_v, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  if (dU) {
    if (lc) HIPCHK(c, hipMemcpyAsync(dU, c->d_u, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
    else memset(dU, 0, sizeof(double) * 3 * n);
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_synchronize(mpcd_gpu_ctx *c) {
  if (!c) return MPCD_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  mpcd_status st = check_launch(c);
  if (st) return st;
  return check_flags(c);
}

extern "C" mpcd_status mpcd_gpu_get_step(const mpcd_gpu_ctx *c, uint32_t *step) {
  if (!c || !step) return MPCD_EINVAL;
  *step = c->step;
  return MPCD_OK;
}

// the draw counter: every draw is keyed by (seed, replica, step, ...), so a
// state saved with its step and restored with it continues the same run
extern "C" mpcd_status mpcd_gpu_set_step(mpcd_gpu_ctx *c, uint32_t step) {
  if (!c) return MPCD_EINVAL;
  if (c->recording || c->tan_on) return fail(c, MPCD_EUNSUPPORTED, "set_step: not while a tape or tangent runs");
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  c->step = step;
  c->pending = false;  // the next step bins the state afresh at this step
  c->hist_ready = false;
  c->x_tables = step ^ 0x80000000u;
  drop_bucket_graph(c);
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_kbtnow(mpcd_gpu_ctx *c, double *kbt_now) {
  if (!c || !kbt_now) return c ? fail(c, MPCD_EINVAL, "kbtnow: null argument") : MPCD_EINVAL;
  return read_kbt(c, kbt_now);
}

extern "C" mpcd_status mpcd_gpu_lc_record(mpcd_gpu_ctx *c, int on) {
  if (!c) return MPCD_EINVAL;
  if (c->lc == MPCD_ISOF) return fail(c, MPCD_EINVAL, "lc_record: the context has no nematic");
  if (on && !c->C.pdec) {
    c->C.pdec = (uint32_t *)dalloc<int32_t>(c, c->ntot);
    c->C.cdec = (uint32_t *)dalloc<int32_t>(c, c->nctot);
    if (!c->C.pdec || !c->C.cdec) return fail(c, MPCD_ENOMEM, "decision records");
    k_fill<<<nblocks(c->ntot, kBlock), kBlock, 0, c->stream>>>((int32_t *)c->C.pdec, 0, c->ntot);
    k_fill<<<nblocks(c->nctot, kBlock), kBlock, 0, c->stream>>>((int32_t *)c->C.cdec, 0, c->nctot);
  }
  if (!on) {
    c->C.pdec = nullptr;  // the buffers stay with the context
    c->C.cdec = nullptr;
  }
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_lc_decisions(mpcd_gpu_ctx *c, int r, uint32_t *part, uint32_t *cell) {
  if (!c || r < 0 || r >= c->nrep) return c ? fail(c, MPCD_EINVAL, "lc_decisions: bad replica") : MPCD_EINVAL;
  if (!c->C.pdec) return fail(c, MPCD_EINVAL, "lc_decisions: recording is off (mpcd_gpu_lc_record)");
  const RepDev &R = c->reps[r];
  if (part)
    HIPCHK(c, hipMemcpyAsync(part, c->C.pdec + R.part_off, sizeof(uint32_t) * R.npart, hipMemcpyDeviceToHost,
                             c->stream));
  if (cell)
    HIPCHK(c, hipMemcpyAsync(cell, c->C.cdec + R.cell_off, sizeof(uint32_t) * R.ncell, hipMemcpyDeviceToHost,
                             c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_get_cells(mpcd_gpu_ctx *c, int r, int32_t *pop, double *vcm, double *S, double *dir) {
  if (!c || r < 0 || r >= c->nrep) return MPCD_EINVAL;
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  if (c->pending) {
    // the reference re-bins and recomputes localPROP after streaming (mpc.c:4936-4942)
    if (c->precision == MPCD_FP64) {
      run_bin<double>(c, 2, true, 0);
      run_props<double>(c, 0);
    } else {
      run_bin<float>(c, 2, true, 0);
      run_props<float>(c, 0);
    }
  }
  const RepDev &R = c->reps[r];
  size_t off = R.cell_off, n = R.ncell;
  if (pop) HIPCHK(c, hipMemcpyAsync(pop, c->C.cell.pop + off, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
  if (vcm) HIPCHK(c, hipMemcpyAsync(vcm, c->C.cell.vcm + 3 * off, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  if (S) HIPCHK(c, hipMemcpyAsync(S, c->C.cell.S + off, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  if (dir) HIPCHK(c, hipMemcpyAsync(dir, c->C.cell.dir + 3 * off, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_flow(mpcd_gpu_ctx *c, int on) {
  if (!c) return MPCD_EINVAL;
  if (c->exact && on) return fail(c, MPCD_EUNSUPPORTED, "the flow field is not summed in the exact-sum FP32 form");
  if ((on != 0) == c->flow_on) return MPCD_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (on && !c->d_flow) {
    c->d_flow = dalloc<double>(c, 3 * (size_t)c->nctot);
    c->d_flow_acc = dalloc<double>(c, 4 * (size_t)c->nctot);
    c->d_flow_n = dalloc<int64_t>(c, c->nrep);
    c->d_flow_part = dalloc<double>(c, kFlowCls * 4 * (size_t)c->nctot);
    if (!c->d_flow || !c->d_flow_acc || !c->d_flow_n || !c->d_flow_part) return fail(c, MPCD_ENOMEM, "flow field");
  }
  c->flow_on = on != 0;
  // recorded graphs hold the step's launches, with or without the flow pair
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  drop_bucket_graph(c);
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_flow_read(mpcd_gpu_ctx *c, int r, double *flow, int64_t *nsteps, int reset) {
  if (!c || r < 0 || r >= c->nrep) return MPCD_EINVAL;
  const RepDev &R = c->reps[r];
  const size_t off = R.cell_off, n = R.ncell;
  if (!c->d_flow) {
    // never switched on: nothing summed
    if (flow) memset(flow, 0, sizeof(double) * 3 * n);
    if (nsteps) *nsteps = 0;
    return MPCD_OK;
  }
  if (flow)
    HIPCHK(c, hipMemcpyAsync(flow, c->d_flow + 3 * off, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  if (nsteps) HIPCHK(c, hipMemcpyAsync(nsteps, c->d_flow_n + r, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  if (reset) {
    HIPCHK(c, hipMemsetAsync(c->d_flow + 3 * off, 0, sizeof(double) * 3 * n, c->stream));
    HIPCHK(c, hipMemsetAsync(c->d_flow_n + r, 0, sizeof(int64_t), c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_get_order(mpcd_gpu_ctx *c, int r, int32_t *order, int32_t *cstart) {
  if (!c || r < 0 || r >= c->nrep) return MPCD_EINVAL;
  mpcd_status st = ensure_init(c);
  if (st) return st;
  st = ensure_dense(c);
  if (st) return st;
  st = ensure_stage(c, c->ntot);
  if (st) return st;
  const RepDev &R = c->reps[r];
  const int32_t *id = (c->precision == MPCD_FP64) ? c->sd[c->cur].id : c->sf[c->cur].id;
  // keys of the current logical order, without moving the state
  if (c->precision == MPCD_FP64) run_bin<double>(c, 2, false, 0);
  else run_bin<float>(c, 2, false, 0);
  k_order_ids<<<nblocks(R.npart, kBlock), kBlock, 0, c->stream>>>(c->C.skey, id, c->d_i32, R.part_off, R.npart);
  if (order) HIPCHK(c, hipMemcpyAsync(order, c->d_i32, sizeof(int32_t) * R.npart, hipMemcpyDeviceToHost, c->stream));
  if (cstart) {
    std::vector<int32_t> h(R.ncell + 1);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->C.start + R.cell_off, sizeof(int32_t) * (R.ncell + 1), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i <= R.ncell; i++) cstart[i] = h[i] - R.part_off;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_launch(c);
}

extern "C" mpcd_status mpcd_gpu_profile(mpcd_gpu_ctx *c, int enable) {
  if (!c) return MPCD_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->prof = enable != 0;
  c->prof_recs.clear();
  c->ev_used = 0;
  return MPCD_OK;
}

extern "C" mpcd_status mpcd_gpu_timing(mpcd_gpu_ctx *c, double *out, int n) {
  if (!c || !out) return MPCD_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::vector<double> tot(F_N, 0.0), cnt(F_N, 0.0);
  for (auto &r : c->prof_recs) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, r.a, r.b));
    tot[r.fam] += ms;
    cnt[r.fam] += 1.0;
  }
  for (int i = 0; i < n; i++) out[i] = 0.0;
  for (int f = 0; f < F_N; f++) {
    if (f < n) out[f] = tot[f];
    if (F_N + f < n) out[F_N + f] = cnt[f];
  }
  return MPCD_OK;
}
