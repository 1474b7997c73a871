/*
 * mpcd_gpu.h -- C-ABI of the MI355X (gfx950) MPCD solvent library.
 *
 * Drop-in boundary for the per-timestep solvent path of NIAMH-MPCD.  The
 * reference has no plugin or FFI layer; its seam is the sequence of calls
 * inside timestep() (mpcd/subroutines/mpc.c:4567-4942) made from main()
 * (mpcd/mpcd.c:268,298).  The entry points below replace those calls:
 *
 *   mpcd_gpu_bin      ranshift + gridShift_all(+)   mpc.c:4567-4575
 *                     bin(shifted=1)                 mpc.c:4596 (bin: 948-992)
 *   mpcd_gpu_collide  acc_all                        mpc.c:4621 (697-700)
 *                     localPROP                      mpc.c:4630 (64-216)
 *                     ghostPart                      mpc.c:4634 (351-567)
 *                     LCcollision, jefferysTorque    mpc.c:4640-4678 (lc.c:427-718)
 *                     MPCcollision per cell          mpc.c:4699-4714 (2777-2812)
 *                     TEMP, scaleT                   mpc.c:4752-4778 (therm.c:52-235)
 *                     gridShift_all(-)               mpc.c:4793-4799
 *   mpcd_gpu_stream   stream_all                     mpc.c:4815-4817 (680-687)
 *                     MPC_BCcollision per particle   mpc.c:4834 (bc.c:1363-1496)
 *                     bin(shifted=0) + localPROP     mpc.c:4936-4942
 *
 * Plain C types only: no C++ or torch types cross this boundary.  All device
 * memory is owned by the context; host arrays are borrowed for the duration
 * of a call.  Calls on one context are stream ordered; a call that returns
 * host data synchronises.  One context may hold many independent replicas
 * (different boxes, densities, walls, seeds) that advance together.
 *
 * Injected draws (parity mode).  When a step is given an mpcd_draws, every
 * random variate is read from host-provided arrays in the keyed layout below
 * instead of the internal counter-based generator (Philox-4x32-10 keyed by
 * seed, replica, step, purpose, key, k).  Keys:
 *   shift   [3 per replica]            x, y, z grid shift in [0,1)   (rand.c:678-696)
 *   cell    [4 per cell]               SRD: 3D ARBAXIS unit axis in [0..2] (ranvec3D),
 *                                      2D ARBAXIS / ORTHAXIS uniform in [0],
 *                                      rotation-sign uniform in [3]  (mpc.c:1286-1291)
 *   ghost   [3 per cell]               ghost-particle normals        (mpc.c:553)
 *   part    [3 per particle slot]      thermal normals of MPC-AT(+a) and LC collisions
 *                                      (mpc.c:1382, 1506; lc.c:1559, 1854)
 *   lc      [lc_stride per slot]       Maier-Saupe sampler variates, slot 0 = 2D angle
 *                                      redraw, 1+2i / 2+2i = trial / acceptance of
 *                                      anneal step i, then phi and sign (3D)  (lc.c:52-411)
 *   escape  [MPCD_ESC_STRIDE per particle and binning]   bin()'s escape repair
 *                                      (mpc.c:970-982); binning b = 0 is the step's shifted
 *                                      bin (mpc.c:4596), b = 1 the unshifted bin after
 *                                      streaming (mpc.c:4936); entry (b * n_total + p) for
 *                                      particle p (replica offset + particle index):
 *                                      [0..2] Maxwell normals (genrand_gaussMB), [3..4]
 *                                      orientation uniforms of genrand_coneNP (2D: the
 *                                      angle; 3D: z, phi), [5 + DIM * t + j] uniform j of
 *                                      position try t of replacePos_WithCheck (init.c:1462)
 *   retherm [MPCD_RETHERM_STRIDE per particle]   wall re-thermalisation after
 *                                      NBOUNCE bounces (push(), bc.c:1476, init.c:1261):
 *                                      [3r + j] variate j of the particle's r-th push of
 *                                      the step (uniforms for VDIST 0 and 3, normals for
 *                                      2 and 4)
 * Cells are numbered x-major over (L0+1) x (L1+1) x (L2+1) (z extent 1 in 2D),
 * concatenated over replicas; particle slots are positions in the collision
 * order (cells x-major, each cell in the reference's linked-list order),
 * concatenated over replicas.  Normals are standard normals (the library
 * multiplies by sqrt(kT/m)); uniforms lie in [0,1).  escape and retherm may
 * be NULL: those rare draws then come from the counter-based generator.
 * The escape and retherm arrays given to a step's mpcd_gpu_bin(shifted = 1)
 * or mpcd_gpu_collide also serve its mpcd_gpu_stream, which takes no draws.
 * A particle that needs more position tries or pushes in one step than its
 * injected entry holds fails the call with MPCD_EINVAL; with the generator
 * the tries are unbounded up to 65536 (a box the walls leave no room in).
 */
#ifndef MPCD_GPU_H
#define MPCD_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPCD_MAXSPEC 10
#define MPCD_MAXWALL 16
#define MPCD_ESC_STRIDE 32      /* keyed escape draws per particle and binning */
#define MPCD_RETHERM_STRIDE 24  /* keyed re-thermalisation draws per particle (8 pushes) */

typedef enum {
    MPCD_OK = 0,
    MPCD_EINVAL = 1,        /* bad argument (reference: printf + exit(1)) */
    MPCD_ENOMEM = 2,        /* device allocation failed */
    MPCD_EDEVICE = 3,       /* HIP runtime error */
    MPCD_EUNSUPPORTED = 4,  /* option outside this path (e.g. mpc.c:2806 "Collision technique unacceptable") */
    MPCD_EESCAPE = 5,       /* particles escaped the box and were re-placed (mpc.c:970-982) */
    MPCD_ECAPACITY = 6      /* a cell exceeded the library's per-cell population limit */
} mpcd_status;

/* collision operators (definitions.h:261-303) supported on this path */
enum { MPCD_ARBAXIS = 0, MPCD_ORTHAXIS = 1, MPCD_MPCAT = 2, MPCD_RAT = 3, MPCD_DIPOLE_DIR_SUM = 16,
       MPCD_DIPOLE_DIR_AV = 17, MPCD_DIPOLE_DIR_SIG = 20, MPCD_DIPOLE_DIR_SIG_SUM = 21 };
/* liquid-crystal modes (definitions.h:194-200) and thermostats (definitions.h:238-246) */
enum { MPCD_ISOF = 0, MPCD_LCL = 1, MPCD_LCG = 2 };
enum { MPCD_NOTHERM = 0, MPCD_VSC = 1, MPCD_BEREND = 2 };
/* state precision.  MPCD_FP64 and MPCD_FP32 step the reference's algorithm in
 * its list order with FP64 cell sums (FP32: positions and velocities stored
 * in FP32).  MPCD_FP32_EXACT (no counterpart in the reference) is the batched
 * FP32 form of the isotropic operators without a thermostat: FP32 state and
 * particle arithmetic, cell sums of terms quantised to 2^-28 and added
 * exactly in FP64, so every sum -- and the step -- is independent of the
 * particles' order; whole steps only (no phase entry points, no injected
 * draws, no flow field). */
enum { MPCD_FP64 = 0, MPCD_FP32 = 1, MPCD_FP32_EXACT = 2 };

/* species -- subset of the reference's spec (SRDclss.h:54-77) */
typedef struct {
    double mass, rfc, tumble, chihi, act, mfpot, damp, sigwidth, sigpos, minactratio;
    int vdist; /* velocity distribution used by wall re-thermalisation (push(), init.c:1261) */
    int pad_;
} mpcd_gpu_species;

/* wall -- subset of the reference's bc (SRDclss.h:128-262) for immobile walls
 * with rule-based collisions (COLL_TYPE 1) and the four-fold rotational
 * symmetry (ROTSYMM 4 4): the surface W(r) = sum_i (A_i (r_i - Q_i))^P_i -
 * R^P_3 (+ the wavy term B, |.| terms if ABS, negated if INV) of calcW
 * (bc.c:55-150).  Axis-aligned planes (P = 1 1 1, B[0] = 0, ABS = 0, one
 * nonzero A) take the planar fast path; periodic planes have DN = box length,
 * MVN = MVT = 1; bounce-back planes DN = 0, MVN = MVT = -1, PHANTOM = 1
 * (read.c:1441-1583).  P, B and ABS are read.c's "P", "wavy" and "abs"
 * (read.c:1286-1301,1394-1409); a P left all zero is taken as 1 1 1 1 (a
 * plane), so zero-initialised records of planar walls need not set it. */
typedef struct {
    double A[3], Q[3], R, DN, DT, DVN, DVT, DVxyz[3], MVN, MVT, MUN, MUT, MUxyz[3], DUxyz[3], V[3], L[3];
    int INV, PHANTOM, COLL_TYPE;
    int inter_mask; /* bit s set: species s interacts with this wall (bc.INTER, SRDclss.h:238) */
    double P[4];    /* exponents of the surface (bc.P) */
    double B[3];    /* wavy-wall amplitude and wave numbers (bc.B) */
    int ABS;        /* absolute values in the surface terms (bc.ABS) */
    int pad_;
} mpcd_gpu_wall;

/* one replica's parameters -- subset of inputList (SRDclss.h:395-424) */
typedef struct {
    int dim;               /* 2 or 3 */
    int L[3];              /* box, cell size 1 (XYZ); L[2] ignored in 2D */
    double dt, kbt;
    double rotang;         /* SRD rotation angle RA; the library uses cos/sin computed on the host */
    double tau;            /* Berendsen relaxation time (TAU) */
    double grav[3];        /* body force (GRAV) */
    double avv[3];         /* reference velocity of TEMP (AVV, therm.c:209) */
    double gndnst;         /* global number density GnDNST (init.c:1002) */
    int rtech, lc, tstech, galinv;
    int true_random_sign;  /* 0: the reference's always-negative SRD angle (mpc.c:1291) */
    int precision;         /* MPCD_FP64, MPCD_FP32 or MPCD_FP32_EXACT particle state */
    int nspec, nwall;
    mpcd_gpu_species sp[MPCD_MAXSPEC];
    mpcd_gpu_wall wall[MPCD_MAXWALL];
} mpcd_gpu_params;

typedef struct {
    const double *shift;   /* [3 * nrep] */
    const double *cell;    /* [4 * ncell_total] */
    const double *ghost;   /* [3 * ncell_total] */
    const double *part;    /* [3 * n_total] */
    const double *lc;      /* [lc_stride * n_total] */
    int lc_stride;
    const double *escape;  /* [2 * n_total * MPCD_ESC_STRIDE] or NULL */
    const double *retherm; /* [n_total * MPCD_RETHERM_STRIDE] or NULL */
} mpcd_draws;

typedef struct mpcd_gpu_ctx mpcd_gpu_ctx;

/* --- lifetime ------------------------------------------------------------ */
/* A context holding `nrep` replicas; params[r] and n_particles[r] describe
 * replica r.  dim, rtech, lc, precision and the species table must agree
 * across replicas.  Replaces the reference's readJson()+initializeSIM()
 * storage (read.c:1699-1719) for the solvent. */
mpcd_status mpcd_gpu_create(const mpcd_gpu_params *params, int nrep, const int64_t *n_particles, uint64_t seed,
                            mpcd_gpu_ctx **out);
void        mpcd_gpu_destroy(mpcd_gpu_ctx *ctx);
const char *mpcd_gpu_last_error(const mpcd_gpu_ctx *ctx);

/* --- state transfer (particleMPC array, SRDclss.h:34-44) ------------------ */
/* Q, V, U: [3*n] in particle-index order; U and spid may be NULL.  Performs
 * the initial binning and cell properties (binin + localPROP,
 * init.c:2114-2118). */
mpcd_status mpcd_gpu_upload(mpcd_gpu_ctx *ctx, int replica, const double *Q, const double *V, const double *U,
                            const int32_t *spid);
mpcd_status mpcd_gpu_download(mpcd_gpu_ctx *ctx, int replica, double *Q, double *V, double *U);
/* The same from / into an array of structures such as the reference's
 * particleMPC (SRDclss.h:34-44, 128 B records): particle i's field is at
 * (char *)base + i * stride + off.  off_q, off_v: double[3]; off_u: double[3]
 * or -1 (none); off_spid: int32 or -1 (none).  A download writes Q, V and
 * (off_u >= 0) U only, leaving every other byte of the records as it was.
 * From C: mpcd_gpu_upload_aos(ctx, 0, P, sizeof(particleMPC),
 * offsetof(particleMPC, Q), offsetof(particleMPC, V), offsetof(particleMPC, U),
 * offsetof(particleMPC, SPID)). */
mpcd_status mpcd_gpu_upload_aos(mpcd_gpu_ctx *ctx, int replica, const void *base, int64_t stride, int64_t off_q,
                                int64_t off_v, int64_t off_u, int64_t off_spid);
mpcd_status mpcd_gpu_download_aos(mpcd_gpu_ctx *ctx, int replica, void *base, int64_t stride, int64_t off_q,
                                  int64_t off_v, int64_t off_u);

/* --- the step, phase by phase (see the table at the top) ------------------ */
/* shifted = 1: draw (or take) the grid shift and bin the shifted positions.
 * shifted = 0: materialise the unshifted binning left by mpcd_gpu_stream.
 * Optional outputs for replica 0..nrep-1 concatenated: cell_out [n_total]
 * cell of every particle (particle-index order per replica), count_out
 * [ncell_total] populations, perm_out [n_total] particle indices in list
 * order. */
mpcd_status mpcd_gpu_bin(mpcd_gpu_ctx *ctx, const mpcd_draws *draws, int shifted, int32_t *cell_out,
                         int32_t *count_out, int32_t *perm_out);
/* kbt_now_out: [nrep] temperature after the thermostat (KBTNOW), may be NULL */
mpcd_status mpcd_gpu_collide(mpcd_gpu_ctx *ctx, const mpcd_draws *draws, double *kbt_now_out);
/* bc_event_counts_out: [4] escapes, particles that exceeded NBOUNCE, rewinds,
 * re-thermalisations (bc.c:1364-1368), may be NULL */
mpcd_status mpcd_gpu_stream(mpcd_gpu_ctx *ctx, int64_t *bc_event_counts_out);

/* For callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc); no counterpart in
 * the reference.  The binning of mpcd_gpu_bin without a context: n particles
 * of one replica (params: dim, L and the walls that make an axis periodic)
 * whose positions q_dev [3n] (FP64, particle-index order) lie in device
 * memory.  shift == NULL: binin (mpc.c:924-934), every list in particle-index
 * order.  shift [3] (host): the positions shifted by it (gridShift_all,
 * mpc.c:728-736) and binned with bin(shifted=1)'s periodic wrap
 * (mpc.c:964-968); given a previous binning (prev_perm_dev [n] its list order,
 * prev_cell_dev [n] each particle's cell), its lists are relinked as bin()
 * does (mpc.c:948-992): in every cell the particles that stayed keep their
 * order, then the arrivals in the previous order; an escaped particle stays
 * in its list (its re-placement is the caller's).  Without a previous binning
 * an escaped particle gets cell -1 and is listed after every cell.  Outputs
 * (device): cell_dev [n] cell of every particle (x-major over
 * (L0+1)(L1+1)(L2+1), z extent 1 in 2D), count_dev [ncell] populations (may
 * be NULL), perm_dev [n] particle indices in list order.  It allocates
 * nothing: its working memory is a static device array of about 42 MB that
 * each GPU keeps until the process ends, shared by every call on that GPU,
 * so calls that overlap on one GPU race; it only launches kernels, on
 * `stream` (a hipStream_t, NULL for the null stream).  n above
 * MPCD_NOALLOC_MAX_PARTICLES or more cells than MPCD_NOALLOC_MAX_CELLS fail
 * with MPCD_EINVAL; errors read through mpcd_gpu_last_error(NULL). */
#define MPCD_NOALLOC_MAX_PARTICLES 2097152
#define MPCD_NOALLOC_MAX_CELLS 1000000
mpcd_status mpcd_gpu_bin_noalloc(const mpcd_gpu_params *params, int64_t n, const double *q_dev, const double *shift,
                                 const int32_t *prev_perm_dev, const int32_t *prev_cell_dev, int32_t *cell_dev,
                                 int32_t *count_dev, int32_t *perm_dev, void *stream);

/* --- whole steps ------------------------------------------------------------ */
/* one full step (bin + collide + stream); draws may be NULL (internal RNG) */
mpcd_status mpcd_gpu_step(mpcd_gpu_ctx *ctx, const mpcd_draws *draws, double *kbt_now_out);
/* nsteps steps with the internal RNG; with use_graph != 0 (fused path) pairs
 * of steps are recorded once in a hipGraph and replayed, the step number
 * being read from device memory.  For the isotropic operators without a
 * thermostat the run uses the bucketed layout (mpcd_gpu_set_layout): the
 * state is moved into per-tile buckets at the start and back into the
 * cell-sorted order at the end (one host synchronisation per call); results
 * are bit-identical to the cell-sorted layout.  A bucket overflow makes the
 * run finish on the cell-sorted path from the last complete step (the
 * boundary-event counters then include the discarded steps' events). */
mpcd_status mpcd_gpu_run(mpcd_gpu_ctx *ctx, int nsteps, int use_graph);
/* Layout of mpcd_gpu_run (no counterpart in the reference): bucketed != 0
 * uses per-tile buckets (the default), 0 the cell-sorted arrays; fill > 0
 * sets the mean bucket population as a fraction of the bucket capacity
 * (default 0.6); capacity > 0 lowers the usable slots per bucket below the
 * kernel's maximum (0 restores it).  A fill near 1 or a low capacity makes
 * overflows, and the fallback, frequent. */
mpcd_status mpcd_gpu_set_layout(mpcd_gpu_ctx *ctx, int bucketed, double fill, int capacity);
/* out[4]: bucketed steps completed, bucket overflows (runs finished on the
 * cell-sorted path), usable slots per bucket, buckets */
mpcd_status mpcd_gpu_layout_stats(const mpcd_gpu_ctx *ctx, int64_t *out);
mpcd_status mpcd_gpu_synchronize(mpcd_gpu_ctx *ctx);

/* The draw counter (no counterpart in the reference, whose generator state a
 * checkpoint does not keep, so its resumes do not repeat a run): every draw
 * is keyed by (seed, replica, step, ...), so a state saved with the step
 * mpcd_gpu_get_step reports and uploaded into a context of the same seed
 * followed by mpcd_gpu_set_step continues the same run.  In the exact-sum
 * form (any particle order) the continuation is bit-identical; the list-order
 * forms bin the uploaded order afresh.  set_step discards a pending binning
 * and is refused while a tape or a tangent runs. */
mpcd_status mpcd_gpu_get_step(const mpcd_gpu_ctx *ctx, uint32_t *step);
mpcd_status mpcd_gpu_set_step(mpcd_gpu_ctx *ctx, uint32_t step);

/* Batched-form tuning (no counterpart in the reference, which has no tiles):
 * runs steps_per_trial real steps (plus one untimed step) for each candidate
 * fill of the layout mpcd_gpu_run uses (six bucket fills, 0.4 to 0.8, of the
 * bucketed layout, else five tile fills of the fused kernel), keeps the fastest and returns its ms
 * per step (0 when the context uses the split path).  The trial steps advance
 * the simulation like mpcd_gpu_run.  Invalidates a graph recorded by
 * mpcd_gpu_run. */
mpcd_status mpcd_gpu_autotune(mpcd_gpu_ctx *ctx, int steps_per_trial, double *ms_per_step_out);
/* the fill in use by mpcd_gpu_run: mean bucket population / bucket capacity
 * (bucketed layout) or mean tile population / LDS capacity (fused kernel) */
double      mpcd_gpu_tile_fill(const mpcd_gpu_ctx *ctx);

/* --- checkpoints and the adjoint ------------------------------------------ */
/* Save / restore everything that carries over between steps (particle state in
 * list order, pending binning, wall positions, step number) in device memory
 * slot `slot`; a restored run repeats the saved one bit for bit.  No
 * counterpart in the reference (its checkpoints are files, pout.c:2325). */
mpcd_status mpcd_gpu_checkpoint(mpcd_gpu_ctx *ctx, int slot);
mpcd_status mpcd_gpu_restore(mpcd_gpu_ctx *ctx, int slot);
/* Pathwise adjoint of the step with cell assignments, draws, wall events and
 * every other discrete decision held fixed (FP64 state, tsTech 0, lc 0 or 1).
 * mpcd_gpu_record(ctx, k) starts a tape of the next k steps taken with the
 * library's generator (the steps run on the unfused path while recording;
 * k = 0 stops recording).  Two backward entry points pop the last taped step:
 *
 * mpcd_gpu_step_backward_state: every operator of the unfused path (SRD,
 * MPC-AT, MPC-AT+a, the nematic collision with Jeffery rotation and dipole
 * activity, ghost particles and anchoring at phantom walls, planar walls).
 * adj_in [9 * n_total] (particle-index order, replicas concatenated; per
 * particle dJ/dQ[3], dJ/dV[3], dJ/dU[3]; may be NULL) is the observable's
 * gradient at the end of that step and is added to the carried adjoint;
 * adj_out [9 * n_total] receives the adjoint at the start of the step (may be
 * NULL); dJ_dtheta [MPCD_NTHETA * nrep] receives per replica the running sums
 * since mpcd_gpu_adjoint_reset of dJ/dGRAV[0..2], 0 (the rotation angle is
 * not a variable here), and dJ/dACT of each species (may be NULL).  The step
 * is replayed from its recorded start state, one kernel at a time, and each
 * kernel's reverse-mode instance (hand-written reverse arithmetic, mpcd_rev.h)
 * turns the adjoints of its outputs into those of its inputs; the context's
 * own state is left as it was.
 *
 * mpcd_gpu_step_backward: SRD and MPC-AT between resting walls, velocities
 * only: adj_in / adj_out [3 * n_total] are dJ/dV; dJ_dtheta [4 * nrep]
 * receives dJ/dGRAV[0..2] and dJ/dRA per replica.
 * No counterpart in the reference. */
#define MPCD_NTHETA (4 + MPCD_MAXSPEC)
mpcd_status mpcd_gpu_record(mpcd_gpu_ctx *ctx, int max_steps);
mpcd_status mpcd_gpu_adjoint_reset(mpcd_gpu_ctx *ctx);
mpcd_status mpcd_gpu_step_backward(mpcd_gpu_ctx *ctx, const double *adj_in, double *adj_out, double *dJ_dtheta);
mpcd_status mpcd_gpu_step_backward_state(mpcd_gpu_ctx *ctx, const double *adj_in, double *adj_out,
                                         double *dJ_dtheta);
/* Observables of the current state for the time-averaged objectives of the
 * adjoint, per replica (value [nrep]) with their gradient (grad [9 * n_total]
 * in the layout of adj_in above, FP64 state; may be NULL):
 *   MPCD_OBS_BAND_VELOCITY, arg = {component, axis, lo, hi}: the mean of
 *     V[component] over the particles whose cell index (int)Q[axis] lies in
 *     [lo, hi) -- a channel's centreline velocity;
 *   MPCD_OBS_MEAN_S (lc 1): the mean scalar order parameter S of the cells
 *     holding more than one particle (localPROP, mpc.c:183-207) in the
 *     unshifted binning that ends the step.
 * The step's pending re-bin is done first, as mpcd_gpu_get_cells does.  No
 * counterpart in the reference. */
#define MPCD_OBS_BAND_VELOCITY 0
#define MPCD_OBS_MEAN_S 1
mpcd_status mpcd_gpu_observable(mpcd_gpu_ctx *ctx, int kind, const double *arg, double *value, double *grad);
/* Forward-mode tangent of the whole step (every operator of the unfused path:
 * SRD, MPC-AT, MPC-AT+a, the nematic LC collision with Jeffery rotation and
 * dipole activity, ghost particles and anchoring at phantom walls, planar
 * walls; FP64 state, tsTech 0, lc 0 or 1) along one direction in parameter
 * space: dgrav [3] = d(GRAV)/d(theta) (may be NULL: zero) and dact [nspec] =
 * d(ACT of species k)/d(theta) (may be NULL: zero).  Cell assignments, draws,
 * wall events and every other discrete decision (Metropolis acceptances,
 * activity side, clamps) are those of the run itself, so the tangent is the
 * pathwise derivative; the values the steps compute are the FP64 values bit
 * for bit.  on != 0 starts it from zero for the current state (a state
 * uploaded or restored later restarts it from zero), on = 0 stops it.  While
 * it is on, steps run on the unfused path without graphs.
 * mpcd_gpu_tangent_download returns d(Q, V, U)/d(theta) of replica `replica`
 * in particle-index order (dU may be NULL).  No counterpart in the reference. */
mpcd_status mpcd_gpu_tangent(mpcd_gpu_ctx *ctx, int on, const double *dgrav, const double *dact);
mpcd_status mpcd_gpu_tangent_download(mpcd_gpu_ctx *ctx, int replica, double *dQ, double *dV, double *dU);

/* --- rest frame ------------------------------------------------------------- */
/* The periodic Galilean transformation to the rest frame that the output
 * step applies when runtime % zeroNetMom == 0 (outputResults pout.c:2557-2562:
 * galileantrans mtools.c:673-718, then zeroExtraDims mtools.c:727-744).  For
 * every replica the mass-weighted mean velocity of its particles less the
 * replica's avv (in the DIM first components) is subtracted from every
 * particle's velocity; in 2D the third components of Q and V are set to 0.
 * Mobile walls and MD atoms are outside this library.  The caller applies
 * the reference's conditions (rFrame on, zeroNetMom > 0, no gravity, more
 * than 2 particles: init.c:2141-2156) and calls it after the step whose
 * runtime is a multiple of zeroNetMom.  The cell values of the step
 * (mpcd_gpu_get_cells) are those computed before it, as in outputResults.
 * MPCD_EUNSUPPORTED while an adjoint tape or a tangent is recording. */
mpcd_status mpcd_gpu_zero_net_momentum(mpcd_gpu_ctx *ctx);

/* --- observables ------------------------------------------------------------ */
/* kbt_now [nrep]: KBTNOW of the last completed step of every replica, the
 * temperature TEMP computes after the collision and after the thermostat's
 * rescaling, before streaming (mpc.c:4752-4760) -- what the reference's
 * energy output and checkpoint write (pout.c:2610, 2325).  0 before the first
 * step. */
mpcd_status mpcd_gpu_kbtnow(mpcd_gpu_ctx *ctx, double *kbt_now);
/* per-cell populations, centre-of-mass velocities [3 per cell], scalar order
 * parameter and director [3 per cell] of replica `replica` (localPROP
 * results, mpc.c:64-216); any pointer may be NULL */
mpcd_status mpcd_gpu_get_cells(mpcd_gpu_ctx *ctx, int replica, int32_t *pop, double *vcm, double *S,
                               double *dir);
/* Parity diagnostics of the nematic path (no counterpart in the reference):
 * with on != 0 every later step records its discrete decisions, so that a
 * comparison with the reference can tell a flipped decision from rounding.
 * mpcd_gpu_lc_decisions returns those of the last step's collision phase for
 * replica `replica`: part [n] one word per particle (particle-index order):
 * bit 0 its cell held more than one particle, bits 1-2 the Maier-Saupe
 * sampler (0 2D Metropolis lc.c:341, 1 3D Gaussian lc.c:85, 2 3D cube root
 * lc.c:192, 3 3D Metropolis lc.c:257), bit 3 the 2D angle redrawn at
 * +-pi/2 (lc.c:352), bit 4 the sign draw gave -1, bits 5-12 annealNum,
 * bits 13-20 accepted moves, bit 21 the dipole force points along +DIR
 * (lc.c:1859-1863); cell [ncell] one word per cell: bit 5 S and the
 * director were computed (more than one particle, mpc.c:184), bits 0-1 the
 * branch of eigenvectors2x2 (mtools.c:1247-1305: 0 m10 != 0, 1 m01 != 0,
 * 2 equal diagonal, 3 other) or the column of eigenvectors3x3's product
 * (mtools.c:1373-1430; 3 none), bit 2 the diagonal branch and bit 3 the
 * clamp of r in eigenvalues3x3 (mtools.c:1307-1371), bit 4 S clamped to 1
 * (mpc.c:207).  Either pointer may be NULL. */
mpcd_status mpcd_gpu_lc_record(mpcd_gpu_ctx *ctx, int on);
mpcd_status mpcd_gpu_lc_decisions(mpcd_gpu_ctx *ctx, int replica, uint32_t *part, uint32_t *cell);
/* list order of replica `replica` after the last phase: order [n] particle
 * indices, cstart [ncell+1] cell offsets */
mpcd_status mpcd_gpu_get_order(mpcd_gpu_ctx *ctx, int replica, int32_t *order, int32_t *cstart);
/* Running flow field of the reference's flow output.  While on, every
 * completed step adds to each cell of the unshifted binning its
 * centre-of-mass velocity, sum m v / sum m (0 in an empty cell), as
 * localFLOW + sumFLOW do at the end of timestep() when flowOut is set
 * (mpc.c:4946-4951, 226-246, 3889-3910).  Replaces those two calls; the
 * caller writes the output, as flowout does (pout.c:1681-1700).  The sums
 * persist while off.  Cost: one pass over the particles and one over the
 * cells per step (FP64 atomics per run of equal cells in the cell-sorted
 * layout, LDS partials per tile in the bucketed one); the atomics add in no
 * fixed order, so a sum equals the reference's list-order sum to round-off.
 * Not rewound by
 * mpcd_gpu_restore. */
mpcd_status mpcd_gpu_flow(mpcd_gpu_ctx *ctx, int on);
/* the sums of replica `replica` (flow: ncell x 3 doubles, cells in the order
 * of mpcd_gpu_get_cells, components >= dim zero) and the number of steps
 * summed; reset != 0 then clears both, as flowout clears FLOW after writing
 * it (pout.c:1696).  Either pointer may be NULL.  Before the first
 * mpcd_gpu_flow(ctx, 1) the sums and the count read as zero. */
mpcd_status mpcd_gpu_flow_read(mpcd_gpu_ctx *ctx, int replica, double *flow, int64_t *nsteps, int reset);
/* number of cells of replica `replica` and the keyed LC stride */
int         mpcd_gpu_ncells(const mpcd_gpu_ctx *ctx, int replica);
int         mpcd_gpu_lc_stride(const mpcd_gpu_ctx *ctx);
/* kernel timing with HIP events recorded on the library's stream around
 * every kernel family while enabled (enable != 0 also clears the records) */
mpcd_status mpcd_gpu_profile(mpcd_gpu_ctx *ctx, int enable);
/* out[f] = total device ms of family f, out[14 + f] = number of timed
 * launches, for the families 0 count (histogram + keys), 1 scan, 2 scatter,
 * 3 in-cell sort, 4 gather, 5 cell properties, 6 ghost particles, 7 LC,
 * 8 collision, 9 temperature/thermostat, 10 stream + walls, 11 fused
 * gather + collision + stream + next binning, 12 bucketed step, 13 flow
 * field; n >= 28 */
mpcd_status mpcd_gpu_timing(mpcd_gpu_ctx *ctx, double *out, int n);

#ifdef __cplusplus
}
#endif

#endif /* MPCD_GPU_H */
