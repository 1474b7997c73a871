/*
 * mpcd_oracle.h -- CPU oracle for the MPCD solvent step (TEST INFRASTRUCTURE).
 *
 * A single-threaded C restatement of the reference's per-timestep solvent
 * path (mpc.c timestep() 4496-4954 restricted to: grid shift, binning,
 * acceleration, cell properties, ghost particles, nematic LC collision and
 * Jeffery torque, the SRD / MPC-AT / MPC-AT+a / LC (+activity) collision
 * switch, temperature and thermostat, streaming and planar walls).  It is a
 * checker for the GPU library and is never linked into it.
 *
 * Random draws go through a draw source with three modes:
 *   ORC_DRAW_SEQ    -- the sequential streams of seqdraw.c in the reference's
 *                      call order (equal, draw for draw, to the reference run
 *                      by oracle/_ref/ref_harness); can record every consumed
 *                      variate into the keyed layout below;
 *   ORC_DRAW_KEYED  -- variates read from keyed arrays (one step's worth);
 *   ORC_DRAW_PHILOX -- variates computed by Philox-4x32-10 from
 *                      (seed, replica, step, purpose, key, k).
 * The keyed layout (purposes, keys and strides) is the one documented in
 * include/mpcd_gpu.h for the GPU library's injected-draw mode.
 */
#ifndef MPCD_ORACLE_H
#define MPCD_ORACLE_H

#include <stdint.h>
#include "seqdraw.h"

#define ORC_MAXSPEC 10
#define ORC_MAXWALL 16

enum { ORC_DRAW_SEQ = 0, ORC_DRAW_KEYED = 1, ORC_DRAW_PHILOX = 2 };

/* draw purposes (shared numbering with include/mpcd_gpu.h) */
enum {
    ORC_P_SHIFT = 1,   /* key 0, k = 0..2 (x, y, z) uniforms            */
    ORC_P_CELL = 2,    /* key = cell, stride 4: SRD axis / uniform, sign  */
    ORC_P_GHOST = 3,   /* key = cell, stride 3: ghost-particle normals    */
    ORC_P_NORMAL = 4,  /* key = slot, stride 3: thermal normals (AT)      */
    ORC_P_LC = 5,      /* key = slot, stride lc_stride: Maier-Saupe draws */
    ORC_P_ESCAPE = 6,  /* key = particle, escape repair in the shifted bin */
    ORC_P_RETHERM = 7, /* key = particle, wall re-thermalisation (rare)    */
    ORC_P_ESCAPE2 = 8  /* key = particle, escape repair in the unshifted bin after streaming */
};

/* keyed layout of the rare draws (MPCD_ESC_STRIDE / MPCD_RETHERM_STRIDE of
 * include/mpcd_gpu.h): escape entries per particle and binning hold [0..2]
 * Maxwell normals, [3..4] orientation uniforms, [5 + DIM*t + j] position
 * uniform j of try t; re-thermalisation entries hold [3r + j], variate j of
 * the particle's r-th push of the step */
#define ORC_ESC_STRIDE 32
#define ORC_ESC_POS0 5
#define ORC_RETH_STRIDE 24

typedef struct {
    double mass, rfc, tumble, chihi, act, mfpot, damp, sigwidth, sigpos, minactratio;
    int vdist;
} orc_species;

/* An axis-aligned planar wall (subset of the reference's bc, SRDclss.h:128-262). */
typedef struct {
    double A[3], Q[3], R, DN, DT, DVN, DVT, DVxyz[3], MVN, MVT, MUN, MUT, MUxyz[3], DUxyz[3], V[3], L[3];
    int INV, PHANTOM, COLL_TYPE;
    int inter[ORC_MAXSPEC];
    /* surface shape (read.c:1264-1300, 1394-1436): exponents P[0..3] (all 0:
     * the plane, 1 1 1 1), wavy amplitudes/wavenumbers B, |.| of the terms */
    double P[4], B[3];
    int ABS;
} orc_wall;

typedef struct {
    int dim, L[3];
    double dt, kbt, rotang, tau, grav[3], avv[3], gndnst;
    int rtech, lc, tstech, galinv, true_random_sign, nspec, nwall;
    orc_species sp[ORC_MAXSPEC];
    orc_wall wall[ORC_MAXWALL];
} orc_params;

typedef struct {
    int mode;
    seqdraw_state seq;
    /* keyed arrays for the current step (ORC_DRAW_KEYED) */
    const double *shift, *cell, *ghost, *part, *lc;
    int lc_stride;
    /* Philox key (ORC_DRAW_PHILOX) */
    uint64_t seed;
    uint32_t replica;
    /* optional recording of consumed sequential variates into keyed arrays */
    double *rec_shift, *rec_cell, *rec_ghost, *rec_part, *rec_lc;
    int rec_lc_stride;
    /* rare draws (escape repair [2 * n * ORC_ESC_STRIDE], re-thermalisation
     * [n * ORC_RETH_STRIDE]): read in keyed mode, written in sequential mode;
     * NULL in keyed mode means the Philox generator serves them */
    double *escape, *retherm;
    int rare_n;
    /* set when a particle needed more escape tries or pushes than a keyed
     * entry holds (the step is then not representable in the keyed layout) */
    int rare_overflow;
} orc_draws;

typedef struct orc_sys orc_sys;

orc_sys *orc_create(const orc_params *p, int n);
void     orc_destroy(orc_sys *s);
/* positions/velocities/orientations [n*3], species ids [n]; performs the
 * initial binning (binin, mpc.c:924) and cell properties (localPROP). */
int      orc_set_state(orc_sys *s, const double *Q, const double *V, const double *U, const int *spid);
void     orc_get_state(const orc_sys *s, double *Q, double *V, double *U);
int      orc_ncells(const orc_sys *s);
void     orc_get_order(const orc_sys *s, int *order, int *cstart);
/* per cell: pop, mass, vcm[3], cm[3], S, dir[3], E[9] */
void     orc_get_cells(const orc_sys *s, int *pop, double *mass, double *vcm, double *cm, double *S,
                       double *dir, double *E);
void     orc_get_walls(const orc_sys *s, double *wq);
double   orc_kbtnow(const orc_sys *s);
double   orc_avs(const orc_sys *s);
void     orc_counters(const orc_sys *s, long *out4);
int      orc_lc_stride(const orc_sys *s);

/* one full time step (mpc.c:4496-4954 subset); returns 0 or an error code */
int      orc_step(orc_sys *s, orc_draws *d);
/* the three phases of a step, matching the GPU library's entry points:
 *   bin    : shift draw + gridShift_all(+) + bin(shifted=1)
 *   collide: acc_all, localPROP, ghostPart, LC, MPCcollision, TEMP/thermostat, gridShift_all(-)
 *   stream : stream_all, MPC_BCcollision, bin(shifted=0), localPROP       */
int      orc_phase_bin(orc_sys *s, orc_draws *d);
int      orc_phase_collide(orc_sys *s, orc_draws *d);
int      orc_phase_stream(orc_sys *s, orc_draws *d);
void     orc_last_shift(const orc_sys *s, double *shift);
/* the rest-frame transformation of an output step (galileantrans +
 * zeroExtraDims, mtools.c:673-744) */
void     orc_zero_net_momentum(orc_sys *s);
/* flow output: with on, every step ends with localFLOW + sumFLOW
 * (mpc.c:226-246, 4946-4951); get returns FLOW [ncell*3] and the steps summed */
void     orc_flow(orc_sys *s, int on);
void     orc_get_flow(orc_sys *s, double *flow, long *nsteps, int reset);

/* draw-source helpers for Python (ctypes) callers */
orc_draws *orc_draws_new(int mode, uint64_t seed, uint32_t replica);
void       orc_draws_free(orc_draws *d);
void       orc_draws_set_keyed(orc_draws *d, const double *shift, const double *cell, const double *ghost,
                               const double *part, const double *lc, int lc_stride);
void       orc_draws_set_record(orc_draws *d, double *shift, double *cell, double *ghost, double *part,
                                double *lc, int lc_stride);
void       orc_draws_counts(const orc_draws *d, uint64_t *out3);
/* the nematic path's discrete decisions of the last collision phase: one
 * word per particle (particle index) and per cell, in the layout of the GPU
 * library's mpcd_gpu_lc_decisions */
void       orc_lc_decisions(const orc_sys *s, uint32_t *part, uint32_t *cell);
/* escape / re-thermalisation arrays (see orc_draws); n = particles */
void       orc_draws_set_rare(orc_draws *d, double *escape, double *retherm, int n);
int        orc_draws_rare_overflow(const orc_draws *d);

/* exposed pieces for closed-form tests */
void     orc_eigen(double *m, int dim, double *eigval); /* m row-major dim x dim -> eigvecs as rows */
void     orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double   orc_philox_uniform(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k);
double   orc_philox_normal(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k);
void     orc_inv3(double inv[3][3], const double m[3][3]);
void     orc_philox_unitvec(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k,
                            double v[3]);
void     orc_philox_unitvec_turn(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key,
                                 uint32_t k, double v[3]);
void     orc_sincos_turn(double u, double *sn, double *cs);
double   orc_log_unit(double u);
double   orc_philox_normal_turn(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k);
void     orc_rodrigues(double *vec, double *axis, double theta);

#endif
