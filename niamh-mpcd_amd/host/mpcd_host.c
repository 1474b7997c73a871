/*
 * mpcd_host.c -- a C host for the MI355X solvent step: reads the reference's
 * input.json, sets up the particles and runs the reference's warm-up and
 * temporal loops (mpcd/mpcd.c:255-300) through the C-ABI of mpcd_gpu.h.
 *
 *   mpcd_host input.json [-o DIR] [--fp32 | --fp32x] [--replicas N] [--every K] [--steps N] [--seed S] [--no-graph]
 *             [--dump-params] [--write-checkpoint FILE] [--binary-checkpoint]
 *
 * JSON keys and defaults follow readJson (mpcd/subroutines/read.c:973-1804)
 * for the options of this path: domain, kbt, dt, simSteps, warmUp, galInv,
 * tsTech, rTech/collOp, lc, tau, rotAng, grav, seed, domainWalls, species
 * (mass, pop/dens, vDist, rfc, tumble, shearSusc, act, mfpot, sigWidth,
 * sigPos, minActRatio, damp) and BC entries, planar, curved or wavy (aInv,
 * Q, P, wavy, abs, rotSym 4 4, R, DN, DT, DVN, DVT, DVxyz, MVN, MVT, MUN,
 * MUT, MUxyz, DUxyz, V, L, inv, phantom, colType, interSRD, homeotropic,
 * planar).  Options outside the path (MD,
 * swimmers, mobile walls) stop with a message, as the reference
 * stops on an unknown option.
 *
 * rFrame (default 1) removes the net momentum of the initial state
 * (galileantrans, init.c:2152); with zeroNetMom = Z > 0, no body force and
 * more than 2 particles the library does it again after the step of every
 * runtime divisible by Z, before that runtime's outputs (outputResults,
 * pout.c:2557-2562; mpcd_gpu_zero_net_momentum).
 *
 * Checkpoints in the reference's file layout (checkpoint.h): checkpointIn
 * recovers one as readchckpnt does (its inputs, particles, loop counters and
 * output flags; no warm-up; the temporal loop from its runtime), and
 * checkpointOut = F writes DIR/checkpoint.dat after the outputs of every
 * runtime divisible by F (mpcd.c:310-312; replica 0).  --write-checkpoint
 * FILE writes the state the run would start from and exits without the GPU.
 * --binary-checkpoint writes DIR/checkpoint.bin (and --write-checkpoint's
 * FILE) in the lossless binary layout instead, with the library's draw
 * counter; "checkpointIn" reads either layout, and a binary one continues
 * the run's draws (bit for bit in the exact-sum form, --fp32x).
 *
 * Output: DIR/solvent.dat with one line per output step: the temporal
 * loop's step counter (runtime, 0..simSteps, mpcd.c:291), the
 * temperature TEMP (therm.c:209-235, about the initial mean velocity AVV),
 * the kinetic energy and the total momentum; a timing line on stdout.
 * Like the reference, each loop runs its step count plus one
 * (mpcd.c:264, 294).
 *
 * flowOut = F > 0 (read.c:1732): after the warm-up, every step adds each
 * cell's centre-of-mass velocity to a running sum on the GPU (localFLOW +
 * sumFLOW, mpc.c:4946-4951), and at every runtime with runtime % F == 0
 * (pout.c:2690) DIR/flowfield.dat receives the sum divided by F for every
 * cell of the domain, in flowout's line format (pout.c:1681-1700), after
 * which the sum starts again.
 * velOut = F > 0: DIR/velfield.dat, every cell's centre-of-mass velocity at
 * every runtime divisible by F (velout, pout.c:1715-1731, 2691).
 * densOut = F > 0: DIR/densityfield.dat, every cell's population, mass and
 * population per species (densityout, pout.c:1744-1757, 2692), counted on
 * the host from the particles (the cell of a particle is (int)Q, as bin()
 * finds it, mpc.c:958-960).
 * energyOut = F > 0: DIR/energy.dat, one line per runtime divisible by F
 * (enout, pout.c:1389-1422, 2609-2612): time, MPCD kinetic energy, the
 * nematic mean-field energy of calcE_LC (therm.c:330-351; 0 without lc),
 * the boundaries' kinetic and rotational energies (0: no displaceable
 * walls here), the total and the temperature of the step's velocities.
 * dirSOut = F > 0 with lc: DIR/directorfield.dat, every cell's director and
 * scalar order parameter, zeros in a cell without particles (orderout,
 * pout.c:2021-2033, 2695).
 * Each field file opens with 13 header lines, as the reference's do (a
 * preamble, a blank line and the reference's column names: flowheader,
 * pout.c:355-357, for both velocity fields, orderheader, pout.c:98-100).
 * Replica r > 0 writes DIR/<name>_r<r>.dat.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "checkpoint.h"
#include "json.h"
#include "mpcd_gpu.h"

/* splitmix64: the host's initial-state generator */
static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t rng_next(void) {
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static double rng_u(void) { return (double)(rng_next() >> 11) * (1.0 / 9007199254740992.0); }
static double rng_n(void) {
    double u1 = 1.0 - rng_u(), u2 = rng_u();
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* integer powers by products, other exponents by pow (smrtPow, mtools.c:38-50) */
static double host_pow(double x, double y) {
    if (y - (int)y == 0 && y <= 10) {
        double r = 1;
        for (int i = 0; i < y; i++) r *= x;
        return r;
    }
    return pow(x, y);
}

/* the surface function of a wall at pos (calcW with its wavy term,
 * bc.c:55-150), W < 0 inside the wall; a curved wall acts through its
 * periodic image nearest pos along the box's axes (shiftBC, bc.c:389-401).
 * Used to place the initial particles and to estimate the accessible
 * volume, which are the host's own; the step itself evaluates walls on the
 * GPU. */
static double host_wall_w(const mpcd_gpu_wall *w, const double *pos, int dim, const int L[3]) {
    double q[3] = {w->Q[0], w->Q[1], w->Q[2]}, Wv = 0.0, W1 = 0.0, W2 = 0.0;
    /* shiftBC moves a surface whose first two exponents both differ from 1 */
    const int shifted = fabs(w->P[0] - 1) > 1e-8 && fabs(w->P[1] - 1) > 1e-8;
    for (int d = 0; d < dim && shifted; d++) {
        if (pos[d] - q[d] >= 0.5 * L[d]) q[d] += L[d];
        else if (pos[d] - q[d] <= -0.5 * L[d]) q[d] -= L[d];
    }
    for (int d = 0; d < dim; d++) {
        double t = w->A[d] * (pos[d] - q[d]);
        Wv += host_pow(w->ABS ? fabs(t) : t, w->P[d]);
    }
    Wv -= host_pow(w->ABS ? fabs(w->R) : w->R, w->P[3]);
    if (fabs(w->B[0]) > 1e-8) {
        int curved = 0, ell = 1, f = -1;
        for (int d = 0; d < dim; d++) {
            curved += fabs(w->P[d] - 1) > 1e-8;
            if (fabs(w->P[d] - 2) > 1e-8 || fabs(w->A[d]) <= 1e-8) ell = 0;
            if (fabs(w->A[d]) <= 1e-8) f = d;
        }
        if (!curved) {
            W1 = (w->A[1] * pos[0] - w->A[0] * pos[1]) / sqrt(w->A[0] * w->A[0] + w->A[1] * w->A[1]);
        } else if (ell) {
            W1 = atan2(w->A[1] * (pos[1] - q[1]), w->A[0] * (pos[0] - q[0]));
            if (dim > 2)
                W2 = atan2(sqrt(pow(w->A[0] * (pos[0] - q[0]), 2) + pow(w->A[1] * (pos[1] - q[1]), 2)),
                           w->A[2] * (pos[2] - q[2]));
        } else if (dim > 2 && f >= 0) {
            const int k = (f + 1) % 3, l = (f + 2) % 3;
            W1 = atan2(w->A[l] * (pos[l] - q[l]), w->A[k] * (pos[k] - q[k]));
        }
        Wv += w->B[0] * cos(w->B[1] * W1) * cos(w->B[2] * W2);
    }
    return w->INV ? -Wv : Wv;
}

/* a position no wall excludes (replacePos_WithCheck, init.c:1462: W <= TOL
 * inside) */
static int host_position_free(const mpcd_gpu_params *p, const double *pos) {
    for (int k = 0; k < p->nwall; k++)
        if (host_wall_w(&p->wall[k], pos, p->dim, p->L) <= 1e-7) return 0;
    return 1;
}

/* accessible volume of species s (accessibleVolume, init.c:954-990): Monte
 * Carlo over the box with the walls the species interacts with, NUMMC = 1000
 * samples per unit volume as the reference takes, at most 2^22 */
static double host_accessible_volume(const mpcd_gpu_params *p, int s) {
    const double cv = (double)p->L[0] * p->L[1] * (p->dim == 3 ? p->L[2] : 1);
    double nd = 1000.0 * cv;
    if (nd > 4194304.0) nd = 4194304.0;
    const long n = (long)nd;
    long fails = 0;
    for (long i = 0; i < n; i++) {
        double pos[3] = {0, 0, 0};
        for (int d = 0; d < p->dim; d++) pos[d] = rng_u() * p->L[d];
        for (int k = 0; k < p->nwall; k++)
            if ((p->wall[k].inter_mask >> s & 1) && host_wall_w(&p->wall[k], pos, p->dim, p->L) < 0.0) {
                fails++;
                break;
            }
    }
    return cv * (1.0 - (double)fails / (double)n);
}

static void die(const char *fmt, const char *arg) {
    fprintf(stderr, "mpcd_host: ");
    fprintf(stderr, fmt, arg);
    fprintf(stderr, "\n");
    exit(1);
}

static void check(mpcd_gpu_ctx *ctx, mpcd_status st, const char *what) {
    if (st == MPCD_OK) return;
    fprintf(stderr, "mpcd_host: %s failed (status %d): %s\n", what, (int)st, ctx ? mpcd_gpu_last_error(ctx) : "");
    exit(1);
}

/* keys this host rejects: they select parts of the reference outside the solvent path */
/* flowheader's column names (pout.c:355-357), for the flow and velocity fields */
#define FLOW_COLUMNS "   t\t   QX\t   QY\t   QZ\tVcmX\t\tVcmY\t\tVcmZ"

/* flowout's lines (pout.c:1681-1700): time, cell indices and the mean
 * centre-of-mass velocity over the interval of every cell of the domain;
 * flow holds the compact cell array (x-major, (L+1) cells per axis, one z
 * layer in 2D) */
static void flow_write(FILE *f, const double *flow, const int L[3], int dim, double t, int interval) {
    const int ny = L[1] + 1, nz = dim == 3 ? L[2] + 1 : 1, kz = dim == 3 ? L[2] : 1;
    for (int i = 0; i < L[0]; i++)
        for (int j = 0; j < L[1]; j++)
            for (int k = 0; k < kz; k++) {
                const double *F = flow + 3 * ((size_t)(i * ny + j) * nz + k);
                double av[3] = {0.0, 0.0, 0.0};
                for (int h = 0; h < dim; h++) av[h] = F[h] / (double)interval;
                fprintf(f, "%12.5e\t%5d\t%5d\t%5d\t%12.5e\t%12.5e\t%12.5e\n", t, i, j, k, av[0], av[1], av[2]);
            }
}

/* velout's lines (pout.c:1715-1731): the instantaneous centre-of-mass
 * velocity of every cell of the domain */
static void vel_write(FILE *f, const double *vcm, const int L[3], int dim, double t) {
    const int ny = L[1] + 1, nz = dim == 3 ? L[2] + 1 : 1, kz = dim == 3 ? L[2] : 1;
    for (int i = 0; i < L[0]; i++)
        for (int j = 0; j < L[1]; j++)
            for (int k = 0; k < kz; k++) {
                const double *v = vcm + 3 * ((size_t)(i * ny + j) * nz + k);
                double vel[3] = {0.0, 0.0, 0.0};
                for (int h = 0; h < dim; h++) vel[h] = v[h];
                fprintf(f, "%12.5e\t%5d\t%5d\t%5d\t%12.5e\t%12.5e\t%12.5e\n", t, i, j, k, vel[0], vel[1], vel[2]);
            }
}

/* orderout's lines (pout.c:2021-2033): director and S of every cell of the
 * domain, zeros where the cell holds no particle */
static void order_write(FILE *f, const int32_t *pop, const double *dir, const double *S, const int L[3], int dim,
                        double t) {
    const int ny = L[1] + 1, nz = dim == 3 ? L[2] + 1 : 1, kz = dim == 3 ? L[2] : 1;
    for (int i = 0; i < L[0]; i++)
        for (int j = 0; j < L[1]; j++)
            for (int k = 0; k < kz; k++) {
                const size_t c = (size_t)(i * ny + j) * nz + k;
                fprintf(f, "%.2f\t%5d\t%5d\t%5d\t", t, i, j, k);
                if (pop[c] == 0) fprintf(f, "%12.5e\t%12.5e\t%12.5e\t%12.5e\n", 0.0, 0.0, 0.0, 0.0);
                else fprintf(f, "%12.5e\t%12.5e\t%12.5e\t%12.5e\n", dir[3 * c], dir[3 * c + 1], dir[3 * c + 2], S[c]);
            }
}

/* densityout's lines (pout.c:1744-1757): population, mass, MPCD / MD /
 * swimmer populations (MD and swimmers: none here) and the population of
 * every species, per cell of the domain; counts and masses per compact cell */
static void density_write(FILE *f, const int32_t *cnt, const double *mass, int nspec, const int L[3], int dim,
                          double t) {
    const int ny = L[1] + 1, nz = dim == 3 ? L[2] + 1 : 1, kz = dim == 3 ? L[2] : 1;
    for (int i = 0; i < L[0]; i++)
        for (int j = 0; j < L[1]; j++)
            for (int k = 0; k < kz; k++) {
                const size_t c = (size_t)(i * ny + j) * nz + k;
                int pop = 0;
                for (int s = 0; s < nspec; s++) pop += cnt[c * nspec + s];
                fprintf(f, "%12.5e\t%5d\t%5d\t%5d\t", t, i, j, k);
                fprintf(f, "%5d\t%12.5e\t%5d\t%5d\t%5d\t", pop, mass[c], pop, 0, 0);
                for (int s = 0; s < nspec; s++) fprintf(f, "\t%5d", cnt[c * nspec + s]);
                fprintf(f, "\n");
            }
}

/* enout's line (pout.c:1389-1422): MPC_K = sum 1/2 m v^2 over the DIM
 * components; wmf = calcE_LC (therm.c:330-351): sum over cells with more than
 * one particle of S (u . n)^2 MFPOT over their particles; KBT = TEMP */
static void energy_write(FILE *f, const mpcd_gpu_params *p, size_t n, const double *Q, const double *V,
                         const double *U, const int32_t *spid, const int32_t *pop, const double *S, const double *dir,
                         double t, double kbt) {
    const int dim = p->dim, ny = p->L[1] + 1, nz = dim == 3 ? p->L[2] + 1 : 1;
    double mpc_k = 0.0, wmf = 0.0;
    for (size_t i = 0; i < n; i++) {
        const int s = p->nspec > 1 ? spid[i] : 0;
        const double m = p->sp[s].mass;
        double e = 0.0;
        for (int j = 0; j < dim; j++) e += V[3 * i + j] * V[3 * i + j];
        mpc_k += 0.5 * m * e;
        if (!p->lc) continue;
        const int a = (int)Q[3 * i], b = (int)Q[3 * i + 1], c = dim == 3 ? (int)Q[3 * i + 2] : 0;
        if (a < 0 || b < 0 || c < 0 || a >= p->L[0] || b >= p->L[1] || (dim == 3 && c >= p->L[2])) continue;
        const size_t k = (size_t)(a * ny + b) * nz + c;
        if (pop[k] <= 1) continue;
        double un = 0.0;
        for (int j = 0; j < dim; j++) un += U[3 * i + j] * dir[3 * k + j];
        wmf += S[k] * un * un * p->sp[s].mfpot;
    }
    fprintf(f, "%lf\t%lf\t%lf\t%lf\t%lf\t%lf\t%lf\n", t, mpc_k, wmf, 0.0, 0.0, mpc_k + wmf, kbt);
}

/* a field file of replica r: DIR/<name>.dat (r = 0) or DIR/<name>_r<r>.dat,
 * with 11 preamble lines, a blank line and the column names: readers of the
 * reference's files skip 13 lines */
static FILE *field_open(const char *outdir, const char *name, const char *what, const char *columns, int r,
                        int nrep, int interval, const int L[3], int dim, int fp32) {
    char path[4096];
    if (r == 0) snprintf(path, sizeof path, "%s/%s.dat", outdir, name);
    else snprintf(path, sizeof path, "%s/%s_r%d.dat", outdir, name, r);
    FILE *f = fopen(path, "w");
    if (!f) die("cannot write %s", path);
    fprintf(f, "# %s\n", what);
    fprintf(f, "# per cell of the domain, one block of cells per output, x-major\n");
    fprintf(f, "# written by mpcd_host through the C-ABI of mpcd_gpu.h\n");
    fprintf(f, "# replica %d of %d\n", r, nrep);
    fprintf(f, "# domain %d x %d x %d, %dD\n", L[0], L[1], dim == 3 ? L[2] : 1, dim);
    fprintf(f, "# output every %d steps\n", interval);
    fprintf(f, "# particle state %s\n", fp32 == 2 ? "FP32, exact FP64 cell sums" : fp32 ? "FP32, cell sums FP64" : "FP64");
    fprintf(f, "# columns: time, cell indices QX QY QZ, then the values\n");
    fprintf(f, "#\n#\n#\n\n");
    fprintf(f, "%s\n", columns);
    return f;
}

/* the runtimes divisible by F end after step runtime + 1: the number of
 * steps to run from `done` completed steps to the next of them */
static int steps_to_next(int done, int F) {
    const int r = ((done + F - 1) / F) * F;
    return r + 1 - done;
}

/* TEMP (therm.c:209-235) of a state: sum m |v - AVV|^2 / (DIM n) */
static double temp_of(const mpcd_gpu_params *p, size_t n, const double *V, const int32_t *spid) {
    double ke = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double m = p->sp[spid[i]].mass;
        for (int d = 0; d < 3; d++) ke += m * (V[3 * i + d] - p->avv[d]) * (V[3 * i + d] - p->avv[d]);
    }
    return n ? ke / (p->dim * (double)n) : 0.0;
}

/* a checkpoint of one replica's state (Q, V, U in particle-index order):
 * KBTNOW is the library's (the last step's temperature before streaming),
 * AVNOW the mean velocity of the state */
static void checkpoint_out(FILE *f, const mpcd_gpu_params *p, int sim_steps, int warm_steps, int runtime, int warmtime,
                           long seed, int flow_out, int vel_out, int dir_out, int ckpt_out, int dens_out, int rframe,
                           int znm, size_t n, const double *Q,
                           const double *V, const double *U, const int32_t *spid, double kbtnow, int binary,
                           uint32_t lib_step) {
    mpcd_ckpt ck;
    memset(&ck, 0, sizeof ck);
    ck.lib_step = lib_step;
    ck.sim_steps = sim_steps;
    ck.warm_steps = warm_steps;
    ck.runtime = runtime;
    ck.warmtime = warmtime;
    ck.seed = seed;
    ck.flow_out = flow_out;
    ck.vel_out = vel_out;
    ck.dir_out = dir_out;
    ck.ckpt_out = ckpt_out;
    ck.dens_out = dens_out;
    ck.rframe = rframe;
    ck.zero_net_mom = znm;
    ck.n = (int64_t)n;
    ck.Q = (double *)Q;
    ck.V = (double *)V;
    ck.U = (double *)U;
    ck.spid = (int32_t *)spid;
    double mom[3] = {0, 0, 0}, mtot = 0.0;
    for (size_t i = 0; i < n; i++) {
        const double m = p->sp[spid[i]].mass;
        for (int d = 0; d < 3; d++) mom[d] += m * V[3 * i + d];
        mtot += m;
    }
    ck.kbtnow = kbtnow;  /* TEMP after the thermostat, mpc.c:4752-4760 */
    for (int d = 0; d < 3; d++) ck.avnow[d] = mtot > 0.0 ? mom[d] / mtot : 0.0;
    if (binary) ckpt_write_bin(f, p, &ck);
    else ckpt_write(f, p, &ck);
}

static void reject_unsupported(const json *J) {
    const char *strs[] = {"mdIn"};
    for (int k = 0; k < 1; k++) {
        const char *v = json_str(J, strs[k], "");
        if (v[0]) die("\"%s\" (outside the solvent path) is not supported", strs[k]);
    }
    const char *ints[] = {"noHI", "incomp", "multiphase"};
    for (int k = 0; k < 3; k++)
        if (json_int(J, ints[k], 0)) die("\"%s\" is not supported", ints[k]);
    const json *sw = json_get(J, "swimmers");
    if (!sw) sw = json_get(J, "swimmer");
    if (sw && sw->type == JSON_ARR && sw->n > 0) die("%s are not supported", "swimmers");
    /* swimmers (read.c:1781) */
    if (json_int(J, "nSwim", 0) > 0) die("%s", "\"nSwim\": swimmers are outside the solvent path");
}

/* BC entry of the JSON (read.c:1136-1440): planes, curved and wavy surfaces */
static void read_wall(const json *o, int nspec, mpcd_gpu_wall *w, int index) {
    char msg[64];
    memset(w, 0, sizeof *w);
    double ainv[3], P[4] = {1, 1, 1, 1}, v3[3];
    if (json_nums(o, "aInv", ainv, 3) != 3) {
        snprintf(msg, sizeof msg, "%d", index);
        die("BC %s: aInv must have 3 components", msg);
    }
    for (int j = 0; j < 3; j++) w->A[j] = (fabs(ainv[j]) >= 1e-8) ? 1.0 / ainv[j] : 0.0;
    if (json_nums(o, "P", P, 4) != 4) {
        snprintf(msg, sizeof msg, "%d", index);
        die("BC %s: P must have 4 elements", msg);
    }
    double B[3] = {0, 0, 0}, rs[2] = {4, 4};
    json_nums(o, "wavy", B, 3);
    /* the general wall engine takes immobile walls of the four-fold symmetry
       (read.c:1268-1284,1370) */
    json_nums(o, "rotSym", rs, 2);
    if (fabs(rs[0] - 4) > 1e-8 || fabs(rs[1] - 4) > 1e-8) {
        snprintf(msg, sizeof msg, "%d", index);
        die("BC %s: rotSym other than 4 4 is outside this path", msg);
    }
    if (json_int(o, "dsplc", 0)) {
        snprintf(msg, sizeof msg, "%d", index);
        die("BC %s: displaceable (mobile) walls are outside this path", msg);
    }
    memcpy(w->P, P, sizeof P);
    memcpy(w->B, B, sizeof B);
    w->ABS = json_int(o, "abs", 0);
    if (json_nums(o, "Q", v3, 3) == 3) memcpy(w->Q, v3, sizeof v3);
    if (json_nums(o, "V", v3, 3) == 3) memcpy(w->V, v3, sizeof v3);
    if (json_nums(o, "L", v3, 3) == 3) memcpy(w->L, v3, sizeof v3);
    if (json_nums(o, "DVxyz", v3, 3) == 3) memcpy(w->DVxyz, v3, sizeof v3);
    /* orientation factors per axis default to 1 (read.c:1344-1348) */
    w->MUxyz[0] = w->MUxyz[1] = w->MUxyz[2] = 1.0;
    if (json_nums(o, "MUxyz", v3, 3) == 3) memcpy(w->MUxyz, v3, sizeof v3);
    if (json_nums(o, "DUxyz", v3, 3) == 3) memcpy(w->DUxyz, v3, sizeof v3);
    w->R = json_num(o, "R", 2.0);
    w->DN = json_num(o, "DN", 1.0);
    w->DT = json_num(o, "DT", 0.0);
    w->DVN = json_num(o, "DVN", 0.0);
    w->DVT = json_num(o, "DVT", 0.0);
    w->MVN = json_num(o, "MVN", 1.0);
    w->MVT = json_num(o, "MVT", 1.0);
    w->MUN = json_num(o, "MUN", 1.0);
    w->MUT = json_num(o, "MUT", 1.0);
    w->INV = json_int(o, "inv", 0);
    w->PHANTOM = json_int(o, "phantom", 0);
    w->COLL_TYPE = json_int(o, "colType", 1);
    if (json_int(o, "homeotropic", 0) == 1) {
        w->MUN = 1;
        w->MUT = 0;
    } else if (json_int(o, "planar", 0) == 1) {
        w->MUN = 0;
        w->MUT = 1;
    }
    double inter[MPCD_MAXSPEC];
    const int ni = json_nums(o, "interSRD", inter, MPCD_MAXSPEC);
    w->inter_mask = 0;
    for (int s = 0; s < nspec; s++)
        if (ni < 0 || (s < ni && inter[s] != 0.0)) w->inter_mask |= 1 << s;
}

/* domainWalls override (read.c:1441-1583): periodic (1) or solid (2) box walls */
static int domain_walls(int dim, const int L[3], int mode, int nspec, mpcd_gpu_wall *out) {
    int n = 0;
    for (int ax = 0; ax < dim; ax++)
        for (int side = 0; side < 2; side++) {
            mpcd_gpu_wall *w = &out[n++];
            memset(w, 0, sizeof *w);
            w->A[ax] = side ? -1.0 : 1.0;
            w->Q[ax] = side ? (double)L[ax] : 0.0;
            w->R = 0.0;
            w->MUN = w->MUT = 1.0;
            w->MUxyz[0] = w->MUxyz[1] = w->MUxyz[2] = 1.0; /* read.c:1558-1559 */
            w->COLL_TYPE = 1;
            w->inter_mask = (1 << nspec) - 1;
            if (mode == 1) {
                w->DN = (double)L[ax];
                w->MVN = w->MVT = 1.0;
            } else {
                w->DN = 0.0;
                w->MVN = w->MVT = -1.0;
                w->PHANTOM = 1;
            }
        }
    return n;
}

int main(int argc, char **argv) {
    const char *input = NULL, *outdir = ".";
    int fp32 = 0, nrep = 1, every = 0, steps_override = -1, use_graph = 1, dump = 0;
    const char *ckpt_now = NULL;
    int ckpt_bin = 0;
    long seed_override = -1;
    for (int a = 1; a < argc; a++) {
        if (!strcmp(argv[a], "-o") && a + 1 < argc) outdir = argv[++a];
        else if (!strcmp(argv[a], "--fp32")) fp32 = 1;
        else if (!strcmp(argv[a], "--fp32x")) fp32 = 2; /* the exact-sum FP32 form */
        else if (!strcmp(argv[a], "--binary-checkpoint")) ckpt_bin = 1;
        else if (!strcmp(argv[a], "--replicas") && a + 1 < argc) nrep = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--every") && a + 1 < argc) every = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--steps") && a + 1 < argc) steps_override = atoi(argv[++a]);
        else if (!strcmp(argv[a], "--seed") && a + 1 < argc) seed_override = atol(argv[++a]);
        else if (!strcmp(argv[a], "--no-graph")) use_graph = 0;
        else if (!strcmp(argv[a], "--dump-params")) dump = 1;
        else if (!strcmp(argv[a], "--write-checkpoint") && a + 1 < argc) ckpt_now = argv[++a];
        else if (argv[a][0] != '-' && !input) input = argv[a];
        else die("unknown argument %s", argv[a]);
    }
    if (!input) {
        fprintf(stderr,
                "usage: mpcd_host input.json [-o DIR] [--fp32 | --fp32x] [--replicas N] [--every K] [--steps N] [--seed S] "
                "[--no-graph]\n");
        return 2;
    }
    if (nrep < 1) die("%s", "--replicas must be at least 1");
    char err[256];
    json *J = json_load(input, err, sizeof err);
    if (!J) die("%s", err);
    if (J->type != JSON_OBJ) die("%s", "the input is not a JSON object");
    reject_unsupported(J);

    /* simulation-scale inputs (read.c:1023-1114) */
    mpcd_gpu_params p;
    memset(&p, 0, sizeof p);
    double dom[3] = {0, 0, 0};
    /* a checkpoint (read.c:1116-1121) replaces the inputs, the particles and
     * the loop counters after readJson, as readchckpnt does (mpcd.c:175-180) */
    const char *ckpt_in = json_str(J, "checkpointIn", "");
    int dim = json_nums(J, "domain", dom, 3);
    if (!ckpt_in[0] && dim != 2 && dim != 3) die("%s", "domain must have 2 or 3 entries");
    if (dim != 2 && dim != 3) dim = 2;
    p.dim = dim;
    for (int d = 0; d < 3; d++) p.L[d] = d < dim ? (int)dom[d] : 1;
    p.kbt = json_num(J, "kbt", 1.0);
    p.dt = json_num(J, "dt", 0.1);
    int sim_steps = json_int(J, "simSteps", 2000);
    int warm = json_int(J, "warmUp", 0);
    int flow_out = json_int(J, "flowOut", 0);
    int vel_out = json_int(J, "velOut", 0);
    int dir_out = json_int(J, "dirSOut", 0);
    int dens_out = json_int(J, "densOut", 0);
    const int en_out = json_int(J, "energyOut", 0);
    int ckpt_out = json_int(J, "checkpointOut", 0);
    p.galinv = json_int(J, "galInv", 1);
    p.tstech = json_int(J, "tsTech", 0);
    /* getJObjIntMultiple (cJson.c:265-279): each tag that differs from the
     * default overrides, so collOp wins over rTech when both are set */
    p.rtech = 2;
    const char *optags[2] = {"rTech", "collOp"};
    for (int k = 0; k < 2; k++) {
        const int v = json_int(J, optags[k], 2);
        if (v != 2) p.rtech = v;
    }
    p.lc = json_int(J, "lc", 0);
    p.tau = json_num(J, "tau", 0.5);
    p.rotang = json_num(J, "rotAng", 1.570796);
    double g[3] = {0, 0, 0};
    json_nums(J, "grav", g, 3);
    memcpy(p.grav, g, sizeof g);
    p.precision = fp32 == 2 ? MPCD_FP32_EXACT : fp32 ? MPCD_FP32 : MPCD_FP64;
    /* rest frame (read.c:1062-1063): rFrame (default 1) removes the net
     * momentum of the initial state; with zeroNetMom = Z > 0 the output step
     * of every runtime divisible by Z does it again */
    int rframe = json_int(J, "rFrame", 1);
    int znm = json_int(J, "zeroNetMom", 0);
    long seed = json_int(J, "seed", 0);
    if (seed_override >= 0) seed = seed_override;
    if (seed == 0) seed = (long)time(NULL);
    rng_state ^= (uint64_t)seed * 0xD1B54A32D192ED03ull;

    /* species (read.c:1598-1680) */
    const json *spa = json_get(J, "species");
    p.nspec = (spa && spa->type == JSON_ARR && spa->n > 0) ? spa->n : 1;
    if (p.nspec > MPCD_MAXSPEC) die("%s", "too many species");
    const double vol = (double)p.L[0] * p.L[1] * (dim == 3 ? p.L[2] : 1);
    int64_t pop[MPCD_MAXSPEC], ntot = 0;
    double dens[MPCD_MAXSPEC];
    int vdist[MPCD_MAXSPEC];
    for (int s = 0; s < p.nspec; s++) {
        const json *o = (spa && spa->type == JSON_ARR && spa->n > 0) ? &spa->items[s] : NULL;
        mpcd_gpu_species *sp = &p.sp[s];
        sp->mass = json_num(o, "mass", 1.0);
        dens[s] = json_num(o, "dens", -1.0);
        pop[s] = dens[s] < 0 ? (int64_t)json_int(o, "pop", (int)(vol * 20)) : 0;
        vdist[s] = json_int(o, "vDist", 0);
        sp->vdist = vdist[s];
        sp->rfc = json_num(o, "rfc", 0.01);
        sp->tumble = json_num(o, "tumble", 2.0);
        sp->chihi = json_num(o, "shearSusc", 0.5);
        sp->act = json_num(o, "act", 0.05);
        sp->mfpot = json_num(o, "mfpot", 10.0);
        sp->sigwidth = json_num(o, "sigWidth", 1.0);
        sp->sigpos = json_num(o, "sigPos", sp->sigwidth);
        sp->minactratio = json_num(o, "minActRatio", 0.0);
        sp->damp = json_num(o, "damp", 0.0);
    }

    /* boundaries: the BC list, then the domain walls of domainWalls = 1
     * (periodic) or 2 (solid) appended after it; without a BC list the
     * domain walls are periodic unless domainWalls = 2 (read.c:1136-1441) */
    const json *bca = json_get(J, "BC");
    int dwalls = json_int(J, "domainWalls", 0);
    p.nwall = 0;
    if (bca && bca->type == JSON_ARR) {
        if (bca->n > MPCD_MAXWALL) die("%s", "too many BC entries");
        for (int w = 0; w < bca->n; w++) read_wall(&bca->items[w], p.nspec, &p.wall[w], w);
        p.nwall = bca->n;
    } else if (dwalls != 2) {
        dwalls = 1;
    }
    if (dwalls == 1 || dwalls == 2) {
        if (p.nwall + 2 * dim > MPCD_MAXWALL) die("%s", "too many BC entries with the domain walls");
        p.nwall += domain_walls(dim, p.L, dwalls, p.nspec, p.wall + p.nwall);
    }
    /* a density sets the population over the accessible volume (read.c:1708-1716) */
    for (int s = 0; s < p.nspec; s++) {
        if (dens[s] >= 0) pop[s] = (int64_t)(host_accessible_volume(&p, s) * dens[s]);
        ntot += pop[s];
    }
    if (ntot <= 0 && !ckpt_in[0]) die("%s", "no particles");

    /* recovering a checkpoint: its inputs, particles, loop counters and
     * output flags replace the input file's; the warm-up is skipped and the
     * temporal loop starts at the checkpoint's runtime (mpcd.c:257, 290) */
    mpcd_ckpt ck;
    memset(&ck, 0, sizeof ck);
    int run0 = 0, warmtime = warm > 0 ? warm + 1 : 0;
    if (ckpt_in[0]) {
        char cerr[512];
        mpcd_gpu_params q;
        if (ckpt_read(ckpt_in, &q, &ck, cerr, sizeof cerr)) die("%s", cerr);
        q.precision = p.precision;
        p = q;
        dim = p.dim;
        sim_steps = ck.sim_steps;
        warm = 0;
        /* the reference's file re-runs its runtime's step (mpcd.c:257, 290);
           a binary checkpoint, which carries the draw counter, continues
           with the next one */
        run0 = ck.binary ? ck.runtime + 1 : ck.runtime;
        warmtime = ck.warmtime;
        flow_out = ck.flow_out;
        vel_out = ck.vel_out;
        dir_out = ck.dir_out;
        ckpt_out = ck.ckpt_out;
        dens_out = ck.dens_out;
        rframe = ck.rframe;
        znm = ck.zero_net_mom;
        if (seed_override < 0 && ck.seed != 0) seed = ck.seed;
        ntot = ck.n;
        for (int s = 0; s < MPCD_MAXSPEC; s++) pop[s] = 0;
        for (int64_t i = 0; i < ck.n; i++) pop[ck.spid[i]]++;
    }
    if (steps_override >= 0) sim_steps = steps_override;
    /* the initial transformation uses rFrame as read; from then on rFrame
     * says whether the periodic one runs: not without zeroNetMom, not with a
     * body force, not for 2 particles or fewer (init.c:2141-2156; a
     * checkpoint holds the flag already reduced) */
    const int rframe_init = rframe;
    if (znm <= 0 || (!ckpt_in[0] && (ntot <= 2 || fabs(p.grav[0]) >= 1e-8 || fabs(p.grav[1]) >= 1e-8 ||
                                      fabs(p.grav[2]) >= 1e-8)))
        rframe = 0;

    if (dump) {
        /* the parameters as read, without touching the GPU */
        printf("{\"dim\": %d, \"L\": [%d, %d, %d], \"dt\": %.17g, \"kbt\": %.17g, \"rotang\": %.17g, "
               "\"tau\": %.17g, \"grav\": [%.17g, %.17g, %.17g], \"rtech\": %d, \"lc\": %d, \"tstech\": %d, "
               "\"galinv\": %d, \"simSteps\": %d, \"warmUp\": %d, \"flowOut\": %d, \"velOut\": %d, \"dirSOut\": %d, "
               "\"densOut\": %d, \"checkpointOut\": %d, \"rFrame\": %d, \"zeroNetMom\": %d, \"runtime\": %d, \"gndnst\": %.17g, \"avv\": [%.17g, %.17g, %.17g], "
               "\"seed\": %ld, \"species\": [",
               p.dim, p.L[0], p.L[1], p.L[2], p.dt, p.kbt, p.rotang, p.tau, p.grav[0], p.grav[1], p.grav[2], p.rtech,
               p.lc, p.tstech, p.galinv, sim_steps, warm, flow_out, vel_out, dir_out, dens_out, ckpt_out, rframe, znm, run0, p.gndnst, p.avv[0],
               p.avv[1], p.avv[2], seed);
        for (int s = 0; s < p.nspec; s++) {
            const mpcd_gpu_species *sp = &p.sp[s];
            printf("%s{\"pop\": %lld, \"mass\": %.17g, \"vdist\": %d, \"rfc\": %.17g, \"tumble\": %.17g, "
                   "\"chihi\": %.17g, \"act\": %.17g, \"mfpot\": %.17g, \"sigwidth\": %.17g, \"sigpos\": %.17g, "
                   "\"minactratio\": %.17g, \"damp\": %.17g}",
                   s ? ", " : "", (long long)pop[s], sp->mass, sp->vdist, sp->rfc, sp->tumble, sp->chihi, sp->act,
                   sp->mfpot, sp->sigwidth, sp->sigpos, sp->minactratio, sp->damp);
        }
        printf("], \"walls\": [");
        for (int w = 0; w < p.nwall; w++) {
            const mpcd_gpu_wall *W = &p.wall[w];
            printf("%s{\"A\": [%.17g, %.17g, %.17g], \"Q\": [%.17g, %.17g, %.17g], \"R\": %.17g, \"DN\": %.17g, "
                   "\"MVN\": %.17g, \"MVT\": %.17g, \"MUN\": %.17g, \"MUT\": %.17g, \"MUxyz\": [%.17g, %.17g, %.17g], "
                   "\"V\": [%.17g, %.17g, %.17g], \"P\": [%.17g, %.17g, %.17g, %.17g], \"B\": [%.17g, %.17g, %.17g], "
                   "\"ABS\": %d, \"INV\": %d, \"PHANTOM\": %d, \"inter_mask\": %d}",
                   w ? ", " : "", W->A[0], W->A[1], W->A[2], W->Q[0], W->Q[1], W->Q[2], W->R, W->DN, W->MVN, W->MVT,
                   W->MUN, W->MUT, W->MUxyz[0], W->MUxyz[1], W->MUxyz[2], W->V[0], W->V[1], W->V[2], W->P[0], W->P[1],
                   W->P[2], W->P[3], W->B[0], W->B[1], W->B[2], W->ABS, W->INV, W->PHANTOM, W->inter_mask);
        }
        printf("]}\n");
        json_free(J);
        free(J);
        return 0;
    }

    /* initial state: uniform positions, velocities by vDist (4 Gaussian,
     * otherwise random direction at the thermal speed), zero net momentum
     * (galileantrans, mtools.c:673), random orientations */
    const size_t n = (size_t)ntot;
    const size_t nall = n * (size_t)nrep;
    double *Q = (double *)calloc(3 * nall, sizeof(double));
    double *V = (double *)calloc(3 * nall, sizeof(double));
    double *U = (double *)calloc(3 * nall, sizeof(double));
    int32_t *spid = (int32_t *)calloc(n, sizeof(int32_t));
    int *ccount = (int *)calloc((size_t)p.L[0] * p.L[1] * p.L[2], sizeof(int));
    mpcd_gpu_params *pr = (mpcd_gpu_params *)calloc((size_t)nrep, sizeof(mpcd_gpu_params));
    int64_t *nr = (int64_t *)calloc((size_t)nrep, sizeof(int64_t));
    if (!Q || !V || !U || !spid || !ccount || !pr || !nr) die("%s", "out of host memory");
    for (int r = 0; r < nrep && ckpt_in[0]; r++) {
        /* every replica starts from the checkpoint's particles */
        memcpy(Q + 3 * n * r, ck.Q, sizeof(double) * 3 * n);
        memcpy(V + 3 * n * r, ck.V, sizeof(double) * 3 * n);
        memcpy(U + 3 * n * r, ck.U, sizeof(double) * 3 * n);
        memcpy(spid, ck.spid, sizeof(int32_t) * n);
        pr[r] = p;
        nr[r] = ntot;
    }
    for (int r = 0; r < nrep && !ckpt_in[0]; r++) {
        double *Qr = Q + 3 * n * r, *Vr = V + 3 * n * r, *Ur = U + 3 * n * r;
        size_t i = 0;
        double mom[3] = {0, 0, 0}, mtot = 0.0;
        memset(ccount, 0, sizeof(int) * (size_t)p.L[0] * p.L[1] * p.L[2]);
        for (int s = 0; s < p.nspec; s++)
            for (int64_t k = 0; k < pop[s]; k++, i++) {
                const double m = p.sp[s].mass, sig = sqrt(p.kbt / m);
                spid[i] = s;
                /* uniform over the space no wall excludes (replacePos_WithCheck, init.c:1462) */
                int tries = 0;
                do {
                    if (++tries > 1000000) die("%s", "no free position for a particle: the walls fill the box");
                    for (int d = 0; d < 3; d++) Qr[3 * i + d] = d < dim ? rng_u() * p.L[d] : 0.0;
                } while (!host_position_free(&p, Qr + 3 * i));
                if (vdist[s] == 4) {
                    for (int d = 0; d < dim; d++) Vr[3 * i + d] = sig * rng_n();
                } else {
                    double u[3] = {0, 0, 0}, nn = 0.0;
                    for (int d = 0; d < dim; d++) {
                        u[d] = rng_n();
                        nn += u[d] * u[d];
                    }
                    nn = sqrt(nn);
                    for (int d = 0; d < dim; d++) Vr[3 * i + d] = sqrt((double)dim) * sig * u[d] / nn;
                }
                double o[3] = {0, 0, 0}, on = 0.0;
                for (int d = 0; d < dim; d++) {
                    o[d] = rng_n();
                    on += o[d] * o[d];
                }
                on = sqrt(on);
                for (int d = 0; d < dim; d++) Ur[3 * i + d] = o[d] / on;
                for (int d = 0; d < 3; d++) mom[d] += m * Vr[3 * i + d];
                mtot += m;
                const int a = (int)Qr[3 * i], b = (int)Qr[3 * i + 1], c = dim == 3 ? (int)Qr[3 * i + 2] : 0;
                ccount[(a * p.L[1] + b) * p.L[2] + c]++;
            }
        /* galileantrans of the initial state when rFrame is on and there are
         * more than 2 particles (init.c:2141-2152) */
        if (rframe_init && n > 2)
            for (size_t k = 0; k < n; k++)
                for (int d = 0; d < 3; d++) Vr[3 * k + d] -= mom[d] / mtot;
        /* GnDNST: mean population of the occupied cells (globalDensities, init.c:995-1018) */
        double occ = 0.0;
        for (int k = 0; k < p.L[0] * p.L[1] * p.L[2]; k++) occ += ccount[k] > 0;
        pr[r] = p;
        pr[r].gndnst = (double)n / occ;
        nr[r] = ntot;
    }
    p.gndnst = pr[0].gndnst;
    if (ckpt_now) {
        /* the state the run would start from, as a checkpoint; no GPU */
        FILE *fc = fopen(ckpt_now, ckpt_bin ? "wb" : "w");
        if (!fc) die("cannot write %s", ckpt_now);
        checkpoint_out(fc, &p, sim_steps, warm, run0, warmtime, seed, flow_out, vel_out, dir_out, ckpt_out, dens_out, rframe,
                       znm, n, Q, V, U,
                       spid, temp_of(&p, n, V, spid), ckpt_bin, ck.lib_step);
        fclose(fc);
        ckpt_free(&ck);
        return 0;
    }
    mpcd_gpu_ctx *ctx = NULL;
    check(NULL, mpcd_gpu_create(pr, nrep, nr, (uint64_t)seed, &ctx), "mpcd_gpu_create");
    const int resume_step = ck.binary; /* a binary checkpoint carries the draw counter */
    for (int r = 0; r < nrep; r++)
        check(ctx, mpcd_gpu_upload(ctx, r, Q + 3 * n * r, V + 3 * n * r, p.lc ? U + 3 * n * r : NULL,
                                   p.nspec > 1 ? spid : NULL),
              "mpcd_gpu_upload");
    if (resume_step) check(ctx, mpcd_gpu_set_step(ctx, ck.lib_step), "mpcd_gpu_set_step");

    /* the reference's loops: warmUp + 1 and simSteps + 1 steps, outputs in between */
    char path[4096];
    mkdir(outdir, 0755);
    snprintf(path, sizeof path, "%s/solvent.dat", outdir);
    FILE *fo = fopen(path, "w");
    if (!fo) die("cannot write %s", path);
    fprintf(fo, "# step  KBT(replica 0)  KE  Px  Py  Pz   (%s state, %d replica%s of %lld particles)\n",
            fp32 == 2 ? "FP32 exact-sum" : fp32 ? "FP32" : "FP64", nrep, nrep > 1 ? "s" : "", (long long)ntot);
    /* steps of the temporal loop: runtime run0 .. simSteps */
    const int total = sim_steps + 1 - run0 > 0 ? sim_steps + 1 - run0 : 0;
    if (every <= 0) every = total > 10 ? total / 10 : 1;
    if (warm > 0) check(ctx, mpcd_gpu_run(ctx, warm + 1, use_graph), "warm-up");
    /* field outputs; the flow sums start after the warm-up (mpc.c:4947:
     * !in.warmupSteps); orderout needs LC (pout.c:2695) */
    const int do_dir = p.lc && dir_out > 0;
    FILE *ff[64] = {NULL}, *fv[64] = {NULL}, *fd[64] = {NULL}, *fn[64] = {NULL}, *fe[64] = {NULL};
    const size_t nc = (size_t)mpcd_gpu_ncells(ctx, 0);
    double *flow = NULL, *cvcm = NULL, *cS = NULL, *cdir = NULL;
    int32_t *cpop = NULL;
    double *kbt_now = (double *)calloc((size_t)nrep, sizeof(double));
    if (!kbt_now) die("out of memory (%s)", "kbt_now");
    if ((flow_out > 0 || vel_out > 0 || do_dir || dens_out > 0 || en_out > 0) && nrep > 64)
        die("%s", "field outputs: at most 64 replicas");
    int32_t *dcnt = NULL;
    double *dmass = NULL;
    if (dens_out > 0) {
        dcnt = (int32_t *)malloc(sizeof(int32_t) * nc * p.nspec);
        dmass = (double *)malloc(sizeof(double) * nc);
        if (!dcnt || !dmass) die("%s", "out of host memory");
        char cols[512];
        int off = snprintf(cols, sizeof cols, "   t\t   QX\t   QY\t   QZ\tpop\t\tmass\t\tpopSRD\t\tpopMD\t\tpopSW");
        for (int s = 0; s < p.nspec && off < (int)sizeof cols; s++)
            off += snprintf(cols + off, sizeof cols - off, "\t\tSP%d", s);
        for (int r = 0; r < nrep; r++)
            fn[r] = field_open(outdir, "densityfield", "density field: population and mass per cell and species", cols, r,
                               nrep, dens_out, p.L, dim, fp32);
    }
    if (flow_out > 0) {
        check(ctx, mpcd_gpu_flow(ctx, 1), "mpcd_gpu_flow");
        flow = (double *)malloc(sizeof(double) * 3 * nc);
        if (!flow) die("%s", "out of host memory");
        for (int r = 0; r < nrep; r++)
            ff[r] = field_open(outdir, "flowfield",
                               "flow field: centre-of-mass velocity averaged over each output interval", FLOW_COLUMNS,
                               r, nrep, flow_out, p.L, dim, fp32);
    }
    for (int r = 0; r < nrep && en_out > 0; r++)
        fe[r] = field_open(outdir, "energy", "energy: kinetic and nematic mean-field energies, temperature",
                           "t\t\tMPCD_kin\t\tMPCD_nem\t\tBC_kin\t\tBC_rot\t\tTotal\t\tKBT", r, nrep, en_out, p.L, dim,
                           fp32);
    if (vel_out > 0 || do_dir || (en_out > 0 && p.lc)) {
        cpop = (int32_t *)malloc(sizeof(int32_t) * nc);
        cvcm = (double *)malloc(sizeof(double) * 3 * nc);
        cS = (double *)malloc(sizeof(double) * nc);
        cdir = (double *)malloc(sizeof(double) * 3 * nc);
        if (!cpop || !cvcm || !cS || !cdir) die("%s", "out of host memory");
    }
    for (int r = 0; r < nrep && vel_out > 0; r++)
        fv[r] = field_open(outdir, "velfield", "velocity field: instantaneous centre-of-mass velocity", FLOW_COLUMNS, r,
                           nrep, vel_out, p.L, dim, fp32);
    for (int r = 0; r < nrep && do_dir; r++)
        fd[r] = field_open(outdir, "directorfield", "director field: director and scalar order parameter",
                           "t\t\tQX\t\tQY\t\tQZ\t\tNX\t\tNY\t\tNZ\t\tS", r, nrep, dir_out, p.L, dim, fp32);
    check(ctx, mpcd_gpu_synchronize(ctx), "synchronize");
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    double busy = 0.0;
    int next_out = every < total ? every : total;
    for (int done = 0; done < total;) {
        int k = next_out - done;
        const int cad[7] = {flow_out, vel_out, do_dir ? dir_out : 0, ckpt_out, dens_out, en_out, rframe ? znm : 0};
        for (int q = 0; q < 7; q++)
            if (cad[q] > 0 && steps_to_next(run0 + done, cad[q]) < k) k = steps_to_next(run0 + done, cad[q]);
        struct timespec a, b;
        clock_gettime(CLOCK_MONOTONIC, &a);
        check(ctx, mpcd_gpu_run(ctx, k, use_graph), "mpcd_gpu_run");
        check(ctx, mpcd_gpu_synchronize(ctx), "synchronize");
        clock_gettime(CLOCK_MONOTONIC, &b);
        busy += (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
        done += k;
        const int rt = run0 + done - 1;
        const double tnow = (double)rt * p.dt;
        /* outputResults removes the net momentum before its outputs
         * (pout.c:2557-2562); the cell values stay the step's */
        if (rframe && rt % znm == 0) check(ctx, mpcd_gpu_zero_net_momentum(ctx), "mpcd_gpu_zero_net_momentum");
        if (flow_out > 0 && rt % flow_out == 0) {
            /* flowout at this runtime (pout.c:2690), then the sums restart */
            for (int r = 0; r < nrep; r++) {
                int64_t nsum = 0;
                check(ctx, mpcd_gpu_flow_read(ctx, r, flow, &nsum, 1), "mpcd_gpu_flow_read");
                flow_write(ff[r], flow, p.L, dim, tnow, flow_out);
            }
        }
        const int vel_due = vel_out > 0 && rt % vel_out == 0, dir_due = do_dir && rt % dir_out == 0;
        if (vel_due || dir_due)
            for (int r = 0; r < nrep; r++) {
                /* the cells of the re-bin that ends the step (localPROP, mpc.c:4942) */
                check(ctx, mpcd_gpu_get_cells(ctx, r, cpop, cvcm, cS, cdir), "mpcd_gpu_get_cells");
                if (vel_due) vel_write(fv[r], cvcm, p.L, dim, tnow);
                if (dir_due) order_write(fd[r], cpop, cdir, cS, p.L, dim, tnow);
            }
        if (dens_out > 0 && rt % dens_out == 0)
            for (int r = 0; r < nrep; r++) {
                double *Qr = Q + 3 * n * r, *Vr = V + 3 * n * r;
                check(ctx, mpcd_gpu_download(ctx, r, Qr, Vr, NULL), "mpcd_gpu_download");
                memset(dcnt, 0, sizeof(int32_t) * nc * p.nspec);
                memset(dmass, 0, sizeof(double) * nc);
                const int ny = p.L[1] + 1, nz = dim == 3 ? p.L[2] + 1 : 1;
                for (size_t i = 0; i < n; i++) {
                    const int a = (int)Qr[3 * i], b = (int)Qr[3 * i + 1], cz = dim == 3 ? (int)Qr[3 * i + 2] : 0;
                    if (a < 0 || b < 0 || cz < 0 || a > p.L[0] || b > p.L[1] || cz >= nz) continue;
                    const size_t c = (size_t)(a * ny + b) * nz + cz;
                    const int s = p.nspec > 1 ? spid[i] : 0;
                    dcnt[c * p.nspec + s]++;
                    dmass[c] += p.sp[s].mass;
                }
                density_write(fn[r], dcnt, dmass, p.nspec, p.L, dim, tnow);
            }
        if ((en_out > 0 && rt % en_out == 0) || (ckpt_out > 0 && rt % ckpt_out == 0))
            check(ctx, mpcd_gpu_kbtnow(ctx, kbt_now), "mpcd_gpu_kbtnow");
        if (en_out > 0 && rt % en_out == 0)
            for (int r = 0; r < nrep; r++) {
                double *Qr = Q + 3 * n * r, *Vr = V + 3 * n * r, *Ur = U + 3 * n * r;
                check(ctx, mpcd_gpu_download(ctx, r, Qr, Vr, p.lc ? Ur : NULL), "mpcd_gpu_download");
                if (p.lc) check(ctx, mpcd_gpu_get_cells(ctx, r, cpop, cvcm, cS, cdir), "mpcd_gpu_get_cells");
                energy_write(fe[r], &p, n, Qr, Vr, Ur, spid, cpop, cS, cdir, tnow, kbt_now[r]);
            }
        if (ckpt_out > 0 && rt % ckpt_out == 0) {
            /* runCheckpoint after the outputs of this runtime (mpcd.c:310-312): DIR/checkpoint.dat, replica 0;
             * the warm-up count is 0 once its loop has run (mpcd.c:278) */
            check(ctx, mpcd_gpu_download(ctx, 0, Q, V, p.lc ? U : NULL), "mpcd_gpu_download");
            char cpath[4096];
            snprintf(cpath, sizeof cpath, "%s/checkpoint.%s", outdir, ckpt_bin ? "bin" : "dat");
            FILE *fc = fopen(cpath, ckpt_bin ? "wb" : "w");
            if (!fc) die("cannot write %s", cpath);
            uint32_t lstep = 0;
            check(ctx, mpcd_gpu_get_step(ctx, &lstep), "mpcd_gpu_get_step");
            checkpoint_out(fc, &p, sim_steps, 0, rt, warmtime, seed, flow_out, vel_out, dir_out, ckpt_out, dens_out, rframe,
                           znm, n, Q, V, U,
                           spid, kbt_now[0], ckpt_bin, lstep);
            fclose(fc);
        }
        if (done < next_out) continue;
        next_out = (next_out + every < total) ? next_out + every : total;
        check(ctx, mpcd_gpu_download(ctx, 0, Q, V, NULL), "mpcd_gpu_download");
        double ke = 0.0, P[3] = {0, 0, 0};
        for (size_t i = 0; i < n; i++) {
            const double m = p.sp[p.nspec > 1 ? spid[i] : 0].mass;
            double v2 = 0.0;
            for (int d = 0; d < 3; d++) {
                v2 += (V[3 * i + d] - p.avv[d]) * (V[3 * i + d] - p.avv[d]);
                P[d] += m * V[3 * i + d];
            }
            ke += m * v2;
        }
        fprintf(fo, "%d %.12g %.12g %.6e %.6e %.6e\n", rt, ke / (dim * (double)n), 0.5 * ke, P[0], P[1], P[2]);
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    fclose(fo);
    for (int r = 0; r < nrep; r++) {
        if (ff[r]) fclose(ff[r]);
        if (fv[r]) fclose(fv[r]);
        if (fd[r]) fclose(fd[r]);
        if (fn[r]) fclose(fn[r]);
        if (fe[r]) fclose(fe[r]);
    }
    free(dcnt);
    free(dmass);
    ckpt_free(&ck);
    free(flow);
    free(cpop);
    free(kbt_now);
    free(cvcm);
    free(cS);
    free(cdir);
    printf("{\"steps\": %d, \"replicas\": %d, \"particles_per_replica\": %lld, \"precision\": \"%s\", "
           "\"wall_s\": %.3f, \"step_s\": %.3f, \"particle_steps_per_s\": %.4g, \"output\": \"%s\"}\n",
           total, nrep, (long long)ntot, fp32 ? "fp32" : "fp64",
           (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec), busy,
           (double)total * (double)ntot * nrep / busy, path);
    mpcd_gpu_destroy(ctx);
    json_free(J);
    free(J);
    free(Q);
    free(V);
    free(U);
    free(spid);
    free(ccount);
    free(pr);
    free(nr);
    return 0;
}
