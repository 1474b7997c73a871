/*
 * ref_harness.c -- runs the reference's own timestep() on an injected initial
 * state and injected draws, and dumps what it computed (TEST INFRASTRUCTURE).
 *
 * Built by oracle/Makefile against objects compiled from the sources under
 * /root/reference (never copied); the output binary lives in oracle/_ref/.
 *
 * The harness builds the reference's own data structures the way
 * read.c/init.c would for the recorded case (cell array allocation as in
 * read.c:1699-1709, wall flags as in read.c:1411-1433, setPBC() for periodic
 * axes, binin()+localPROP()+globalDensities() as in initializeSIM()
 * init.c:2114-2128, in->C/in->S as in initvar() init.c:1194-1195), then calls
 * the reference's timestep() (mpc.c:4496) once per step with every output
 * switched off.  Random draws come from ref_rand.c (sequential streams).
 *
 * Usage:  ref_harness CASE.txt STATE.bin OUT.bin [timing | flow | ckpt]
 *
 * CASE.txt is a list of "key value..." lines (see tests/golden/make_golden.py).
 * STATE.bin holds N records of 9 doubles (Q, V, U) followed by N int32 SPIDs.
 * OUT.bin receives, after the initial binning and after every step:
 *   int32 tag, double kbtnow, uint64 draw counts[3],
 *   N*9 doubles (Q,V,U in particle-index order),
 *   ncell int32 cell populations (reference x-major layout over XYZ_P1),
 *   N int32 particle indices in list order (cells x-major, lists head->tail),
 *   ncell*(1+3+3) doubles: S, DIR[3], VCM[3] per cell.
 * With a 4th argument "timing" nothing is dumped and the wall-clock time of
 * the step loop is printed (used as the CPU baseline).  With "flow" the
 * reference's flow output is switched on (outFlags.FLOWOUT = OUT, so every
 * timestep() ends with localFLOW + sumFLOW, mpc.c:4946-4951), and after the
 * last step OUT.bin receives int32 tag -2 and ncell*3 doubles: FLOW per cell.
 * With "ckpt" the reference's own checkpoint() (pout.c:2325-2404) writes
 * OUT.bin.ckpt after the last step (runtime = the last step, warmtime 0,
 * simSteps and warmupSteps as given by the case's "simsteps"/"warmup" keys).
 * A case key "zeronetmom Z" (Z > 0) follows every step s with s % Z == 0 by
 * what outputResults does for rFrame (pout.c:2540-2562): bin(unshifted),
 * localPROP, galileantrans and zeroExtraDims.
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "definitions.h"
#include "globals.h"
#include "SRDclss.h"
#include "mpc.h"
#include "init.h"
#include "read.h"
#include "therm.h"
#include "lc.h"
#include "pout.h"
#include "mtools.h"

void ref_rand_seed(uint64_t seed);
void ref_rand_counts(uint64_t out[3]);
void setPBC(bc *WALL);
void globalDensities(cell ***CL, spec SP[]);

#define MAXW 16

typedef struct {
    int dim, L[3], N, steps, rtech, lc, tstech, galinv, nspec, nwall, simsteps, warmup, zeronetmom;
    double dt, kbt, rotang, tau, grav[3], avv[3];
    unsigned long long seed;
    double spv[MAXSPECI][12];
    double wv[MAXW][32];
    double shape[MAXW][8];  /* P[4] B[3] ABS of a curved or wavy wall ("shape" after its "wall" line) */
    int has_shape[MAXW];
} hcase;

static void die(const char *m)
{
    fprintf(stderr, "ref_harness: %s\n", m);
    exit(2);
}

static void read_case(const char *path, hcase *c)
{
    FILE *f = fopen(path, "r");
    char key[64];
    int i, ns = 0, nw = 0;
    if (!f) die("cannot open case");
    memset(c, 0, sizeof(*c));
    while (fscanf(f, "%63s", key) == 1) {
        if (!strcmp(key, "dim")) { if (fscanf(f, "%d", &c->dim) != 1) die("dim"); }
        else if (!strcmp(key, "L")) { if (fscanf(f, "%d %d %d", &c->L[0], &c->L[1], &c->L[2]) != 3) die("L"); }
        else if (!strcmp(key, "N")) { if (fscanf(f, "%d", &c->N) != 1) die("N"); }
        else if (!strcmp(key, "steps")) { if (fscanf(f, "%d", &c->steps) != 1) die("steps"); }
        else if (!strcmp(key, "simsteps")) { if (fscanf(f, "%d", &c->simsteps) != 1) die("simsteps"); }
        else if (!strcmp(key, "warmup")) { if (fscanf(f, "%d", &c->warmup) != 1) die("warmup"); }
        else if (!strcmp(key, "zeronetmom")) { if (fscanf(f, "%d", &c->zeronetmom) != 1) die("zeronetmom"); }
        else if (!strcmp(key, "rtech")) { if (fscanf(f, "%d", &c->rtech) != 1) die("rtech"); }
        else if (!strcmp(key, "lc")) { if (fscanf(f, "%d", &c->lc) != 1) die("lc"); }
        else if (!strcmp(key, "tstech")) { if (fscanf(f, "%d", &c->tstech) != 1) die("tstech"); }
        else if (!strcmp(key, "galinv")) { if (fscanf(f, "%d", &c->galinv) != 1) die("galinv"); }
        else if (!strcmp(key, "dt")) { if (fscanf(f, "%lf", &c->dt) != 1) die("dt"); }
        else if (!strcmp(key, "kbt")) { if (fscanf(f, "%lf", &c->kbt) != 1) die("kbt"); }
        else if (!strcmp(key, "rotang")) { if (fscanf(f, "%lf", &c->rotang) != 1) die("rotang"); }
        else if (!strcmp(key, "tau")) { if (fscanf(f, "%lf", &c->tau) != 1) die("tau"); }
        else if (!strcmp(key, "grav")) { if (fscanf(f, "%lf %lf %lf", &c->grav[0], &c->grav[1], &c->grav[2]) != 3) die("grav"); }
        else if (!strcmp(key, "avv")) { if (fscanf(f, "%lf %lf %lf", &c->avv[0], &c->avv[1], &c->avv[2]) != 3) die("avv"); }
        else if (!strcmp(key, "seed")) { if (fscanf(f, "%llu", &c->seed) != 1) die("seed"); }
        else if (!strcmp(key, "species")) {
            /* mass rfc tumble chihi act mfpot damp sigwidth sigpos minactratio vdist len */
            if (ns >= MAXSPECI) die("too many species");
            for (i = 0; i < 12; i++) if (fscanf(f, "%lf", &c->spv[ns][i]) != 1) die("species");
            ns++;
        }
        else if (!strcmp(key, "wall")) {
            /* A[3] Q[3] R DN DT DVN DVT DVxyz[3] MVN MVT MUN MUT MUxyz[3] DUxyz[3] V[3] INV PHANTOM COLL_TYPE */
            if (nw >= MAXW) die("too many walls");
            for (i = 0; i < 30; i++) if (fscanf(f, "%lf", &c->wv[nw][i]) != 1) die("wall");
            nw++;
        }
        else if (!strcmp(key, "shape")) {
            /* P[4] B[3] ABS of the wall on the line before */
            if (nw < 1) die("shape before wall");
            for (i = 0; i < 8; i++) if (fscanf(f, "%lf", &c->shape[nw - 1][i]) != 1) die("shape");
            c->has_shape[nw - 1] = 1;
        }
        else die("unknown key");
    }
    c->nspec = ns;
    c->nwall = nw;
    fclose(f);
}

static void dump(FILE *o, int tag, double kbtnow, particleMPC *pp, cell ***CL)
{
    uint64_t cnt[3];
    int a, b, k, i, d;
    double x[9];
    ref_rand_counts(cnt);
    fwrite(&tag, sizeof(int), 1, o);
    fwrite(&kbtnow, sizeof(double), 1, o);
    fwrite(cnt, sizeof(uint64_t), 3, o);
    for (i = 0; i < GPOP; i++) {
        for (d = 0; d < 3; d++) {
            x[d] = pp[i].Q[d];
            x[3 + d] = pp[i].V[d];
            x[6 + d] = pp[i].U[d];
        }
        fwrite(x, sizeof(double), 9, o);
    }
    for (a = 0; a < XYZ_P1[0]; a++) for (b = 0; b < XYZ_P1[1]; b++) for (k = 0; k < XYZ_P1[2]; k++) {
        int n = 0;
        particleMPC *p = CL[a][b][k].pp;
        while (p) { n++; p = p->next; }
        fwrite(&n, sizeof(int), 1, o);
    }
    for (a = 0; a < XYZ_P1[0]; a++) for (b = 0; b < XYZ_P1[1]; b++) for (k = 0; k < XYZ_P1[2]; k++) {
        particleMPC *p = CL[a][b][k].pp;
        while (p) {
            int idx = (int)(p - pp);
            fwrite(&idx, sizeof(int), 1, o);
            p = p->next;
        }
    }
    for (a = 0; a < XYZ_P1[0]; a++) for (b = 0; b < XYZ_P1[1]; b++) for (k = 0; k < XYZ_P1[2]; k++) {
        cell *c = &CL[a][b][k];
        double v[7] = {c->S, c->DIR[0], c->DIR[1], c->DIR[2], c->VCM[0], c->VCM[1], c->VCM[2]};
        fwrite(v, sizeof(double), 7, o);
    }
}

int main(int argc, char **argv)
{
    hcase hc;
    int i, j, s;
    FILE *fs, *o = NULL;
    cell ***CL;
    particleMPC *pp;
    spec *SP;
    bc *WALL;
    inputList in;
    outputFlagsList of;
    outputFilesList ofl;
    specSwimmer SS;
    double AVNOW[3] = {0, 0, 0}, AVV[3], avDIR[3] = {0, 0, 0}, KBTNOW = 0.0, AVS = 0.0;
    int timing, flow, ckpt;

    if (argc < 4) die("usage: ref_harness CASE STATE OUT [timing]");
    timing = (argc > 4 && !strcmp(argv[4], "timing"));
    flow = (argc > 4 && !strcmp(argv[4], "flow"));
    ckpt = (argc > 4 && !strcmp(argv[4], "ckpt"));
    read_case(argv[1], &hc);

    DBUG = -1; /* below DBGRUN: silence the reference's diagnostic printing */
    DIM = hc.dim;
    MDmode = 0;
    NS = 0;
    for (i = 0; i < 3; i++) {
        XYZ[i] = (i < DIM) ? hc.L[i] : 1;
        XYZ_P1[i] = XYZ[i] + 1;
    }
    VOL = 1.0;
    for (i = 0; i < DIM; i++) VOL *= XYZ[i];
    GPOP = hc.N;
    NSPECI = hc.nspec;
    NBC = hc.nwall;

    SP = (spec *)calloc(NSPECI, sizeof(spec));
    for (i = 0; i < NSPECI; i++) {
        SP[i].MASS = hc.spv[i][0];
        SP[i].RFC = hc.spv[i][1];
        SP[i].TUMBLE = hc.spv[i][2];
        SP[i].CHIHI = hc.spv[i][3];
        SP[i].ACT = hc.spv[i][4];
        SP[i].MFPOT = hc.spv[i][5];
        SP[i].DAMP = hc.spv[i][6];
        SP[i].SIGWIDTH = hc.spv[i][7];
        SP[i].SIGPOS = hc.spv[i][8];
        SP[i].MINACTRATIO = hc.spv[i][9];
        SP[i].VDIST = (int)hc.spv[i][10];
        SP[i].LEN = hc.spv[i][11];
    }
    WALL = (bc *)calloc(NBC > 0 ? NBC : 1, sizeof(bc));
    for (i = 0; i < NBC; i++) {
        double *w = hc.wv[i];
        bc *W = &WALL[i];
        for (j = 0; j < 3; j++) {
            W->A[j] = w[j];
            W->AINV[j] = (w[j] != 0.0) ? 1.0 / w[j] : 0.0;
            W->Q[j] = w[3 + j];
            W->DVxyz[j] = w[11 + j];
            W->MUxyz[j] = w[18 + j];
            W->DUxyz[j] = w[21 + j];
            W->V[j] = w[24 + j];
        }
        W->R = w[6];
        W->DN = w[7];
        W->DT = w[8];
        W->DVN = w[9];
        W->DVT = w[10];
        W->MVN = w[14];
        W->MVT = w[15];
        W->MUN = w[16];
        W->MUT = w[17];
        W->INV = (int)w[27];
        W->PHANTOM = (int)w[28];
        W->COLL_TYPE = (int)w[29];
        W->E = -1.0;
        W->KBT = 1.0;
        W->MASS = 1.0;
        for (j = 0; j < 4; j++) W->P[j] = hc.has_shape[i] ? hc.shape[i][j] : 1.0;
        for (j = 0; j < 3; j++) W->B[j] = hc.has_shape[i] ? hc.shape[i][4 + j] : 0.0;
        W->ABS = hc.has_shape[i] ? (int)hc.shape[i][7] : 0;
        W->ROTSYMM[0] = W->ROTSYMM[1] = 4.0;
        for (j = 0; j < MAXSPECI + 2; j++) W->INTER[j] = BCON;
        /* planar flag as read.c:1423-1434 sets it; immobile, unrotated walls (init.c:2063-2065) */
        W->PLANAR = 0;
        if (feq(W->B[0], 0.0) && feq(W->P[0], 1.0) && feq(W->P[1], 1.0) && feq(W->P[2], 1.0)) {
            if (fneq(W->A[0], 0.0) && feq(W->A[1], 0.0) && feq(W->A[2], 0.0)) W->PLANAR = 1;
            else if (feq(W->A[0], 0.0) && fneq(W->A[1], 0.0) && feq(W->A[2], 0.0)) W->PLANAR = 1;
            else if (feq(W->A[0], 0.0) && feq(W->A[1], 0.0) && fneq(W->A[2], 0.0)) W->PLANAR = 1;
        }
        W->REORIENT = 0;
    }
    for (i = 0; i < 3; i++) XYZPBC[i] = 0;
    for (i = 0; i < NBC; i++) setPBC(&WALL[i]);

    CL = (cell ***)malloc(XYZ_P1[0] * sizeof(cell **));
    for (i = 0; i < XYZ_P1[0]; i++) {
        CL[i] = (cell **)malloc(XYZ_P1[1] * sizeof(cell *));
        for (j = 0; j < XYZ_P1[1]; j++) CL[i][j] = (cell *)malloc(XYZ_P1[2] * sizeof(cell));
    }
    pp = (particleMPC *)calloc(GPOP, sizeof(particleMPC));
    fs = fopen(argv[2], "rb");
    if (!fs) die("cannot open state");
    for (i = 0; i < GPOP; i++) {
        double x[9];
        if (fread(x, sizeof(double), 9, fs) != 9) die("short state");
        for (j = 0; j < 3; j++) {
            pp[i].Q[j] = x[j];
            pp[i].V[j] = x[3 + j];
            pp[i].U[j] = x[6 + j];
        }
        pp[i].S_flag = STREAM;
    }
    for (i = 0; i < GPOP; i++) {
        int sp;
        if (fread(&sp, sizeof(int), 1, fs) != 1) die("short spid");
        pp[i].SPID = sp;
    }
    fclose(fs);

    memset(&in, 0, sizeof(in));
    in.KBT = hc.kbt;
    in.dt = hc.dt;
    in.RA = hc.rotang;
    in.C = cos(in.RA);
    in.S = sin(in.RA);
    for (j = 0; j < 3; j++) in.GRAV[j] = hc.grav[j];
    in.GRAV_FLAG = (fneq(in.GRAV[0], 0.0) || fneq(in.GRAV[1], 0.0) || fneq(in.GRAV[2], 0.0)) ? 1 : 0;
    in.TSTECH = hc.tstech;
    in.TAU = hc.tau;
    in.RTECH = hc.rtech;
    in.LC = hc.lc;
    in.GALINV = hc.galinv;
    in.FRICCO = 1.0;
    in.stepsMD = 20;
    memset(&of, 0, sizeof(of));
    if (flow) of.FLOWOUT = OUT;
    memset(&ofl, 0, sizeof(ofl));
    memset(&SS, 0, sizeof(SS));
    for (j = 0; j < 3; j++) AVV[j] = hc.avv[j];

    ref_rand_seed((uint64_t)hc.seed);
    zerocell(CL);
    binin(pp, CL);
    localPROP(CL, SP, SS, in.RTECH, in.LC);
    globalDensities(CL, SP);

    if (!timing) {
        o = fopen(argv[3], "wb");
        if (!o) die("cannot open out");
        fwrite(&GnDNST, sizeof(double), 1, o);
        fwrite(XYZPBC, sizeof(int), 3, o);
        dump(o, -1, 0.0, pp, CL);
    }
    {
        struct timespec t0, t1;
        clock_gettime(CLOCK_MONOTONIC, &t0);
        for (s = 0; s < hc.steps; s++) {
            timestep(CL, pp, SP, WALL, NULL, &SS, NULL, AVNOW, AVV, avDIR, in, &KBTNOW, &AVS, s, 0, of, ofl);
            if (hc.zeronetmom > 0 && s % hc.zeronetmom == 0) {
                /* the rest-frame part of outputResults (pout.c:2540-2562) at runtime s */
                bin(CL, SP, WALL, in.KBT, in.LC, 0);
                localPROP(CL, SP, SS, in.RTECH, in.LC);
                galileantrans(pp, WALL, NULL, SP, in.KBT, AVV, GPOP, NBC, 0, DIM);
                zeroExtraDims(pp, WALL, NULL, GPOP, NBC, 0, DIM);
            }
            if (!timing) dump(o, s, KBTNOW, pp, CL);
        }
        clock_gettime(CLOCK_MONOTONIC, &t1);
        if (timing) {
            double sec = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
            printf("{\"steps\": %d, \"particles\": %d, \"seconds\": %.6f}\n", hc.steps, GPOP, sec);
        }
    }
    if (ckpt) {
        char path[4096];
        FILE *fc;
        kinTheory *th = (kinTheory *)calloc(NSPECI, sizeof(kinTheory)), thg;
        memset(&thg, 0, sizeof thg);
        snprintf(path, sizeof path, "%s.ckpt", argv[3]);
        fc = fopen(path, "w");
        if (!fc) die("cannot open ckpt");
        in.simSteps = hc.simsteps;
        in.warmupSteps = hc.warmup;
        in.seed = hc.seed;
        checkpoint(fc, in, SP, pp, 0, WALL, of, hc.steps - 1, 0, 0.0, 0.0, avDIR, 0.0, 0.0, KBTNOW, AVV, AVNOW, th, thg,
                   SS, NULL);
        fclose(fc);
        free(th);
    }
    if (o && flow) {
        int a, b, k, tag = -2;
        fwrite(&tag, sizeof(int), 1, o);
        for (a = 0; a < XYZ_P1[0]; a++) for (b = 0; b < XYZ_P1[1]; b++) for (k = 0; k < XYZ_P1[2]; k++)
            fwrite(CL[a][b][k].FLOW, sizeof(double), 3, o);
    }
    if (o) fclose(o);
    return 0;
}
