/*
 * checkpoint.c -- reads and writes the reference's checkpoint file (the
 * layout of checkpoint(), mpcd/subroutines/pout.c:2325-2404).
 *
 * The reader follows the order in which checkpoint() writes; the reference's
 * own reader takes the output-flag record in a different order
 * (read.c:722), so the writer's order is the file's.  Every value is read as
 * a double, which also parses the integers.
 */
#include "checkpoint.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    FILE *f;
    int ok;
} rd_t;

static double rd(rd_t *r) {
    double v = 0.0;
    if (r->ok && fscanf(r->f, "%lf", &v) != 1) r->ok = 0;
    return v;
}

static void rdn(rd_t *r, double *v, int n) {
    for (int k = 0; k < n; k++) v[k] = rd(r);
}

static void skip(rd_t *r, int n) {
    for (int k = 0; k < n; k++) (void)rd(r);
}

#define FAIL(...)                              \
    do {                                       \
        snprintf(err, errlen, __VA_ARGS__);    \
        if (f) fclose(f);                      \
        ckpt_free(ck);                         \
        return -1;                             \
    } while (0)

static const char kBinMagic[8] = {'M', 'P', 'C', 'D', 'C', 'K', 'B', '1'};

void ckpt_write_bin(FILE *f, const mpcd_gpu_params *p, const mpcd_ckpt *ck) {
    const int32_t iv[11] = {ck->sim_steps, ck->warm_steps, ck->runtime, ck->warmtime, ck->flow_out, ck->vel_out,
                            ck->dir_out,   ck->ckpt_out,   ck->dens_out, ck->rframe, ck->zero_net_mom};
    const int64_t seed = ck->seed;
    fwrite(kBinMagic, 1, sizeof kBinMagic, f);
    fwrite(p, sizeof *p, 1, f);
    fwrite(iv, sizeof iv, 1, f);
    fwrite(&seed, sizeof seed, 1, f);
    fwrite(&ck->kbtnow, sizeof(double), 1, f);
    fwrite(ck->avnow, sizeof(double), 3, f);
    fwrite(&ck->lib_step, sizeof(uint32_t), 1, f);
    fwrite(&ck->n, sizeof(int64_t), 1, f);
    fwrite(ck->Q, sizeof(double), 3 * (size_t)ck->n, f);
    fwrite(ck->V, sizeof(double), 3 * (size_t)ck->n, f);
    fwrite(ck->U, sizeof(double), 3 * (size_t)ck->n, f);
    fwrite(ck->spid, sizeof(int32_t), (size_t)ck->n, f);
}

static int ckpt_read_bin(FILE *f, mpcd_gpu_params *p, mpcd_ckpt *ck, char *err, size_t errlen) {
    int32_t iv[11];
    int64_t seed;
    int ok = fread(p, sizeof *p, 1, f) == 1 && fread(iv, sizeof iv, 1, f) == 1 && fread(&seed, sizeof seed, 1, f) == 1 &&
             fread(&ck->kbtnow, sizeof(double), 1, f) == 1 && fread(ck->avnow, sizeof(double), 3, f) == 3 &&
             fread(&ck->lib_step, sizeof(uint32_t), 1, f) == 1 && fread(&ck->n, sizeof(int64_t), 1, f) == 1;
    if (!ok || ck->n <= 0 || ck->n >= (1ll << 31) || p->nspec < 1 || p->nspec > MPCD_MAXSPEC || p->nwall < 0 ||
        p->nwall > MPCD_MAXWALL) {
        snprintf(err, errlen, "binary checkpoint: the header is cut short or out of range");
        return -1;
    }
    ck->sim_steps = iv[0];
    ck->warm_steps = iv[1];
    ck->runtime = iv[2];
    ck->warmtime = iv[3];
    ck->flow_out = iv[4];
    ck->vel_out = iv[5];
    ck->dir_out = iv[6];
    ck->ckpt_out = iv[7];
    ck->dens_out = iv[8];
    ck->rframe = iv[9];
    ck->zero_net_mom = iv[10];
    ck->seed = (long)seed;
    ck->binary = 1;
    const size_t n = (size_t)ck->n;
    ck->Q = (double *)calloc(3 * n, sizeof(double));
    ck->V = (double *)calloc(3 * n, sizeof(double));
    ck->U = (double *)calloc(3 * n, sizeof(double));
    ck->spid = (int32_t *)calloc(n, sizeof(int32_t));
    if (!ck->Q || !ck->V || !ck->U || !ck->spid) {
        snprintf(err, errlen, "binary checkpoint: out of host memory");
        return -1;
    }
    if (fread(ck->Q, sizeof(double), 3 * n, f) != 3 * n || fread(ck->V, sizeof(double), 3 * n, f) != 3 * n ||
        fread(ck->U, sizeof(double), 3 * n, f) != 3 * n || fread(ck->spid, sizeof(int32_t), n, f) != n) {
        snprintf(err, errlen, "binary checkpoint: the particle arrays are cut short");
        return -1;
    }
    for (size_t i = 0; i < n; i++)
        if (ck->spid[i] < 0 || ck->spid[i] >= p->nspec) {
            snprintf(err, errlen, "binary checkpoint: particle %zu has species %d", i, (int)ck->spid[i]);
            return -1;
        }
    return 0;
}

int ckpt_read(const char *path, mpcd_gpu_params *p, mpcd_ckpt *ck, char *err, size_t errlen) {
    memset(ck, 0, sizeof *ck);
    FILE *f = fopen(path, "rb");
    if (!f) FAIL("cannot open the checkpoint %s", path);
    char magic[8] = {0};
    if (fread(magic, 1, sizeof magic, f) == sizeof magic && !memcmp(magic, kBinMagic, sizeof magic)) {
        const int r = ckpt_read_bin(f, p, ck, err, errlen);
        fclose(f);
        return r;
    }
    rewind(f);
    rd_t R = {f, 1};
    double v[32];
    memset(p, 0, sizeof *p);
    ck->sim_steps = (int)rd(&R);
    ck->warm_steps = (int)rd(&R);
    p->dt = rd(&R);
    ck->seed = (long)rd(&R);
    rdn(&R, v, 6);  /* DIM, XYZ[3], KBT, KBTNOW */
    p->dim = (int)v[0];
    for (int d = 0; d < 3; d++) p->L[d] = (int)v[1 + d];
    p->kbt = v[4];
    ck->kbtnow = v[5];
    rdn(&R, v, 6);  /* RFRAME, zeroNetMom, GALINV, TSTECH, RTECH, LC */
    ck->rframe = (int)v[0];
    ck->zero_net_mom = (int)v[1];
    p->galinv = (int)v[2];
    p->tstech = (int)v[3];
    p->rtech = (int)v[4];
    p->lc = (int)v[5];
    rdn(&R, v, 3);  /* TAU, RA, FRICCO */
    p->tau = v[0];
    p->rotang = v[1];
    rdn(&R, v, 4);  /* noHI, inCOMP, MULTIPHASE, MFPLAYERH */
    if (R.ok && (v[0] != 0.0 || v[1] != 0.0 || v[2] != 0.0)) FAIL("checkpoint: noHI / inCOMP / multiphase runs are outside this path");
    rdn(&R, p->grav, 3);
    skip(&R, 3);  /* MAG */
    rdn(&R, v, 2);  /* MD mode, stepsMD */
    if (R.ok && v[0] != 0.0) FAIL("checkpoint: MD coupling is outside this path");
    rdn(&R, v, 2);  /* GPOP, NSPECI */
    ck->n = (int64_t)v[0];
    p->nspec = (int)v[1];
    if (R.ok && (p->nspec < 1 || p->nspec > MPCD_MAXSPEC)) FAIL("checkpoint: %d species", p->nspec);
    if (R.ok && (p->dim < 2 || p->dim > 3 || ck->n <= 0)) FAIL("checkpoint: dimension %d, %lld particles", p->dim, (long long)ck->n);
    rdn(&R, v, 7);  /* runtime, warmtime, C, S, GRAV_FLAG, MAG_FLAG, tolD */
    ck->runtime = (int)v[0];
    ck->warmtime = (int)v[1];
    skip(&R, 8);  /* AVVEL, AVS, avDIR[3], S4, stdN, VOL */
    rdn(&R, p->avv, 3);
    rdn(&R, ck->avnow, 3);
    skip(&R, 6);  /* theory (global) */
    skip(&R, 2);  /* chckpntIn, MDmode */
    rdn(&R, v, 32);  /* output flags, in checkpoint()'s order */
    ck->flow_out = (int)v[4];
    ck->vel_out = (int)v[6];
    ck->dens_out = (int)v[7];
    ck->dir_out = (int)v[10];
    ck->ckpt_out = (int)v[29];
    skip(&R, 2 + 7 + 5 + 3);  /* SPOUT PRESOUT, histograms, spectra/topology, swimmer outputs */
    for (int s = 0; s < p->nspec && R.ok; s++) {
        mpcd_gpu_species *sp = &p->sp[s];
        rdn(&R, v, 5);  /* MASS, POP, QDIST, VDIST, ODIST */
        sp->mass = v[0];
        sp->vdist = (int)v[3];
        rdn(&R, v, 10);  /* RFC, LEN, TUMBLE, CHIHI, CHIA, ACT, MFPOT, SIGWIDTH, SIGPOS, DAMP */
        sp->rfc = v[0];
        sp->tumble = v[2];
        sp->chihi = v[3];
        sp->act = v[5];
        sp->mfpot = v[6];
        sp->sigwidth = v[7];
        sp->sigpos = v[8];
        sp->damp = v[9];
        rdn(&R, v, 5);  /* VOL, nDNST, mDNST, MINACTRATIO, BS */
        sp->minactratio = v[3];
        skip(&R, p->nspec + 6);  /* binary-fluid M[], species theory */
    }
    rdn(&R, v, 4);  /* GnDNST, GmDNST, GMASS, maxXYZ */
    p->gndnst = v[0];
    p->nwall = (int)rd(&R);
    if (R.ok && (p->nwall < 0 || p->nwall > MPCD_MAXWALL)) FAIL("checkpoint: %d boundaries", p->nwall);
    for (int w = 0; w < p->nwall && R.ok; w++) {
        mpcd_gpu_wall *W = &p->wall[w];
        rdn(&R, v, 12);  /* COLL_TYPE, PHANTOM, E, Q[3], V[3], O[3] */
        W->COLL_TYPE = (int)v[0];
        W->PHANTOM = (int)v[1];
        memcpy(W->Q, v + 3, sizeof W->Q);
        memcpy(W->V, v + 6, sizeof W->V);
        const double O[3] = {v[9], v[10], v[11]};
        rdn(&R, v, 20);  /* L[3], G[3], A[3], AINV[3], P[4], R, B[3] */
        memcpy(W->L, v, sizeof W->L);
        memcpy(W->A, v + 6, sizeof W->A);
        memcpy(W->P, v + 12, sizeof W->P);
        W->R = v[16];
        memcpy(W->B, v + 17, sizeof W->B);
        const int turned = fabs(O[0]) + fabs(O[1]) + fabs(O[2]) > 0.0;
        rdn(&R, v, 14);  /* DN, DT, DVN, DVT, DVxyz[3], MVN, MVT, MUN, MUT, MUxyz[3] */
        W->DN = v[0];
        W->DT = v[1];
        W->DVN = v[2];
        W->DVT = v[3];
        memcpy(W->DVxyz, v + 4, sizeof W->DVxyz);
        W->MVN = v[7];
        W->MVT = v[8];
        W->MUN = v[9];
        W->MUT = v[10];
        memcpy(W->MUxyz, v + 11, sizeof W->MUxyz);
        rdn(&R, v, 7);  /* DUxyz[3], KBT, DSPLC, INV, MASS */
        memcpy(W->DUxyz, v, sizeof W->DUxyz);
        const int dsplc = (int)v[4];
        W->INV = (int)v[5];
        skip(&R, 17);  /* W, VOL, Q_old, O_old, I */
        rdn(&R, v, 5); /* PLANAR, REORIENT, ABS, ROTSYMM */
        W->ABS = (int)v[2];
        const int sym4 = fabs(v[3] - 4.0) <= 1e-8 && fabs(v[4] - 4.0) <= 1e-8;
        skip(&R, 6);   /* dV, dL */
        rdn(&R, v, CKPT_MAXSPECI + 2);
        W->inter_mask = 0;
        for (int s = 0; s < p->nspec; s++)
            if (v[s] == 1.0) W->inter_mask |= 1 << s;  /* BCON = 1: interacts (bc.c:1263) */
        if (R.ok && (turned || dsplc || !sym4))
            FAIL("checkpoint: boundary %d is turned, displaceable or not four-fold symmetric (outside this path)", w);
    }
    if (!R.ok) FAIL("checkpoint %s: the file ends or breaks before the particles", path);
    ck->Q = (double *)calloc(3 * (size_t)ck->n, sizeof(double));
    ck->V = (double *)calloc(3 * (size_t)ck->n, sizeof(double));
    ck->U = (double *)calloc(3 * (size_t)ck->n, sizeof(double));
    ck->spid = (int32_t *)calloc((size_t)ck->n, sizeof(int32_t));
    if (!ck->Q || !ck->V || !ck->U || !ck->spid) FAIL("checkpoint: out of host memory");
    for (int64_t i = 0; i < ck->n && R.ok; i++) {
        rdn(&R, v, 15);  /* S_flag, SPID, q, Q[3], V[3], U[3], T[3] */
        ck->spid[i] = (int32_t)v[1];
        memcpy(ck->Q + 3 * i, v + 3, 3 * sizeof(double));
        memcpy(ck->V + 3 * i, v + 6, 3 * sizeof(double));
        memcpy(ck->U + 3 * i, v + 9, 3 * sizeof(double));
        if (ck->spid[i] < 0 || ck->spid[i] >= p->nspec) FAIL("checkpoint: particle %lld of species %d", (long long)i, ck->spid[i]);
    }
    const int ns = (int)rd(&R);
    if (!R.ok) FAIL("checkpoint %s: the file ends or breaks in the particles", path);
    if (ns != 0) FAIL("checkpoint: swimmers are outside this path");
    fclose(f);
    return 0;
}

void ckpt_write(FILE *f, const mpcd_gpu_params *p, const mpcd_ckpt *ck) {
    const int dim = p->dim;
    const int X = p->L[0], Y = p->L[1], Z = dim == 3 ? p->L[2] : 1;
    const double vol = (double)X * Y * Z;
    int64_t pop[MPCD_MAXSPEC] = {0};
    double gmass = 0.0;
    for (int64_t i = 0; i < ck->n; i++) {
        pop[ck->spid[i]]++;
        gmass += p->sp[ck->spid[i]].mass;
    }
    int grav_flag = 0;
    for (int d = 0; d < 3; d++) grav_flag |= p->grav[d] != 0.0;
    fprintf(f, "%d\n", ck->sim_steps);
    fprintf(f, "%d %lf\n", ck->warm_steps, p->dt);
    fprintf(f, "%ld\n", ck->seed);
    fprintf(f, "%d %d %d %d %lf %lf\n", dim, X, Y, Z, p->kbt, ck->kbtnow);
    fprintf(f, "%d %d %d %d %d %d\n", ck->rframe, ck->zero_net_mom, p->galinv, p->tstech, p->rtech, p->lc);
    fprintf(f, "%lf %lf %lf\n", p->tau, p->rotang, 1.0);
    fprintf(f, "%d %d %d %d\n", 0, 0, 0, 0);
    fprintf(f, "%lf %lf %lf\n", p->grav[0], p->grav[1], p->grav[2]);
    fprintf(f, "%lf %lf %lf\n", 0.0, 0.0, 0.0);
    fprintf(f, "%d %d\n", 0, 20);
    fprintf(f, "%lld %d\n", (long long)ck->n, p->nspec);
    fprintf(f, "%d %d %lf %lf %d %d %lf\n", ck->runtime, ck->warmtime, cos(p->rotang), sin(p->rotang), grav_flag, 0, 0.0);
    fprintf(f, "%lf %lf %lf %lf %lf %lf %lf %lf\n", 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, vol);
    fprintf(f, "%lf %lf %lf %lf %lf %lf\n", p->avv[0], p->avv[1], p->avv[2], ck->avnow[0], ck->avnow[1], ck->avnow[2]);
    fprintf(f, "%lf %lf %lf %lf %lf %lf\n", 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
    fprintf(f, "%d %d\n", 0, 0);
    /* output flags in checkpoint()'s order: DBUG, TRAJOUT, printSP, COAROUT,
     * FLOWOUT, SWFLOWOUT, VELOUT, DENSITYOUT, AVVELOUT, AVORIOUT, ORDEROUT,
     * ..., CHCKPNT (30th), CHCKPNTrcvr, CHCKPNTTIMER */
    int fl[31] = {0};
    fl[0] = -1;
    fl[4] = ck->flow_out;
    fl[6] = ck->vel_out;
    fl[7] = ck->dens_out;
    fl[10] = ck->dir_out;
    fl[29] = ck->ckpt_out;
    for (int k = 0; k < 31; k++) fprintf(f, "%d ", fl[k]);
    fprintf(f, "%f\n", 0.0);
    fprintf(f, "%d %d\n", 0, 0);
    fprintf(f, "%d %d %d %d %d %d %d\n", 0, 0, 0, 0, 0, 0, 0);
    fprintf(f, "%d %d %d %d %d\n", 0, 0, 0, 0, 0);
    fprintf(f, "%d %d %d\n", 0, 0, 0);
    for (int s = 0; s < p->nspec; s++) {
        const mpcd_gpu_species *sp = &p->sp[s];
        const double n_dens = (double)pop[s] / vol;
        fprintf(f, "%lf %i %i %i %i ", sp->mass, (int)pop[s], 0, sp->vdist, 0);
        fprintf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf ", sp->rfc, 0.007, sp->tumble, sp->chihi, 0.0, sp->act,
                sp->mfpot, sp->sigwidth, sp->sigpos, sp->damp);
        fprintf(f, "%lf %lf %lf %lf %lf\n", vol, n_dens, n_dens * sp->mass, sp->minactratio, 0.0);
        for (int j = 0; j < p->nspec; j++) fprintf(f, "%lf ", 0.0);
        fprintf(f, "\n");
        fprintf(f, "%lf %lf %lf %lf %lf %lf\n", 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
    }
    const int maxxyz = X > Y ? (X > Z ? X : Z) : (Y > Z ? Y : Z);
    fprintf(f, "%lf %lf %lf %d\n", p->gndnst, ck->n > 0 ? p->gndnst * gmass / (double)ck->n : 0.0, gmass, maxxyz);
    fprintf(f, "%d\n", p->nwall);
    for (int w = 0; w < p->nwall; w++) {
        const mpcd_gpu_wall *W = &p->wall[w];
        double ainv[3];
        for (int d = 0; d < 3; d++) ainv[d] = W->A[d] != 0.0 ? 1.0 / W->A[d] : 0.0;
        fprintf(f, "%d %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", W->COLL_TYPE, W->PHANTOM, -1.0, W->Q[0], W->Q[1],
                W->Q[2], W->V[0], W->V[1], W->V[2], 0.0, 0.0, 0.0);
        fprintf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", W->L[0], W->L[1],
                W->L[2], 0.0, 0.0, 0.0, W->A[0], W->A[1], W->A[2], ainv[0], ainv[1], ainv[2], W->P[0], W->P[1],
                W->P[2], W->P[3], W->R, W->B[0], W->B[1], W->B[2]);
        fprintf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", W->DN, W->DT, W->DVN, W->DVT,
                W->DVxyz[0], W->DVxyz[1], W->DVxyz[2], W->MVN, W->MVT, W->MUN, W->MUT, W->MUxyz[0], W->MUxyz[1],
                W->MUxyz[2]);
        fprintf(f, "%lf %lf %lf %lf %d %d %lf\n", W->DUxyz[0], W->DUxyz[1], W->DUxyz[2], 1.0, 0, W->INV, 1.0);
        for (int k = 0; k < 17; k++) fprintf(f, k ? " %lf" : "%lf", 0.0);  /* W, VOL, Q_old, O_old, I */
        fprintf(f, "\n");
        /* PLANAR (read.c:1423-1434), REORIENT, ABS, ROTSYMM */
        int naxis = 0;
        for (int d = 0; d < 3; d++) naxis += fabs(W->A[d]) >= 1e-8;
        const int plane = naxis == 1 && fabs(W->B[0]) <= 1e-8 && fabs(W->P[0] - 1) <= 1e-8 &&
                          fabs(W->P[1] - 1) <= 1e-8 && fabs(W->P[2] - 1) <= 1e-8;
        fprintf(f, "%d %d %d %lf %lf\n", plane, 0, W->ABS, 4.0, 4.0);
        fprintf(f, "%lf %lf %lf %lf %lf %lf\n", 0.0, 0.0, 0.0, 0.0, 0.0, 0.0);
        /* species slots past nspec, MD and swimmers keep the default BCON = 1 (read.c:1375) */
        for (int j = 0; j < CKPT_MAXSPECI + 2; j++) fprintf(f, "%i ", j < p->nspec ? (W->inter_mask >> j) & 1 : 1);
        fprintf(f, "\n");
    }
    for (int64_t i = 0; i < ck->n; i++) {
        const double *q = ck->Q + 3 * i, *v = ck->V + 3 * i, *u = ck->U + 3 * i;
        fprintf(f, "%d %d %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", 1, ck->spid[i], 0.0, q[0], q[1], q[2],
                v[0], v[1], v[2], u[0], u[1], u[2], 0.0, 0.0, 0.0);
    }
    fprintf(f, "%d %d %d %d %d %d %lf %lf %d %d\n", 0, 0, 0, 0, 0, 0, 0.0, 0.0, 0, 0);
    fprintf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %lf %d %lf %lf %lf %lf %d %lf\n", 0.0, 0.0, 0.0, 0.0, 0.0,
            0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0, 0.0, 0.0, 0, 0.0);
}

void ckpt_free(mpcd_ckpt *ck) {
    free(ck->Q);
    free(ck->V);
    free(ck->U);
    free(ck->spid);
    ck->Q = ck->V = ck->U = NULL;
    ck->spid = NULL;
}
