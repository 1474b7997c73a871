/* exact_step.c -- TEST INFRASTRUCTURE (oracle): a sequential C restatement
 * of one step of the library's exact-sum FP32 form (precision
 * MPCD_FP32_EXACT) for one replica whose walls are the periodic faces of its
 * box.  Only tests/ use it, as the checker of the GPU kernels; nothing in the
 * product links it.
 *
 * The form follows the reference's step (mpc.c:4567-4942) with the
 * reference's operators -- grid shift (gridShift_all, mpc.c:720-766), body
 * force (acc_all, mpc.c:697), cell sums (localPROP, mpc.c:64-216), SRD
 * rotation (rotate/stochrotMPC, mpc.c:1181-1333), MPC-AT (andersenMPC,
 * mpc.c:1353-1444) and MPC-AT+a (andersenROT, mpc.c:1461-1657), TEMP
 * (therm.c:209-235), stream_all (mpc.c:680) -- but in FP32 particle
 * arithmetic, with cell sums of terms quantised to 2^-28 (exact in FP64, so
 * the particles here are added in index order and the result is that of
 * any order), and with the counter-based generator keyed by cell (SRD axis)
 * and by particle (MPC-AT normals).  On a box of periodic faces a drifted
 * particle wraps by the box length.
 *
 * Every operation is written as the kernels write it (mpcd_exact.h), in the
 * same order, so that the comparison is bit for bit.  Build: gcc with
 * -ffp-contract=off (no fused multiply-adds), as the library is built.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mpcd_oracle.h"

#define XQ 268435456.0f
#define XQD 268435456.0
#define XQINV (1.0 / 268435456.0)
#define XP_SHIFT 1
#define XP_CELL 2
#define XP_NORMAL 4

typedef struct {
    int dim, L[3], rtech, true_sign, galinv, nspec;
    int cell_off; /* first cell of the replica in the context (the draws' cell key) */
    double dt, kbt, cosA, sinA, grav[3], avv[3];
    double mass[10], damp[10];
} orc_xcfg;

static double xq(float x) { return (double)rintf(x * XQ); }
static double xqd(double x) { return rint(x * XQD); }

static int xcell(const orc_xcfg *p, const float q[3])
{
    int a = (int)q[0], b = (int)q[1], c = (p->dim == 3) ? (int)q[2] : 0;
    const int Lz = (p->dim == 3) ? p->L[2] : 1, ny = p->L[1] + 1, nz = (p->dim == 3) ? p->L[2] + 1 : 1;
    /* shifted bin on periodic axes: a == L wraps to 0 (mpc.c:964-968) */
    if (a == p->L[0]) a = 0;
    if (b == p->L[1]) b = 0;
    if (p->dim == 3 && c == Lz) c = 0;
    if (a > p->L[0] || b > p->L[1] || a < 0 || b < 0 || (p->dim == 3 && (c > Lz || c < 0))) return -1;
    return (a * ny + b) * nz + c;
}

/* one step; x, v: [3 n] FP32 (unshifted positions), sp: species (NULL: 0),
 * ids: the particles' ids, the key of their draws (NULL: the index).
 * Returns the number of particles outside the box (none on this path) and
 * writes KBTNOW. */
int orc_exact_step(const orc_xcfg *p, int n, float *x, float *v, const int *sp, const int *ids, uint64_t seed,
                   uint32_t replica, uint32_t step, double *kbt_out)
{
    const int dim = p->dim, op = (p->rtech == 2) ? 1 : (p->rtech == 3) ? 2 : 0;
    const int ns = op == 0 ? 3 : op == 1 ? 6 : 18;
    const int ny = p->L[1] + 1, nz = (dim == 3) ? p->L[2] + 1 : 1, nc = (p->L[0] + 1) * ny * nz;
    const int onesp = p->nspec == 1;
    float s[3] = {0.0f, 0.0f, 0.0f}, L[3], dtg[3], avv[3];
    int bad = 0, gravflag = 0;
    for (int j = 0; j < 3; j++) {
        if (p->galinv && j < dim) s[j] = (float)orc_philox_uniform(seed, replica, step, XP_SHIFT, 0, (uint32_t)j);
        L[j] = (float)p->L[j];
        dtg[j] = (float)(p->dt * p->grav[j]);
        avv[j] = (float)p->avv[j];
        if (fabs(p->grav[j]) >= 1e-8) gravflag = 1;
    }
    int *cell = (int *)malloc(sizeof(int) * (size_t)n), *cnt = (int *)calloc((size_t)nc, sizeof(int));
    int *nsp = (int *)calloc((size_t)nc * 10, sizeof(int));
    double *sum = (double *)calloc((size_t)nc * (size_t)ns, sizeof(double));
    float *cd = (float *)calloc((size_t)nc * 12, sizeof(float)), *rt = (float *)calloc((size_t)n * 3, sizeof(float));
    /* 1. shift, body force, thermal velocities, quantised sums */
    for (int i = 0; i < n; i++) {
        float *q = &x[3 * i], *w = &v[3 * i];
        const int spid = sp ? sp[i] : 0;
        const double m = p->mass[spid];
        for (int j = 0; j < dim; j++) {
            const float t = q[j] + s[j];
            q[j] = (t > L[j] && t <= L[j] + s[j]) ? t - L[j] : t;
        }
        const int c = xcell(p, q);
        cell[i] = c;
        if (c < 0) {
            bad++;
            continue;
        }
        float rv[3] = {0.0f, 0.0f, 0.0f};
        if (op != 0) {
            const double sig = sqrt(p->kbt / m);
            const uint32_t key = (uint32_t)(ids ? ids[i] : i);
            rv[0] = (float)(sig * orc_philox_normal_turn(seed, replica, step, XP_NORMAL, key, 0));
            rv[1] = (float)(sig * orc_philox_normal_turn(seed, replica, step, XP_NORMAL, key, 1));
            rv[2] = (dim == 3) ? (float)(sig * orc_philox_normal_turn(seed, replica, step, XP_NORMAL, key, 2)) : 0.0f;
            memcpy(&rt[3 * i], rv, sizeof rv);
        }
        const double mm = onesp ? 1.0 : m;
        double t[18];
        for (int d = 0; d < 3; d++) t[d] = (mm == 1.0) ? xq(w[d]) : xqd(mm * (double)w[d]);
        if (op != 0)
            for (int d = 0; d < 3; d++) t[3 + d] = xqd(mm * (double)rv[d]);
        if (op == 2) {
            const int ca = c / (ny * nz), cb = (c / nz) % ny, cc = c % nz;
            const float r[3] = {q[0] - (float)ca, q[1] - (float)cb, (dim == 3) ? q[2] - (float)cc : 0.0f};
            const double rd[3] = {r[0], r[1], r[2]};
            for (int d = 0; d < 3; d++) t[6 + d] = xqd(mm * rd[d]);
            t[9] = xqd(mm * (rd[0] * rd[0]));
            t[10] = xqd(mm * (rd[1] * rd[1]));
            t[11] = xqd(mm * (rd[2] * rd[2]));
            t[12] = xqd(mm * (rd[0] * rd[1]));
            t[13] = xqd(mm * (rd[0] * rd[2]));
            t[14] = xqd(mm * (rd[1] * rd[2]));
            const double dv[3] = {mm * ((double)w[0] - (double)rv[0]), mm * ((double)w[1] - (double)rv[1]),
                                  mm * ((double)w[2] - (double)rv[2])};
            t[15] = xqd(rd[1] * dv[2] - rd[2] * dv[1]);
            t[16] = xqd(rd[2] * dv[0] - rd[0] * dv[2]);
            t[17] = xqd(rd[0] * dv[1] - rd[1] * dv[0]);
        }
        cnt[c]++;
        nsp[c * 10 + spid]++;
        for (int d = 0; d < ns; d++) sum[(size_t)c * ns + d] += t[d];
    }
    /* 2. the collision data of each cell */
    for (int c = 0; c < nc; c++) {
        const int np_ = cnt[c];
        if (np_ == 0) continue;
        double mass = 0.0, sm[18], vcm[3] = {0.0, 0.0, 0.0};
        float *o = &cd[(size_t)c * 12];
        if (onesp) mass = (double)np_ * p->mass[0];
        else
            for (int k = 0; k < p->nspec; k++) mass += (double)nsp[c * 10 + k] * p->mass[k];
        for (int j = 0; j < ns; j++) sm[j] = sum[(size_t)c * ns + j];
        if (onesp)
            for (int j = 0; j < (op == 0 ? 3 : ns); j++) sm[j] *= p->mass[0];
        const double im = XQINV / mass; /* mass > 0: the cell holds a particle */
        /* the body force's kick enters once per cell, in FP64: o[0..2] the
           kicked centre-of-mass velocity (SRD: the offset b of v' = M v + b) */
        double g[3] = {0.0, 0.0, 0.0};
        if (gravflag)
            for (int d = 0; d < dim; d++) g[d] = p->dt * p->grav[d];
        for (int d = 0; d < dim; d++) vcm[d] = sm[d] * im + g[d];
        for (int d = 0; d < 3; d++) o[d] = (float)vcm[d];
        if (np_ <= 1) continue;
        if (op == 0) {
            double RV[3] = {0.0, 0.0, 0.0}, M[9];
            int CA = 0;
            if (p->rtech == 0) {
                if (dim == 3) orc_philox_unitvec_turn(seed, replica, step, XP_CELL, (uint32_t)(c + p->cell_off), 0, RV);
            } else {
                CA = (int)((int)dim * orc_philox_uniform(seed, replica, step, XP_CELL, (uint32_t)(c + p->cell_off), 0));
            }
            const int pos = p->true_sign ? (orc_philox_uniform(seed, replica, step, XP_CELL, (uint32_t)(c + p->cell_off), 3) >= 0.5) : 0;
            const double Cc = p->cosA, Sn = pos ? p->sinA : -1. * p->sinA;
            if (dim == 3) {
                const double *l = RV;
                if (p->rtech == 0) {
                    M[0] = l[0] * l[0] + (1. - l[0] * l[0]) * Cc;
                    M[1] = l[0] * l[1] * (1. - Cc) - l[2] * Sn;
                    M[2] = l[0] * l[2] * (1. - Cc) + l[1] * Sn;
                    M[3] = l[0] * l[1] * (1. - Cc) + l[2] * Sn;
                    M[4] = l[1] * l[1] + (1. - l[1] * l[1]) * Cc;
                    M[5] = l[1] * l[2] * (1. - Cc) - l[0] * Sn;
                    M[6] = l[0] * l[2] * (1. - Cc) - l[1] * Sn;
                    M[7] = l[1] * l[2] * (1. - Cc) + l[0] * Sn;
                    M[8] = l[2] * l[2] + (1. - l[2] * l[2]) * Cc;
                } else if (CA == 0) {
                    double m0[9] = {1., 0., 0., 0., Cc, Sn, 0., -.1 * Sn, Cc};
                    memcpy(M, m0, sizeof m0);
                } else if (CA == 1) {
                    double m1[9] = {Cc, 0., Sn, 0., 1., 0., -1. * Sn, 0., Cc};
                    memcpy(M, m1, sizeof m1);
                } else {
                    double m2[9] = {Cc, Sn, 0., -1. * Sn, Cc, 0., 0., 0., 1.};
                    memcpy(M, m2, sizeof m2);
                }
            } else {
                double m3[9] = {Cc, -1. * Sn, 0., Sn, Cc, 0., 0., 0., 0.};
                memcpy(M, m3, sizeof m3);
            }
            for (int k = 0; k < 9; k++) o[3 + k] = (float)M[k];
            /* v' = vcm + M (v + g dt - vcm) = M v + b with the unkicked v */
            for (int d = 0; d < 3; d++) {
                double b = vcm[d];
                for (int k = 0; k < 3; k++) b -= M[3 * d + k] * (vcm[k] - g[k]);
                o[d] = (float)b;
            }
            continue;
        }
        for (int d = 0; d < 3; d++) o[3 + d] = (float)((d < dim) ? sm[3 + d] * im : 0.0);
        if (op != 2) continue;
        double cm[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < dim; d++) cm[d] = sm[6 + d] * im;
        const double Sxx = sm[9] * XQINV - mass * cm[0] * cm[0], Syy = sm[10] * XQINV - mass * cm[1] * cm[1];
        const double Szz = sm[11] * XQINV - mass * cm[2] * cm[2];
        const double Sxy = sm[12] * XQINV - mass * cm[0] * cm[1], Sxz = sm[13] * XQINV - mass * cm[0] * cm[2];
        const double Syz = sm[14] * XQINV - mass * cm[1] * cm[2];
        const double I[6] = {Syy + Szz, Sxx + Szz, Sxx + Syy, -Sxy, -Sxz, -Syz};
        const double P[3] = {(sm[0] - sm[3]) * XQINV, (sm[1] - sm[4]) * XQINV, (sm[2] - sm[5]) * XQINV};
        double Lc[3] = {sm[15] * XQINV, sm[16] * XQINV, sm[17] * XQINV};
        Lc[0] -= cm[1] * P[2] - cm[2] * P[1];
        Lc[1] -= cm[2] * P[0] - cm[0] * P[2];
        Lc[2] -= cm[0] * P[1] - cm[1] * P[0];
        double II[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (dim == 3) {
            if (np_ >= 3) {
                /* invert3x3 (mtools.c:1580-1612): adjugate over the determinant */
                const double a[3][3] = {{I[0], I[3], I[4]}, {I[3], I[1], I[5]}, {I[4], I[5], I[2]}};
                orc_inv3(II, a);
            }
        } else {
            II[2][2] = 1. / I[2];
        }
        for (int i = 0; i < 3; i++) {
            double w = 0.;
            for (int j = 0; j < 3; j++) w += II[i][j] * Lc[j];
            o[6 + i] = (float)w;
            o[9 + i] = (float)cm[i];
        }
    }
    /* 3. collision, kinetic energy, shift back, drift with the periodic wrap */
    long long ke = 0;
    const float dt = (float)p->dt;
    for (int i = 0; i < n; i++) {
        const int c = cell[i];
        if (c < 0) continue;
        float *q = &x[3 * i], *w = &v[3 * i];
        const int spid = sp ? sp[i] : 0;
        const int np_ = cnt[c];
        if (np_ > 1) {
            const float *o = &cd[(size_t)c * 12];
            const float vc[3] = {o[0], o[1], o[2]};
            if (op == 0) {
                /* v' = M v + b */
                const float *M = o + 3;
                const float d0 = w[0], d1 = w[1], d2 = w[2];
                if (dim == 3) {
                    float t0 = M[0] * d0;
                    t0 += M[1] * d1;
                    t0 += M[2] * d2;
                    float t1 = M[3] * d0;
                    t1 += M[4] * d1;
                    t1 += M[5] * d2;
                    float t2 = M[6] * d0;
                    t2 += M[7] * d1;
                    t2 += M[8] * d2;
                    w[0] = vc[0] + t0;
                    w[1] = vc[1] + t1;
                    w[2] = vc[2] + t2;
                } else {
                    float t0 = M[0] * d0;
                    t0 += M[1] * d1;
                    float t1 = M[3] * d0;
                    t1 += M[4] * d1;
                    w[0] = vc[0] + t0;
                    w[1] = vc[1] + t1;
                }
            } else if (op == 1) {
                const float inv = 1.0f / (float)np_, damp = (float)p->damp[spid];
                for (int d = 0; d < dim; d++) w[d] = vc[d] + rt[3 * i + d] - o[3 + d] - damp * vc[d] * inv;
            } else {
                const int ca = c / (ny * nz), cb = (c / nz) % ny, cc = c % nz;
                const float r[3] = {q[0] - (float)ca, q[1] - (float)cb, (dim == 3) ? q[2] - (float)cc : 0.0f};
                const float rr[3] = {r[0] - o[9], r[1] - o[10], (dim == 3) ? r[2] - o[11] : 0.0f};
                const float W[3] = {o[6], o[7], o[8]};
                const float at[3] = {W[1] * rr[2] - W[2] * rr[1], W[2] * rr[0] - W[0] * rr[2],
                                     W[0] * rr[1] - W[1] * rr[0]};
                const float damp = (float)p->damp[spid];
                for (int d = 0; d < dim; d++) w[d] = vc[d] + rt[3 * i + d] - o[3 + d] - damp * vc[d] + at[d];
            }
        } else if (gravflag) {
            /* a particle alone in its cell takes the kick itself */
            for (int d = 0; d < dim; d++) w[d] = w[d] + dtg[d];
        }
        {
            const float a = w[0] - avv[0], b = w[1] - avv[1], cz = w[2] - avv[2];
            float e = a * a;
            e += b * b;
            e += cz * cz;
            ke += (long long)rint((double)e * p->mass[spid] * XQD);
        }
        for (int j = 0; j < dim; j++) {
            const float sg = -s[j];
            const float t = q[j] + sg;
            q[j] = (t < 0.0f && t >= sg) ? t + L[j] : t;
        }
        for (int j = 0; j < dim; j++) {
            const float t = q[j] + dt * w[j];
            q[j] = (t < 0.0f) ? t + L[j] : ((t >= L[j]) ? t - L[j] : t);
        }
    }
    *kbt_out = ((double)ke * XQINV) / ((double)n * (double)dim);
    free(cell);
    free(cnt);
    free(nsp);
    free(sum);
    free(cd);
    free(rt);
    return bad;
}
