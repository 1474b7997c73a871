/*
 * mpcd_oracle.c -- CPU oracle for the MPCD solvent step (TEST INFRASTRUCTURE).
 *
 * This file is the checker for the GPU library: tests, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it; the product never does.
 *
 * It restates, in plain single-threaded C, the reference's per-step solvent
 * path.  Each function names the reference function (file:line) it follows.
 * The reference walks per-cell doubly linked lists; here the lists are an
 * explicit "order" array (cells x-major, each list head to tail) plus per-cell
 * start offsets, and bin() is restated as the stable partition that the
 * reference's tail-append relinking produces (mpc.c:948-992 + addlink
 * 1069-1087): in every cell, the particles that stayed keep their previous
 * relative order and are followed by the arrivals in previous global order.
 * Floating-point expressions keep the reference's operation order so that,
 * given the same draws, results agree bit for bit with the reference build.
 *
 * Parity pins (see tests/): every function here is exercised against the
 * golden fixtures recorded from the reference's own timestep() by
 * oracle/_ref/ref_harness with identical sequential draws; closed-form tests
 * additionally pin the eigen-solver, Rodrigues rotation, Philox and the
 * conservation laws.  Parts with no fixture are listed in the header of
 * tests/test_oracle_golden.py.
 */
#include "mpcd_oracle.h"
#include "philox_ref.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_TOL 1e-8          /* definitions.h:57 */
#define ORC_NBOUNCE 50        /* definitions.h:63 */
#define ORC_PI 3.14159265358979323846
#define ORC_MCINT 5           /* definitions.h:355 */
#define ORC_MCSLOPE 0.1       /* definitions.h:353 */
#define ORC_BUSMIN 0.5        /* definitions.h:179 */
#define ORC_BUSMAX 5.0        /* definitions.h:181 */

/* collision operator ids (definitions.h:261-303) */
enum { R_ARBAXIS = 0, R_ORTHAXIS = 1, R_MPCAT = 2, R_RAT = 3, R_VICSEK = 7, R_CHATE = 8, R_CHATE_MPCAT = 13,
       R_CHATE_LANG = 14, R_DIPOLE_VCM = 15, R_DIPOLE_DIR_SUM = 16, R_DIPOLE_DIR_AV = 17, R_DIPOLE_DIR_SIG = 20,
       R_DIPOLE_DIR_SIG_SUM = 21 };
enum { LC_ISOF = 0, LC_LCL = 1, LC_LCG = 2 };
enum { TS_NOTHERM = 0, TS_VSC = 1, TS_BEREND = 2 };

struct orc_sys {
    orc_params p;
    int N, nx, ny, nz, nc;
    int pbc[3];
    double *Q, *V, *U, *T;
    int *spid;
    int *order, *cstart, *cellof;
    int *pop, *spcnt;
    double *mass, *vcm, *cm, *I, *E, *S, *dir;
    double wq[ORC_MAXWALL][3];
    int planar[ORC_MAXWALL];  /* read.c:1423-1434 */
    double shift[3];
    double kbtnow, avs, avdir[3];
    long step;
    long n_escape, n_bcfail, n_rewind, n_retherm;
    int lc_stride;
    /* flow output (outFlags.FLOWOUT): FLOW per cell and the steps summed */
    int flow_on;
    double *flow;
    long flow_n;
    /* scratch */
    int *tmp_order, *newcell, *tmp_cnt, *tmp_seen;
    int *slot_of; /* particle -> slot in the collision order */
    /* discrete decisions of the last collision phase (nematics; the words of
     * orc_lc_decisions) and whether local_prop records the cell words */
    uint32_t *pdec, *cdec;
    int rec_cdec;
};

static double wall_W(const orc_sys *s, int w, const double *q);
static int wall_planar(const orc_wall *W);

/* ------------------------------------------------------------------------ */
/* small vector helpers, restating mtools.c                                  */
/* ------------------------------------------------------------------------ */

static int feq(double x, double y) { return fabs(x - y) <= ORC_TOL; }   /* mtools.c:61 */
static int fneq(double x, double y) { return fabs(x - y) >= ORC_TOL; }  /* mtools.c:74 */

static double dotprod(const double *x, const double *y, int n)        /* mtools.c:111 */
{
    double r = 0.;
    int i;
    for (i = 0; i < n; i++) r += x[i] * y[i];
    return r;
}

static void crossprod(const double x[3], const double y[3], double r[3])  /* mtools.c:177 */
{
    r[0] = x[1] * y[2] - x[2] * y[1];
    r[1] = x[2] * y[0] - x[0] * y[2];
    r[2] = x[0] * y[1] - x[1] * y[0];
}

static double length(const double *x, int n)                           /* mtools.c:227 */
{
    double r = 0.;
    int i;
    for (i = 0; i < n; i++) r += x[i] * x[i];
    return sqrt(r);
}

static void vnorm(double *x, int n)                                    /* mtools.c:246 */
{
    double l = length(x, n);
    int i;
    if (fneq(l, 0.0))
        for (i = 0; i < n; i++) x[i] = x[i] / l;
}

static void proj(const double *v, const double *nrm, double *vn, int n) /* mtools.c:293 */
{
    double x = dotprod(v, nrm, n);
    int i;
    for (i = 0; i < n; i++) vn[i] = x * nrm[i];
}

static void tang(const double *v, const double *vn, double *vt, int n)  /* mtools.c:310 */
{
    int i;
    for (i = 0; i < n; i++) vt[i] = v[i] - vn[i];
}

/* the reference's own atan2 (mtools.c:342-356), which overrides libm's */
static double ref_atan2(double y, double x)
{
    double at = 0.0;
    if (x > 0.0) at = atan(y / x);
    else if (x < 0.0 && y >= 0.0) at = atan(y / x) + ORC_PI;
    else if (x < 0.0 && y < 0.0) at = atan(y / x) - ORC_PI;
    else if (feq(x, 0.0) && y > 0.0) at = 0.5 * ORC_PI;
    else if (feq(x, 0.0) && y < 0.0) at = -0.5 * ORC_PI;
    return at;
}

static double cosang(const double *a, const double *b, int n)          /* mtools.c:325 */
{
    double c = dotprod(a, b, n);
    c /= length(a, n);
    c /= length(b, n);
    return c;
}

static double absAngle(const double *a, const double *b, int n) { return acos(cosang(a, b, n)); } /* mtools.c:365 */

static double signedAngle(const double *v1, const double *v2, int n)   /* mtools.c:380 */
{
    double A[3] = {0, 0, 0}, B[3] = {0, 0, 0}, c3[3];
    int i;
    for (i = 0; i < n; i++) { A[i] = v1[i]; B[i] = v2[i]; }
    crossprod(A, B, c3);
    return ref_atan2(length(c3, 3), dotprod(A, B, n));
}

/* mtools.c:1572-1584 */
void orc_rodrigues(double *vec, double *rotAx, double theta)
{
    double cp[3], dp, oldl, newl;
    int i;
    vnorm(rotAx, 3);
    oldl = length(vec, 3);
    crossprod(rotAx, vec, cp);
    dp = dotprod(rotAx, vec, 3);
    for (i = 0; i < 3; i++) vec[i] = vec[i] * cos(theta) + cp[i] * sin(theta) + rotAx[i] * dp * (1.0 - cos(theta));
    newl = length(vec, 3);
    for (i = 0; i < 3; i++) vec[i] *= (oldl / newl);
}

static double det3x3(double m[3][3])                                   /* mtools.c:857 */
{
    double c1 = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]);
    double c2 = m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]);
    double c3 = m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    return (c1 - c2 + c3);
}

static double cofactor3x3(double m[3][3], int i, int j)                /* mtools.c:972 */
{
    /* minor of (i,j): rows != i, columns != j, in increasing order */
    int r0 = (i == 0) ? 1 : 0, r1 = (i == 2) ? 1 : 2;
    int c0 = (j == 0) ? 1 : 0, c1 = (j == 2) ? 1 : 2;
    int c = (2 * ((i + j) / 2) == (i + j)) ? 1 : -1;
    return c * (m[r0][c0] * m[r1][c1] - m[r0][c1] * m[r1][c0]);
}

static void invert3x3(double inv[3][3], double m[3][3])                /* mtools.c:1031 */
{
    double det = det3x3(m);
    int i, j;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) inv[j][i] = cofactor3x3(m, i, j) / det;
}

void orc_inv3(double inv[3][3], const double m[3][3])
{
    double a[3][3];
    memcpy(a, m, sizeof a);
    invert3x3(inv, a);
}

static void matvec3(double M[3][3], const double *v, double *r)        /* mtools.c:130 */
{
    int i, j;
    for (i = 0; i < 3; i++) r[i] = 0.;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) r[i] += M[i][j] * v[j];
}

/* eigen-solvers, mtools.c:1230-1476.  m is row-major; on return the rows of
 * m are the (normalised) eigenvectors, eigval[0] the largest eigenvalue. */
static int eig2(double m[2][2], double ev[2])
{
    double vec[2][2];
    int br;
    double det = m[0][0] * m[1][1] - m[1][0] * m[0][1];   /* determinant(), mtools.c:900 */
    double tr = m[0][0] + m[1][1];
    double sq = sqrt(tr * tr * 0.25 - det);
    int i, j;
    ev[0] = tr * 0.5 + sq;
    ev[1] = tr * 0.5 - sq;
    vec[0][0] = vec[0][1] = vec[1][0] = vec[1][1] = 0.0;
    br = fneq(m[1][0], 0.0) ? 0 : fneq(m[0][1], 0.0) ? 1 : feq(m[0][0], m[1][1]) ? 2 : 3;
    if (fneq(m[1][0], 0.0)) {
        vec[0][0] = ev[0] - m[1][1];
        vec[0][1] = m[1][0];
        vnorm(vec[0], 2);
        vec[1][0] = ev[1] - m[1][1];
        vec[1][1] = m[1][0];
        vnorm(vec[1], 2);
    } else if (fneq(m[0][1], 0.0)) {
        vec[0][0] = m[0][1];
        vec[0][1] = ev[0] - m[0][0];
        vnorm(vec[0], 2);
        vec[1][0] = m[0][1];
        vec[1][1] = ev[1] - m[0][0];
        vnorm(vec[1], 2);
    } else {
        if (feq(m[0][0], m[1][1])) {
            vec[0][0] = sqrt(2.);
            vec[0][1] = vec[0][0];
            vec[1][0] = vec[0][0];
            vec[1][1] = vec[0][0];
        } else if (feq(m[0][0], ev[0])) {
            vec[0][0] = 1.0; vec[0][1] = 0.0; vec[1][0] = 0.0; vec[1][1] = 1.0;
        } else if (feq(m[1][1], ev[0])) {
            /* mtools.c:1288-1293 writes eigvec[0] twice; eigvec[1] is left unset */
            vec[0][0] = 1.0; vec[0][1] = 0.0;
        }
    }
    for (i = 0; i < 2; i++)
        for (j = 0; j < 2; j++) m[i][j] = vec[i][j];
    return br;
}

static int eig3(double m[3][3], double ev[3])
{
    double vec[3][3];
    double q, p1, p2, p, ip, r, phi, B[3][3];
    int i, j, k, row, col, br = 0;
    p1 = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
    if (feq(p1, 0.0)) {
        br |= 4;
        for (i = 0; i < 3; i++) ev[i] = m[i][i];
        if (ev[1] > ev[0]) { q = ev[0]; ev[0] = ev[1]; ev[1] = q; }
        if (ev[2] > ev[0]) { q = ev[0]; ev[0] = ev[2]; ev[2] = q; }
        if (ev[2] > ev[1]) { q = ev[1]; ev[1] = ev[2]; ev[2] = q; }
    } else {
        p2 = 0.;
        q = (m[0][0] + m[1][1] + m[2][2]) / 3.;
        for (i = 0; i < 3; i++) p2 += (m[i][i] - q) * (m[i][i] - q);
        p2 += 2. * p1;
        p = sqrt(p2 / 6.);
        ip = 1. / p;
        for (i = 0; i < 3; i++)
            for (j = 0; j < 3; j++) B[i][j] = m[i][j];
        for (i = 0; i < 3; i++) B[i][i] -= q;
        for (i = 0; i < 3; i++)
            for (j = 0; j < 3; j++) B[i][j] *= ip;
        r = det3x3(B) * 0.5;
        if (r <= -1. || r >= 1.) br |= 8;
        if (r <= -1.) phi = ORC_PI / 3.;
        else if (r >= 1.) phi = 0.;
        else phi = acos(r) / 3.;
        ev[0] = q + 2. * p * cos(phi);
        ev[2] = q + 2. * p * cos(phi + (2. * ORC_PI / 3.));
        ev[1] = 3. * q - ev[0] - ev[2];
    }
    /* eigenvectors by Cayley-Hamilton column products, mtools.c:1373-1430 */
    for (k = 0; k < 3; k++) {
        double ea = (k == 0) ? ev[1] : ev[0];
        double eb = (k == 2) ? ev[1] : ev[2];
        if (k == 0) br |= 3; /* the column of the first eigenvector; 3: none */
        for (col = 0; col < 3; col++) {
            for (row = 0; row < 3; row++) vec[k][row] = 0.;
            for (row = 0; row < 3; row++)
                for (i = 0; i < 3; i++) {
                    double a = m[row][i], b;
                    if (i == row) a -= ea;
                    b = m[i][col];
                    if (i == col) b -= eb;
                    vec[k][row] += a * b;
                }
            if (!(feq(vec[k][0], 0.0) && feq(vec[k][1], 0.0) && feq(vec[k][2], 0.0))) {
                if (k == 0) br = (br & ~3) | col;
                break;
            }
        }
    }
    for (k = 0; k < 3; k++) vnorm(vec[k], 3);
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) m[i][j] = vec[i][j];
    return br;
}

void orc_eigen(double *m, int dim, double *eigval)
{
    int i, j;
    if (dim == 2) {
        double a[2][2];
        for (i = 0; i < 2; i++) for (j = 0; j < 2; j++) a[i][j] = m[i * 2 + j];
        eig2(a, eigval);
        for (i = 0; i < 2; i++) for (j = 0; j < 2; j++) m[i * 2 + j] = a[i][j];
    } else {
        double a[3][3];
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) a[i][j] = m[i * 3 + j];
        eig3(a, eigval);
        for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) m[i * 3 + j] = a[i][j];
    }
}

/* ------------------------------------------------------------------------ */
/* draws                                                                     */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    philox_ref_4x32_10(ctr, key, out);
}

static void philox_block(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t blk,
                         uint32_t w[4])
{
    uint32_t ctr[4], k2[2];
    ctr[0] = key;
    ctr[1] = ((uint32_t)purpose << 24) | (blk & 0xFFFFFFu);
    ctr[2] = step;
    ctr[3] = replica;
    k2[0] = (uint32_t)seed;
    k2[1] = (uint32_t)(seed >> 32);
    philox_ref_4x32_10(ctr, k2, w);
}

static double u53(uint32_t a, uint32_t b)
{
    return (double)(((uint64_t)a << 21) | (uint64_t)(b >> 11)) * (1.0 / 9007199254740992.0);
}

double orc_philox_uniform(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k)
{
    uint32_t w[4];
    philox_block(seed, replica, step, purpose, key, k, w);
    return u53(w[0], w[1]);
}

double orc_philox_normal(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k)
{
    uint32_t w[4];
    double u1, u2, r;
    philox_block(seed, replica, step, purpose, key, k >> 1, w);
    u1 = 1.0 - u53(w[0], w[1]);
    u2 = u53(w[2], w[3]);
    r = sqrt(-2.0 * log(u1));
    return (k & 1) ? r * sin(6.283185307179586 * u2) : r * cos(6.283185307179586 * u2);
}

static void philox_unitvec(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k,
                           double v[3])
{
    uint32_t w[4];
    double z, s, phi;
    philox_block(seed, replica, step, purpose, key, k, w);
    z = 1.0 - 2.0 * u53(w[0], w[1]);
    phi = 6.283185307179586 * u53(w[2], w[3]);
    s = sqrt((1.0 - z) * (1.0 + z));
    v[0] = s * cos(phi);
    v[1] = s * sin(phi);
    v[2] = z;
}

void orc_philox_unitvec(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k,
                        double v[3])
{
    philox_unitvec(seed, replica, step, purpose, key, k, v);
}

/* sine and cosine of 2 pi u for u = n 2^-53 in [0, 1) without a library
   sine: octant of u (8u, its fraction f and 1 - f are exact), angle g pi/4 in
   the octant, Taylor polynomials of degree 13 and 14 on [0, pi/4].  Plain
   FP64 adds and multiplies in a fixed order (no contraction under -std=c99),
   so the exact-sum form's device code gives the same bits. */
void orc_sincos_turn(double u, double *sn, double *cs)
{
    const double t = 8.0 * u;
    const int o = (int)t;
    const double f = t - (double)o;
    const double g = (o & 1) ? 1.0 - f : f;
    const double th = g * 0.78539816339744830962;
    const double x = th * th;
    double p = 1.0 / 6227020800.0, q = -1.0 / 87178291200.0, s, c, a, b;
    int swap;
    p = p * x + -1.0 / 39916800.0;
    p = p * x + 1.0 / 362880.0;
    p = p * x + -1.0 / 5040.0;
    p = p * x + 1.0 / 120.0;
    p = p * x + -1.0 / 6.0;
    s = th + (th * x) * p;
    q = q * x + 1.0 / 479001600.0;
    q = q * x + -1.0 / 3628800.0;
    q = q * x + 1.0 / 40320.0;
    q = q * x + -1.0 / 720.0;
    q = q * x + 1.0 / 24.0;
    q = q * x + -0.5;
    c = 1.0 + x * q;
    swap = ((o + 1) >> 1) & 1;
    a = swap ? s : c;
    b = swap ? c : s;
    *cs = ((o + 2) & 4) ? -a : a;
    *sn = (o & 4) ? -b : b;
}

/* natural log of u in (0, 1] without a library log: u = m 2^e with m in
   [1/sqrt 2, sqrt 2) (frexp and a doubling, exact), log m = 2 atanh s with
   s = (m - 1) / (m + 1) from its series to s^19, plus e ln 2; the same
   operations in the same order as the exact-sum form's device code */
double orc_log_unit(double u)
{
    int e;
    double m = frexp(u, &e), s, x, p;
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    s = (m - 1.0) / (m + 1.0);
    x = s * s;
    p = 1.0 / 19.0;
    p = p * x + 1.0 / 17.0;
    p = p * x + 1.0 / 15.0;
    p = p * x + 1.0 / 13.0;
    p = p * x + 1.0 / 11.0;
    p = p * x + 1.0 / 9.0;
    p = p * x + 1.0 / 7.0;
    p = p * x + 1.0 / 5.0;
    p = p * x + 1.0 / 3.0;
    p = p * x + 1.0;
    return (double)e * 0.69314718055994530942 + 2.0 * (s * p);
}

/* orc_philox_normal with orc_log_unit and orc_sincos_turn */
double orc_philox_normal_turn(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k)
{
    uint32_t w[4];
    double r, sn, cs;
    philox_block(seed, replica, step, purpose, key, k >> 1, w);
    r = sqrt(-2.0 * orc_log_unit(1.0 - u53(w[0], w[1])));
    orc_sincos_turn(u53(w[2], w[3]), &sn, &cs);
    return (k & 1) ? r * sn : r * cs;
}

/* philox_unitvec with the sine and cosine of orc_sincos_turn */
void orc_philox_unitvec_turn(uint64_t seed, uint32_t replica, uint32_t step, int purpose, uint32_t key, uint32_t k,
                             double v[3])
{
    uint32_t w[4];
    double z, s, sn, cs;
    philox_block(seed, replica, step, purpose, key, k, w);
    z = 1.0 - 2.0 * u53(w[0], w[1]);
    orc_sincos_turn(u53(w[2], w[3]), &sn, &cs);
    s = sqrt((1.0 - z) * (1.0 + z));
    v[0] = s * cs;
    v[1] = s * sn;
    v[2] = z;
}

static const double *keyed_base(const orc_draws *d, int purpose, int *stride)
{
    switch (purpose) {
    case ORC_P_SHIFT: *stride = 3; return d->shift;
    case ORC_P_CELL: *stride = 4; return d->cell;
    case ORC_P_GHOST: *stride = 3; return d->ghost;
    case ORC_P_NORMAL: *stride = 3; return d->part;
    case ORC_P_LC: *stride = d->lc_stride; return d->lc;
    default: *stride = 0; return NULL;
    }
}

static double *rec_base(const orc_draws *d, int purpose, int *stride)
{
    switch (purpose) {
    case ORC_P_SHIFT: *stride = 3; return d->rec_shift;
    case ORC_P_CELL: *stride = 4; return d->rec_cell;
    case ORC_P_GHOST: *stride = 3; return d->rec_ghost;
    case ORC_P_NORMAL: *stride = 3; return d->rec_part;
    case ORC_P_LC: *stride = d->rec_lc_stride; return d->rec_lc;
    default: *stride = 0; return NULL;
    }
}

/* the rare draws' keyed entry: escape repair of binning 0 (ORC_P_ESCAPE) or
 * 1 (ORC_P_ESCAPE2), re-thermalisation (ORC_P_RETHERM); NULL when absent */
static double *rare_base(const orc_draws *d, int purpose, int *stride)
{
    if ((purpose == ORC_P_ESCAPE || purpose == ORC_P_ESCAPE2) && d->escape) {
        *stride = ORC_ESC_STRIDE;
        return d->escape + (purpose == ORC_P_ESCAPE2 ? (size_t)d->rare_n * ORC_ESC_STRIDE : 0);
    }
    if (purpose == ORC_P_RETHERM && d->retherm) {
        *stride = ORC_RETH_STRIDE;
        return d->retherm;
    }
    *stride = 0;
    return NULL;
}

static void record(orc_draws *d, int purpose, long key, int k, double x)
{
    int stride;
    double *b = rec_base(d, purpose, &stride);
    if (!b) b = rare_base(d, purpose, &stride);
    if (!b) return;
    if (k < stride) b[key * stride + k] = x;
    else if (purpose >= ORC_P_ESCAPE) d->rare_overflow = 1;
}

/* keyed read of a rare draw; returns 0 when the generator serves it */
static int rare_keyed(orc_draws *d, int purpose, long key, int k, double *x)
{
    int stride;
    const double *b;
    if (d->mode != ORC_DRAW_KEYED || purpose < ORC_P_ESCAPE) return 0;
    b = rare_base(d, purpose, &stride);
    if (!b) return 0;
    if (k >= stride) {
        d->rare_overflow = 1;
        *x = 0.5;
        return 1;
    }
    *x = b[key * stride + k];
    return 1;
}

/* uniform on [0,1) -- the reference's genrand_real() (rand.c:365) */
static double dr_uniform(orc_sys *s, orc_draws *d, int purpose, long key, int k)
{
    double x;
    if (d->mode == ORC_DRAW_SEQ) {
        x = seqdraw_uniform(&d->seq);
        record(d, purpose, key, k, x);
        return x;
    }
    if (d->mode == ORC_DRAW_KEYED && purpose <= ORC_P_LC) {
        int stride;
        const double *b = keyed_base(d, purpose, &stride);
        return b[key * stride + k];
    }
    if (rare_keyed(d, purpose, key, k, &x)) return x;
    x = orc_philox_uniform(d->seed, d->replica, (uint32_t)s->step, purpose, (uint32_t)key, (uint32_t)k);
    record(d, purpose, key, k, x);
    return x;
}

/* standard normal -- the reference's genrand_gauss() (rand.c:473), a float */
static double dr_normal(orc_sys *s, orc_draws *d, int purpose, long key, int k)
{
    double x;
    if (d->mode == ORC_DRAW_SEQ) {
        x = (double)seqdraw_gauss(&d->seq);
        record(d, purpose, key, k, x);
        return x;
    }
    if (d->mode == ORC_DRAW_KEYED && purpose <= ORC_P_LC) {
        int stride;
        const double *b = keyed_base(d, purpose, &stride);
        return b[key * stride + k];
    }
    if (rare_keyed(d, purpose, key, k, &x)) return x;
    x = orc_philox_normal(d->seed, d->replica, (uint32_t)s->step, purpose, (uint32_t)key, (uint32_t)k);
    record(d, purpose, key, k, x);
    return x;
}

/* unit vector on the sphere -- the reference's ranvec3D() (rand.c:708) */
static void dr_unitvec(orc_sys *s, orc_draws *d, int purpose, long key, int k, double v[3])
{
    if (d->mode == ORC_DRAW_SEQ) {
        seqdraw_unitvec3(&d->seq, v);
        record(d, purpose, key, k, v[0]);
        record(d, purpose, key, k + 1, v[1]);
        record(d, purpose, key, k + 2, v[2]);
        return;
    }
    if (d->mode == ORC_DRAW_KEYED && purpose <= ORC_P_LC) {
        int stride;
        const double *b = keyed_base(d, purpose, &stride);
        v[0] = b[key * stride + k];
        v[1] = b[key * stride + k + 1];
        v[2] = b[key * stride + k + 2];
        return;
    }
    philox_unitvec(d->seed, d->replica, (uint32_t)s->step, purpose, (uint32_t)key, (uint32_t)k, v);
    record(d, purpose, key, k, v[0]);
    record(d, purpose, key, k + 1, v[1]);
    record(d, purpose, key, k + 2, v[2]);
}

static double dr_gaussMB(orc_sys *s, orc_draws *d, double kbt, double m, int purpose, long key, int k)
{
    double sigma = sqrt(kbt / m);                       /* rand.c:510-513 */
    return sigma * dr_normal(s, d, purpose, key, k);
}

/* ------------------------------------------------------------------------ */
/* system setup                                                              */
/* ------------------------------------------------------------------------ */

static int cidx(const orc_sys *s, int a, int b, int c) { return (a * s->ny + b) * s->nz + c; }

/* setPBC, read.c:584-603, for the axis-aligned planar walls used here */
static void set_pbc(orc_sys *s)
{
    int i, w;
    for (i = 0; i < 3; i++) s->pbc[i] = 0;
    for (w = 0; w < s->p.nwall; w++) {
        const orc_wall *W = &s->p.wall[w];
        for (i = 0; i < 3; i++) {
            double ainv = (W->A[i] != 0.0) ? 1.0 / W->A[i] : 0.0;
            if (feq(fabs(ainv), 1.0) && W->MVN > 0.0 && W->MVT > 0.0 && (fneq(W->DN, 0.0) || fneq(W->DT, 0.0)))
                s->pbc[i] = 1;
        }
    }
}

static int anneal_max(const orc_params *p)
{
    double effm = 0.0;
    int i;
    for (i = 0; i < p->nspec; i++)
        if (p->sp[i].mfpot > effm) effm = p->sp[i].mfpot;
    effm = effm * 0.5 * (double)p->dim;
    return (int)(ORC_MCINT + ORC_MCSLOPE * effm * 1.0 / p->kbt) + 1;
}

orc_sys *orc_create(const orc_params *p, int n)
{
    orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
    int w;
    if (!s) return NULL;
    s->p = *p;
    s->N = n;
    for (w = 0; w < p->nwall; w++) {
        orc_wall *W = &s->p.wall[w];
        /* P all zero: the plane (1 1 1 1) */
        if (W->P[0] == 0.0 && W->P[1] == 0.0 && W->P[2] == 0.0 && W->P[3] == 0.0)
            W->P[0] = W->P[1] = W->P[2] = W->P[3] = 1.0;
        s->planar[w] = wall_planar(W);
    }
    s->nx = p->L[0] + 1;
    s->ny = p->L[1] + 1;
    s->nz = (p->dim == 3) ? p->L[2] + 1 : 1;
    s->nc = s->nx * s->ny * s->nz;
    set_pbc(s);
    s->Q = (double *)calloc((size_t)n * 3, sizeof(double));
    s->V = (double *)calloc((size_t)n * 3, sizeof(double));
    s->U = (double *)calloc((size_t)n * 3, sizeof(double));
    s->T = (double *)calloc((size_t)n * 3, sizeof(double));
    s->spid = (int *)calloc((size_t)n, sizeof(int));
    s->order = (int *)calloc((size_t)n, sizeof(int));
    s->tmp_order = (int *)calloc((size_t)n, sizeof(int));
    s->newcell = (int *)calloc((size_t)n, sizeof(int));
    s->cellof = (int *)calloc((size_t)n, sizeof(int));
    s->slot_of = (int *)calloc((size_t)n, sizeof(int));
    s->pdec = (uint32_t *)calloc((size_t)n, sizeof(uint32_t));
    s->cdec = (uint32_t *)calloc((size_t)s->nc, sizeof(uint32_t));
    s->cstart = (int *)calloc((size_t)s->nc + 1, sizeof(int));
    s->tmp_cnt = (int *)calloc((size_t)s->nc + 1, sizeof(int));
    s->tmp_seen = (int *)calloc((size_t)s->nc + 1, sizeof(int));
    s->pop = (int *)calloc((size_t)s->nc, sizeof(int));
    s->spcnt = (int *)calloc((size_t)s->nc * ORC_MAXSPEC, sizeof(int));
    s->mass = (double *)calloc((size_t)s->nc, sizeof(double));
    s->vcm = (double *)calloc((size_t)s->nc * 3, sizeof(double));
    s->flow = (double *)calloc((size_t)s->nc * 3, sizeof(double));
    s->cm = (double *)calloc((size_t)s->nc * 3, sizeof(double));
    s->I = (double *)calloc((size_t)s->nc * 9, sizeof(double));
    s->E = (double *)calloc((size_t)s->nc * 9, sizeof(double));
    s->S = (double *)calloc((size_t)s->nc, sizeof(double));
    s->dir = (double *)calloc((size_t)s->nc * 3, sizeof(double));
    for (w = 0; w < p->nwall; w++) {
        s->wq[w][0] = p->wall[w].Q[0];
        s->wq[w][1] = p->wall[w].Q[1];
        s->wq[w][2] = p->wall[w].Q[2];
    }
    s->lc_stride = 3 + 2 * anneal_max(p);
    return s;
}

void orc_destroy(orc_sys *s)
{
    if (!s) return;
    free(s->Q); free(s->V); free(s->U); free(s->T); free(s->spid);
    free(s->order); free(s->tmp_order); free(s->newcell); free(s->cellof); free(s->slot_of);
    free(s->cstart); free(s->tmp_cnt); free(s->tmp_seen);
    free(s->pop); free(s->spcnt); free(s->mass); free(s->vcm); free(s->flow); free(s->cm); free(s->I); free(s->E);
    free(s->S); free(s->dir); free(s->pdec); free(s->cdec);
    free(s);
}

static int cell_of_pos(const orc_sys *s, const double *q, int shifted, int *escaped)
{
    int a = (int)q[0], b = (int)q[1], c = (int)q[2];
    const int *L = s->p.L;
    int Lz = (s->p.dim == 3) ? L[2] : 1;
    if (shifted) {                                   /* mpc.c:964-968 */
        if (s->pbc[0] && a == L[0]) a = 0;
        if (s->pbc[1] && b == L[1]) b = 0;
        if (s->pbc[2] && c == Lz) c = 0;
    }
    if (a > L[0] || b > L[1] || c > Lz || a < 0 || b < 0 || c < 0) {
        *escaped = 1;
        return -1;
    }
    *escaped = 0;
    if (s->p.dim < 3) c = 0;
    return cidx(s, a, b, c);
}

/* ------------------------------------------------------------------------ */
/* binning                                                                   */
/* ------------------------------------------------------------------------ */

/* rebuild cstart/cellof from an order array and per-particle cells */
static void rebuild_lists(orc_sys *s)
{
    int c, t;
    for (c = 0; c <= s->nc; c++) s->cstart[c] = 0;
    for (t = 0; t < s->N; t++) s->cstart[s->cellof[s->order[t]] + 1]++;
    for (c = 0; c < s->nc; c++) s->cstart[c + 1] += s->cstart[c];
}

/* binin, mpc.c:924-934: lists in particle-index order */
static int binin(orc_sys *s)
{
    int i, c, esc, t;
    for (c = 0; c <= s->nc; c++) s->tmp_cnt[c] = 0;
    for (i = 0; i < s->N; i++) {
        c = cell_of_pos(s, &s->Q[i * 3], 0, &esc);
        if (esc) return 1;
        s->cellof[i] = c;
        s->tmp_cnt[c + 1]++;
    }
    for (c = 0; c < s->nc; c++) s->tmp_cnt[c + 1] += s->tmp_cnt[c];
    for (c = 0; c <= s->nc; c++) s->tmp_seen[c] = 0;
    for (i = 0; i < s->N; i++) {
        c = s->cellof[i];
        t = s->tmp_cnt[c] + s->tmp_seen[c]++;
        s->order[t] = i;
    }
    rebuild_lists(s);
    return 0;
}

static void push_velocity(orc_sys *s, orc_draws *d, int i, int purpose, int kbase);

/* bin()'s escape repair, mpc.c:970-982, of particle i in binning b (0: the
 * shifted bin, 1: the unshifted bin after streaming): replacePos_WithCheck
 * (init.c:1462-1483) redraws a uniform position in the box (replace,
 * init.c:1240-1243) while any wall has calcW <= TOL at the walls' current
 * positions; then DIM Maxwell velocities (genrand_gaussMB,
 * rand.c:510, for every VDIST) and, for nematics, genrand_coneNP(U, pi, DIM)
 * (rand.c:397-425).  Draws are consumed in that order. */
static void escape_repair(orc_sys *s, orc_draws *d, int i, int b)
{
    const int purpose = b ? ORC_P_ESCAPE2 : ORC_P_ESCAPE, dim = s->p.dim;
    const double m = s->p.sp[s->spid[i]].mass;
    int flag = 1, w, j, t = 0;
    while (flag) {
        for (j = 0; j < dim; j++)
            s->Q[i * 3 + j] = (double)s->p.L[j] * dr_uniform(s, d, purpose, i, ORC_ESC_POS0 + dim * t + j);
        t++;
        flag = 0;
        for (w = 0; w < s->p.nwall; w++)
            if (wall_W(s, w, &s->Q[i * 3]) <= ORC_TOL) flag = 1;
        if (d->rare_overflow) break; /* a keyed entry ran out of tries */
    }
    for (j = 0; j < dim; j++) s->V[i * 3 + j] = dr_gaussMB(s, d, s->p.kbt, m, purpose, i, j);
    if (s->p.lc > LC_ISOF) {
        double *u = &s->U[i * 3];
        if (dim == 3) {
            const double ct = cos(ORC_PI);
            const double z = dr_uniform(s, d, purpose, i, 3) * (1.0 - ct) + ct;
            const double phi = dr_uniform(s, d, purpose, i, 4) * 2.0 * ORC_PI;
            u[0] = z;
            u[1] = sqrt(1.0 - z * z) * cos(phi);
            u[2] = sqrt(1.0 - z * z) * sin(phi);
        } else {
            const double phi = ORC_PI * (1. - 2. * dr_uniform(s, d, purpose, i, 3));
            u[0] = cos(phi);
            u[1] = sin(phi);
        }
    }
}

/* bin, mpc.c:948-992, restated as the stable partition the linked lists produce */
static void bin(orc_sys *s, orc_draws *d, int shifted)
{
    int t, i, c, esc;
    for (c = 0; c <= s->nc; c++) { s->tmp_cnt[c] = 0; s->tmp_seen[c] = 0; }
    for (t = 0; t < s->N; t++) {
        i = s->order[t];
        c = cell_of_pos(s, &s->Q[i * 3], shifted, &esc);
        if (esc) {
            /* mpc.c:970-982: the particle stays in the list it was in */
            s->n_escape++;
            escape_repair(s, d, i, shifted ? 0 : 1);
            c = s->cellof[i];
        }
        s->newcell[t] = c;
        s->tmp_cnt[c + 1]++;
        if (c == s->cellof[i]) s->tmp_seen[c]++; /* stayers per cell */
    }
    for (c = 0; c < s->nc; c++) s->tmp_cnt[c + 1] += s->tmp_cnt[c];
    /* stayers go to [start, start+stay), arrivals follow; both in previous order */
    {
        int *stay = s->tmp_seen;
        int *arr_seen = (int *)calloc((size_t)s->nc, sizeof(int));
        int *stay_seen = (int *)calloc((size_t)s->nc, sizeof(int));
        for (t = 0; t < s->N; t++) {
            i = s->order[t];
            c = s->newcell[t];
            if (c == s->cellof[i]) s->tmp_order[s->tmp_cnt[c] + stay_seen[c]++] = i;
            else s->tmp_order[s->tmp_cnt[c] + stay[c] + arr_seen[c]++] = i;
        }
        free(arr_seen);
        free(stay_seen);
    }
    for (t = 0; t < s->N; t++) s->order[t] = s->tmp_order[t];
    for (c = 0; c < s->nc; c++)
        for (t = s->tmp_cnt[c]; t < s->tmp_cnt[c + 1]; t++) s->cellof[s->order[t]] = c;
    for (c = 0; c <= s->nc; c++) s->cstart[c] = s->tmp_cnt[c];
}

/* ------------------------------------------------------------------------ */
/* cell properties                                                           */
/* ------------------------------------------------------------------------ */

static int compute_cm(const orc_params *p)            /* mpc.c:73 */
{
    return p->rtech == R_RAT || p->lc != LC_ISOF || p->rtech == R_DIPOLE_VCM || p->rtech == R_DIPOLE_DIR_SUM ||
           p->rtech == R_DIPOLE_DIR_AV;
}

/* tensOrderParam, lc.c:1125-1148, over the particles of cell c */
static void tens_order(orc_sys *s, int c, double Sm[3][3])
{
    int dim = s->p.dim, i, j, t;
    double fDIM = (double)dim, cc = 1. / (fDIM - 1.), POPinv;
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) Sm[i][j] = 0.0;
    if (s->cstart[c + 1] == s->cstart[c]) return;
    POPinv = 1. / ((double)s->pop[c]);
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        const double *u = &s->U[s->order[t] * 3];
        for (i = 0; i < dim; i++)
            for (j = i; j < dim; j++) Sm[i][j] += u[i] * u[j];
    }
    for (i = 0; i < dim; i++)
        for (j = i; j < dim; j++) Sm[i][j] *= cc * fDIM * POPinv;
    for (i = 0; i < dim; i++) Sm[i][i] -= cc;
    for (i = 0; i < dim; i++)
        for (j = i + 1; j < dim; j++) Sm[j][i] = Sm[i][j];
}

/* S and director from an order tensor, mpc.c:184-207 */
static void order_from_tensor(orc_sys *s, int c, double Sm[3][3])
{
    int dim = s->p.dim, i, br, clamped;
    double ev[3] = {0, 0, 0};
    if (dim == 3) {
        br = eig3(Sm, ev);
        s->S[c] = -1. * (ev[1] + ev[2]);
    } else {
        double a[2][2] = {{Sm[0][0], Sm[0][1]}, {Sm[1][0], Sm[1][1]}};
        br = eig2(a, ev);
        Sm[0][0] = a[0][0]; Sm[0][1] = a[0][1];
        s->S[c] = ev[0];
    }
    for (i = 0; i < dim; i++) s->dir[c * 3 + i] = Sm[0][i];
    clamped = s->S[c] > 1.0;
    if (s->S[c] > 1.0) s->S[c] = 1.0;
    if (s->rec_cdec) s->cdec[c] = 32u | (uint32_t)(br & 15) | (clamped ? 16u : 0u);
}

/* velGradD2Q9 (lc.c:2130-2162) and velGradD3Q15 (lc.c:2070-2119), including
 * the edge copies whose corners read values not yet updated this call. */
static void vel_grad(orc_sys *s)
{
    int a, b, c, i, j, k;
    int X = s->p.L[0], Y = s->p.L[1];
    double *E = s->E;
    if (s->p.dim == 2) {
        static const int uv[8][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
        const double w1 = 1. / 9., w5 = 1. / 36.;
        const double w[8] = {w1, w1, w1, w1, w5, w5, w5, w5};
        for (a = 1; a < X; a++)
            for (b = 1; b < Y; b++)
                for (i = 0; i < 2; i++)
                    for (j = 0; j < 2; j++) {
                        int ce = cidx(s, a, b, 0);
                        E[ce * 9 + i * 3 + j] = 0.;
                        for (k = 0; k < 8; k++) {
                            int cn = cidx(s, a + uv[k][0], b + uv[k][1], 0);
                            E[ce * 9 + i * 3 + j] += w[k] * uv[k][j] * s->vcm[cn * 3 + i];
                        }
                        E[ce * 9 + i * 3 + j] *= 3.;
                    }
        for (b = 0; b < Y; b++)
            for (i = 0; i < 2; i++)
                for (j = 0; j < 2; j++) E[cidx(s, 0, b, 0) * 9 + i * 3 + j] = E[cidx(s, 1, b, 0) * 9 + i * 3 + j];
        for (b = 0; b < Y; b++)
            for (i = 0; i < 2; i++)
                for (j = 0; j < 2; j++) E[cidx(s, X, b, 0) * 9 + i * 3 + j] = E[cidx(s, X - 1, b, 0) * 9 + i * 3 + j];
        for (a = 0; a < X; a++)
            for (i = 0; i < 2; i++)
                for (j = 0; j < 2; j++) E[cidx(s, a, 0, 0) * 9 + i * 3 + j] = E[cidx(s, a, 1, 0) * 9 + i * 3 + j];
        for (a = 0; a < X; a++)
            for (i = 0; i < 2; i++)
                for (j = 0; j < 2; j++) E[cidx(s, a, Y, 0) * 9 + i * 3 + j] = E[cidx(s, a, Y - 1, 0) * 9 + i * 3 + j];
    } else {
        static const int uv[14][3] = {{1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0},  {0, 0, 1},
                                      {0, 0, -1}, {1, 1, 1},  {-1, 1, 1}, {1, -1, 1},  {-1, -1, 1},
                                      {1, 1, -1}, {-1, 1, -1}, {1, -1, -1}, {-1, -1, -1}};
        const double w1 = 1. / 9., w7 = 1. / 72.;
        double w[14];
        int Z = s->p.L[2];
        for (k = 0; k < 14; k++) w[k] = (k < 6) ? w1 : w7;
        for (a = 1; a < X; a++)
            for (b = 1; b < Y; b++)
                for (c = 1; c < Z; c++)
                    for (i = 0; i < 3; i++)
                        for (j = 0; j < 3; j++) {
                            int ce = cidx(s, a, b, c);
                            E[ce * 9 + i * 3 + j] = 0.;
                            for (k = 0; k < 14; k++) {
                                int cn = cidx(s, a + uv[k][0], b + uv[k][1], c + uv[k][2]);
                                E[ce * 9 + i * 3 + j] += w[k] * uv[k][j] * s->vcm[cn * 3 + i];
                            }
                            E[ce * 9 + i * 3 + j] *= 3.;
                        }
#define CP(ad, bd, cd, as, bs, cs)                                                                          \
    for (i = 0; i < 3; i++)                                                                                 \
        for (j = 0; j < 3; j++) E[cidx(s, ad, bd, cd) * 9 + i * 3 + j] = E[cidx(s, as, bs, cs) * 9 + i * 3 + j];
        for (b = 0; b < Y; b++) for (c = 0; c < Z; c++) { CP(0, b, c, 1, b, c) }
        for (b = 0; b < Y; b++) for (c = 0; c < Z; c++) { CP(X, b, c, X - 1, b, c) }
        for (a = 0; a < X; a++) for (c = 0; c < Z; c++) { CP(a, 0, c, a, 1, c) }
        for (a = 0; a < X; a++) for (c = 0; c < Z; c++) { CP(a, Y, c, a, Y - 1, c) }
        for (a = 0; a < X; a++) for (b = 0; b < Y; b++) { CP(a, b, 0, a, b, 1) }
        for (a = 0; a < X; a++) for (b = 0; b < Y; b++) { CP(a, b, Z, a, b, Z - 1) }
#undef CP
    }
}

/* localPROP, mpc.c:64-216, with localMomInertiaTensor (mpc.c:4209-4299) */
static void local_prop(orc_sys *s)
{
    const orc_params *p = &s->p;
    int dim = p->dim, c, t, d, i;
    int ccm = compute_cm(p);
    for (c = 0; c < s->nc; c++) {
        double *vcm = &s->vcm[c * 3], *cm = &s->cm[c * 3];
        s->pop[c] = 0;
        s->mass[c] = 0.0;
        for (d = 0; d < dim; d++) vcm[d] = 0.0;
        for (d = 0; d < p->nspec; d++) s->spcnt[c * ORC_MAXSPEC + d] = 0;
        if (ccm) for (d = 0; d < dim; d++) cm[d] = 0.0;
        for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
            int pi = s->order[t], id = s->spid[pi];
            double m = p->sp[id].mass;
            s->pop[c]++;
            s->mass[c] += m;
            s->spcnt[c * ORC_MAXSPEC + id]++;
            for (d = 0; d < dim; d++) vcm[d] += s->V[pi * 3 + d] * m;
            if (ccm) for (d = 0; d < dim; d++) cm[d] += s->Q[pi * 3 + d] * m;
        }
        if (s->mass[c] > 0.0)
            for (d = 0; d < dim; d++) {
                vcm[d] /= s->mass[c];
                if (ccm) cm[d] /= s->mass[c];
            }
    }
    if (p->rtech == R_RAT || p->lc != LC_ISOF) {
        for (c = 0; c < s->nc; c++) {
            double *I = &s->I[c * 9];
            for (i = 0; i < 9; i++) I[i] = 0.0;
            for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
                int pi = s->order[t];
                double m = p->sp[s->spid[pi]].mass, q[3] = {0, 0, 0};
                for (d = 0; d < dim; d++) q[d] = s->Q[pi * 3 + d] - s->cm[c * 3 + d];
                I[0] += m * (q[1] * q[1] + q[2] * q[2]);
                I[4] += m * (q[0] * q[0] + q[2] * q[2]);
                I[8] += m * (q[0] * q[0] + q[1] * q[1]);
                I[1] += m * q[0] * q[1];
                I[2] += m * q[0] * q[2];
                I[5] += m * q[1] * q[2];
            }
            I[1] *= -1.; I[2] *= -1.; I[5] *= -1.;
            I[3] = I[1]; I[6] = I[2]; I[7] = I[5];
        }
    }
    /* with lc = 0 the order tensor is only needed by operators outside this path */
    if (p->lc != LC_ISOF) {
        for (c = 0; c < s->nc; c++)
            if (s->pop[c] > 1) {
                double Sm[3][3];
                tens_order(s, c, Sm);
                order_from_tensor(s, c, Sm);
            }
    }
    if (p->lc != LC_ISOF) vel_grad(s);
}

/* ------------------------------------------------------------------------ */
/* walls                                                                     */
/* ------------------------------------------------------------------------ */

/* smrtPow, mtools.c:38-50: integer powers up to 10 by repeated products */
static double smrt_pow(double x, double y)
{
    if ((y - (int)y == 0) && (y <= 10)) {
        int i;
        double r = 1;
        for (i = 0; i < y; i++) r *= x;
        return r;
    }
    return pow(x, y);
}

/* calcWavyW, bc.c:94-150, for the wall at wq */
static double calc_wavy_w(const orc_wall *W, const double *wq, const double *pos, double Wv, int dim)
{
    double W1 = 0.0, W2 = 0.0;
    int i, flag = 0;
    for (i = 0; i < dim; i++) if (!feq(W->P[i], 1.0)) flag += 1;
    if (!flag) {
        if (!feq(W->B[1], 0.0))
            W1 = (W->A[1] * pos[0] - W->A[0] * pos[1]) / sqrt(W->A[0] * W->A[0] + W->A[1] * W->A[1]);
        if (dim > 2 && !feq(W->B[2], 0.0)) {
            W2 = W->A[0] * W->A[2] * pos[0] + W->A[1] * W->A[2] * pos[1] -
                 (W->A[0] * W->A[0] + W->A[1] * W->A[1]) * pos[2];
            W2 /= sqrt(pow(W->A[0], 4) + pow(W->A[1], 4) + pow(W->A[0], 2) * pow(W->A[2], 2) +
                       2.0 * pow(W->A[0], 2) * pow(W->A[1], 2) + pow(W->A[1], 2) * pow(W->A[2], 2));
        }
    } else {
        flag = 0;
        for (i = 0; i < dim; i++) if (!feq(W->P[i], 2.0) || feq(W->A[i], 0.0)) flag += 1;
        if (!flag) {
            if (!feq(W->B[1], 0.0)) W1 = ref_atan2(W->A[1] * (pos[1] - wq[1]), W->A[0] * (pos[0] - wq[0]));
            if (dim > 2 && !feq(W->B[2], 0.0))
                W2 = ref_atan2(sqrt(pow(W->A[0] * (pos[0] - wq[0]), 2) + pow(W->A[1] * (pos[1] - wq[1]), 2)),
                               W->A[2] * (pos[2] - wq[2]));
        } else if (dim > 2 && flag == 1) {
            int f = -1;
            for (i = 0; i < dim; i++) if (feq(W->A[i], 0.0)) f = i;
            if (f != -1) {
                int j, k, l;
                if (f == 0) { j = 0; k = 1; l = 2; }
                else if (f == 1) { j = 1; k = 2; l = 0; }
                else { j = 2; k = 0; l = 1; }
                if (!feq(W->B[1], 0.0)) W1 = ref_atan2(W->A[l] * (pos[l] - wq[l]), W->A[k] * (pos[k] - wq[k]));
                if (!feq(W->B[2], 0.0)) W2 = ref_atan2(sqrt(Wv + pow(W->R, W->P[3])), sqrt(pos[j]));
            }
        }
    }
    return W->B[0] * cos(W->B[1] * W1) * cos(W->B[2] * W2);
}

/* calcW, bc.c:55-80 (= surf_func, mtools.c:1196-1220), four-fold symmetric
 * surfaces, for the wall at wq */
static double calc_w(const orc_wall *W, const double *wq, const double *pos, int dim)
{
    double terms, Wv = 0.0;
    int i;
    for (i = 0; i < dim; i++) {
        terms = W->A[i] * (pos[i] - wq[i]);
        if (W->ABS) terms = fabs(terms);
        terms = smrt_pow(terms, W->P[i]);
        Wv += terms;
    }
    terms = W->R;
    if (W->ABS) terms = fabs(terms);
    terms = smrt_pow(terms, W->P[3]);
    Wv -= terms;
    if (!feq(W->B[0], 0.0)) Wv += calc_wavy_w(W, wq, pos, Wv, dim);
    if (W->INV) Wv *= -1.0;
    return Wv;
}

/* calcW_PLANE, bc.c:306-317 */
static double calc_w_plane(const orc_wall *W, const double *wq, const double *pos, int dim)
{
    double Wv = 0.;
    int i;
    for (i = 0; i < dim; i++) Wv += W->A[i] * (pos[i] - wq[i]);
    Wv -= W->R;
    if (W->INV) Wv *= -1.;
    return Wv;
}

/* the wall's W at its current position: calcW (bc.c:55) */
static double wall_W(const orc_sys *s, int w, const double *q)
{
    return calc_w(&s->p.wall[w], s->wq[w], q, s->p.dim);
}

/* read.c:1423-1434: an axis-aligned plane that is not wavy */
static int wall_planar(const orc_wall *W)
{
    if (!feq(W->B[0], 0.0)) return 0;
    if (!(feq(W->P[0], 1.0) && feq(W->P[1], 1.0) && feq(W->P[2], 1.0))) return 0;
    if (fneq(W->A[0], 0.0) && feq(W->A[1], 0.0) && feq(W->A[2], 0.0)) return 1;
    if (feq(W->A[0], 0.0) && fneq(W->A[1], 0.0) && feq(W->A[2], 0.0)) return 1;
    if (feq(W->A[0], 0.0) && feq(W->A[1], 0.0) && fneq(W->A[2], 0.0)) return 1;
    return 0;
}

/* secant_time, bc.c:329-373: the time back along -V at which the surface
 * (wall at wq) was crossed */
static double secant_time(const orc_wall *W, const double *wq, const double *Q, const double *V, int dim,
                          double t_step)
{
    double Qi[3] = {0, 0, 0}, QiM1[3] = {0, 0, 0}, ti, tiM1, root, fi, fiM1;
    int i, iter = 0;
    for (i = 0; i < dim; i++) {
        QiM1[i] = Q[i];
        Qi[i] = Q[i] + t_step * (-V[i]);
    }
    tiM1 = 0.0;
    ti = t_step;
    root = ti;
    do {
        iter++;
        fi = calc_w(W, wq, Qi, dim);
        fiM1 = calc_w(W, wq, QiM1, dim);
        if (!feq(fi, fiM1)) root = ti - fi * (ti - tiM1) / (fi - fiM1);
        if (root < ORC_TOL / 100.) {
            ti = tiM1;
            tiM1 = root;
        } else {
            tiM1 = ti;
            ti = root;
        }
        for (i = 0; i < dim; i++) {
            QiM1[i] = Q[i] + tiM1 * (-V[i]);
            Qi[i] = Q[i] + ti * (-V[i]);
        }
    } while (iter < 10 && fabs(ti - tiM1) > ORC_TOL / 100. && fabs(fi - fiM1) > ORC_TOL / 100.);
    return root;
}

/* crosstimeReverse, bc.c:251-298 */
static void crosstime_reverse(const orc_wall *W, int planar, const double *wq, const double *Q, const double *V,
                              int dim, double t_step, double *tc_pos, double *tc_neg)
{
    double a = 0.0, b = 0.0, c = 0.0;
    int i;
    if (planar || (feq(W->P[0], 1.0) && feq(W->P[1], 1.0) && feq(W->P[2], 1.0) && feq(W->P[3], 1.0) &&
                   feq(W->B[0], 0.0))) {
        *tc_pos = W->R;
        for (i = 0; i < dim; i++) *tc_pos += W->A[i] * (wq[i] - Q[i]);
        *tc_pos /= (W->A[0] * (-V[0]) + W->A[1] * (-V[1]) + W->A[2] * (-V[2]));
        *tc_neg = *tc_pos;
    } else if ((dim == 2 && feq(W->P[0], 2.0) && feq(W->P[1], 2.0) && feq(W->B[0], 0.0)) ||
               (dim > 2 && feq(W->P[0], 2.0) && feq(W->P[1], 2.0) && feq(W->P[2], 2.0) && feq(W->B[0], 0.0))) {
        for (i = 0; i < dim; i++) {
            a += W->A[i] * W->A[i] * (-V[i]) * (-V[i]);
            b += W->A[i] * W->A[i] * (-V[i]) * (Q[i] - wq[i]);
            c += W->A[i] * W->A[i] * (Q[i] * Q[i] - 2. * Q[i] * wq[i] + wq[i] * wq[i]);
        }
        b *= 2.0;
        c -= smrt_pow(W->R, W->P[3]);
        *tc_neg = -b - sqrt(b * b - 4. * a * c);
        *tc_neg /= (2. * a);
        *tc_pos = -b + sqrt(b * b - 4. * a * c);
        *tc_pos /= (2. * a);
    } else {
        *tc_pos = secant_time(W, wq, Q, V, dim, t_step);
        if (*tc_pos < 0 || *tc_pos > t_step) *tc_pos = t_step;
        *tc_neg = *tc_pos;
    }
}

/* chooseT, bc.c:849-880 */
static double choose_t(double tstep, double tp, double tn)
{
    double tc = tstep, zero = -ORC_TOL, step = tstep + ORC_TOL;
    if (tp > step && tn > step) tc = 0.;
    else if (tp < zero && tn < zero) tc = 0.;
    else if (tp < zero) tc = tn;
    else if (tn < zero) tc = tp;
    else if (tn < tp) tc = tn;
    else tc = tp;
    if (tc < zero) tc = 0.;
    if (tc > step) tc = 0.;
    return tc;
}

/* shiftBC, bc.c:389-405: the periodic image of a curved wall nearest the
 * particle; the wall position is moved in place (and moved back by
 * shift_back_bc), as the reference does */
static void shift_bc(orc_sys *s, int w, const double *q, double *shift)
{
    const orc_wall *W = &s->p.wall[w];
    int k, dim = s->p.dim;
    for (k = 0; k < dim; k++) shift[k] = 0.0;
    if (fneq(W->P[0], 1.0) && fneq(W->P[1], 1.0)) {
        for (k = 0; k < dim; k++) {
            if (q[k] - s->wq[w][k] >= 0.5 * (double)s->p.L[k]) shift[k] = (double)s->p.L[k];
            else if (q[k] - s->wq[w][k] <= -0.5 * (double)s->p.L[k]) shift[k] = -1.0 * (double)s->p.L[k];
            s->wq[w][k] += shift[k];
        }
    }
}

static void shift_back_bc(orc_sys *s, int w, const double *shift)   /* shiftbackBC, bc.c:416-422 */
{
    int i;
    for (i = 0; i < s->p.dim; i++) s->wq[w][i] -= shift[i];
}

/* normalWavy, bc.c:1568-1640 */
static void wall_normal_wavy(const orc_wall *W, const double *wq, const double *point, int dimension, double *n)
{
    double W1 = 0.0, W2 = 0.0, div, dw1[3] = {0, 0, 0}, dw2[3] = {0, 0, 0};
    int i;
    if (feq(W->P[0], 1.0) && feq(W->P[1], 1.0) && feq(W->P[2], 1.0)) {
        if (!feq(W->B[1], 0.0)) {
            W1 = (W->A[1] * point[0] - W->A[0] * point[1]) / sqrt(W->A[0] * W->A[0] + W->A[1] * W->A[1]);
            dw1[0] = W->A[1] / sqrt(W->A[0] * W->A[0] + W->A[1] * W->A[1]);
            dw1[1] = -W->A[0] / sqrt(W->A[0] * W->A[0] + W->A[1] * W->A[1]);
            dw1[2] = 0;
        }
        if (dimension > 2 && !feq(W->B[2], 0.0)) {
            div = sqrt(pow(W->A[0], 4) + pow(W->A[1], 4) + pow(W->A[0], 2) * pow(W->A[2], 2) +
                       2.0 * pow(W->A[0], 2) * pow(W->A[1], 2) + pow(W->A[1], 2) * pow(W->A[2], 2));
            W2 = (W->A[0] * W->A[2] * point[0] + W->A[1] * W->A[2] * point[1] -
                  (W->A[0] * W->A[0] + W->A[1] * W->A[1]) * point[2]) / div;
            dw2[0] = (W->A[0] * W->A[2]) / div;
            dw2[1] = (W->A[1] * W->A[2]) / div;
            dw2[2] = -(W->A[0] * W->A[0] + W->A[1] * W->A[1]) / div;
        }
    } else {
        int flag = 0;
        for (i = 0; i < dimension; i++) if (!feq(W->P[i], 2.0) || feq(W->A[i], 0.0)) flag += 1;
        if (!flag) {
            div = pow(W->A[0] * (point[0] - wq[0]), 2) + pow(W->A[1] * (point[1] - wq[1]), 2);
            W1 = ref_atan2(W->A[1] * (-wq[1] + point[1]), W->A[0] * (-wq[0] + point[0]));
            if (!feq(W->B[1], 0.0)) {
                dw1[0] = W->A[0] * W->A[1] * (wq[1] - point[1]) / div;
                dw1[1] = W->A[0] * W->A[1] * (point[0] - wq[0]) / div;
                dw1[2] = 0;
            }
            if (dimension > 2 && !feq(W->B[2], 0.0)) {
                dw2[0] = W->A[2] * W->A[0] * (point[0] - wq[0]) * (point[2] - wq[2]) /
                         (sqrt(div) * (div + pow(W->A[2] * (point[2] - wq[2]), 2)));
                dw2[1] = W->A[2] * W->A[1] * (point[1] - wq[1]) * (point[2] - wq[2]) /
                         (sqrt(div) * (div + pow(W->A[2] * (point[2] - wq[2]), 2)));
                dw2[2] = -W->A[2] * sqrt(div) / (div + pow(W->A[2] * (point[2] - wq[2]), 2));
            }
        } else if (dimension > 2 && flag == 1) {
            int f = -1;
            for (i = 0; i < dimension; i++) if (feq(W->A[i], 0.0)) f = i;
            if (f != -1) {
                int j, k, l;
                if (f == 0) { j = 0; k = 1; l = 2; }
                else if (f == 1) { j = 1; k = 2; l = 0; }
                else { j = 2; k = 0; l = 1; }
                div = pow(W->A[k] * (point[k] - wq[k]), 2) + pow(W->A[l] * (point[l] - wq[l]), 2);
                if (!feq(W->B[1], 0.0)) {
                    dw1[k] = W->A[k] * W->A[l] * (wq[l] - point[l]) / div;
                    dw1[l] = W->A[k] * W->A[l] * (point[k] - wq[k]) / div;
                    dw1[j] = 0;
                }
                if (!feq(W->B[2], 0.0)) {
                    dw2[k] = W->A[k] * W->A[k] * (point[k] - wq[k]) * point[j] / (sqrt(div) * (div + point[k]));
                    dw2[l] = W->A[l] * W->A[l] * (point[l] - wq[l]) * point[j] / (sqrt(div) * (div + point[l]));
                    dw2[j] = -sqrt(div) / (div + point[j]);
                }
            }
        }
    }
    for (i = 0; i < 3; i++)
        n[i] -= W->B[0] * (W->B[1] * dw1[i] * sin(W->B[1] * W1) * cos(W->B[2] * W2) +
                           W->B[2] * dw2[i] * cos(W->B[1] * W1) * sin(W->B[2] * W2));
}

/* normal, bc.c:1516-1556 (four-fold symmetric surfaces), at point, wall at wq */
static void wall_normal(const orc_wall *W, const double *wq, const double *point, int dimension, double *n)
{
    int i;
    if (feq(W->P[0], 1.0) && feq(W->P[1], 1.0) && feq(W->P[2], 1.0))
        for (i = 0; i < dimension; i++) n[i] = W->A[i];
    else if ((dimension == 2 && feq(W->P[0], 2.0) && feq(W->P[1], 2.0)) ||
             (dimension > 2 && feq(W->P[0], 2.0) && feq(W->P[1], 2.0) && feq(W->P[2], 2.0)))
        for (i = 0; i < dimension; i++) n[i] = 2.0 * W->A[i] * W->A[i] * (point[i] - wq[i]);
    else if ((dimension == 2 && feq(W->P[0], 4.0) && feq(W->P[1], 4.0)) ||
             (dimension > 2 && feq(W->P[0], 4.0) && feq(W->P[1], 4.0) && feq(W->P[2], 4.0)))
        for (i = 0; i < dimension; i++)
            n[i] = 4.0 * W->A[i] * W->A[i] * W->A[i] * W->A[i] * (point[i] - wq[i]) * (point[i] - wq[i]) *
                   (point[i] - wq[i]);
    else if ((dimension == 2 && feq(W->P[0], 6.0) && feq(W->P[1], 6.0)) ||
             (dimension > 2 && feq(W->P[0], 6.0) && feq(W->P[1], 6.0) && feq(W->P[2], 6.0)))
        for (i = 0; i < dimension; i++) {
            n[i] = 4.0 * W->A[i] * W->A[i] * W->A[i] * W->A[i] * W->A[i] * W->A[i];
            n[i] *= (point[i] - wq[i]) * (point[i] - wq[i]) * (point[i] - wq[i]) * (point[i] - wq[i]) *
                    (point[i] - wq[i]);
        }
    else if (W->ABS)
        for (i = 0; i < dimension; i++)
            n[i] = W->A[i] * W->A[i] * W->P[i] * (point[i] - wq[i]) *
                   smrt_pow(fabs(W->A[i] * (point[i] - wq[i])), W->P[i] - 2.0);
    else
        for (i = 0; i < dimension; i++)
            n[i] = W->P[i] * smrt_pow(W->A[i], W->P[i]) * smrt_pow(point[i] - wq[i], W->P[i] - 1.0);
    if (!feq(W->B[0], 0.0)) wall_normal_wavy(W, wq, point, dimension, n);
}

static int wall_anchoring(const orc_wall *W) { return feq(W->MUN, 0.0) || feq(W->MUT, 0.0); }

/* oriBC, lc.c:1290-1393 (immobile walls) */
static void ori_bc(orc_sys *s, int pi, const orc_wall *W, const double *n)
{
    int dim = s->p.dim, i;
    double *u = &s->U[pi * 3];
    double UN[3] = {0, 0, 0}, UT[3] = {0, 0, 0};
    vnorm(u, dim);
    if (dotprod(u, n, dim) >= 0.0) { for (i = 0; i < dim; i++) u[i] *= 1.; }
    else { for (i = 0; i < dim; i++) u[i] *= -1.; }
    proj(u, n, UN, dim);
    tang(u, UN, UT, dim);
    if (feq(W->MUN, 0.0) && feq(W->MUT, 0.0)) {
        for (i = 0; i < dim; i++) u[i] = W->DUxyz[i];
    } else {
        for (i = 0; i < dim; i++) { UN[i] *= W->MUN; UT[i] *= W->MUT; }
        for (i = 0; i < dim; i++) u[i] = UN[i] + UT[i];
        u[0] *= W->MUxyz[0];
        if (dim >= 2) u[1] *= W->MUxyz[1];
        if (dim >= 3) u[2] *= W->MUxyz[2];
    }
    vnorm(u, dim);
}

/* ghostPart, mpc.c:351-567, for immobile planar walls */
static void ghost_part(orc_sys *s, orc_draws *d)
{
    const orc_params *p = &s->p;
    int dim = p->dim, a, b, c, j, w, k, nco = (dim == 3) ? 8 : 4;
    int X = p->L[0], Y = p->L[1], Zmax = (dim == 3) ? p->L[2] : 0;
    double tq[8][3];
    for (a = 0; a <= X; a++)
        for (b = 0; b <= Y; b++)
            for (c = 0; c <= Zmax; c++) {
                int ce = cidx(s, a, b, c);
                int numBC = 0, flagW = 0, wallindex = -1;
                for (j = 0; j < 8; j++) {
                    tq[j][0] = (double)a + (double)(j & 1);
                    tq[j][1] = (double)b + (double)((j >> 1) & 1);
                    tq[j][2] = (double)c + (double)((j >> 2) & 1);
                }
                if (s->pop[ce] <= 0) continue;
                for (w = 0; w < p->nwall; w++)
                    if (p->wall[w].PHANTOM && wall_anchoring(&p->wall[w])) {
                        int fw = 0;
                        for (j = 0; j < nco; j++)
                            if (wall_W(s, w, tq[j]) < 0.0) { fw = 1; wallindex = w; }
                        if (fw) numBC++;
                    }
                flagW = 0;
                for (w = 0; w < p->nwall; w++)
                    if (p->wall[w].PHANTOM) {
                        const orc_wall *W = &p->wall[w];
                        for (j = 0; j < nco; j++)
                            if (wall_W(s, w, tq[j]) < 0.0) flagW = 1;
                        if (p->lc != LC_ISOF && flagW && numBC == 1 && wallindex == w && wall_anchoring(W)) {
                            int t;
                            for (t = s->cstart[ce]; t < s->cstart[ce + 1]; t++) {
                                int pi = s->order[t];
                                if (W->inter[s->spid[pi]] == 1) {
                                    double n[3] = {0, 0, 0};
                                    wall_normal(W, s->wq[w], &s->Q[pi * 3], dim, n);   /* normal, bc.c:1516 */
                                    vnorm(n, dim);
                                    ori_bc(s, pi, W, n);
                                }
                            }
                            if ((feq(W->P[0], 1.0) && feq(W->P[1], 1.0) && feq(W->P[2], 1.0)) || s->pop[ce] == 1) {
                                /* a plane, or one particle: S = 1, director = first particle's orientation */
                                s->S[ce] = 1.0;
                                for (k = 0; k < dim; k++)
                                    s->dir[ce * 3 + k] = s->U[s->order[s->cstart[ce]] * 3 + k];
                                vnorm(&s->dir[ce * 3], dim);
                            } else if (s->pop[ce] > 1) {
                                /* a curved wall: the order tensor of the anchored orientations */
                                double Sm[3][3];
                                const int rec = s->rec_cdec;
                                s->rec_cdec = 0;
                                tens_order(s, ce, Sm);
                                order_from_tensor(s, ce, Sm);
                                s->rec_cdec = rec;
                            }
                        }
                    }
                if (flagW && s->pop[ce] < p->gndnst && s->pop[ce] > 0) {
                    double invN = 1.0 / (p->gndnst - (double)s->pop[ce]);
                    for (k = 0; k < dim; k++) {
                        double R = dr_gaussMB(s, d, p->kbt, invN, ORC_P_GHOST, ce, k);
                        s->vcm[ce * 3 + k] += R;
                        s->vcm[ce * 3 + k] /= p->gndnst;
                    }
                }
            }
}

/* ------------------------------------------------------------------------ */
/* nematic LC                                                                */
/* ------------------------------------------------------------------------ */

/* genrand_maierSaupeMetropolis_2D, lc.c:341-411 */
/* the decision word of a Maier-Saupe draw: bit 0 set, bits 1-2 the sampler,
 * bit 3 the 2D redraw, bit 4 pm = -1, bits 5-12 annealNum, bits 13-20 the
 * accepted moves (the GPU library's lc_decision_word) */
static uint32_t dec_word(int sampler, int redraw, int neg, int anneal, int accepted)
{
    uint32_t a = (uint32_t)(anneal < 0 ? 0 : anneal > 255 ? 255 : anneal);
    uint32_t k = (uint32_t)(accepted < 0 ? 0 : accepted > 255 ? 255 : accepted);
    return 1u | ((uint32_t)sampler << 1) | (redraw ? 8u : 0u) | (neg ? 16u : 0u) | (a << 5) | (k << 13);
}

static uint32_t ms_metropolis_2d(orc_sys *s, orc_draws *d, long slot, const double *DIR, double *rotAx,
                                 double rotAngle, double *U, double KBT, double S, double effM)
{
    double newU[3] = {0, 0, 0}, R, ang0, uz0, w0, ang1, uz1, w1, dw;
    int i, dd, annealNum, accept, cnt = 0, redraw;
    for (dd = 0; dd < 2; dd++) newU[dd] = U[dd];
    ang0 = signedAngle(DIR, U, 2);
    redraw = feq(ang0, 0.5 * ORC_PI) || feq(ang0, -0.5 * ORC_PI);
    if (redraw) ang0 = ORC_PI * (0.5 - dr_uniform(s, d, ORC_P_LC, slot, 0));
    uz0 = dotprod(DIR, U, 2);
    annealNum = (int)(ORC_MCINT + ORC_MCSLOPE * effM * S / KBT);
    for (i = 0; i < annealNum; i++) {
        accept = 0;
        w0 = effM * S * uz0 * uz0 / KBT;
        ang1 = ORC_PI * (0.5 - dr_uniform(s, d, ORC_P_LC, slot, 1 + 2 * i));
        uz1 = cos(ang1);
        w1 = effM * S * uz1 * uz1 / KBT;
        dw = w1 - w0;
        if (dw >= 0.) accept = 1;
        else {
            R = dr_uniform(s, d, ORC_P_LC, slot, 2 + 2 * i);
            if (R <= exp(dw)) accept = 1;
        }
        if (accept) {
            cnt += 1;
            ang0 = ang1;
            uz0 = uz1;
        }
    }
    newU[0] = uz0;
    newU[1] = sin(ang0);
    if (cnt > 0) {
        orc_rodrigues(newU, rotAx, rotAngle);
        for (dd = 0; dd < 2; dd++) U[dd] = newU[dd];
    }
    return dec_word(0, redraw, 0, annealNum, cnt);
}

/* genrand_maierSaupeMetropolis_3D, lc.c:257-324 */
static uint32_t ms_metropolis_3d(orc_sys *s, orc_draws *d, long slot, const double *DIR, double *rotAx,
                                 double rotAngle, double *U, double KBT, double S, double effM)
{
    double newU[3], phi, R, uz0, w0, uz1, w1, dw, pm;
    int i, dd, annealNum, accept, cnt = 0;
    for (dd = 0; dd < 3; dd++) newU[dd] = U[dd];
    uz0 = dotprod(DIR, U, 3);
    annealNum = (int)(ORC_MCINT + ORC_MCSLOPE * effM * S / KBT);
    for (i = 0; i < annealNum; i++) {
        accept = 0;
        w0 = effM * S * uz0 * uz0 / KBT;
        uz1 = dr_uniform(s, d, ORC_P_LC, slot, 1 + 2 * i);
        w1 = effM * S * uz1 * uz1 / KBT;
        dw = w1 - w0;
        if (dw >= 0.) accept = 1;
        else {
            R = dr_uniform(s, d, ORC_P_LC, slot, 2 + 2 * i);
            if (R <= exp(dw)) accept = 1;
        }
        if (accept) {
            cnt += 1;
            uz0 = uz1;
        }
    }
    phi = dr_uniform(s, d, ORC_P_LC, slot, 1 + 2 * annealNum) * 2.0 * ORC_PI;
    pm = (dr_uniform(s, d, ORC_P_LC, slot, 2 + 2 * annealNum) <= 0.5) ? -1.0 : 1.0;
    newU[0] = pm * uz0;
    newU[1] = sqrt(1.0 - uz0 * uz0) * cos(phi);
    newU[2] = sqrt(1.0 - uz0 * uz0) * sin(phi);
    if (cnt > 0) {
        orc_rodrigues(newU, rotAx, rotAngle);
        for (dd = 0; dd < 3; dd++) U[dd] = newU[dd];
    }
    return dec_word(3, 0, pm < 0.0, annealNum, cnt);
}

/* genrand_maierSaupeGAUSS_3D, lc.c:85-126 */
static uint32_t ms_gauss_3d(orc_sys *s, orc_draws *d, long slot, double *rotAx, double rotAngle, double *U,
                            double KBT, double S, double effM)
{
    double std, phi, theta, uz, pm, newU[3];
    int dd;
    for (dd = 0; dd < 3; dd++) newU[dd] = U[dd];
    std = 1. / sqrt(2. * effM * S / KBT);
    theta = std * dr_normal(s, d, ORC_P_LC, slot, 0) + 0.0;
    uz = 1. - 0.5 * theta * theta;
    pm = (dr_uniform(s, d, ORC_P_LC, slot, 1) <= 0.5) ? -1.0 : 1.0;
    newU[0] = pm * uz;
    phi = dr_uniform(s, d, ORC_P_LC, slot, 2) * 2.0 * ORC_PI;
    newU[1] = sqrt(1.0 - uz * uz) * cos(phi);
    newU[2] = sqrt(1.0 - uz * uz) * sin(phi);
    orc_rodrigues(newU, rotAx, rotAngle);
    for (dd = 0; dd < 3; dd++) U[dd] = newU[dd];
    return dec_word(1, 0, pm < 0.0, 0, 0);
}

/* genrand_maierSaupeEXP_3D, lc.c:192-241 */
static uint32_t ms_exp_3d(orc_sys *s, orc_draws *d, long slot, double *rotAx, double rotAngle, double *U,
                          double KBT, double S, double effM)
{
    double BUS, c, R, num, denom, phi, uz, pm, newU[3];
    int dd;
    for (dd = 0; dd < 3; dd++) newU[dd] = U[dd];
    BUS = effM * S / KBT;
    if (BUS < ORC_TOL) BUS = ORC_TOL;
    R = dr_uniform(s, d, ORC_P_LC, slot, 0) * (1. + BUS / 3.);
    c = sqrt(9. * BUS * BUS * BUS * BUS * R * R + 4. * BUS * BUS * BUS);
    num = pow(3. * BUS * BUS * R + c, 1. / 3.);
    denom = pow(2., 1. / 3.);
    uz = num / (denom * BUS) - denom / num;
    pm = (dr_uniform(s, d, ORC_P_LC, slot, 1) <= 0.5) ? -1.0 : 1.0;
    newU[0] = pm * uz;
    phi = dr_uniform(s, d, ORC_P_LC, slot, 2) * 2.0 * ORC_PI;
    newU[1] = sqrt(1.0 - uz * uz) * cos(phi);
    newU[2] = sqrt(1.0 - uz * uz) * sin(phi);
    orc_rodrigues(newU, rotAx, rotAngle);
    for (dd = 0; dd < 3; dd++) U[dd] = newU[dd];
    return dec_word(2, 0, pm < 0.0, 0, 0);
}

/* LCcollision, lc.c:427-544 */
static void lc_collision(orc_sys *s, orc_draws *d, int c, double SG)
{
    const orc_params *p = &s->p;
    int dim = p->dim, i, t;
    double DIR[3] = {0, 0, 0}, dU[3] = {0, 0, 0}, rotAx[3], xaxis[3] = {1.0, 0.0, 0.0}, rotAngle;
    double S = (p->lc == LC_LCL) ? s->S[c] : SG;
    double avMFPOT = 0.0, effM;
    for (i = 0; i < dim; i++) DIR[i] = s->dir[c * 3 + i];
    crossprod(DIR, xaxis, rotAx);
    rotAngle = -1. * absAngle(DIR, xaxis, dim);
    vnorm(rotAx, 3);
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) avMFPOT += p->sp[s->spid[s->order[t]]].mfpot;
    if (s->pop[c] > 1) avMFPOT /= (float)(s->pop[c]);
    effM = avMFPOT * 0.5 * (double)dim;
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t];
        double *U = &s->U[pi * 3], *T = &s->T[pi * 3];
        double rfc = p->sp[s->spid[pi]].rfc;
        for (i = 0; i < 3; i++) T[i] = 0.;
        for (i = 0; i < dim; i++) dU[i] = U[i];
        if (dim == 3) {
            double BUS = effM * S / p->kbt;
            if (BUS > ORC_BUSMAX) s->pdec[pi] = ms_gauss_3d(s, d, t, rotAx, rotAngle, U, p->kbt, S, effM);
            else if (BUS < ORC_BUSMIN) s->pdec[pi] = ms_exp_3d(s, d, t, rotAx, rotAngle, U, p->kbt, S, effM);
            else s->pdec[pi] = ms_metropolis_3d(s, d, t, DIR, rotAx, rotAngle, U, p->kbt, S, effM);
        } else {
            s->pdec[pi] = ms_metropolis_2d(s, d, t, DIR, rotAx, rotAngle, U, p->kbt, S, effM);
        }
        for (i = 0; i < dim; i++) dU[i] = (U[i] - dU[i]) / p->dt;
        crossprod(U, dU, T);
        for (i = 0; i < 3; i++) T[i] *= rfc;
    }
}

/* jefferysTorque + larsonRotRate, lc.c:559-718 */
static void jeffery(orc_sys *s, int c)
{
    const orc_params *p = &s->p;
    const double *E = &s->E[c * 9];
    double D[3][3], W[3][3];
    int i, j, k, t;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) {
            D[j][i] = 0.5 * (E[i * 3 + j] + E[j * 3 + i]);
            W[j][i] = 0.5 * (E[i * 3 + j] - E[j * 3 + i]);
        }
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t];
        const orc_species *sp = &p->sp[s->spid[pi]];
        double u[3], dudt[3], w[3], uuuD[3] = {0, 0, 0}, uw[3] = {0, 0, 0}, uD[3] = {0, 0, 0}, theta;
        for (i = 0; i < 3; i++) u[i] = s->U[pi * 3 + i];
        for (j = 0; j < 3; j++) {
            for (i = 0; i < 3; i++) {
                for (k = 0; k < 3; k++) uuuD[j] += u[j] * (u[i] * D[i][k] * u[k]);
                uD[j] += u[i] * D[i][j];
                uw[j] += u[i] * W[i][j];
            }
            dudt[j] = uw[j] + sp->tumble * (uD[j] - uuuD[j]);
        }
        crossprod(u, dudt, w);
        for (i = 0; i < 3; i++) w[i] *= sp->chihi;
        for (i = 0; i < 3; i++) w[i] *= p->dt;
        theta = length(w, 3);
        vnorm(w, 3);
        orc_rodrigues(&s->U[pi * 3], w, theta);
    }
}

/* avOrderParam, lc.c:982-1039 (global S for lc = 2) */
static double av_order_param(orc_sys *s)
{
    int dim = s->p.dim, i, j, k;
    double Sm[3][3] = {{0}}, U[3] = {0, 0, 0}, fDIM = (double)dim, cc = 1. / (fDIM - 1.);
    double GPOPinv = 1. / ((double)s->N), ev[3] = {0, 0, 0}, avS;
    for (i = 0; i < s->N; i++) {
        for (j = 0; j < dim; j++) U[j] = s->U[i * 3 + j];
        for (j = 0; j < dim; j++)
            for (k = j; k < dim; k++) Sm[j][k] += U[j] * U[k];
    }
    for (i = 0; i < dim; i++)
        for (j = i; j < dim; j++) Sm[i][j] *= cc * fDIM * GPOPinv;
    for (i = 0; i < dim; i++) Sm[i][i] -= cc;
    for (i = 0; i < dim; i++)
        for (j = i + 1; j < dim; j++) Sm[j][i] = Sm[i][j];
    if (dim == 3) {
        eig3(Sm, ev);
        avS = -1. * (ev[1] + ev[2]);
    } else {
        double a[2][2] = {{Sm[0][0], Sm[0][1]}, {Sm[1][0], Sm[1][1]}};
        eig2(a, ev);
        Sm[0][0] = a[0][0]; Sm[0][1] = a[0][1];
        avS = ev[0];
    }
    for (i = 0; i < dim; i++) s->avdir[i] = Sm[0][i];
    return avS;
}

/* ------------------------------------------------------------------------ */
/* collision operators                                                       */
/* ------------------------------------------------------------------------ */

/* rotate + stochrotMPC, mpc.c:1181-1333 */
static void srd_collide(orc_sys *s, orc_draws *d, int c)
{
    const orc_params *p = &s->p;
    int dim = p->dim, i, j, t, CA = 0;
    double RV[3] = {0., 0., 0.}, R[3][3], C = cos(p->rotang), Sn, sgnu;
    const double *vcm = &s->vcm[c * 3];
    /* in->C and in->S are cos(RA), sin(RA) (init.c:1194-1195) */
    double Cc = C, Ss = sin(p->rotang);
    if (p->rtech == R_ARBAXIS) {
        if (dim == 3) dr_unitvec(s, d, ORC_P_CELL, c, 0, RV);
        else { double x = dr_uniform(s, d, ORC_P_CELL, c, 0); RV[0] = x; RV[1] = sqrt(1. - x * x); }
    }
    if (p->rtech == R_ORTHAXIS) CA = (int)((int)dim * dr_uniform(s, d, ORC_P_CELL, c, 0));
    sgnu = dr_uniform(s, d, ORC_P_CELL, c, 3);
    /* mpc.c:1291: SIGN is a long holding genrand_real(), hence always 0 and
     * the angle always -RA, unless the true-random-sign option is set */
    if (p->true_random_sign ? (sgnu >= 0.5) : ((long)sgnu > 0.5)) Sn = Ss;
    else Sn = -1. * Ss;
    if (dim == 3) {
        const double *L = RV;
        if (p->rtech == R_ARBAXIS) {
            R[0][0] = L[0] * L[0] + (1. - L[0] * L[0]) * Cc;
            R[0][1] = L[0] * L[1] * (1. - Cc) - L[2] * Sn;
            R[0][2] = L[0] * L[2] * (1. - Cc) + L[1] * Sn;
            R[1][0] = L[0] * L[1] * (1. - Cc) + L[2] * Sn;
            R[1][1] = L[1] * L[1] + (1. - L[1] * L[1]) * Cc;
            R[1][2] = L[1] * L[2] * (1. - Cc) - L[0] * Sn;
            R[2][0] = L[0] * L[2] * (1. - Cc) - L[1] * Sn;
            R[2][1] = L[1] * L[2] * (1. - Cc) + L[0] * Sn;
            R[2][2] = L[2] * L[2] + (1. - L[2] * L[2]) * Cc;
        } else if (CA == 0) {
            /* mpc.c:1208-1216, including R[2][1] = -.1*S */
            R[0][0] = 1.; R[0][1] = 0.; R[0][2] = 0.;
            R[1][0] = 0.; R[1][1] = Cc; R[1][2] = Sn;
            R[2][0] = 0.; R[2][1] = -.1 * Sn; R[2][2] = Cc;
        } else if (CA == 1) {
            R[0][0] = Cc; R[0][1] = 0.; R[0][2] = Sn;
            R[1][0] = 0.; R[1][1] = 1.; R[1][2] = 0.;
            R[2][0] = -1. * Sn; R[2][1] = 0.; R[2][2] = Cc;
        } else {
            R[0][0] = Cc; R[0][1] = Sn; R[0][2] = 0.;
            R[1][0] = -1. * Sn; R[1][1] = Cc; R[1][2] = 0.;
            R[2][0] = 0.; R[2][1] = 0.; R[2][2] = 1.;
        }
    } else {
        R[0][0] = Cc; R[0][1] = -1. * Sn;
        R[1][0] = Sn; R[1][1] = Cc;
    }
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t];
        double V[3], TEMP[3] = {0.0, 0.0, 0.0};
        for (j = 0; j < dim; j++) V[j] = s->V[pi * 3 + j] - vcm[j];
        for (i = 0; i < dim; i++)
            for (j = 0; j < dim; j++) TEMP[i] += R[i][j] * V[j];
        for (j = 0; j < dim; j++) s->V[pi * 3 + j] = vcm[j] + TEMP[j];
    }
}

/* andersenMPC, mpc.c:1353-1444 */
static void at_collide(orc_sys *s, orc_draws *d, int c)
{
    const orc_params *p = &s->p;
    int dim = p->dim, j, t, n = s->pop[c];
    double RS[3] = {0, 0, 0};
    const double *vcm = &s->vcm[c * 3];
    double *RV = (double *)calloc((size_t)n * 3, sizeof(double));
    double *DV = (double *)calloc((size_t)n * 3, sizeof(double));
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t], i = t - s->cstart[c];
        const orc_species *sp = &p->sp[s->spid[pi]];
        for (j = 0; j < dim; j++) RV[i * 3 + j] = dr_gaussMB(s, d, p->kbt, sp->mass, ORC_P_NORMAL, t, j);
        for (j = 0; j < dim; j++) RS[j] += sp->mass * RV[i * 3 + j];
        for (j = 0; j < dim; j++) DV[i * 3 + j] = (sp->damp) * (vcm[j]) / ((double)n);
    }
    for (j = 0; j < dim; j++) RS[j] /= s->mass[c];
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t], i = t - s->cstart[c];
        for (j = 0; j < dim; j++) s->V[pi * 3 + j] = vcm[j] + RV[i * 3 + j] - RS[j] - DV[i * 3 + j];
    }
    free(RV);
    free(DV);
}

static void inertia_inverse(const orc_sys *s, int c, double II[3][3])
{
    double Im[3][3];
    int i, j;
    for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) II[i][j] = 0.;
    if (s->p.dim == 3) {
        if (s->pop[c] >= 3) {
            for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) Im[i][j] = s->I[c * 9 + i * 3 + j];
            invert3x3(II, Im);
        }
    } else {
        II[2][2] = 1. / s->I[c * 9 + 8];
    }
}

/* andersenROT (mpc.c:1461-1657), andersenROT_LC (lc.c:1515-1728) and
 * dipoleAndersenROT_LC (lc.c:1749-2043): Andersen thermostat with angular
 * momentum conservation; `mode` 0 = plain RAT, 1 = LC, 2 = LC + activity */
static void rat_collide(orc_sys *s, orc_draws *d, int c, int mode)
{
    const orc_params *p = &s->p;
    int dim = p->dim, j, t, n = s->pop[c], i;
    const double *vcm = &s->vcm[c * 3], *cm = &s->cm[c * 3];
    double RS[3] = {0, 0, 0}, AS[3] = {0, 0, 0}, L[3] = {0, 0, 0}, Llm[3] = {0, 0, 0}, Lnm[3] = {0, 0, 0};
    double W[3] = {0, 0, 0}, VCM[3], II[3][3], ACT = 0., sigW = 0., sigP = 0.;
    double PA[3] = {0, 0, 0}, PQ[3] = {0, 0, 0};
    double *RV = (double *)calloc((size_t)n * 3, sizeof(double));
    double *DV = (double *)calloc((size_t)n * 3, sizeof(double));
    double *AV = (double *)calloc((size_t)n * 3, sizeof(double));
    double *relQ = (double *)calloc((size_t)n * 3, sizeof(double));
    for (j = 0; j < 3; j++) VCM[j] = vcm[j];
    if (mode == 2) {
        for (j = 0; j < dim; j++) { PA[j] = s->dir[c * 3 + j]; PQ[j] = cm[j]; }
        for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
            int id = s->spid[s->order[t]];
            const orc_species *sp = &p->sp[id];
            if (((double)sp->minactratio == 0.0) || (((double)sp->minactratio) * p->gndnst < (double)s->spcnt[c * ORC_MAXSPEC + id])) {
                ACT += (double)sp->act / (double)sp->mass;
                sigW += (double)sp->sigwidth;
                sigP += (double)sp->sigpos;
            }
        }
        if (p->rtech == R_DIPOLE_DIR_AV || p->rtech == R_DIPOLE_DIR_SIG) ACT *= p->gndnst / ((double)n);
        if (p->rtech == R_DIPOLE_DIR_SIG || p->rtech == R_DIPOLE_DIR_SIG_SUM) {
            double fall;
            sigW /= (double)n;
            sigP /= (double)n;
            fall = (1 - tanh(((double)n - p->gndnst * (1 + sigP)) / (p->gndnst * sigW)));
            ACT *= fall / 2;
        }
        ACT *= p->dt;
    }
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t];
        const orc_species *sp = &p->sp[s->spid[pi]];
        i = t - s->cstart[c];
        for (j = 0; j < dim; j++) RV[i * 3 + j] = dr_gaussMB(s, d, p->kbt, sp->mass, ORC_P_NORMAL, t, j);
        for (j = 0; j < dim; j++) RS[j] += sp->mass * RV[i * 3 + j];
        if (mode != 1)
            for (j = 0; j < dim; j++) DV[i * 3 + j] = (sp->damp) * (vcm[j]);
        if (mode == 2 && fneq(ACT, 0.0)) {
            /* calcW of the plane through CM with normal DIR, lc.c:1791-1863 */
            double pW = 0.0, pm;
            for (j = 0; j < dim; j++) pW += PA[j] * (s->Q[pi * 3 + j] - PQ[j]);
            pW -= 0.0;
            pm = (pW <= 0) ? -1. : 1.;
            if (pm > 0.0) s->pdec[pi] |= 1u << 21;
            for (j = 0; j < dim; j++) AV[i * 3 + j] = pm * ACT * PA[j];
            for (j = 0; j < dim; j++) AS[j] += sp->mass * AV[i * 3 + j];
        }
    }
    for (j = 0; j < dim; j++) RS[j] /= s->mass[c];
    if (mode == 2) for (j = 0; j < dim; j++) AS[j] /= s->mass[c];
    inertia_inverse(s, c, II);
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t];
        const orc_species *sp = &p->sp[s->spid[pi]];
        double diffV[3] = {0, 0, 0}, am[3];
        i = t - s->cstart[c];
        for (j = 0; j < dim; j++) relQ[i * 3 + j] = s->Q[pi * 3 + j] - cm[j];
        if (mode == 2) for (j = 0; j < dim; j++) diffV[j] = sp->mass * (s->V[pi * 3 + j] - RV[i * 3 + j] - AV[i * 3 + j]);
        else for (j = 0; j < dim; j++) diffV[j] = sp->mass * (s->V[pi * 3 + j] - RV[i * 3 + j]);
        if (dim < 3) { relQ[i * 3 + 2] = 0.; diffV[2] = 0.; }
        crossprod(&relQ[i * 3], diffV, am);
        if (mode == 0) for (j = 0; j < 3; j++) L[j] += am[j];
        else {
            for (j = 0; j < 3; j++) Llm[j] += am[j];
            for (j = 0; j < 3; j++) Lnm[j] += s->T[pi * 3 + j] * p->dt;
        }
    }
    if (mode != 0) for (j = 0; j < 3; j++) L[j] = Llm[j] + Lnm[j];
    matvec3(II, L, W);
    for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
        int pi = s->order[t];
        double at[3];
        i = t - s->cstart[c];
        crossprod(W, &relQ[i * 3], at);
        for (j = 0; j < dim; j++) {
            if (mode == 0) s->V[pi * 3 + j] = VCM[j] + RV[i * 3 + j] - RS[j] - DV[i * 3 + j] + at[j];
            else if (mode == 1) s->V[pi * 3 + j] = VCM[j] + RV[i * 3 + j] - RS[j] + at[j];
            else s->V[pi * 3 + j] = VCM[j] + RV[i * 3 + j] - RS[j] + at[j] + AV[i * 3 + j] - AS[j] - DV[i * 3 + j];
        }
    }
    free(RV); free(DV); free(AV); free(relQ);
}

static int is_dipole_lc(int r)
{
    return r == R_DIPOLE_DIR_SUM || r == R_DIPOLE_DIR_AV || r == R_DIPOLE_DIR_SIG || r == R_DIPOLE_DIR_SIG_SUM;
}

/* MPCcollision switch, mpc.c:2777-2812 */
static int mpc_collision(orc_sys *s, orc_draws *d, int c)
{
    const orc_params *p = &s->p;
    if (p->lc != LC_ISOF) {
        if (is_dipole_lc(p->rtech)) rat_collide(s, d, c, 2);
        else rat_collide(s, d, c, 1);
        return 0;
    }
    if (p->rtech == R_MPCAT) at_collide(s, d, c);
    else if (p->rtech == R_RAT) rat_collide(s, d, c, 0);
    else if (p->rtech == R_ORTHAXIS || p->rtech == R_ARBAXIS) srd_collide(s, d, c);
    else return 4; /* collision operator outside this path */
    return 0;
}

/* ------------------------------------------------------------------------ */
/* temperature                                                               */
/* ------------------------------------------------------------------------ */

static double temp_now(const orc_sys *s)                               /* therm.c:209-235 */
{
    double K = 0., tmp;
    const double *A = s->p.avv;
    int i;
    for (i = 0; i < s->N; i++) {
        const double *v = &s->V[i * 3];
        tmp = (v[0] - A[0]) * (v[0] - A[0]);
        tmp += (v[1] - A[1]) * (v[1] - A[1]);
        tmp += (v[2] - A[2]) * (v[2] - A[2]);
        K += s->p.sp[s->spid[i]].mass * tmp;
    }
    K /= ((double)((s->N + 0) * s->p.dim));
    return K;
}

static void scale_t(orc_sys *s, int tstech)                            /* therm.c:52-143 */
{
    double g, KBT = s->p.kbt, K = s->kbtnow;
    int i, j;
    if (tstech == TS_VSC) g = sqrt(KBT / K);
    else if (tstech == TS_BEREND) g = sqrt(1. + (KBT / K - 1.) * s->p.dt / s->p.tau);
    else return;
    for (i = 0; i < s->N; i++)
        for (j = 0; j < s->p.dim; j++) s->V[i * 3 + j] *= g;
}

/* ------------------------------------------------------------------------ */
/* streaming and walls                                                       */
/* ------------------------------------------------------------------------ */

/* push(), init.c:1261-1293: RANDVEL (0), HOMO (2), HOMOAXIS (3), GAUSS (4);
 * READ (1) needs a file and is not on this path */
static void push_velocity(orc_sys *s, orc_draws *d, int i, int purpose, int kbase)
{
    const orc_species *sp = &s->p.sp[s->spid[i]];
    const int dim = s->p.dim;
    double *V = &s->V[i * 3];
    int j;
    if (sp->vdist == 0) {
        for (j = 0; j < dim; j++) V[j] = sqrt(s->p.kbt / sp->mass) * (2. * dr_uniform(s, d, purpose, i, kbase + j) - 1.);
    } else if (sp->vdist == 3) {
        for (j = 0; j < dim; j++) {
            V[j] = sqrt(dim * s->p.kbt / sp->mass);
            V[j] *= (dr_uniform(s, d, purpose, i, kbase + j) <= 0.5) ? -1.0 : 1.0; /* genrand_pmOne, rand.c:378 */
        }
    } else {
        for (j = 0; j < dim; j++) V[j] = dr_gaussMB(s, d, s->p.kbt, sp->mass, purpose, i, kbase + j);
        if (sp->vdist == 2) {
            double l = 0.0, nrm;
            for (j = 0; j < dim; j++) l += V[j] * V[j];
            l = sqrt(l);
            if (fabs(l - 0.0) >= ORC_TOL) for (j = 0; j < dim; j++) V[j] = V[j] / l;   /* norm, mtools.c:246 */
            nrm = sqrt((double)dim * s->p.kbt / sp->mass);
            for (j = 0; j < dim; j++) V[j] *= nrm;
        }
    }
}

/* MPC_BCcollision, bc.c:1363-1496, with chooseBC (1252-1328: planes, wavy
 * planes, and curved surfaces at their periodic image nearest the particle),
 * crosstimeReverse (251-298), chooseT (849-880), normal (1516-1556),
 * velBC's rule branch (656-672), posBC (760-774) and oriBC for LC;
 * immobile walls with four-fold symmetric surfaces */
static void bc_collision(orc_sys *s, orc_draws *d, int pi)
{
    const orc_params *p = &s->p;
    int dim = p->dim, i, w, chosen = 0, flag = 1, cnt = 0, cnt2 = 0, flagCNT = 0, retherm_k = 0;
    double t_step = p->dt, t_left = t_step, t_coll = 0.0, W = 1.0;
    double n[3] = {0, 0, 0};
    double *Q = &s->Q[pi * 3], *V = &s->V[pi * 3];
    const orc_species *sp = &p->sp[s->spid[pi]];
    while (flag) {
        /* chooseBC */
        double tmin = t_left, chW = 1.0;
        for (w = 0; w < p->nwall; w++) {
            const orc_wall *Wl = &p->wall[w];
            double tw, t1, t2, tc, shift[3];
            if (Wl->inter[s->spid[pi]] != 1) continue;
            if (s->planar[w]) {
                tw = calc_w_plane(Wl, s->wq[w], Q, dim);
                if (tw < 0.0) {
                    crosstime_reverse(Wl, 1, s->wq[w], Q, V, dim, t_left, &t1, &t2);
                    tc = t_left - t1;
                    if (tc < tmin) { tmin = tc; chosen = w; chW = tw; }
                }
            } else if (!feq(Wl->B[0], 0.0) && feq(Wl->P[0], 1.0) && feq(Wl->P[1], 1.0) && feq(Wl->P[2], 1.0)) {
                tw = calc_w(Wl, s->wq[w], Q, dim);
                if (tw < 0.0) {
                    crosstime_reverse(Wl, 0, s->wq[w], Q, V, dim, t_left, &t1, &t2);
                    tc = choose_t(t_left, t_left - t1, t_left - t2);
                    if (tc < tmin) { tmin = tc; chosen = w; chW = tw; }
                }
            } else {
                shift_bc(s, w, Q, shift);
                tw = calc_w(Wl, s->wq[w], Q, dim);
                if (tw < 0.0) {
                    crosstime_reverse(Wl, 0, s->wq[w], Q, V, dim, t_left, &t1, &t2);
                    tc = choose_t(t_left, t_left - t1, t_left - t2);
                    if (tc < tmin) { tmin = tc; chosen = w; chW = tw; }
                }
                shift_back_bc(s, w, shift);
            }
        }
        t_coll = tmin;
        W = chW;
        if (W > -ORC_TOL) { flag = 0; break; }
        cnt++;
        {
            const orc_wall *Wl = &p->wall[chosen];
            double shift[3];
            shift_bc(s, chosen, Q, shift);
            if (t_coll > -ORC_TOL) {
                double R[3] = {0, 0, 0}, VR[3], Vr[3] = {0, 0, 0}, VN[3] = {0, 0, 0}, VT[3] = {0, 0, 0};
                double IMpart, PN[3] = {0, 0, 0}, PT[3] = {0, 0, 0};
                for (i = 0; i < dim; i++) Q[i] = Q[i] - (t_left - t_coll) * V[i];      /* rewind_P */
                wall_normal(Wl, s->wq[chosen], Q, dim, n);                              /* normal, bc.c:1516 */
                vnorm(n, dim);
                /* velBC, rule-based collisions (bc.c:598-672) */
                for (i = 0; i < dim; i++) R[i] = Q[i] - s->wq[chosen][i];
                crossprod(Wl->L, R, VR);
                for (i = 0; i < dim; i++) Vr[i] = V[i] - Wl->V[i] - VR[i];
                proj(Vr, n, VN, dim);
                tang(Vr, VN, VT, dim);
                IMpart = 1.0 / sp->mass;
                for (i = 0; i < dim; i++) {
                    VN[i] *= Wl->MVN;
                    VN[i] += Wl->DVN * n[i];
                    VT[i] *= Wl->MVT;
                    VT[i] += Wl->DVT;
                }
                for (i = 0; i < dim; i++) Vr[i] = VN[i] + VT[i];
                for (i = 0; i < dim; i++) Vr[i] += Wl->DVxyz[i];
                for (i = 0; i < dim; i++) Vr[i] += Wl->V[i] + VR[i];
                for (i = 0; i < dim; i++) {
                    Vr[i] -= V[i];
                    Vr[i] *= sp->mass;
                }
                for (i = 0; i < dim; i++) V[i] += Vr[i] * 1.0 * IMpart;
                /* posBC */
                proj(Q, n, PN, dim);
                tang(Q, PN, PT, dim);
                for (i = 0; i < dim; i++) {
                    PN[i] += Wl->DN * n[i] * (1. - 2. * ORC_TOL);
                    PT[i] += Wl->DT;
                }
                for (i = 0; i < dim; i++) Q[i] = PN[i] + PT[i];
                if (p->lc != LC_ISOF) ori_bc(s, pi, Wl, n);
                t_left -= t_coll;
                for (i = 0; i < dim; i++) Q[i] = Q[i] + t_left * V[i];                /* stream_P */
                if (wall_W(s, chosen, Q) < -ORC_TOL)
                    for (i = 0; i < dim; i++) Q[i] = Q[i] - t_left * V[i];
            } else {
                for (i = 0; i < dim; i++) Q[i] = Q[i] - (t_left - t_coll) * V[i];
            }
            shift_back_bc(s, chosen, shift);
        }
        if (cnt > ORC_NBOUNCE) {
            flagCNT = 1;
            if (cnt < 2 * ORC_NBOUNCE && cnt2 == 0) {
                s->n_rewind++;
                if (t_left > ORC_TOL) t_left *= 1.1;
                else {
                    t_left = 0.9 * t_step;
                    t_coll = 0.1 * t_step;
                }
            } else {
                s->n_retherm++;
                push_velocity(s, d, pi, ORC_P_RETHERM, retherm_k);
                retherm_k += 3;
                t_left = t_step;
                t_coll = 0.0;
                cnt2++;
                cnt = 0;
            }
        }
    }
    s->n_bcfail += flagCNT;
}

/* ------------------------------------------------------------------------ */
/* step                                                                      */
/* ------------------------------------------------------------------------ */

/* ranshift + gridShift_all, rand.c:678-696 and mpc.c:720-766 */
static void grid_shift(orc_sys *s, int back)
{
    int i, j, w, dim = s->p.dim;
    double sg[3];
    for (j = 0; j < dim; j++) sg[j] = back ? -1.0 * s->shift[j] : s->shift[j];
    for (i = 0; i < s->N; i++)
        for (j = 0; j < dim; j++) {
            double *q = &s->Q[i * 3 + j];
            *q += sg[j];
            if (back == 0 && *q > s->p.L[j] && *q <= s->p.L[j] + sg[j]) *q -= s->p.L[j];
            else if (back == 1 && *q < 0.0 && *q >= sg[j]) *q += s->p.L[j];
        }
    for (w = 0; w < s->p.nwall; w++)
        for (j = 0; j < dim; j++) s->wq[w][j] += sg[j];
}

int orc_phase_bin(orc_sys *s, orc_draws *d)
{
    int t;
    s->kbtnow = 0.;
    s->avs = 0.;
    if (s->p.galinv) {
        int dim = s->p.dim;
        double x, y, z;
        z = (dim >= 3) ? dr_uniform(s, d, ORC_P_SHIFT, 0, 2) : 0.0;
        y = (dim >= 2) ? dr_uniform(s, d, ORC_P_SHIFT, 0, 1) : 0.0;
        x = dr_uniform(s, d, ORC_P_SHIFT, 0, 0);
        s->shift[0] = x; s->shift[1] = y; s->shift[2] = z;
        grid_shift(s, 0);
    } else {
        s->shift[0] = s->shift[1] = s->shift[2] = 0.0;
    }
    bin(s, d, 1);
    for (t = 0; t < s->N; t++) s->slot_of[s->order[t]] = t;
    return 0;
}

int orc_phase_collide(orc_sys *s, orc_draws *d)
{
    const orc_params *p = &s->p;
    int c, i, j, rc;
    if (fneq(p->grav[0], 0.0) || fneq(p->grav[1], 0.0) || fneq(p->grav[2], 0.0))
        for (i = 0; i < s->N; i++)
            for (j = 0; j < p->dim; j++) s->V[i * 3 + j] = s->V[i * 3 + j] + p->dt * p->grav[j]; /* acc_all */
    memset(s->pdec, 0, sizeof(uint32_t) * (size_t)s->N);
    memset(s->cdec, 0, sizeof(uint32_t) * (size_t)s->nc);
    s->rec_cdec = 1;
    local_prop(s);
    s->rec_cdec = 0;
    ghost_part(s, d);
    if (p->lc != LC_ISOF) {
        if (p->lc == LC_LCG) s->avs = av_order_param(s);
        for (c = 0; c < s->nc; c++)
            if (s->pop[c] > 1) lc_collision(s, d, c, s->avs);
        for (c = 0; c < s->nc; c++)
            if (s->pop[c] > 1) jeffery(s, c);
    }
    for (c = 0; c < s->nc; c++)
        if (s->pop[c] > 1) {
            rc = mpc_collision(s, d, c);
            if (rc) return rc;
        }
    s->kbtnow = temp_now(s);
    if (p->tstech != TS_NOTHERM) {
        if (p->tstech != TS_VSC && p->tstech != TS_BEREND) return 4;
        scale_t(s, p->tstech);
        s->kbtnow = temp_now(s);
    }
    if (p->rtech >= R_VICSEK && p->rtech <= R_DIPOLE_DIR_AV && s->kbtnow > 50.0 * p->kbt) {
        scale_t(s, TS_VSC);
        s->kbtnow = temp_now(s);
    }
    if (p->galinv) grid_shift(s, 1);
    return 0;
}

/* localFLOW + sumFLOW (mpc.c:226-246), as timestep() ends with them when
 * the flow output is on (mpc.c:4946-4951): each cell's centre-of-mass
 * velocity by localMPCVCM (mpc.c:3889-3910, list order, divided by 1 in an
 * empty cell) is added to FLOW, components < DIM */
static void sum_flow(orc_sys *s)
{
    const orc_params *p = &s->p;
    int c, t, d;
    for (c = 0; c < s->nc; c++) {
        double vcm[3] = {0., 0., 0.}, summ = 0.0;
        for (t = s->cstart[c]; t < s->cstart[c + 1]; t++) {
            int pi = s->order[t];
            double mass = p->sp[s->spid[pi]].mass;
            summ += mass;
            for (d = 0; d < p->dim; d++) vcm[d] += s->V[pi * 3 + d] * mass;
        }
        for (d = 0; d < p->dim; d++) vcm[d] /= (summ ? summ : 1.);
        for (d = 0; d < p->dim; d++) s->flow[c * 3 + d] += vcm[d];
    }
    s->flow_n++;
}

int orc_phase_stream(orc_sys *s, orc_draws *d)
{
    int i, j;
    for (i = 0; i < s->N; i++)
        for (j = 0; j < s->p.dim; j++) s->Q[i * 3 + j] = s->Q[i * 3 + j] + s->p.dt * s->V[i * 3 + j];
    for (i = 0; i < s->N; i++) bc_collision(s, d, i);
    bin(s, d, 0);
    local_prop(s);
    if (s->flow_on) sum_flow(s);
    s->step++;
    return 0;
}

void orc_flow(orc_sys *s, int on) { s->flow_on = on != 0; }

void orc_get_flow(orc_sys *s, double *flow, long *nsteps, int reset)
{
    if (flow) memcpy(flow, s->flow, sizeof(double) * 3 * (size_t)s->nc);
    if (nsteps) *nsteps = s->flow_n;
    if (reset) {
        memset(s->flow, 0, sizeof(double) * 3 * (size_t)s->nc);
        s->flow_n = 0;
    }
}

int orc_step(orc_sys *s, orc_draws *d)
{
    int rc = orc_phase_bin(s, d);
    if (!rc) rc = orc_phase_collide(s, d);
    if (!rc) rc = orc_phase_stream(s, d);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* state access                                                              */
/* ------------------------------------------------------------------------ */

int orc_set_state(orc_sys *s, const double *Q, const double *V, const double *U, const int *spid)
{
    memcpy(s->Q, Q, sizeof(double) * 3 * (size_t)s->N);
    memcpy(s->V, V, sizeof(double) * 3 * (size_t)s->N);
    if (U) memcpy(s->U, U, sizeof(double) * 3 * (size_t)s->N);
    if (spid) memcpy(s->spid, spid, sizeof(int) * (size_t)s->N);
    if (binin(s)) return 1;
    local_prop(s);
    return 0;
}

void orc_get_state(const orc_sys *s, double *Q, double *V, double *U)
{
    if (Q) memcpy(Q, s->Q, sizeof(double) * 3 * (size_t)s->N);
    if (V) memcpy(V, s->V, sizeof(double) * 3 * (size_t)s->N);
    if (U) memcpy(U, s->U, sizeof(double) * 3 * (size_t)s->N);
}

int orc_ncells(const orc_sys *s) { return s->nc; }

void orc_get_order(const orc_sys *s, int *order, int *cstart)
{
    if (order) memcpy(order, s->order, sizeof(int) * (size_t)s->N);
    if (cstart) memcpy(cstart, s->cstart, sizeof(int) * ((size_t)s->nc + 1));
}

void orc_get_cells(const orc_sys *s, int *pop, double *mass, double *vcm, double *cm, double *S, double *dir,
                   double *E)
{
    if (pop) memcpy(pop, s->pop, sizeof(int) * (size_t)s->nc);
    if (mass) memcpy(mass, s->mass, sizeof(double) * (size_t)s->nc);
    if (vcm) memcpy(vcm, s->vcm, sizeof(double) * 3 * (size_t)s->nc);
    if (cm) memcpy(cm, s->cm, sizeof(double) * 3 * (size_t)s->nc);
    if (S) memcpy(S, s->S, sizeof(double) * (size_t)s->nc);
    if (dir) memcpy(dir, s->dir, sizeof(double) * 3 * (size_t)s->nc);
    if (E) memcpy(E, s->E, sizeof(double) * 9 * (size_t)s->nc);
}

void orc_get_walls(const orc_sys *s, double *wq)
{
    int w;
    for (w = 0; w < s->p.nwall; w++) {
        wq[w * 3 + 0] = s->wq[w][0];
        wq[w * 3 + 1] = s->wq[w][1];
        wq[w * 3 + 2] = s->wq[w][2];
    }
}

/* galileantrans (mtools.c:673-718) on the particles, then zeroExtraDims
 * (mtools.c:727-744): the mass-weighted mean velocity less AVV, summed in
 * particle-index order, is subtracted in the DIM first components; in 2D the
 * third components of Q and V are zeroed.  Called by outputResults
 * (pout.c:2557-2562) after the timestep of a runtime divisible by zeroNetMom. */
void orc_zero_net_momentum(orc_sys *s)
{
    double net[3] = {0., 0., 0.}, totm = 0.;
    int i, j, dim = s->p.dim;
    for (i = 0; i < s->N; i++) {
        double m = s->p.sp[s->spid[i]].mass;
        totm += m;
        for (j = 0; j < dim; j++) net[j] += m * s->V[3 * i + j];
    }
    for (j = 0; j < dim; j++) net[j] /= totm;
    for (j = 0; j < dim; j++) net[j] -= s->p.avv[j];
    for (i = 0; i < s->N; i++)
        for (j = 0; j < dim; j++) s->V[3 * i + j] -= net[j];
    if (dim < 3)
        for (i = 0; i < s->N; i++) s->Q[3 * i + 2] = s->V[3 * i + 2] = 0.;
}

double orc_kbtnow(const orc_sys *s) { return s->kbtnow; }
double orc_avs(const orc_sys *s) { return s->avs; }
int orc_lc_stride(const orc_sys *s) { return s->lc_stride; }
void orc_last_shift(const orc_sys *s, double *shift) { shift[0] = s->shift[0]; shift[1] = s->shift[1]; shift[2] = s->shift[2]; }

void orc_counters(const orc_sys *s, long *out4)
{
    out4[0] = s->n_escape;
    out4[1] = s->n_bcfail;
    out4[2] = s->n_rewind;
    out4[3] = s->n_retherm;
}

orc_draws *orc_draws_new(int mode, uint64_t seed, uint32_t replica)
{
    orc_draws *d = (orc_draws *)calloc(1, sizeof(orc_draws));
    d->mode = mode;
    d->seed = seed;
    d->replica = replica;
    seqdraw_init(&d->seq, seed);
    return d;
}

void orc_draws_free(orc_draws *d) { free(d); }

void orc_draws_set_keyed(orc_draws *d, const double *shift, const double *cell, const double *ghost,
                         const double *part, const double *lc, int lc_stride)
{
    d->shift = shift; d->cell = cell; d->ghost = ghost; d->part = part; d->lc = lc; d->lc_stride = lc_stride;
}

void orc_draws_set_record(orc_draws *d, double *shift, double *cell, double *ghost, double *part, double *lc,
                          int lc_stride)
{
    d->rec_shift = shift; d->rec_cell = cell; d->rec_ghost = ghost; d->rec_part = part; d->rec_lc = lc;
    d->rec_lc_stride = lc_stride;
}

void orc_lc_decisions(const orc_sys *s, uint32_t *part, uint32_t *cell)
{
    if (part) memcpy(part, s->pdec, sizeof(uint32_t) * (size_t)s->N);
    if (cell) memcpy(cell, s->cdec, sizeof(uint32_t) * (size_t)s->nc);
}

void orc_draws_set_rare(orc_draws *d, double *escape, double *retherm, int n)
{
    d->escape = escape;
    d->retherm = retherm;
    d->rare_n = n;
    d->rare_overflow = 0;
}

int orc_draws_rare_overflow(const orc_draws *d) { return d->rare_overflow; }

void orc_draws_counts(const orc_draws *d, uint64_t *out3)
{
    out3[0] = d->seq.nU;
    out3[1] = d->seq.nG;
    out3[2] = d->seq.nA;
}
