/*
 * ref_rand.c -- draw injection for the reference build (TEST INFRASTRUCTURE).
 *
 * oracle/Makefile compiles the reference's own sources from /root/reference
 * (all of mpcd/subroutines, md/ and cJSON, except mpcd/subroutines/rand.c and
 * the main() in mpcd/mpcd.c) and links them with this file instead of
 * rand.c.  Every symbol declared in the reference's mpcd/headers/rand.h is
 * defined here with the reference's variate semantics, but the raw
 * generator is the deterministic sequential stream of seqdraw.c, so the
 * reference's step consumes exactly the variates the oracle's sequential
 * mode consumes.  This is how the reference's step is run "on identical
 * injected draws".
 *
 * Semantics kept from the reference (rand.c line numbers):
 *   genrand_real 365-368, genrand_pmOne 378-382, genrand_coneNP 397-425,
 *   genrand_gaussMB 510-513, genrand_gaussGen 526-528, ranshift 678-696
 *   (z, y, x draw order), ranvec2D 735-742, ranvec 754-758.
 * Replaced: ranvec3D (Marsaglia rejection, 708-722) serves one unit vector
 * from stream A; genrand_gauss (polar Box-Muller with a cross-call spare,
 * 473-498) serves one float gaussian from stream G.
 */
#include <math.h>
#include <stdint.h>
#include "seqdraw.h"

#define REF_PI 3.14159265358979323846

static seqdraw_state g_seq;

void ref_rand_seed(uint64_t seed) { seqdraw_init(&g_seq, seed); }
void ref_rand_counts(uint64_t out[3])
{
    out[0] = g_seq.nU;
    out[1] = g_seq.nG;
    out[2] = g_seq.nA;
}

unsigned long RandomSeedSRD(unsigned long seed) { return seed; }
void init_genrand(unsigned long s) { (void)s; }
void init_by_array(unsigned long init_key[], int key_length) { (void)init_key; (void)key_length; }

unsigned long genrand_int32(void)
{
    return (unsigned long)(seqdraw_uniform(&g_seq) * 4294967296.0);
}
long genrand_int31(void) { return (long)(genrand_int32() >> 1); }

double genrand_real(void) { return seqdraw_uniform(&g_seq); }

double genrand_pmOne(void)
{
    double r = genrand_real();
    if (r <= 0.5) return -1.0;
    return 1.0;
}

void genrand_coneNP(double vec[], double theta, int dimension)
{
    double z, phi, ct = cos(theta);
    if (dimension == 3) {
        z = genrand_real() * (1.0 - ct) + ct;
        phi = genrand_real() * 2.0 * REF_PI;
        vec[0] = z;
        vec[1] = sqrt(1.0 - z * z) * cos(phi);
        vec[2] = sqrt(1.0 - z * z) * sin(phi);
    } else if (dimension == 2) {
        phi = theta * (1. - 2. * genrand_real());
        vec[0] = cos(phi);
        vec[1] = sin(phi);
    } else if (dimension == 1) {
        phi = theta * 0.5 * REF_PI * (1. - 2. * genrand_real());
        vec[0] = (phi > 1.0 || phi < -1.0) ? -1.0 : 1.0;
    }
}

/* Only reached from swimmer/MD code paths that the recorded cases never
 * enable; kept so the reference links. */
void genrand_cone(double axis[], double vecOut[], double theta, int dimension)
{
    double v[3] = {0.0, 0.0, 0.0};
    int i;
    (void)axis;
    genrand_coneNP(v, theta, dimension);
    for (i = 0; i < dimension; i++) vecOut[i] = v[i];
}

float genrand_gauss(void) { return seqdraw_gauss(&g_seq); }

double genrand_gaussMB(double KBT, double M)
{
    double sigma = sqrt(KBT / M);
    return sigma * genrand_gauss();
}

double genrand_gaussGen(double mu, double sigma) { return sigma * genrand_gauss() + mu; }

double genrand_exp(double lambda) { return -log(1.0 - genrand_real()) / lambda; }

int genrand_poisson(double lambda)
{
    double L = exp(-lambda), p = 1.0;
    int k = 0;
    do { k++; p *= genrand_real(); } while (p > L);
    return k - 1;
}

float genrand_rayleigh(float std)
{
    return (float)(std * sqrt(-2.0 * log(1.0 - genrand_real())));
}

void genrand_sphere(double vec[], int dimension)
{
    if (dimension == 3) seqdraw_unitvec3(&g_seq, vec);
    else genrand_coneNP(vec, REF_PI, dimension);
}

double *ranvec3D(double *v)
{
    seqdraw_unitvec3(&g_seq, v);
    return v;
}

double *ranvec2D(double *v)
{
    double x = genrand_real();
    v[0] = x;
    v[1] = sqrt(1. - x * x);
    return v;
}

double *ranvec(double *v, int dimension)
{
    if (dimension == 3) ranvec3D(v);
    if (dimension == 2) ranvec2D(v);
    return v;
}

double *ranshift(double *v, int doShift, int dimension)
{
    double x, y, z;
    if (doShift) {
        z = (dimension >= 3) ? genrand_real() : 0.0;
        y = (dimension >= 2) ? genrand_real() : 0.0;
        x = genrand_real();
        v[0] = x;
        v[1] = y;
        v[2] = z;
    } else {
        v[0] = v[1] = v[2] = 0.0;
    }
    return v;
}

int rand_particle(int POP) { return (int)(genrand_real() * (double)POP); }
