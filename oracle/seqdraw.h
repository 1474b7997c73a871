/*
 * seqdraw.h -- deterministic "sequential" variate streams (TEST INFRASTRUCTURE).
 *
 * These streams replace the reference's global xoshiro generator when the
 * reference's own C step (compiled by oracle/Makefile into oracle/_ref) is
 * run to record golden fixtures, and the oracle's sequential draw mode
 * consumes exactly the same streams in the same call order.  Both sides
 * call the functions below, compiled from this one file with the same
 * flags, so identical call sequences give bit-identical variates.
 *
 * Variate semantics follow the reference's draw sites:
 *   uniform  : genrand_real()   rand.c:365-368 (a 32-bit integer / 2^32)
 *   gauss    : genrand_gauss()  rand.c:473-498 (returned as float)
 *   unit vec : ranvec3D()       rand.c:708-722 (uniform on the sphere)
 * Only the generator behind them is replaced; the reference's consumers
 * (mpc.c, lc.c, bc.c) are unchanged.
 *
 * Three independent streams are kept so that one kind of draw never shifts
 * another: U (uniforms), G (gaussians), A (3D unit vectors).
 */
#ifndef ORACLE_SEQDRAW_H
#define ORACLE_SEQDRAW_H

#include <stdint.h>

typedef struct {
    uint64_t seed;
    uint64_t nU, nG, nA; /* variates consumed so far from each stream */
} seqdraw_state;

void   seqdraw_init(seqdraw_state *s, uint64_t seed);
double seqdraw_uniform(seqdraw_state *s);          /* [0,1), multiple of 2^-32 */
float  seqdraw_gauss(seqdraw_state *s);            /* N(0,1) rounded to float */
void   seqdraw_unitvec3(seqdraw_state *s, double v[3]);

#endif
