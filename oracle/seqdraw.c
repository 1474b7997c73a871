/*
 * seqdraw.c -- deterministic sequential variate streams (TEST INFRASTRUCTURE).
 * See seqdraw.h.  The raw generator is SplitMix64 (Steele, Lea, Flood 2014)
 * applied to (stream key + counter), so any variate is a pure function of
 * (seed, stream, index).
 */
#include "seqdraw.h"
#include <math.h>

static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static uint64_t raw(const seqdraw_state *s, uint64_t stream, uint64_t k)
{
    uint64_t key = splitmix64(s->seed ^ (stream * 0xD1B54A32D192ED03ULL));
    return splitmix64(key + k);
}

void seqdraw_init(seqdraw_state *s, uint64_t seed)
{
    s->seed = seed;
    s->nU = s->nG = s->nA = 0;
}

double seqdraw_uniform(seqdraw_state *s)
{
    uint64_t r = raw(s, 1, s->nU++);
    return (double)(uint32_t)(r >> 32) * (1.0 / 4294967296.0);
}

float seqdraw_gauss(seqdraw_state *s)
{
    uint64_t r1 = raw(s, 2, 2 * s->nG);
    uint64_t r2 = raw(s, 2, 2 * s->nG + 1);
    s->nG++;
    double u1 = ((double)(r1 >> 11) + 1.0) * (1.0 / 9007199254740992.0); /* (0,1] */
    double u2 = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);         /* [0,1) */
    double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    return (float)z;
}

void seqdraw_unitvec3(seqdraw_state *s, double v[3])
{
    uint64_t r1 = raw(s, 3, 2 * s->nA);
    uint64_t r2 = raw(s, 3, 2 * s->nA + 1);
    s->nA++;
    double u1 = (double)(r1 >> 11) * (1.0 / 9007199254740992.0);
    double u2 = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);
    double z = 1.0 - 2.0 * u1;
    double rr = sqrt((1.0 - z) * (1.0 + z));
    double phi = 6.283185307179586 * u2;
    v[0] = rr * cos(phi);
    v[1] = rr * sin(phi);
    v[2] = z;
}
