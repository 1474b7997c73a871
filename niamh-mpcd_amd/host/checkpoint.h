/*
 * checkpoint.h -- the reference's checkpoint file for the C host.
 *
 * The file is the text that checkpoint() writes (mpcd/subroutines/pout.c:
 * 2325-2404) and readchckpnt() reads (read.c:669-860): global inputs,
 * output flags, species, boundaries, every particle and the swimmers, one
 * record after another, numbers separated by white space.  ckpt_read takes
 * a file of that layout apart (the path's options only: no MD coupling, no
 * swimmers, fixed walls); ckpt_write writes one in the same
 * layout, with the reference's defaults for what this path does not carry.
 */
#ifndef MPCD_CHECKPOINT_H
#define MPCD_CHECKPOINT_H

#include <stdint.h>
#include <stdio.h>

#include "mpcd_gpu.h"

/* species interaction flags per wall: MAXSPECI species, then MD and swimmers
 * (definitions.h MAXSPECI = 10, SRDclss.h bc.INTER[MAXSPECI+2]) */
#define CKPT_MAXSPECI 10

typedef struct {
    int sim_steps, warm_steps;   /* in.simSteps, in.warmupSteps */
    int runtime, warmtime;       /* loop counters at the checkpoint */
    long seed;
    double kbtnow;               /* KBTNOW */
    double avnow[3];             /* AVNOW: mean velocity now */
    int flow_out, vel_out, dir_out, ckpt_out, dens_out;  /* FLOWOUT, VELOUT, ORDEROUT, CHCKPNT, DENSITYOUT */
    int rframe, zero_net_mom;    /* RFRAME (after initializeSIM: the periodic reset is on), zeroNetMom */
    uint32_t lib_step;           /* the library's draw counter (binary checkpoints only) */
    int binary;                  /* read from a binary checkpoint */
    int64_t n;                   /* particles (GPOP) */
    double *Q, *V, *U;           /* n x 3 each, particle-index order */
    int32_t *spid;               /* n */
} mpcd_ckpt;

/* Fills p (domain, time step, temperature, collision, LC, thermostat,
 * body force, AVV, GnDNST, species, walls) and ck (loop state, output
 * flags, particles; arrays malloc'ed) from the file at path.  Returns 0, or
 * -1 with a message in err. */
int ckpt_read(const char *path, mpcd_gpu_params *p, mpcd_ckpt *ck, char *err, size_t errlen);

/* Writes p and ck in the reference's layout. */
void ckpt_write(FILE *f, const mpcd_gpu_params *p, const mpcd_ckpt *ck);

/* The lossless binary variant (no counterpart in the reference): the
 * parameters as the library takes them, the loop state, the library's draw
 * counter and the particles in FP64, so a resume continues the run (bit for
 * bit in the exact-sum form).  ckpt_read recognises it by its magic. */
void ckpt_write_bin(FILE *f, const mpcd_gpu_params *p, const mpcd_ckpt *ck);

void ckpt_free(mpcd_ckpt *ck);

#endif
