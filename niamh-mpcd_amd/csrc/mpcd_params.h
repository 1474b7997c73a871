// mpcd_params.h -- host helpers on the replica parameters of the C-ABI, shared
// by the library's translation units.
#pragma once

#include "mpcd_math.h"
#include "../../include/mpcd_gpu.h"

namespace mpcd {

// setPBC (read.c:584-603) for the axis-aligned planar walls of a replica
inline void set_pbc(const mpcd_gpu_params &p, int pbc[3]) {
  for (int i = 0; i < 3; i++) pbc[i] = 0;
  for (int w = 0; w < p.nwall; w++) {
    const mpcd_gpu_wall &W = p.wall[w];
    for (int i = 0; i < 3; i++) {
      double ainv = (W.A[i] != 0.0) ? 1.0 / W.A[i] : 0.0;
      if (feq(fabs(ainv), 1.0) && W.MVN > 0.0 && W.MVT > 0.0 && (fneq(W.DN, 0.0) || fneq(W.DT, 0.0))) pbc[i] = 1;
    }
  }
}

// the last error of a call without a context (mpcd_gpu_bin_noalloc), read
// by mpcd_gpu_last_error(NULL)
const char *noalloc_error();

}  // namespace mpcd
