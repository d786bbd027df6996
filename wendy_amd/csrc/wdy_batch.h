// Shared by the batched-ensemble handle (wdy_batch) and the serving pool
// (wdy_pool): the per-simulation descriptor and the launch of the kernel that
// advances every described simulation by one output interval.
#pragma once
#include "wdy_internal.h"

namespace wdy {

struct BatchDesc {
  u64 off;        // first arena slot
  u32 n;          // sheets
  u32 flags;      // 1 equal masses, 2 harmonic term
  int cur;        // current buffer set
  int nleap;      // leapfrog steps this interval (0: skip)
  u32 eslot;      // error word of this simulation: err[eslot]
  u32 pad_;
  double h;       // leapfrog step
  double mq, totmass, omega2, fscale, finv;
};

// Arena: two buffer sets of SoA arrays.  idx holds a sheet index that orders
// exact ties (any index consistent within one simulation); mass[] is null
// when every simulation has equal masses.
struct BatchArrays {
  u64 *key[2];
  double *vel[2];
  i32 *idx[2];
  double *mass[2];
};

// One workgroup per descriptor; descriptors in d_descs[0..count).  A
// simulation that fails its on-device check sets err[its eslot].
int batch_launch(hipStream_t st, const BatchDesc *d_descs, int count,
                 const BatchArrays &arr, u32 *err);

// Descriptor fields derived from a simulation's masses (equal-mass flag,
// fixed-point scale of the exact unequal-mass prefix).
void batch_desc_fill(BatchDesc &d, const double *m, u32 n, double totmass,
                     double omega2);

}  // namespace wdy
