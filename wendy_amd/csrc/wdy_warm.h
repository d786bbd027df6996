// Shared declarations of the warm-start leapfrog step (wdy_warm.hip).
#pragma once

#include "wdy_internal.h"

namespace wdy {

constexpr int WARM_WMAX = 64;          // largest window (buckets either side)
// Tile shapes the warm kernel is compiled for: threads per workgroup and the
// most sheets one bucket may hold (its LDS capacity).  The default bucket
// target leaves 1/5 of the capacity, at least 256 sheets, as headroom
// (warm_default_target): bucket counts drift with the net flow of sheets
// across the stale splitters; over 2400 steps of the N = 2^24 disk, 1536-
// sheet buckets never overfilled at targets up to 1260 and did at 1280.
struct WarmShape {
  int threads, cap, nq;  // nq: 32-bit words of the sort's bins
  // widest window the partition's per-wave segment counters (nq u32 shared
  // with the sort's bins) can hold: waves * (2W + 2) <= nq
  int wmax() const {
    const int w = (nq / (threads / 64) - 2) / 2;
    return w < WARM_WMAX ? w : WARM_WMAX;
  }
};
constexpr int WARM_NSHAPES = 10;
// default tile shape (measured at N = 2^24 with the pipelined kernel, equal
// and unequal masses): 256 threads per workgroup, buckets of up to 1536
// sheets, target 1229, window 6; four workgroups per CU with equal masses
// (38 KB of LDS, at most 128 VGPRs), three with unequal ones.  When sheets cross
// many buckets per step the window fitting moves to 2048-sheet buckets
// (WARM_SHAPE_WIDE: 256 threads, 1024 bins), then to 4096-sheet ones.
constexpr int WARM_SHAPE_DEFAULT = 8;
constexpr int WARM_SHAPE_WIDE = 7;
constexpr int WARM_WINDOW_DEFAULT = 6;
const WarmShape &warm_shape(int v);
// The tile shape a simulation of n sheets starts from when it has the GPU
// to itself.  While one step's buckets do not fill the GPU (N below ~2^22 at
// the default's 1229-sheet buckets) a step lasts about as long as its
// slowest bucket, and larger workgroups finish a bucket sooner.  Timed for
// every shape on the sech^2 disk, equal and lognormal masses, one output of
// 20 steps (per-interval ms, default -> best): 2^14-2^16 shape 3 (512
// threads, 1024 sheets; 0.44 -> 0.34 equal, 0.49 -> 0.38 lognormal at
// 2^14), 2^17-2^18 shape 1 (512 threads, 2048 sheets; 0.57 -> 0.51 at
// 2^18), 2^19-2^20 shape 5 (512 threads, 3072 sheets) with equal masses and
// shape 6 (768 threads, 3072 sheets) with unequal ones (1.46 -> 0.95 equal,
// 1.57 -> 1.09 lognormal at 2^20); 2^21-2^22 shape 0 (1024 threads, 4096
// sheets) with equal masses (2^21: 1.40 with shape 6 -> 1.21; 2^22: 2.36 ->
// 1.93), 2^21 shape 6 and 2^22-2^23 shape 7 (256 threads, 2048 sheets) with
// unequal ones (2^22: 3.56 -> 2.67), 2^23 shape 7 with equal ones (3.67 ->
// 3.34), from 2^24 the default.
int warm_initial_shape(int64_t n, bool equal_mass);
inline u32 warm_default_target(int cap) {
  const u32 c = (u32)cap, room = c / 5u > 256u ? c / 5u : 256u;
  return c > 2u * room ? c - room : c / 2u;
}
constexpr u32 WARM_OVF_CAP = 1u << 16; // far movers one step may carry
constexpr u32 WARM_T_MIN = 32;

// Segment table and overflow list written by one step, read by the next.
// Source bucket t occupies [boff[t], boff[t] + count) of the state arrays,
// partitioned into S+1 contiguous segments: segment d (< S) holds the sheets
// whose next bucket is t + d - W; segment S holds the far movers, which are
// also copied into the overflow list with their destination bucket.
struct WarmMeta {
  u32 *segcnt = nullptr;  // [nb * (S + 1)]
  u32 *segoff = nullptr;  // [nb * (S + 1)] offset of the segment in its bucket
  u32 *boff = nullptr;    // [nb + 1] region start of each bucket
  u64 *spl = nullptr;     // [nb + 1] splitters this step produced
  u64 *okey = nullptr;    // [WARM_OVF_CAP] overflow records
  double *ov = nullptr;
  i32 *oid = nullptr;
  double *om = nullptr;
  u32 *odst = nullptr;
};

struct WarmArgs {
  u32 n, nb, T, W, S;
  const u64 *kin;
  const double *vin;
  const i32 *iin;
  const double *min;
  u64 *kout;
  double *vout;
  i32 *iout;
  double *mout;
  WarmMeta rd, wr;
  u64 *status;  // [nb * 5] look-back words
  // per-launch counters, rings of 4 indexed by epoch: this launch's ticket
  // counter and overflow count, the overflow count of the list it reads,
  // and the slots launch epoch+2 will use (cleared by ticket 0)
  u32 *tick, *tick_next2, *ovf_r, *ovf_w, *ovf_next2;
  u32 *err;     // [2] epoch of the first failing launch, failure kinds
  u32 epoch;
  double mq, totmass, omega2, h, pre, hd, hp, fscale, finv;
  ExtForce ext;  // external force (kind 0: none) evaluated at time t
  double t;
  u64 *prof;    // null, or [nb * 8] phase timestamps (wall clock) per bucket
};

// failure kinds (err[1] bits)
constexpr u32 WARM_ERR_CAPACITY = 1;  // a bucket exceeded the LDS capacity
constexpr u32 WARM_ERR_OVF = 2;       // the overflow list was full
constexpr u32 WARM_ERR_LOVF = 4;      // too many far movers into one bucket
constexpr u32 WARM_ERR_STALL = 8;     // look-back gave up (bounded wait)

// mode 0: one workgroup per bucket; 1: partition a radix-sorted array;
// 2: mode 0 by persistent workgroups that prefetch the next bucket
int warm_step_launch(hipStream_t st, const WarmArgs &a, bool equal_mass,
                     bool harm, int mode, int shape);
// zero the ticket ring, the overflow-count ring except slot `keep` (4: none),
// and the error word
int warm_reset_launch(hipStream_t st, u32 *ring, u32 keep, u32 *err);
// shapes the gang kernel is compiled for (the default, the wide one, shape 0:
// every shape the automatic window fitting and back-off choose)
bool warm_gang_shape(int shape);
// most simulations one gang launch can hold (one ballot over their ticket
// starts)
constexpr int WARM_GANG_MAX = 64;

}  // namespace wdy
