// Small-N path: one workgroup runs a whole output interval (all nleap
// leapfrog steps) in a single launch, with the sorted positions in LDS.
//
// Layout (small mode of a sim):
//   key[0][i], idx[0][i]   sorted (position key, sheet index) pairs, the
//                          reference's xi[] array (wendy/wendy.h:12-16)
//   vel[0][j], mass[0][j]  velocity and mass*twopiG of sheet j, in original
//                          sheet order (the reference's v[] and m[])
// LDS holds the (key, index) pairs: 12 B per sheet.  Each step re-sorts them
// with odd-even transposition rounds until a round swaps nothing: the input is
// the previous step's order, and sheets move only a few ranks per step, so a
// few rounds suffice (any input still ends sorted; it just takes more rounds).
// Pairs compare by (key, index), so exactly tied positions are ordered by
// sheet index.  The mass prefix is a single-thread sequential sum for
// n <= SEQ_SCAN_MAX (the reference's exact summation order,
// wendy/wendy.c:359-360) and a blocked parallel scan above.
#include "wdy_internal.h"

namespace wdy {
namespace {

constexpr u32 SEQ_SCAN_MAX = 512;

__device__ __forceinline__ bool pair_less(u64 ka, i32 ia, u64 kb, i32 ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// Sorts (ks, is)[0..n) in LDS by odd-even transposition until no swap.
__device__ void lds_transposition_sort(u64 *ks, i32 *is, u32 n) {
  const u32 t = threadIdx.x, T = blockDim.x;
  while (true) {
    int sw = 0;
#pragma unroll
    for (u32 ph = 0; ph < 2; ++ph) {
      for (u32 p = t; 2 * p + ph + 1 < n; p += T) {
        const u32 a = 2 * p + ph;
        const u64 ka = ks[a], kb = ks[a + 1];
        const i32 ia = is[a], ib = is[a + 1];
        if (pair_less(kb, ib, ka, ia)) {
          ks[a] = kb;
          ks[a + 1] = ka;
          is[a] = ib;
          is[a + 1] = ia;
          sw = 1;
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(sw)) break;
  }
}

// Exclusive prefix of mass[is[i]] over sorted slots into cum[], blocked over
// threads (each thread sums a contiguous chunk sequentially).
__device__ void block_mass_prefix(const i32 *is, const double *mass,
                                  double *cum, u32 n, double *s_part) {
  const u32 t = threadIdx.x, T = blockDim.x;
  const u32 per = (n + T - 1) / T;
  const u32 b = t * per;
  double acc = 0.0;
  for (u32 j = 0; j < per && b + j < n; ++j) acc = acc + mass[is[b + j]];
  s_part[t] = acc;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (u32 q = 0; q < T; ++q) {
      const double v = s_part[q];
      s_part[q] = run;
      run = run + v;
    }
  }
  __syncthreads();
  double run = s_part[t];
  for (u32 j = 0; j < per && b + j < n; ++j) {
    cum[b + j] = run;
    run = run + mass[is[b + j]];
  }
}

template <bool EQ, bool HARM>
__global__ __launch_bounds__(1024) void k_small_advance(
    u32 n, u64 *__restrict__ key_g, i32 *__restrict__ id_g,
    double *__restrict__ vel, const double *__restrict__ mass,
    double *__restrict__ cum, double mq, double totmass, double omega2,
    double h, int nleap, double *__restrict__ acc_out,
    double *__restrict__ cum_out, ExtForce ext, double t0) {
  extern __shared__ u64 smem[];
  __shared__ double s_part[1024];
  u64 *ks = smem;
  i32 *is = (i32 *)(smem + n);
  const u32 t = threadIdx.x, T = blockDim.x;
  const double hh = h / 2.;
  for (u32 i = t; i < n; i += T) {
    const i32 id = id_g[i];
    is[i] = id;
    // half drift (wendy/wendy.c:318-323,398)
    ks[i] = encode_pos(decode_pos(key_g[i]) + hh * vel[id]);
  }
  __syncthreads();
  double tf = t0;  // time of this step's force (wendy/wendy.c:401-408)
  for (int step = 0; step < nleap; ++step) {
    const bool last = step == nleap - 1;
    const double hd = last ? hh : h;
    lds_transposition_sort(ks, is, n);
    if (!EQ) {
      if (n <= SEQ_SCAN_MAX) {
        if (t == 0) {
          double c = 0.0;
          for (u32 i = 0; i < n; ++i) {
            cum[i] = c;
            c = c + mass[is[i]];
          }
        }
      } else {
        block_mass_prefix(is, mass, cum, n, s_part);
      }
      __syncthreads();
    }
    for (u32 i = t; i < n; i += T) {
      const i32 id = is[i];
      double x = decode_pos(ks[i]);
      const double C = EQ ? (double)i * mq : cum[i];
      const double m = EQ ? mq : mass[id];
      double F = (totmass - 2.0 * C) - m;
      if (HARM) F = F - omega2 * x;
      const double a = (ext.kind ? ext_accel(ext, x, tf) : 0.0) + F;
      const double v = vel[id] + h * a;
      vel[id] = v;
      x = x + hd * v;
      ks[i] = encode_pos(x);
      if (last && acc_out) acc_out[id] = a;
      if (last && cum_out) cum_out[i] = C;
    }
    tf = tf + h;
    __syncthreads();
  }
  for (u32 i = t; i < n; i += T) {
    key_g[i] = ks[i];
    id_g[i] = is[i];
  }
}

}  // namespace

u32 small_max_n(int device) {
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                            device) != hipSuccess)
    lds = 65536;
  // 12 B per sheet of dynamic LDS, plus 8 KiB of static LDS
  const long cap = ((long)lds - 8192 - 64) / 12;
  return cap > 0 ? (u32)cap : 0u;
}

int launch_small_advance(wdy_sim *s, double h, int nleap, double *acc_out,
                         double *cum_out) {
  const u32 n = (u32)s->n;
  u32 threads = ((n / 2 + 63) / 64) * 64;
  if (threads < 64) threads = 64;
  if (threads > 1024) threads = 1024;
  const size_t lds = (size_t)n * 12 + 16;
  const bool harm = s->omega2 >= 0.0;
  if (s->timing) timer_begin(s->timer, s->stream, KC_LEAP);
#define WDY_SMALL(EQ, HARM)                                                   \
  do {                                                                        \
    auto fn = k_small_advance<EQ, HARM>;                                      \
    if (lds > 65536)                                                          \
      WDY_HIP(hipFuncSetAttribute((const void *)fn,                           \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, \
                                  (int)lds));                                 \
    fn<<<1, threads, lds, s->stream>>>(n, s->key[0], s->idx[0], s->vel[0],    \
                                       s->mass[0], s->cum, s->m_equal,        \
                                       s->totmass, s->omega2, h, nleap,       \
                                       acc_out, cum_out, s->ext, s->t0);      \
  } while (0)
  if (s->equal_mass) {
    if (harm) WDY_SMALL(true, true); else WDY_SMALL(true, false);
  } else {
    if (harm) WDY_SMALL(false, true); else WDY_SMALL(false, false);
  }
#undef WDY_SMALL
  if (s->timing) timer_end(s->timer, s->stream, KC_LEAP);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

}  // namespace wdy
