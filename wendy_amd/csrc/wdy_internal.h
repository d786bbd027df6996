// Internal declarations shared by the translation units of libwendy_amd.so.
//
// Device state of one simulation ("sim") is structure-of-arrays, kept in
// sorted position order between leapfrog steps:
//
//   key[s][i]   uint64  order-preserving encoding of the position of the
//                       sheet at sorted slot i (the encoding is a bijection on
//                       non-NaN doubles with -0.0 folded into +0.0, so the key
//                       IS the position: no separate x array exists)
//   vel[s][i]   double  velocity of that sheet
//   idx[s][i]   int32   original index of that sheet (only read at output)
//   mass[s][i]  double  mass * twopiG (unequal-mass runs only)
//
// s in {0,1} selects one of two buffer sets; a step reads one set and writes
// the other, so no kernel updates an array another thread of it still reads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/wendy_amd.h"

namespace wdy {

using u64 = uint64_t;
using u32 = uint32_t;
using i32 = int32_t;

constexpr u64 SIGN = 0x8000000000000000ull;

// Order-preserving map double -> uint64: non-negative values get the sign bit
// set, negative values are bit-inverted.  -0.0 is folded into +0.0 first so
// that the two zeros compare equal, as the reference comparator
// (wendy/wendy.c:19-31) treats them.
__host__ __device__ inline u64 encode_pos(double x) {
  u64 b;
  __builtin_memcpy(&b, &x, 8);
  if (b == SIGN) b = 0;
  return (b & SIGN) ? ~b : (b | SIGN);
}

__host__ __device__ inline double decode_pos(u64 k) {
  u64 b = (k & SIGN) ? (k & ~SIGN) : ~k;
  double x;
  __builtin_memcpy(&x, &b, 8);
  return x;
}

// External force evaluated on the device: the counterpart of the reference's
// ext_force callback (wendy/wendy.py:389-420, called at wendy/wendy.c:362-370)
// for a fixed menu of forms.  a = E(x, t) + F, as the reference stores the
// callback's value in a[] and then adds the self-gravity term.
struct ExtForce {
  int kind = 0;  // WDY_EXT_* of include/wendy_amd.h; 0: none
  double p0 = 0.0, p1 = 0.0, p2 = 0.0;
  // EXT_TABLE: acceleration of sheet j in tab[j], filled by the host from a
  // caller's function before each step (radix-path steps only)
  const double *tab = nullptr;
};
constexpr int EXT_TABLE = 4;

__device__ inline double ext_accel(const ExtForce &f, double x, double t) {
  switch (f.kind) {
    case 1:  // harmonic: -omega**2 * x, p0 = omega**2
      return -f.p0 * x;
    case 2:  // isothermal slab: amp * tanh(0.5 * x / zh) / zh, (amp, zh)
      return (f.p0 * tanh((0.5 * x) / f.p1)) / f.p1;
    case 3: {  // uniform pulse: amp * exp(-0.5 * ((t - tc) / tau)**2)
      const double u = (t - f.p1) / f.p2;
      return f.p0 * exp(-0.5 * (u * u));
    }
    default:
      return 0.0;
  }
}

// The same for a kernel that knows the sheet index (the per-sheet table).
__device__ inline double ext_accel_id(const ExtForce &f, double x, double t,
                                      i32 id) {
  return f.kind == EXT_TABLE ? f.tab[id] : ext_accel(f, x, t);
}

// Error reporting: every C-ABI entry point returns one of the WDY_E* codes and
// leaves a message for wdy_last_error() (thread-local).
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

#define WDY_HIP(call)                                                        \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::wdy::fail(e_ == hipErrorOutOfMemory ? WDY_ENOMEM : WDY_ERUNTIME, \
                         "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                \
  } while (0)

#define WDY_CHECK_LAUNCH()                                                   \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess)                                                    \
      return ::wdy::fail(WDY_ERUNTIME, "kernel launch failed: %s (%s:%d)",  \
                         hipGetErrorString(e_), __FILE__, __LINE__);         \
  } while (0)

// --------------------------------------------------------------------------
// Kernel timing: when enabled on a sim, every launch of a timed kernel class
// is bracketed by two HIP events on the sim's stream.
// --------------------------------------------------------------------------
enum KernelClass {
  KC_SORT_HIST = 0,   // radix: all-digit histogram
  KC_SORT_PASS = 1,   // radix: one onesweep scatter pass
  KC_LEAP = 2,        // fused force + kick + drift
  KC_SCAN = 3,        // mass prefix
  KC_OTHER = 4,       // drift-half, de-sort, tie/verify helpers
  KC_WARM = 5,        // warm-start bucket sort + fused leap
  KC_COUNT = 6
};

struct Timer;
Timer *timer_create();
void timer_destroy(Timer *t);
void timer_begin(Timer *t, hipStream_t s, int kc);
void timer_end(Timer *t, hipStream_t s, int kc);
int timer_collect(Timer *t, int kc, double *ms_total, int64_t *count);

// --------------------------------------------------------------------------
// Radix sort of (uint64 key, uint32 payload) pairs: LSD, 8 passes of 8 bits,
// one all-digit histogram pass plus one onesweep scatter pass per digit.
// --------------------------------------------------------------------------
constexpr u32 RADIX_SPIN_LIMIT = 1u << 20;  // probes of one look-back wait
// keys one scatter workgroup ranks: 24 per thread.  Measured per scatter
// pass at N=2^24: 2048 keys 216.8 us, 3072 188.3, 4096 173.5, 5120 158.1,
// 6144 147.5 (fewer tiles, fewer look-back links; 6144 is the largest whose
// 78,848 B of LDS still fit two workgroups in a CU), 8192 (one per CU) 182.9
constexpr u32 RADIX_TILE_KEYS = 6144;
struct RadixWork {
  u32 n = 0;
  u32 ntiles = 0;
  u32 *hist = nullptr;       // [8][256] digit counts, then exclusive offsets
  u32 *counters = nullptr;   // [64]: [0..8) tile tickets, one per pass, and
                             // the error word (radix_err_word)
  u64 *status = nullptr;     // [ntiles][256] lookback words
  u32 epoch = 0;             // sequence number of the next pass
  // clear the error word at the start of every sort (else it accumulates
  // until its owner clears it)
  bool clear_err = false;
};

// sizes of a sort's scratch for n keys (for callers that provide it)
u32 radix_tiles(u32 n);
size_t radix_hist_bytes();
size_t radix_counter_bytes();
size_t radix_status_bytes(u32 n);  // look-back words: zero before first use
// zeroes its counters and look-back words on stream st
int radix_alloc(RadixWork *w, u32 n, hipStream_t st);
// Source slot of sorted position i.  A sort whose look-back wait gave up can
// leave slots of its payload unwritten (arbitrary values); clamping keeps
// every access that follows inside the arrays until the host reports the
// failure.  In normal operation perm[i] < n always.
__device__ __forceinline__ u32 sorted_slot(const u32 *perm, u32 i, u32 n) {
  const u32 s = perm[i];
  return s < n ? s : n - 1;
}
// Device word set nonzero when a look-back wait of a sort gave up (its
// results are then wrong; every kernel still finishes).
u32 *radix_err_word(RadixWork *w);
void radix_free(RadixWork *w);
// Sorts keys[0..n) ascending, stably; keys ping-pong between k0 and k1 and
// end in k0 (8 passes).  The payload ends in p0: if identity_payload, the
// first pass generates p = source position and p0's prior contents are
// ignored; p1 is scratch.
int radix_sort_pairs(RadixWork *w, hipStream_t s, Timer *tm, u64 *k0, u64 *k1,
                     u32 *p0, u32 *p1, bool identity_payload);

// --------------------------------------------------------------------------
// Simulation state
// --------------------------------------------------------------------------
struct WarmCtx;  // warm-start step metadata (wdy_warm.hip)
}  // namespace wdy

struct wdy_sim {
  int64_t n = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  // output copies: device-to-host on their own stream so that they overlap
  // the next interval (wdy_read_and_advance)
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_out = nullptr, ev_copied = nullptr;
  bool equal_mass = false;
  double m_equal = 0.0;      // mass * twopiG of every sheet (equal-mass runs)
  double totmass = 0.0;      // numpy.sum of mass * twopiG, supplied by the host
  double omega2 = -1.0;      // < 0: no harmonic term
  int cur = 0;               // buffer set holding the current state
  bool sorted_now = false;   // state is sorted (no drift since the last sort)
  // small mode: one workgroup advances a whole output interval in LDS;
  // key[0]/idx[0] sorted pairs, vel[0]/mass[0] in original sheet order.
  // large mode: three buffer sets; the current one holds every sheet once,
  // in sorted order after a radix step or in the warm-step segment layout.
  bool small = false;

  wdy::u64 *key[3] = {nullptr, nullptr, nullptr};
  double *vel[3] = {nullptr, nullptr, nullptr};
  wdy::i32 *idx[3] = {nullptr, nullptr, nullptr};
  double *mass[3] = {nullptr, nullptr, nullptr};
  double *cum = nullptr;         // exclusive mass prefix in sorted order
  wdy::u32 *perm[2] = {nullptr, nullptr};
  wdy::u64 *skey = nullptr;      // sort ping-pong partner of the key array
  double *tile_sums = nullptr;   // mass-prefix tile partials
  double *red = nullptr;         // reduction partials (energy, momentum)
  double *out_x = nullptr;       // de-sorted output staging
  double *out_v = nullptr;
  wdy::RadixWork radix;
  wdy::WarmCtx *warm = nullptr;  // null: the large path uses radix steps only

  // external force and its clock: t0 advances by h after every force
  // evaluation while an external force is set (wendy/wendy.c:401-402,407-408)
  wdy::ExtForce ext;
  double t0 = 0.0;
  // external force given as a host function with the reference's callback
  // signature (wendy/wendy.c:362-370), evaluated before every step on the
  // positions copied out; ext_tab holds its values on the device (and the
  // positions it is called on)
  double (*host_force)(int, double *, double, double *) = nullptr;
  double *ext_tab = nullptr;
  // ranks the fastest sheets cross per unit time, estimated from a sample
  // at creation (sizes the warm-start window once h is known)
  double disp_rate = 0.0;

  int64_t warm_steps = 0, cold_steps = 0, redos = 0;
  // warm intervals that failed, by kind: bucket over capacity, overflow list
  // full (or too many far movers into one bucket), bounded wait exceeded
  int64_t fail_capacity = 0, fail_overflow = 0, fail_stall = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  wdy::Timer *timer = nullptr;
  bool timing = false;
  // the radix sort's error word (a bounded look-back wait gave up), copied
  // to h_rerr behind the steps that sorted and checked by the next call;
  // once set the state is lost and every later call fails
  wdy::u32 *h_rerr = nullptr;
  hipEvent_t ev_rerr = nullptr;
  bool rerr_pending = false;
  bool broken = false;
};

namespace wdy {
// Step kernels (wdy_step.hip)
int launch_drift(wdy_sim *s, double hd);
int launch_sort_state(wdy_sim *s);   // sort state -> perm/keys, no payload move
// radix-path leapfrog step: sort, force, kick, drift into the next buffer
// set; spl_out (may be null) receives the sorted keys at ranks k*T
int launch_leap(wdy_sim *s, double hk, double hd, double *acc_out,
                double *cum_out, u64 *spl_out, u32 T, u32 nb, double t);
// de-sort the current state: x_out[j], v_out[j] (v_out may be null, and is
// ignored in small mode, whose velocities are already in sheet order),
// acc_out[j] = acc[slot of j] when acc is given, m_out[j] (large unequal-mass
// mode only) when m_out is given
int launch_desort(wdy_sim *s, double *x_out, double *v_out, const double *acc,
                  double *acc_out, double *m_out);
// cold sort at creation: builds the mode's layout from staged arrays
int launch_build_state(wdy_sim *s, const double *x_slot, const i32 *id_slot,
                       const double *v_id, const double *m_id);
// total energy of sheets given by device arrays in any order (m_id null:
// every mass is mq); masses are mscale * m
int energy_core(hipStream_t st, u32 n, const double *x, const double *v,
                const double *m, double mq, double mscale, double twopiG,
                double omega2, double *e_out);
// The same, enqueued only: scratch is carved from ws (at least
// energy_scratch_bytes(n, m != null) bytes, 256-byte aligned); status holds
// the sort's look-back words (radix_status_bytes(n)), zero or left by earlier
// sorts with other epochs; *epoch is the next sort epoch and is advanced.
// The result goes to e_dev (device memory); it is NaN when a look-back wait
// of the sort gave up, which also sets the device word *err_word (if
// err_word is given).  No allocation, no host synchronisation.
size_t energy_scratch_bytes(u32 n, bool masses);
int energy_enqueue(hipStream_t st, u32 n, const double *x, const double *v,
                   const double *m, double mq, double mscale, double twopiG,
                   double omega2, double *e_dev, u64 *status, void *ws,
                   u32 *epoch, const u32 **err_word);
// small mode (wdy_small.hip)
u32 small_max_n(int device);
// Largest n the one-workgroup path takes by default (further capped by its
// LDS capacity).  It runs a whole interval in one workgroup, so its time
// grows with n while the large path's stays near its floor of ~0.33 ms per
// 20-step interval up to ~2^16 sheets: measured, equal masses 0.21 vs 0.32
// ms at 2^12 and 0.43 vs 0.32 at 2^13; unequal masses (whose mass prefix
// the small path runs sequentially) 0.23 vs 0.40 at 2^10 and 0.43 vs 0.33
// at 2^11.
constexpr u64 SMALL_MAX_EQUAL = 6144;
constexpr u64 SMALL_MAX_UNEQUAL = 1536;
int launch_small_advance(wdy_sim *s, double h, int nleap, double *acc_out,
                         double *cum_out);
int launch_energy(wdy_sim *s, double twopiG, double omega2, double *e_out);
// warm-start step (wdy_warm.hip)
int warm_configure(wdy_sim *s, int shape, u32 T, u32 W);
void warm_free(wdy_sim *s);
bool warm_ready_for(wdy_sim *s, double h);
u64 *warm_cold_splitters(wdy_sim *s, u32 *T, u32 *nb);
int warm_partition_after_cold(wdy_sim *s, double hp);
// Before the first warm interval: pick the window (and, if it would exceed
// the shape's limit, the largest buckets) from s->disp_rate * h.  Does
// nothing after wdy_tune or once done.
int warm_fit(wdy_sim *s, double h);
void warm_mark_tuned(wdy_sim *s);
// the state was advanced outside the warm path: its layout must be rebuilt
void warm_invalidate(wdy_sim *s);
// enqueue one warm output interval (checked later by warm_settle)
int warm_interval(wdy_sim *s, double h, int nleap);
// wait for and check the warm interval in flight; *failed: it was rewound
// and (h, nleap) must be redone on the radix path
int warm_settle(wdy_sim *s, bool *failed, double *h, int *nleap);
int warm_current_shape(wdy_sim *s);
// the warm path's tile shape, bucket target and window; false without one
bool warm_config(wdy_sim *s, int *shape, u32 *T, u32 *W);
// wait (without checking) for the warm interval in flight, if any
void warm_wait(wdy_sim *s);
// the events bracketing the last interval when it ran in a gang
bool warm_gang_events(wdy_sim *s, hipEvent_t *start, hipEvent_t *done);
// the next interval runs outside a gang: drop the settled gang record
void warm_gang_release(wdy_sim *s);
// gangs: several simulations' warm intervals, one launch per leapfrog step
// for all of them, on the gang's own stream (wdy_warm.hip)
struct WarmGang;
int warm_gang_create(int device, WarmGang **out);
void warm_gang_destroy(WarmGang *g);
int warm_gang_device(const WarmGang *g);
// the sim's next interval of step h can run in a gang: it is warm-ready for
// h, settled, on a gang-compiled shape and has no host-function force
bool warm_gang_eligible(wdy_sim *s, double h);
int warm_gang_interval(WarmGang *g, wdy_sim *const *sims, int count,
                       const double *h, const int *nleap);
int launch_energy_arrays(int64_t n, const double *x, const double *v,
                         const double *m, double twopiG, double omega2,
                         double *e_out);
int launch_momentum_arrays(int64_t n, const double *v, const double *m,
                           double *p_out);
int launch_potential_arrays(int64_t ny, const double *y, int64_t n,
                            const double *x, const double *m, double twopiG,
                            double omega2, double *out);
int launch_energy_individual_arrays(int64_t n, const double *x, const double *v,
                                    const double *m, double twopiG,
                                    double omega2, double *out);
}  // namespace wdy
