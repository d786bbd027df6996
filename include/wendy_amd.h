/*
 * wendy_amd.h -- C ABI of libwendy_amd.so, the MI355X (gfx950) implementation
 * of wendy's approximate-integration path (drift-kick-drift leapfrog with
 * exact 1-D self-gravity).
 *
 * Plain C types only: int64_t sizes, double/int arrays in host memory,
 * opaque handles.  Every entry point returns 0 (WDY_OK) on success or a
 * negative WDY_E* code, leaves a message for wdy_last_error(), never throws
 * across the ABI and never exits the process.  A wdy_sim owns its device
 * buffers and one HIP stream; calls on one handle must be serialised by the
 * caller, different handles may be driven from different threads.
 *
 * Reference counterparts (paths relative to the reference repository):
 *   the whole sim-handle API replaces the per-call state hand-over of
 *   _wendy_nbody_approx_onestep (wendy/wendy.c:385-418, prototype
 *   wendy/wendy.h:29-34), which wendy/wendy.py:76-93 binds with ctypes and
 *   wendy/wendy.py:424-433 calls once per output.
 */
#ifndef WENDY_AMD_H
#define WENDY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WDY_ABI_VERSION 1

/* return codes */
#define WDY_OK 0
#define WDY_EINVAL (-1)   /* invalid argument */
#define WDY_ENOMEM (-2)   /* device allocation failed */
#define WDY_ERUNTIME (-3) /* HIP runtime / launch failure */
#define WDY_EVERIFY (-4)  /* on-device verification failed */
#define WDY_ENODEV (-5)   /* no usable GPU */
#define WDY_EUNSUPPORTED (-6)

/* wdy_create flags */
#define WDY_EQUAL_MASS 1  /* every m[i] is the same value: force from rank */
/* Path selection.  By default n up to the per-workgroup LDS capacity
 * (12 B per sheet) runs as one persistent workgroup per output interval
 * ("small" path), larger n as a multi-kernel step over the whole GPU
 * ("large" path).  These flags force one path, e.g. so that tests reach the
 * large path's code at small n. */
#define WDY_FORCE_LARGE 2
#define WDY_FORCE_SMALL 4
/* The large path re-sorts each leapfrog step with one warm-start kernel
 * (value buckets seeded by the previous steps' order) and falls back to the
 * full radix sort when a bucket outgrows its limits.  This flag keeps the
 * large path on the radix sort alone (for comparison and testing). */
#define WDY_NO_WARM 8
/* The simulation will share the GPU with others, advanced in gangs
 * (wdy_gang_advance): the large path keeps its throughput tile shape (1536-
 * sheet buckets) at every size.  Without the flag a simulation below ~2^22
 * sheets starts from larger workgroups, which shorten a step when its buckets
 * alone do not fill the GPU. */
#define WDY_SHARED_GPU 16

typedef struct wdy_sim wdy_sim;

/* Version of this ABI (WDY_ABI_VERSION).  No reference counterpart. */
int wdy_abi_version(void);

/* Message for the last nonzero return on this thread.  Replaces the
 * reference's warnings.warn/RuntimeError text (wendy/wendy.py:30-35,195-202). */
const char *wdy_last_error(void);

/* Number of visible GPUs (0 and WDY_ENODEV when there is none). */
int wdy_device_count(int *count);

/* Create a simulation on the current HIP device and upload its state.
 *   x, v, m   host arrays of n doubles; m is already multiplied by twopiG,
 *             exactly as wendy/wendy.py:371 does before its C call
 *   totmass   numpy.sum(m) computed by the host (wendy/wendy.py:383)
 *   omega2    omega**2, or any negative value for no harmonic term
 *             (the -1 sentinel of wendy/wendy.py:363-366)
 *   flags     WDY_EQUAL_MASS when all m are equal
 * The caller keeps ownership of the host arrays.  Replaces the setup of
 * wendy/wendy.py:363-387 (xi array, copies, cumulmass). */
int wdy_create(int64_t n, const double *x, const double *v, const double *m,
               double totmass, double omega2, int flags, wdy_sim **out);

/* Enqueue one output interval: half drift, nleap x (sort, force, kick,
 * drift), last drift halved -- the body of _wendy_nbody_approx_onestep
 * (wendy/wendy.c:394-411) with h = dt/nleap.  Asynchronous: it returns
 * once the interval is enqueued.  The device's verdict on a warm-start
 * interval (a bucket that outgrew its capacity, too many far movers) is
 * checked by the next call on the handle that reads or extends the state
 * (wdy_advance, wdy_read, wdy_read_and_advance, wdy_energy, wdy_tune, the
 * ext-force setters and the statistics), which then redoes the interval from
 * its untouched start on the radix path before doing its own work; so a
 * result never depends on whether a redo happened. */
int wdy_advance(wdy_sim *sim, double h, int nleap);

/* De-sort the state into original sheet order and copy it to the host:
 * x_out[j], v_out[j] for sheet j (wendy/wendy.c:413-415).  Blocks until the
 * copy is done.  seconds_out (may be NULL) receives the device time of the
 * last wdy_advance, the counterpart of time_elapsed (wendy/wendy.c:396,416).
 * Either array pointer may be NULL. */
int wdy_read(wdy_sim *sim, double *x_out, double *v_out, double *seconds_out);

/* One generator step in a single call: de-sort the current state, start the
 * next output interval (h, nleap) on the device, and copy the de-sorted x, v
 * to the host while that interval runs; returns once the copy is complete
 * (the interval may still be running).  seconds_out: device time of the
 * interval just read.  With host buffers from wdy_host_alloc the copy
 * overlaps the interval; with pageable memory the result is the same but the
 * copy may not overlap.  The counterpart of one next() of the reference
 * generator (wendy/wendy.py:424-437) with the following interval started
 * ahead of the caller's request. */
int wdy_read_and_advance(wdy_sim *sim, double *x_out, double *v_out, double h,
                         int nleap, double *seconds_out);

/* Benchmark calibration: `launches` launches of a plain device copy (read
 * `bytes`, write `bytes`, 16 B per thread per iteration) between two freshly
 * allocated buffers; *ms_out (may be NULL) receives the mean milliseconds per
 * launch.  Its HBM counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) scale those
 * of the kernels above to bytes.  No reference counterpart. */
int wdy_copy_probe(int64_t bytes, int launches, double *ms_out);

/* Page-locked host memory for output buffers (hipHostMalloc). */
int wdy_host_alloc(int64_t bytes, void **out);
int wdy_host_free(void *ptr);

/* Total energy of the current device state, computed on the device without
 * copying the state out: twopiG * sum_{i<j} m_i m_j |x_i - x_j| +
 * sum m v^2/2 (+ sum m omega^2 x^2/2 when omega2 >= 0), with m the masses
 * given to wdy_create divided by twopiG.  Counterpart of wendy.energy
 * (wendy/wendy.py:439-475) evaluated on the generator's live state. */
int wdy_energy(wdy_sim *sim, double twopiG, double omega2, double *e_out);

/* Free the simulation's device buffers and stream. */
int wdy_destroy(wdy_sim *sim);

/* Gangs: several simulations advanced together, one kernel launch per
 * leapfrog step for all of them, so that many medium simulations fill the
 * GPU the way one large one does (serving, ensembles of large members).
 * No reference counterpart: the reference advances one simulation per
 * generator (wendy/wendy.py:424-437). */
typedef struct wdy_gang wdy_gang;

/* A gang on the current HIP device, with its own stream. */
int wdy_gang_create(wdy_gang **out);

/* Waits for the gang's work and frees it.  Its members stay valid. */
int wdy_gang_destroy(wdy_gang *gang);

/* Enqueue one output interval of each of sims[0..count) (1 <= count <= 64,
 * distinct, on the gang's device): nleap[i] leapfrog steps of h[i], with the
 * results and the error handling of wdy_advance(sims[i], h[i], nleap[i]),
 * bit for bit.  Members whose next interval runs on the warm-start kernel
 * (large path, already partitioned for h[i], one of the automatically
 * chosen tile shapes, no host-function external force) run in the gang's
 * launches; the others run their interval on their own streams.
 * Asynchronous like wdy_advance: each member's interval is checked by the
 * next call on that member. */
int wdy_gang_advance(wdy_gang *gang, wdy_sim *const *sims, int count,
                     const double *h, const int *nleap);

/* Warm-start tuning of a large-path sim.
 *   shape          tile shape of the warm kernel (-1 keeps the current one):
 *                  0: 1024 threads, buckets of up to 4096 sheets
 *                  1: 512 threads, up to 2048;  2: 256 threads, up to 1024;
 *                  3: 512 threads, up to 1024;  4: 512 threads, up to 2048
 *                  with 1024 position bins;  5: 512 threads, up to 3072
 *                  with 1024 bins;  6: 768 threads, up to 3072 with 768
 *                  bins;  7: 256 threads, up to 2048 with 1024 bins;
 *                  8: 256 threads, up to 1536 with 2048 bins (the
 *                  default);  9: 256 threads, up to 1792 with 1024 bins
 *   bucket_target  sheets per value bucket, WDY_BUCKET_MIN .. the shape's
 *                  capacity; 0 leaves 1/5 of the capacity (at least 256
 *                  sheets) free
 *   window         neighbouring buckets a sheet may move into per step
 *                  (0 .. WDY_WINDOW_MAX, default 6, widened at the first
 *                  interval when sampled velocities need more; at most 31
 *                  for shape 6, 63 for shapes 3-5); farther movers travel
 *                  through an overflow list
 * Takes effect from the next output interval, which runs on the radix path
 * to rebuild the bucket layout.  No reference counterpart (the reference
 * re-sorts from scratch each step). */
#define WDY_BUCKET_MIN 32
#define WDY_WINDOW_MAX 64
int wdy_tune(wdy_sim *sim, int shape, int bucket_target, int window);

/* External force evaluated on the device, the counterpart of nbody's
 * ext_force callback (wendy/wendy.py:389-420; the C side adds it at
 * wendy/wendy.c:362-370 and advances *t0 by h after every force evaluation,
 * :401-402,407-408).  Evaluated inside the step kernels, so the state never
 * leaves the device, for a menu of forms (any other function goes through
 * wdy_set_ext_force_host), at the sorted position x and the step's force
 * time t (t0 at the first force of an output interval, then +h per step, the
 * reference's clock exactly):
 *   WDY_EXT_HARMONIC    -p0 * x                              (p0 = omega**2)
 *   WDY_EXT_ISOTHERMAL  (p0 * tanh((0.5 * x) / p1)) / p1     (amp, zh)
 *   WDY_EXT_PULSE       p0 * exp(-0.5 * ((t - p1) / p2)**2)  (amp, tc, tau)
 * params holds p0, p1, p2 (ignored for WDY_EXT_NONE); t0 sets the clock.
 * The acceleration becomes a = E(x, t) + F; without a force t0 is frozen,
 * as in the reference.  Applies from the next wdy_advance. */
#define WDY_EXT_NONE 0
#define WDY_EXT_HARMONIC 1
#define WDY_EXT_ISOTHERMAL 2
#define WDY_EXT_PULSE 3
int wdy_set_ext_force(wdy_sim *sim, int kind, const double *params, double t0);
/* External force given as a host function with the reference's callback
 * signature, double f(int N, double *x, double t, double *a): the
 * counterpart of an arbitrary ext_force callable (wendy/wendy.py:389-420,
 * called at wendy/wendy.c:362-370).  Before every leapfrog step the
 * positions of the force evaluation are copied to the host and the function
 * is called the way the reference calls it -- above 10 sheets once with the
 * de-sorted positions x[N] (it writes a[N]); for up to 10 sheets once per
 * sheet in sorted order with N = 1, a = NULL, the return value being that
 * sheet's acceleration -- then the values are uploaded and added to the
 * self-gravity (a = E + F).  t0 sets the clock, which advances by h after
 * every evaluation.  Each step costs a device-host round trip; the state
 * stays on the device.  The simulation must be on the large path
 * (WDY_FORCE_LARGE, or n above the small-path capacity): otherwise
 * WDY_EUNSUPPORTED.  NULL removes the force.  Replaces any force set by
 * wdy_set_ext_force, and the reverse. */
int wdy_set_ext_force_host(wdy_sim *sim,
                           double (*force)(int, double *, double, double *),
                           double t0);
/* The simulation clock t0 (advanced only while an external force is set). */
int wdy_time(wdy_sim *sim, double *t_out);

/* Leapfrog steps taken so far by the warm-start kernel and by the radix
 * path, and the number of output intervals redone on the radix path after a
 * warm-start failure.  No reference counterpart. */
int wdy_path_stats(wdy_sim *sim, int64_t *warm_steps, int64_t *cold_steps,
                   int64_t *redos);

/* Warm intervals redone so far, by cause: a bucket outgrew its capacity,
 * far movers overflowed their list, or a bounded wait of the look-back gave
 * up (the last disables the warm path for the sim).  No reference
 * counterpart. */
int wdy_warm_failures(wdy_sim *sim, int64_t *capacity, int64_t *overflow,
                      int64_t *stall);

/* The warm path's current tile shape (see wdy_tune), bucket target and
 * window, as chosen by the defaults, the window fitting at the first
 * interval, failure adaptation or wdy_tune; shape -1 when the sim has no
 * warm path (small path, WDY_NO_WARM, or disabled after a stall).  No
 * reference counterpart. */
int wdy_warm_config(wdy_sim *sim, int *shape, int *bucket_target, int *window);

/* Per-kernel-class device time, for benchmarks.  wdy_set_timing(sim, 1)
 * brackets every subsequent launch with HIP events on the sim's stream;
 * wdy_kernel_time returns the summed milliseconds and launch count of one
 * class since timing was enabled (classes: 0 radix histogram, 1 radix
 * scatter pass, 2 fused leap, 3 mass prefix, 4 other, 5 warm-start sort).
 * No reference counterpart. */
int wdy_set_timing(wdy_sim *sim, int on);
int wdy_kernel_time(wdy_sim *sim, int kernel_class, double *ms_total,
                    int64_t *count);

/* ---- stateless diagnostics on host arrays (uploaded, computed on the GPU) */

/* wendy.energy(x, v, m, twopiG, individual=False, omega)
 * (wendy/wendy.py:439-475); omega2 < 0: no harmonic term. */
int wdy_energy_arrays(int64_t n, const double *x, const double *v,
                      const double *m, double twopiG, double omega2,
                      double *e_out);

/* wendy.energy(..., individual=True) (wendy/wendy.py:458-466): out[n]. */
int wdy_energy_individual_arrays(int64_t n, const double *x, const double *v,
                                 const double *m, double twopiG, double omega2,
                                 double *out);

/* Allocation-free form of wdy_energy_arrays, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  x, v, m (raw masses, as for
 * wdy_energy_arrays) and e_out are DEVICE pointers; the computation is
 * enqueued on `stream` (a hipStream_t; NULL: the default stream) and the
 * call returns without waiting: *e_out holds the energy, bit for bit the
 * value wdy_energy_arrays returns for the same arrays, once the stream has
 * reached it (NaN if a bounded wait of its sort gave up, which does not
 * happen in normal operation).  Every scratch buffer comes from one static
 * device array of
 * about 51 MiB that each GPU the library runs on holds until the process
 * ends; nothing is allocated or freed.  n must be 1 .. WDY_NOALLOC_MAX_N
 * (WDY_EINVAL beyond).  All calls on one GPU share that array, so calls that
 * overlap on one GPU (other streams, other threads) race: serialise them. */
#define WDY_NOALLOC_MAX_N (1 << 20)
int wdy_energy_device_noalloc(int64_t n, const double *x, const double *v,
                              const double *m, double twopiG, double omega2,
                              double *e_out, void *stream);

/* wendy.momentum(v, m) (wendy/wendy.py:477-491). */
int wdy_momentum_arrays(int64_t n, const double *v, const double *m,
                        double *p_out);

/* wendy.potential(y, x, v, m, twopiG, omega) (wendy/wendy.py:493-517):
 * out[ny]. */
int wdy_potential_arrays(int64_t ny, const double *y, int64_t n,
                         const double *x, const double *m, double twopiG,
                         double omega2, double *out);

/* ---- batched ensemble ------------------------------------------------------
 * Many independent simulations (uneven n) advanced together: one kernel
 * launch per output interval for all of them.  No reference counterpart
 * (the reference runs one simulation per generator); each simulation's
 * arithmetic is that of wdy_create/wdy_advance.
 *   offsets   nsims + 1 increasing values, offsets[0] = 0: simulation k owns
 *             sheets [offsets[k], offsets[k+1]) of x, v, m
 *   m         masses already multiplied by twopiG (>= 0)
 *   totmass   per simulation, numpy.sum of its m
 *   omega2    per simulation, omega**2 or negative for none */
typedef struct wdy_batch wdy_batch;
int wdy_batch_create(int64_t nsims, const int64_t *offsets, const double *x,
                     const double *v, const double *m, const double *totmass,
                     const double *omega2, wdy_batch **out);
/* One output interval of every simulation (same h and nleap for all). */
int wdy_batch_advance(wdy_batch *batch, double h, int nleap);
/* De-sorted x, v of all simulations, concatenated as at creation; blocks.
 * seconds_out: device time of the last wdy_batch_advance. */
int wdy_batch_read(wdy_batch *batch, double *x_out, double *v_out,
                   double *seconds_out);
int wdy_batch_destroy(wdy_batch *batch);

/* ---- serving pool ----------------------------------------------------------
 * Continuous batching: simulations join and leave while others run.  The
 * pool owns an arena of capacity_sheets sheets (56 B of HBM per sheet: two
 * buffer sets of key, velocity, index and mass) and max_sims slots; every
 * wdy_pool_step advances a chosen subset of live slots, each by its own h and
 * step count (one output interval: half drift, nleap x (sort, force, kick,
 * drift), last drift halved -- as wdy_advance), in one launch.  Uploads by
 * wdy_pool_add run on a second stream and overlap a running step.  A
 * simulation's result does not depend on what else shares the pool.  No
 * reference counterpart.  Calls on one pool must be serialised. */
typedef struct wdy_pool wdy_pool;
/* max_n: largest simulation the pool accepts (sizes the staging buffers). */
int wdy_pool_create(int64_t capacity_sheets, int max_sims, int64_t max_n,
                    wdy_pool **out);
/* Admit one simulation (m already x twopiG, totmass = numpy.sum(m), omega2 <
 * 0 for none).  WDY_ENOMEM when no slot or no free arena range of n sheets is
 * left: the caller keeps the request queued and retries after a removal.
 * The host arrays must stay valid until the next wdy_pool_sync or
 * wdy_pool_read returns. */
int wdy_pool_add(wdy_pool *pool, int64_t n, const double *x, const double *v,
                 const double *m, double totmass, double omega2, int *slot_out);
/* Asynchronous: slot slots[k] advances by one interval of nleap[k] steps of
 * size h[k] (nleap[k] = 0: untouched). */
int wdy_pool_step(wdy_pool *pool, int count, const int *slots, const double *h,
                  const int *nleap);
/* De-sorted x, v of one slot (either may be NULL); blocks until every step
 * enqueued before it is done.  WDY_EVERIFY when this slot's simulation
 * failed its on-device check (each slot has its own error word, cleared
 * when the slot is next admitted to). */
int wdy_pool_read(wdy_pool *pool, int slot, double *x_out, double *v_out);
/* Free a slot and its arena range (reused only after the steps enqueued so
 * far have finished). */
int wdy_pool_remove(wdy_pool *pool, int slot);
/* Wait for all uploads and steps; seconds_out: device time of the last step. */
int wdy_pool_sync(wdy_pool *pool, double *seconds_out);
int wdy_pool_stats(wdy_pool *pool, int *live, int64_t *used_sheets,
                   int64_t *largest_free);
int wdy_pool_destroy(wdy_pool *pool);

/* ---- exact event-driven solver (CPU) ---------------------------------------
 * The reference's approx=False path (wendy/wendy.c:33-315 with the set-up of
 * wendy/wendy.py:158-279): sheets fly on parabolas (or, with omega > 0, on
 * harmonic orbits about their equilibrium offsets) between crossings of
 * position neighbours; each crossing swaps the pair and updates the forces.
 * Runs on the host, as a machine-precision check of the GPU leapfrog.
 *   m       masses already multiplied by twopiG and, with omega > 0, divided
 *           by omega**2 (wendy/wendy.py:170-174); all > 0
 *   omega   external harmonic frequency, or <= 0 for none
 * wdy_exact_advance runs one output interval dt; *ncoll receives its number
 * of crossings (the reference's ncoll_c) and *hit_max is set when maxcoll
 * were processed (the reference's err = -2), in which case the remaining
 * crossings of the interval were skipped, as in the reference;
 * wdy_exact_reset then re-derives order, forces and crossing times from the
 * current positions (the warn_maxcoll recovery of wendy/wendy.py:225-232).
 * Ties in x at set-up keep the lower sheet index first. */
typedef struct wdy_exact wdy_exact;
int wdy_exact_create(int64_t n, const double *x, const double *v, const double *m,
                     double omega, wdy_exact **out);
int wdy_exact_advance(wdy_exact *ex, double dt, int64_t maxcoll, int64_t *ncoll,
                      int *hit_max, double *seconds_out);
int wdy_exact_reset(wdy_exact *ex);
int wdy_exact_read(wdy_exact *ex, double *x_out, double *v_out);
int wdy_exact_destroy(wdy_exact *ex);

/* ---- reference-signature exact-solver entry points -------------------------
 * Same names, arguments and meaning as the reference's, so that its loader
 * (wendy/wendy.py:40-72,94-101 binds all four at import) finds them in this
 * library.  The caller keeps the solver state between calls exactly as the
 * reference's set-up (wendy/wendy.py:243-279) lays it out: x, v, a per sheet
 * (a: force, or with omega the equilibrium offset), m (already x twopiG, and
 * / omega**2 with omega), sindx (the sheet at each position rank), tcoll[N-1]
 * (crossing times of neighbouring ranks), *cindx / *next_tcoll (the earliest
 * crossing).  One call runs one output interval dt on the host CPU with the
 * same event queue as wdy_exact_*: x, v, a, sindx, tcoll, *cindx and
 * *next_tcoll are updated in place, with times shifted to the interval's
 * end; *ncoll receives the crossings processed, *time_elapsed the CPU time of
 * the crossing loop, and *err is set to -2 when maxcoll crossings were
 * processed (left untouched otherwise), as in the reference.  Crossing times
 * that are exactly equal are taken in pair-index order.
 *   _wendy_nbody_onestep       replaces wendy/wendy.c:53-165
 *   _wendy_nbody_harm_onestep  replaces wendy/wendy.c:167-315 (omega > 0;
 *                              otherwise *err = WDY_EINVAL, nothing moves)
 *   _solve_coll_quad           replaces wendy/wendy.c:33-38: the root
 *                              0.5 * (r + sqrt(r*r - 4 c0/c2)), r = -c1/c2
 *   _solve_coll_harm           replaces wendy/wendy.c:39-52 */
void _wendy_nbody_onestep(int N, double *x, double *v, double *a, double *m,
                          int *sindx, int *cindx, double *next_tcoll,
                          double *tcoll, double dt, int maxcoll, int *err,
                          int *ncoll, double *time_elapsed);
void _wendy_nbody_harm_onestep(int N, double *x, double *v, double *a,
                               double *m, int *sindx, int *cindx,
                               double *next_tcoll, double *tcoll, double dt,
                               int maxcoll, int *err, int *ncoll,
                               double *time_elapsed, double omega);
double _solve_coll_quad(double c0, double c1, double c2);
double _solve_coll_harm(double c0, double c1, double c2, double omega);

/* ---- reference-signature compatibility entry point ----------------------
 * Same name, arguments and meaning as the reference's
 * _wendy_nbody_approx_onestep (wendy/wendy.c:385-393, wendy/wendy.h:29-34),
 * so the reference's own ctypes binding (wendy/wendy.py:76-93) can load this
 * library instead of its C extension.  Each call uploads xi/v/m, runs the
 * nleap leapfrog steps on the GPU, and writes back xi (sorted), x
 * (de-sorted), v, a (last force, original order) and cumulmass (last mass
 * prefix, sorted order) and *time_elapsed (device seconds).  *t0 is not
 * touched unless ext_force is given: then the function is called before
 * every step as wdy_set_ext_force_host describes (the reference's own
 * calling convention) and *t0 advances by dt per step, as in the reference.
 * Every sort_type maps to the device sort; unlike the reference, which
 * silently skips the sort, an unknown sort_type (not 0..4) sets
 * *err = WDY_EINVAL without computing.  Other failures set *err to the
 * WDY_E* code; after any failure x, v and xi[].val hold NaN and a message
 * goes to stderr (the reference's caller never reads err). */
struct array_w_index {
  int idx;
  double val;
};
void _wendy_nbody_approx_onestep(int N, struct array_w_index *xi, double *x,
                                 double *v, double *m, double *a,
                                 double totmass, double dt, int nleap,
                                 double *t0, double omega2,
                                 double (*ext_force)(int, double *, double,
                                                     double *),
                                 int sort_type, int *err,
                                 double *time_elapsed, double *cumulmass);

#ifdef __cplusplus
}
#endif

#endif /* WENDY_AMD_H */
