// C ABI of libwendy_amd.so (declared in include/wendy_amd.h).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <new>
#include <vector>

#include "wdy_internal.h"
#include "wdy_warm.h"

namespace wdy {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// ---- kernel timing ----------------------------------------------------------
struct Timer {
  struct Pair {
    hipEvent_t a, b;
  };
  std::vector<Pair> open[KC_COUNT];  // recorded, not yet folded
  std::vector<hipEvent_t> pool;
  double ms[KC_COUNT] = {};
  int64_t count[KC_COUNT] = {};
  hipEvent_t pending[KC_COUNT] = {};
};

static hipEvent_t ev_get(Timer *t) {
  if (!t->pool.empty()) {
    hipEvent_t e = t->pool.back();
    t->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

static void fold(Timer *t, int kc) {
  for (auto &p : t->open[kc]) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess &&
        hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      t->ms[kc] += ms;
      t->count[kc] += 1;
    }
    t->pool.push_back(p.a);
    t->pool.push_back(p.b);
  }
  t->open[kc].clear();
}

Timer *timer_create() { return new (std::nothrow) Timer(); }

void timer_destroy(Timer *t) {
  if (!t) return;
  for (int kc = 0; kc < KC_COUNT; ++kc) {
    for (auto &p : t->open[kc]) {
      (void)hipEventDestroy(p.a);
      (void)hipEventDestroy(p.b);
    }
    if (t->pending[kc]) (void)hipEventDestroy(t->pending[kc]);
  }
  for (auto e : t->pool) (void)hipEventDestroy(e);
  delete t;
}

void timer_begin(Timer *t, hipStream_t s, int kc) {
  hipEvent_t e = ev_get(t);
  if (!e) return;
  (void)hipEventRecord(e, s);
  t->pending[kc] = e;
}

void timer_end(Timer *t, hipStream_t s, int kc) {
  hipEvent_t a = t->pending[kc];
  if (!a) return;
  t->pending[kc] = nullptr;
  hipEvent_t b = ev_get(t);
  if (!b) {
    t->pool.push_back(a);
    return;
  }
  (void)hipEventRecord(b, s);
  t->open[kc].push_back({a, b});
  if (t->open[kc].size() >= 4096) fold(t, kc);
}

int timer_collect(Timer *t, int kc, double *ms_total, int64_t *count) {
  fold(t, kc);
  *ms_total = t->ms[kc];
  *count = t->count[kc];
  return WDY_OK;
}

// ---- sim lifecycle ----------------------------------------------------------
__global__ void k_iota(u32 n, i32 *__restrict__ idx) {
  const u32 i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) idx[i] = (i32)i;
}

static void free_sim(wdy_sim *s) {
  if (!s) return;
  for (int b = 0; b < 3; ++b) {
    if (s->key[b]) (void)hipFree(s->key[b]);
    if (s->vel[b]) (void)hipFree(s->vel[b]);
    if (s->idx[b]) (void)hipFree(s->idx[b]);
    if (s->mass[b]) (void)hipFree(s->mass[b]);
  }
  for (int b = 0; b < 2; ++b)
    if (s->perm[b]) (void)hipFree(s->perm[b]);
  warm_free(s);
  if (s->skey) (void)hipFree(s->skey);
  if (s->cum) (void)hipFree(s->cum);
  if (s->tile_sums) (void)hipFree(s->tile_sums);
  if (s->red) (void)hipFree(s->red);
  if (s->out_x) (void)hipFree(s->out_x);
  if (s->out_v) (void)hipFree(s->out_v);
  if (s->ext_tab) (void)hipFree(s->ext_tab);
  radix_free(&s->radix);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->ev_rerr) (void)hipEventDestroy(s->ev_rerr);
  if (s->h_rerr) (void)hipHostFree(s->h_rerr);
  if (s->ev_out) (void)hipEventDestroy(s->ev_out);
  if (s->ev_copied) (void)hipEventDestroy(s->ev_copied);
  if (s->copy_stream) (void)hipStreamDestroy(s->copy_stream);
  if (s->timer) timer_destroy(s->timer);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// Ranks per unit time crossed by the fastest sheets: n * (peak density of a
// strided sample of up to 4096 positions, 32-neighbour estimate) * (largest
// |v| of the sample).  A sizing hint only.
static double displacement_rate(int64_t n, const double *x_slot, const double *v_id,
                                const i32 *id_slot) {
  const int64_t ns = n < 4096 ? n : 4096, stride = n / ns;
  if (ns < 64) return 0.0;
  std::vector<double> xs((size_t)ns);
  double vmax = 0.0;
  for (int64_t i = 0; i < ns; ++i) {
    const int64_t k = i * stride;
    xs[(size_t)i] = x_slot[k];
    const double vk = std::fabs(v_id[id_slot ? id_slot[k] : k]);
    vmax = vk > vmax ? vk : vmax;
  }
  std::sort(xs.begin(), xs.end());
  const int64_t K = 32;
  double rho = 0.0;
  for (int64_t i = 0; i + K < ns; ++i) {
    const double w = xs[(size_t)(i + K)] - xs[(size_t)i];
    if (w > 0.0) rho = std::max(rho, (double)K / ((double)ns * w));
  }
  return (double)n * rho * vmax;
}

// Creates a sim from positions given per slot (x_slot[i] is the position of
// sheet id_slot[i]; id_slot null means the identity) and velocities/masses
// given per sheet (v_id[j], m_id[j]).  The positions are sorted on the device
// at once, so the state is sorted before the first step.
static int create_impl(int64_t n, const double *x_slot, const i32 *id_slot,
                       const double *v_id, const double *m_id, double totmass,
                       double omega2, int flags, wdy_sim **out) {
  if (!out) return fail(WDY_EINVAL, "wdy_create: out is NULL");
  *out = nullptr;
  if (n < 1 || n >= (int64_t(1) << 30))
    return fail(WDY_EINVAL, "wdy_create: n=%lld outside [1, 2^30)", (long long)n);
  if (!x_slot || !v_id || !m_id) return fail(WDY_EINVAL, "wdy_create: NULL array");
  if (flags & ~(WDY_EQUAL_MASS | WDY_FORCE_LARGE | WDY_FORCE_SMALL |
                WDY_NO_WARM | WDY_SHARED_GPU))
    return fail(WDY_EINVAL, "wdy_create: unknown flags 0x%x", flags);
  if ((flags & WDY_FORCE_LARGE) && (flags & WDY_FORCE_SMALL))
    return fail(WDY_EINVAL, "wdy_create: WDY_FORCE_LARGE and WDY_FORCE_SMALL");
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(x_slot[i]) || !std::isfinite(v_id[i]) ||
        !std::isfinite(m_id[i]))
      return fail(WDY_EINVAL, "wdy_create: non-finite input at sheet %lld",
                  (long long)i);
  const bool eq = flags & WDY_EQUAL_MASS;
  if (eq)
    for (int64_t i = 1; i < n; ++i)
      if (m_id[i] != m_id[0])
        return fail(WDY_EINVAL,
                    "wdy_create: WDY_EQUAL_MASS set but m[%lld] != m[0]",
                    (long long)i);
  if (!std::isfinite(totmass) || !std::isfinite(omega2))
    return fail(WDY_EINVAL, "wdy_create: non-finite totmass or omega2");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(WDY_ENODEV, "wdy_create: no GPU visible");

  wdy_sim *s = new (std::nothrow) wdy_sim();
  if (!s) return fail(WDY_ENOMEM, "wdy_create: host allocation failed");
  int rc = [&]() -> int {
    WDY_HIP(hipGetDevice(&s->device));
    const u32 smax = small_max_n(s->device);
    if ((flags & WDY_FORCE_SMALL) && (u64)n > smax)
      return fail(WDY_EINVAL, "wdy_create: WDY_FORCE_SMALL but n=%lld > %u",
                  (long long)n, smax);
    const u64 lim = eq ? SMALL_MAX_EQUAL : SMALL_MAX_UNEQUAL;
    const u64 slim = smax < lim ? smax : lim;
    s->small = (flags & WDY_FORCE_SMALL) ||
               (!(flags & WDY_FORCE_LARGE) && (u64)n <= slim);
    WDY_HIP(hipStreamCreate(&s->stream));
    WDY_HIP(hipStreamCreate(&s->copy_stream));
    WDY_HIP(hipEventCreate(&s->ev0));
    WDY_HIP(hipEventCreate(&s->ev1));
    WDY_HIP(hipEventCreateWithFlags(&s->ev_out, hipEventDisableTiming));
    WDY_HIP(hipEventCreateWithFlags(&s->ev_copied, hipEventDisableTiming));
    WDY_HIP(hipEventCreateWithFlags(&s->ev_rerr, hipEventDisableTiming));
    WDY_HIP(hipHostMalloc((void **)&s->h_rerr, 8, hipHostMallocDefault));
    *s->h_rerr = 0;
    s->timer = timer_create();
    if (!s->timer) return fail(WDY_ENOMEM, "wdy_create: timer allocation failed");
    s->n = n;
    s->equal_mass = eq;
    s->m_equal = m_id[0];
    s->totmass = totmass;
    s->omega2 = omega2 < 0.0 ? -1.0 : omega2;
    const size_t N = (size_t)n;
    for (int b = 0; b < 3; ++b) {
      if (s->small && b == 2) break;
      WDY_HIP(hipMalloc(&s->key[b], N * sizeof(u64)));
      WDY_HIP(hipMalloc(&s->idx[b], N * sizeof(i32)));
      if (b == 0 || !s->small) {
        WDY_HIP(hipMalloc(&s->vel[b], N * sizeof(double)));
        if (!eq) WDY_HIP(hipMalloc(&s->mass[b], N * sizeof(double)));
      }
    }
    for (int b = 0; b < 2; ++b)
      WDY_HIP(hipMalloc(&s->perm[b], N * sizeof(u32)));
    WDY_HIP(hipMalloc(&s->skey, N * sizeof(u64)));
    WDY_HIP(hipMalloc(&s->cum, N * sizeof(double)));
    WDY_HIP(hipMalloc(&s->tile_sums, (N / 4096 + 2) * sizeof(double)));
    WDY_HIP(hipMalloc(&s->red, (N / 256 + 4) * sizeof(double)));
    WDY_HIP(hipMalloc(&s->out_x, N * sizeof(double)));
    WDY_HIP(hipMalloc(&s->out_v, N * sizeof(double)));
    int r = radix_alloc(&s->radix, (u32)n, s->stream);
    if (r) return r;
    // staging: x -> out_x, v -> out_v, m -> cum, sheet ids -> idx[1]
    WDY_HIP(hipMemcpyAsync(s->out_x, x_slot, N * 8, hipMemcpyHostToDevice,
                           s->stream));
    WDY_HIP(hipMemcpyAsync(s->out_v, v_id, N * 8, hipMemcpyHostToDevice,
                           s->stream));
    if (!eq)
      WDY_HIP(hipMemcpyAsync(s->cum, m_id, N * 8, hipMemcpyHostToDevice,
                             s->stream));
    if (id_slot) {
      WDY_HIP(hipMemcpyAsync(s->idx[1], id_slot, N * 4, hipMemcpyHostToDevice,
                             s->stream));
    } else {
      k_iota<<<(unsigned)((N + 255) / 256), 256, 0, s->stream>>>((u32)n,
                                                                  s->idx[1]);
      WDY_CHECK_LAUNCH();
    }
    r = launch_build_state(s, s->out_x, s->idx[1], s->out_v, s->cum);
    if (r) return r;
    // warm-start step for the large path, unless disabled; it needs
    // non-negative masses (fixed-point prefix)
    bool warm = !s->small && !(flags & WDY_NO_WARM);
    for (int64_t i = 0; warm && i < n; ++i) warm = m_id[i] >= 0.0;
    if (warm) {
      const int shape0 = (flags & WDY_SHARED_GPU) ? WARM_SHAPE_DEFAULT
                                                  : warm_initial_shape(n, eq);
      r = warm_configure(s, shape0, 0u, (u32)WARM_WINDOW_DEFAULT);
      if (r) return r;
      s->disp_rate = displacement_rate(n, x_slot, v_id, id_slot);
    }
    WDY_HIP(hipMemcpyAsync(s->h_rerr, radix_err_word(&s->radix), 4,
                           hipMemcpyDeviceToHost, s->stream));
    WDY_HIP(hipStreamSynchronize(s->stream));
    if (*s->h_rerr)
      return fail(WDY_ERUNTIME, "wdy_create: a look-back wait of the initial "
                  "sort gave up");
    return WDY_OK;
  }();
  if (rc) {
    free_sim(s);
    return rc;
  }
  *out = s;
  return WDY_OK;
}

// Inputs of the number of sheets up to which the reference evaluates an
// external force sheet by sheet in sorted order rather than on the whole
// de-sorted array (EXTERNAL_SWITCH, wendy/wendy.h:9-11, wendy/wendy.c:362-370).
constexpr int64_t EXT_SWITCH = 10;

// One output interval with the external force of a host function: the
// radix-path steps, with the positions of each step's force evaluation
// copied out, the function called on them as the reference calls its
// callback (wendy/wendy.c:362-370: the whole de-sorted array above
// EXT_SWITCH sheets, else one sheet at a time in sorted order), and its
// values uploaded to the per-sheet table the step kernel adds to the
// self-gravity (a = E + F).  The clock t0 advances by h after every force
// evaluation (wendy/wendy.c:401-402,407-408).
static int advance_host_force(wdy_sim *s, double h, int nleap,
                              double *acc_last, double *cum_last) {
  const size_t N = (size_t)s->n;
  // [0, N): the force table; [N, 2N): positions handed to the function (not
  // out_x, which an overlapped output copy may still be reading)
  if (!s->ext_tab) WDY_HIP(hipMalloc(&s->ext_tab, 2 * N * sizeof(double)));
  double *const xpos = s->ext_tab + N;
  std::vector<double> hx(N), ha(N);
  std::vector<u64> hk;
  std::vector<i32> hi;
  ExtForce f;
  f.kind = EXT_TABLE;
  f.tab = s->ext_tab;
  const ExtForce saved = s->ext;
  s->ext = f;
  int rc = [&]() -> int {
    if (int r = launch_drift(s, h / 2.)) return r;
    double t = s->t0;
    for (int ii = 0; ii < nleap; ++ii) {
      const bool last = ii == nleap - 1;
      if ((int64_t)N > EXT_SWITCH) {
        if (int r = launch_desort(s, xpos, nullptr, nullptr, nullptr, nullptr))
          return r;
        WDY_HIP(hipMemcpyAsync(hx.data(), xpos, N * 8, hipMemcpyDeviceToHost,
                               s->stream));
        WDY_HIP(hipStreamSynchronize(s->stream));
        (void)s->host_force((int)N, hx.data(), t, ha.data());
      } else {
        hk.resize(N);
        hi.resize(N);
        WDY_HIP(hipMemcpyAsync(hk.data(), s->key[s->cur], N * 8,
                               hipMemcpyDeviceToHost, s->stream));
        WDY_HIP(hipMemcpyAsync(hi.data(), s->idx[s->cur], N * 4,
                               hipMemcpyDeviceToHost, s->stream));
        WDY_HIP(hipStreamSynchronize(s->stream));
        std::vector<size_t> o(N);
        for (size_t q = 0; q < N; ++q) o[q] = q;
        std::sort(o.begin(), o.end(), [&](size_t p, size_t q) {
          return hk[p] < hk[q] || (hk[p] == hk[q] && hi[p] < hi[q]);
        });
        for (size_t q : o) {
          double xq = decode_pos(hk[q]);
          ha[(size_t)hi[q]] = s->host_force(1, &xq, t, nullptr);
        }
      }
      WDY_HIP(hipMemcpyAsync(s->ext_tab, ha.data(), N * 8, hipMemcpyHostToDevice,
                             s->stream));
      if (int r = launch_leap(s, h, last ? h / 2. : h, last ? acc_last : nullptr,
                              last ? cum_last : nullptr, nullptr, 0u, 0u, t))
        return r;
      t = t + h;
    }
    // the host arrays are reused by the next step: wait for the upload
    WDY_HIP(hipStreamSynchronize(s->stream));
    s->t0 = t;
    s->cold_steps += nleap;
    warm_invalidate(s);  // the warm layout no longer describes the state
    return WDY_OK;
  }();
  s->ext = saved;
  return rc;
}

// One output interval on the radix path (a full sort every step), then,
// when the sim has a warm path, the partition that hands the next interval
// to it.
static int cold_interval(wdy_sim *s, double h, int nleap, double *acc_last,
                         double *cum_last) {
  // size the warm-start layout for this h before the splitters of the
  // next partition are written (a resize reallocates their buffers)
  if (s->warm) {
    if (int rc = warm_fit(s, h)) return rc;
  }
  int rc = launch_drift(s, h / 2.);
  if (rc) return rc;
  u32 T = 0, nb = 0;
  u64 *spl = s->warm ? warm_cold_splitters(s, &T, &nb) : nullptr;
  double t = s->t0;  // force time of each step (wendy/wendy.c:401-408)
  for (int ii = 0; ii < nleap; ++ii) {
    const bool last = ii == nleap - 1;
    rc = launch_leap(s, h, last ? h / 2. : h, last ? acc_last : nullptr,
                     last ? cum_last : nullptr, last ? spl : nullptr, T, nb, t);
    if (rc) return rc;
    t = t + h;
  }
  if (s->ext.kind) s->t0 = t;
  s->cold_steps += nleap;
  if (s->warm) {
    rc = warm_partition_after_cold(s, h / 2.);
    if (rc) return rc;
  }
  return WDY_OK;
}

// Completes the warm interval in flight, if any: waits for its error word
// and, when the device reported a failure (a bucket over capacity, too many
// far movers, a bounded wait exceeded), redoes the interval from its
// untouched start on the radix path.  Every entry point that reads the
// state or extends it calls this first.
// Copies the radix sort's error word to the host behind the work enqueued
// so far (checked by the next settle).
static int radix_check_enqueue(wdy_sim *s) {
  WDY_HIP(hipMemcpyAsync(s->h_rerr, radix_err_word(&s->radix), 4,
                         hipMemcpyDeviceToHost, s->stream));
  WDY_HIP(hipEventRecord(s->ev_rerr, s->stream));
  s->rerr_pending = true;
  return WDY_OK;
}

static int settle(wdy_sim *s) {
  if (s->broken)
    return fail(WDY_ERUNTIME, "a radix sort of this simulation gave up a "
                "bounded wait; its state is lost");
  WDY_HIP(hipSetDevice(s->device));
  if (s->warm) {
    bool failed = false;
    double h = 0.0;
    int nleap = 0;
    if (int rc = warm_settle(s, &failed, &h, &nleap)) return rc;
    if (failed) {
      s->redos += 1;
      if (int rc = cold_interval(s, h, nleap, nullptr, nullptr)) return rc;
      WDY_HIP(hipEventRecord(s->ev1, s->stream));
      if (int rc = radix_check_enqueue(s)) return rc;
    }
  }
  if (s->rerr_pending) {
    WDY_HIP(hipEventSynchronize(s->ev_rerr));
    s->rerr_pending = false;
    if (*s->h_rerr) {
      s->broken = true;
      return fail(WDY_ERUNTIME, "a look-back wait of a radix sort gave up; "
                  "the simulation's state is lost");
    }
  }
  return WDY_OK;
}

static int advance_impl(wdy_sim *s, double h, int nleap, double *acc_last,
                        double *cum_last) {
  WDY_HIP(hipSetDevice(s->device));
  if (int rc = settle(s)) return rc;
  warm_gang_release(s);  // ev0/ev1 bracket this interval
  WDY_HIP(hipEventRecord(s->ev0, s->stream));
  if (s->host_force) {
    if (int rc = advance_host_force(s, h, nleap, acc_last, cum_last)) return rc;
    if (int rc = radix_check_enqueue(s)) return rc;
  } else if (s->small) {
    int rc = launch_small_advance(s, h, nleap, acc_last, cum_last);
    if (rc) return rc;
    if (s->ext.kind)  // the kernel's clock, replayed (wendy/wendy.c:401-408)
      for (int ii = 0; ii < nleap; ++ii) s->t0 = s->t0 + h;
  } else if (!acc_last && !cum_last && warm_ready_for(s, h)) {
    // warm interval when the state is partitioned for this h and nothing
    // asks for the last force: enqueued and checked by the next settle,
    // which redoes it on the radix path after a device-side failure
    if (int rc = warm_interval(s, h, nleap)) return rc;
  } else {
    if (int rc = cold_interval(s, h, nleap, acc_last, cum_last)) return rc;
    if (int rc = radix_check_enqueue(s)) return rc;
  }
  WDY_HIP(hipEventRecord(s->ev1, s->stream));
  return WDY_OK;
}

// Device seconds of the last interval: its own events, or its gang
// interval's when it ran in a gang.
static float interval_ms(wdy_sim *s) {
  hipEvent_t a = s->ev0, b = s->ev1;
  (void)warm_gang_events(s, &a, &b);
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) ms = 0.f;
  return ms;
}

static int check_sim(wdy_sim *s, const char *fn) {
  if (!s) return fail(WDY_EINVAL, "%s: sim is NULL", fn);
  return WDY_OK;
}

// check_sim, then settle the warm interval in flight
static int check_settled(wdy_sim *s, const char *fn) {
  if (int rc = check_sim(s, fn)) return rc;
  return settle(s);
}

}  // namespace wdy

using namespace wdy;

extern "C" {

int wdy_abi_version(void) { return WDY_ABI_VERSION; }

const char *wdy_last_error(void) { return g_err; }

int wdy_device_count(int *count) {
  if (!count) return fail(WDY_EINVAL, "wdy_device_count: count is NULL");
  *count = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
    return fail(WDY_ENODEV, "no GPU visible");
  *count = n;
  return WDY_OK;
}

int wdy_create(int64_t n, const double *x, const double *v, const double *m,
               double totmass, double omega2, int flags, wdy_sim **out) {
  return create_impl(n, x, nullptr, v, m, totmass, omega2, flags, out);
}

int wdy_advance(wdy_sim *sim, double h, int nleap) {
  if (int rc = check_sim(sim, "wdy_advance")) return rc;
  if (nleap < 1) return fail(WDY_EINVAL, "wdy_advance: nleap=%d < 1", nleap);
  if (!std::isfinite(h)) return fail(WDY_EINVAL, "wdy_advance: h not finite");
  return advance_impl(sim, h, nleap, nullptr, nullptr);
}

int wdy_read(wdy_sim *sim, double *x_out, double *v_out, double *seconds_out) {
  if (int rc = check_settled(sim, "wdy_read")) return rc;
  WDY_HIP(hipSetDevice(sim->device));
  const size_t N = (size_t)sim->n;
  if (x_out || v_out) {
    if (int rc = launch_desort(sim, sim->out_x, sim->out_v, nullptr, nullptr,
                               nullptr))
      return rc;
    if (x_out)
      WDY_HIP(hipMemcpyAsync(x_out, sim->out_x, N * 8, hipMemcpyDeviceToHost,
                             sim->stream));
    if (v_out)
      WDY_HIP(hipMemcpyAsync(v_out, sim->small ? sim->vel[0] : sim->out_v,
                             N * 8, hipMemcpyDeviceToHost, sim->stream));
  }
  WDY_HIP(hipStreamSynchronize(sim->stream));
  if (seconds_out) *seconds_out = interval_ms(sim) * 1e-3;
  return WDY_OK;
}

int wdy_read_and_advance(wdy_sim *sim, double *x_out, double *v_out, double h,
                         int nleap, double *seconds_out) {
  if (int rc = check_settled(sim, "wdy_read_and_advance")) return rc;
  if (nleap < 1) return fail(WDY_EINVAL, "wdy_read_and_advance: nleap=%d < 1", nleap);
  if (!std::isfinite(h)) return fail(WDY_EINVAL, "wdy_read_and_advance: h not finite");
  if (sim->small) {  // velocities live in the state buffer: read, then advance
    if (int rc = wdy_read(sim, x_out, v_out, seconds_out)) return rc;
    return wdy_advance(sim, h, nleap);
  }
  WDY_HIP(hipSetDevice(sim->device));
  const size_t N = (size_t)sim->n;
  WDY_HIP(hipEventSynchronize(sim->ev1));  // the interval being read
  if (seconds_out) *seconds_out = interval_ms(sim) * 1e-3;
  if (x_out || v_out) {
    if (int rc = launch_desort(sim, sim->out_x, sim->out_v, nullptr, nullptr, nullptr))
      return rc;
    WDY_HIP(hipEventRecord(sim->ev_out, sim->stream));
    WDY_HIP(hipStreamWaitEvent(sim->copy_stream, sim->ev_out, 0));
    if (x_out)
      WDY_HIP(hipMemcpyAsync(x_out, sim->out_x, N * 8, hipMemcpyDeviceToHost,
                             sim->copy_stream));
    if (v_out)
      WDY_HIP(hipMemcpyAsync(v_out, sim->out_v, N * 8, hipMemcpyDeviceToHost,
                             sim->copy_stream));
    WDY_HIP(hipEventRecord(sim->ev_copied, sim->copy_stream));
  }
  // the next interval never touches the output staging buffers
  const int rc = advance_impl(sim, h, nleap, nullptr, nullptr);
  if (x_out || v_out) WDY_HIP(hipEventSynchronize(sim->ev_copied));
  return rc;
}

// Plain streaming copy: each thread moves 16 B per iteration of a grid-
// stride loop (dwordx4 loads and stores).
__global__ __launch_bounds__(256) void k_copy_probe(
    const ulonglong2 *__restrict__ src, ulonglong2 *__restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

int wdy_copy_probe(int64_t bytes, int launches, double *ms_out) {
  if (bytes < 16 || bytes % 16 != 0 || launches < 1)
    return fail(WDY_EINVAL, "wdy_copy_probe: bytes must be a positive multiple "
                "of 16 and launches >= 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(WDY_ENODEV, "wdy_copy_probe: no GPU visible");
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = [&]() -> int {
    WDY_HIP(hipMalloc(&a, (size_t)bytes));
    WDY_HIP(hipMalloc(&b, (size_t)bytes));
    WDY_HIP(hipMemset(a, 1, (size_t)bytes));
    WDY_HIP(hipEventCreate(&e0));
    WDY_HIP(hipEventCreate(&e1));
    int dev = 0, cus = 0;
    WDY_HIP(hipGetDevice(&dev));
    WDY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t n = (size_t)bytes / 16;
    const unsigned grid = (unsigned)std::max(1, cus) * 8u;
    WDY_HIP(hipDeviceSynchronize());
    WDY_HIP(hipEventRecord(e0, nullptr));
    for (int k = 0; k < launches; ++k) {
      k_copy_probe<<<grid, 256>>>((const ulonglong2 *)a, (ulonglong2 *)b, n);
      WDY_CHECK_LAUNCH();
    }
    WDY_HIP(hipEventRecord(e1, nullptr));
    WDY_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    WDY_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms_out) *ms_out = ms / launches;
    return WDY_OK;
  }();
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

int wdy_host_alloc(int64_t bytes, void **out) {
  if (!out || bytes < 0) return fail(WDY_EINVAL, "wdy_host_alloc: bad arguments");
  *out = nullptr;
  WDY_HIP(hipHostMalloc(out, (size_t)(bytes > 0 ? bytes : 1), hipHostMallocDefault));
  return WDY_OK;
}

int wdy_host_free(void *p) {
  if (p) WDY_HIP(hipHostFree(p));
  return WDY_OK;
}

int wdy_energy(wdy_sim *sim, double twopiG, double omega2, double *e_out) {
  if (int rc = check_settled(sim, "wdy_energy")) return rc;
  if (!e_out) return fail(WDY_EINVAL, "wdy_energy: e_out is NULL");
  if (!(twopiG != 0.0) || !std::isfinite(twopiG))
    return fail(WDY_EINVAL, "wdy_energy: twopiG must be finite and nonzero");
  WDY_HIP(hipSetDevice(sim->device));
  return launch_energy(sim, twopiG, omega2, e_out);
}

int wdy_destroy(wdy_sim *sim) {
  if (!sim) return WDY_OK;
  (void)hipSetDevice(sim->device);
  (void)hipStreamSynchronize(sim->stream);
  warm_wait(sim);  // an interval may run on a gang's stream
  (void)hipEventSynchronize(sim->ev1);
  free_sim(sim);
  return WDY_OK;
}

int wdy_gang_create(wdy_gang **out) {
  if (!out) return fail(WDY_EINVAL, "wdy_gang_create: out is NULL");
  *out = nullptr;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess)
    return fail(WDY_ENODEV, "wdy_gang_create: no GPU visible");
  WarmGang *g = nullptr;
  if (int rc = warm_gang_create(dev, &g)) return rc;
  *out = (wdy_gang *)g;
  return WDY_OK;
}

int wdy_gang_destroy(wdy_gang *gang) {
  warm_gang_destroy((WarmGang *)gang);
  return WDY_OK;
}

int wdy_gang_advance(wdy_gang *gang, wdy_sim *const *sims, int count,
                     const double *h, const int *nleap) {
  if (!gang || !sims || !h || !nleap)
    return fail(WDY_EINVAL, "wdy_gang_advance: NULL argument");
  if (count < 1 || count > WARM_GANG_MAX)
    return fail(WDY_EINVAL, "wdy_gang_advance: count=%d outside [1, %d]",
                count, WARM_GANG_MAX);
  WarmGang *g = (WarmGang *)gang;
  const int dev = warm_gang_device(g);
  for (int i = 0; i < count; ++i) {
    if (!sims[i]) return fail(WDY_EINVAL, "wdy_gang_advance: sims[%d] is NULL", i);
    if (nleap[i] < 1)
      return fail(WDY_EINVAL, "wdy_gang_advance: nleap[%d]=%d < 1", i, nleap[i]);
    if (!std::isfinite(h[i]))
      return fail(WDY_EINVAL, "wdy_gang_advance: h[%d] not finite", i);
    if (sims[i]->device != dev)
      return fail(WDY_EINVAL, "wdy_gang_advance: sims[%d] is on device %d, "
                  "the gang on %d", i, sims[i]->device, dev);
    for (int q = 0; q < i; ++q)
      if (sims[q] == sims[i])
        return fail(WDY_EINVAL, "wdy_gang_advance: sims[%d] repeats sims[%d]",
                    i, q);
  }
  // settle every member (a failed interval is redone now), then run the
  // members whose next interval is not warm on their own streams
  std::vector<wdy_sim *> gs;
  std::vector<double> gh;
  std::vector<int> gn;
  for (int i = 0; i < count; ++i) {
    wdy_sim *s = sims[i];
    if (int rc = settle(s)) return rc;
    if (warm_gang_eligible(s, h[i])) {
      gs.push_back(s);
      gh.push_back(h[i]);
      gn.push_back(nleap[i]);
    } else if (int rc = advance_impl(s, h[i], nleap[i], nullptr, nullptr)) {
      return rc;
    }
  }
  if (gs.empty()) return WDY_OK;
  WDY_HIP(hipSetDevice(dev));
  return warm_gang_interval(g, gs.data(), (int)gs.size(), gh.data(), gn.data());
}

int wdy_tune(wdy_sim *sim, int shape, int bucket_target, int window) {
  if (int rc = check_settled(sim, "wdy_tune")) return rc;
  if (sim->small || !sim->warm)
    return fail(WDY_EUNSUPPORTED, "wdy_tune: the sim has no warm-start path");
  if (shape < -1 || shape >= WARM_NSHAPES)
    return fail(WDY_EINVAL, "wdy_tune: shape %d out of range", shape);
  const WarmShape &sh = warm_shape(shape >= 0 ? shape : warm_current_shape(sim));
  if ((bucket_target != 0 &&
       (bucket_target < (int)WDY_BUCKET_MIN || bucket_target > sh.cap)) ||
      window < 0 || window > sh.wmax())
    return fail(WDY_EINVAL, "wdy_tune: bucket_target %d / window %d out of range",
                bucket_target, window);
  WDY_HIP(hipSetDevice(sim->device));
  WDY_HIP(hipStreamSynchronize(sim->stream));
  warm_mark_tuned(sim);
  return warm_configure(sim, shape, (u32)bucket_target, (u32)window);
}

int wdy_set_ext_force(wdy_sim *sim, int kind, const double *params, double t0) {
  if (int rc = check_settled(sim, "wdy_set_ext_force")) return rc;
  if (kind < WDY_EXT_NONE || kind > WDY_EXT_PULSE)
    return fail(WDY_EINVAL, "wdy_set_ext_force: unknown kind %d", kind);
  if (kind != WDY_EXT_NONE && !params)
    return fail(WDY_EINVAL, "wdy_set_ext_force: params is NULL");
  if (!std::isfinite(t0)) return fail(WDY_EINVAL, "wdy_set_ext_force: t0 not finite");
  ExtForce f;
  f.kind = kind;
  if (kind != WDY_EXT_NONE) {
    for (int i = 0; i < 3; ++i)
      if (!std::isfinite(params[i]))
        return fail(WDY_EINVAL, "wdy_set_ext_force: params[%d] not finite", i);
    f.p0 = params[0];
    f.p1 = params[1];
    f.p2 = params[2];
    if ((kind == WDY_EXT_ISOTHERMAL && f.p1 == 0.0) ||
        (kind == WDY_EXT_PULSE && f.p2 == 0.0))
      return fail(WDY_EINVAL, "wdy_set_ext_force: zero scale length / width");
  }
  WDY_HIP(hipSetDevice(sim->device));
  WDY_HIP(hipStreamSynchronize(sim->stream));
  sim->ext = f;
  sim->host_force = nullptr;
  sim->t0 = t0;
  return WDY_OK;
}

int wdy_set_ext_force_host(wdy_sim *sim,
                           double (*force)(int, double *, double, double *),
                           double t0) {
  if (int rc = check_settled(sim, "wdy_set_ext_force_host")) return rc;
  if (!std::isfinite(t0))
    return fail(WDY_EINVAL, "wdy_set_ext_force_host: t0 not finite");
  if (force && sim->small)
    return fail(WDY_EUNSUPPORTED,
                "wdy_set_ext_force_host: the simulation runs on the small path; "
                "create it with WDY_FORCE_LARGE");
  WDY_HIP(hipSetDevice(sim->device));
  WDY_HIP(hipStreamSynchronize(sim->stream));
  sim->ext = ExtForce();
  sim->host_force = force;
  sim->t0 = t0;
  return WDY_OK;
}

int wdy_time(wdy_sim *sim, double *t_out) {
  if (int rc = check_settled(sim, "wdy_time")) return rc;
  if (!t_out) return fail(WDY_EINVAL, "wdy_time: t_out is NULL");
  *t_out = sim->t0;
  return WDY_OK;
}

int wdy_path_stats(wdy_sim *sim, int64_t *warm_steps, int64_t *cold_steps,
                   int64_t *redos) {
  if (int rc = check_settled(sim, "wdy_path_stats")) return rc;
  if (warm_steps) *warm_steps = sim->warm_steps;
  if (cold_steps) *cold_steps = sim->cold_steps;
  if (redos) *redos = sim->redos;
  return WDY_OK;
}

int wdy_warm_failures(wdy_sim *sim, int64_t *capacity, int64_t *overflow,
                      int64_t *stall) {
  if (int rc = check_settled(sim, "wdy_warm_failures")) return rc;
  if (capacity) *capacity = sim->fail_capacity;
  if (overflow) *overflow = sim->fail_overflow;
  if (stall) *stall = sim->fail_stall;
  return WDY_OK;
}

int wdy_warm_config(wdy_sim *sim, int *shape, int *bucket_target, int *window) {
  if (int rc = check_settled(sim, "wdy_warm_config")) return rc;
  int sh = -1;
  u32 T = 0, W = 0;
  if (!warm_config(sim, &sh, &T, &W)) sh = -1;
  if (shape) *shape = sh;
  if (bucket_target) *bucket_target = (int)T;
  if (window) *window = (int)W;
  return WDY_OK;
}

int wdy_set_timing(wdy_sim *sim, int on) {
  if (int rc = check_sim(sim, "wdy_set_timing")) return rc;
  sim->timing = on != 0;
  return WDY_OK;
}

int wdy_kernel_time(wdy_sim *sim, int kernel_class, double *ms_total,
                    int64_t *count) {
  if (int rc = check_sim(sim, "wdy_kernel_time")) return rc;
  if (kernel_class < 0 || kernel_class >= KC_COUNT || !ms_total || !count)
    return fail(WDY_EINVAL, "wdy_kernel_time: bad arguments");
  WDY_HIP(hipSetDevice(sim->device));
  return timer_collect(sim->timer, kernel_class, ms_total, count);
}

static int check_diag(int64_t n, const char *fn) {
  if (n < 0 || n >= (int64_t(1) << 30))
    return fail(WDY_EINVAL, "%s: n=%lld outside [0, 2^30)", fn, (long long)n);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(WDY_ENODEV, "%s: no GPU visible", fn);
  return WDY_OK;
}

int wdy_energy_arrays(int64_t n, const double *x, const double *v,
                      const double *m, double twopiG, double omega2,
                      double *e_out) {
  if (int rc = check_diag(n, "wdy_energy_arrays")) return rc;
  if (!e_out || (n > 0 && (!x || !v || !m)))
    return fail(WDY_EINVAL, "wdy_energy_arrays: NULL pointer");
  if (n == 0) {
    *e_out = 0.0;
    return WDY_OK;
  }
  return launch_energy_arrays(n, x, v, m, twopiG, omega2, e_out);
}

int wdy_energy_individual_arrays(int64_t n, const double *x, const double *v,
                                 const double *m, double twopiG, double omega2,
                                 double *out) {
  if (int rc = check_diag(n, "wdy_energy_individual_arrays")) return rc;
  if (n == 0) return WDY_OK;
  if (!x || !v || !m || !out)
    return fail(WDY_EINVAL, "wdy_energy_individual_arrays: NULL pointer");
  return launch_energy_individual_arrays(n, x, v, m, twopiG, omega2, out);
}

int wdy_momentum_arrays(int64_t n, const double *v, const double *m,
                        double *p_out) {
  if (int rc = check_diag(n, "wdy_momentum_arrays")) return rc;
  if (!p_out || (n > 0 && (!v || !m)))
    return fail(WDY_EINVAL, "wdy_momentum_arrays: NULL pointer");
  if (n == 0) {
    *p_out = 0.0;
    return WDY_OK;
  }
  return launch_momentum_arrays(n, v, m, p_out);
}

int wdy_potential_arrays(int64_t ny, const double *y, int64_t n,
                         const double *x, const double *m, double twopiG,
                         double omega2, double *out) {
  if (int rc = check_diag(n, "wdy_potential_arrays")) return rc;
  if (ny < 0 || ny >= (int64_t(1) << 31))
    return fail(WDY_EINVAL, "wdy_potential_arrays: bad ny");
  if (ny == 0) return WDY_OK;
  if (!y || !out || (n > 0 && (!x || !m)))
    return fail(WDY_EINVAL, "wdy_potential_arrays: NULL pointer");
  return launch_potential_arrays(ny, y, n, x, m, twopiG, omega2, out);
}

void _wendy_nbody_approx_onestep(int N, struct array_w_index *xi, double *x,
                                 double *v, double *m, double *a,
                                 double totmass, double dt, int nleap,
                                 double *t0, double omega2,
                                 double (*ext_force)(int, double *, double,
                                                     double *),
                                 int sort_type, int *err,
                                 double *time_elapsed, double *cumulmass) {
  // The reference's signature has no error path on this call (err is never
  // written there): a failure here sets *err, leaves a message for
  // wdy_last_error() and is also reported on stderr, so that a caller that
  // ignores err still sees it.
  // The positions and velocities handed back after a failure are NaN, so
  // that a run that goes on regardless does not continue from a state that
  // silently stopped advancing.
  auto set_err = [&](int code) {
    if (err) *err = code;
    if (!code) return;
    fprintf(stderr, "wendy_amd: _wendy_nbody_approx_onestep failed (%d): %s\n",
            code, wdy_last_error());
    const double nan = std::nan("");
    for (int i = 0; i < N; ++i) {
      if (x) x[i] = nan;
      if (v) v[i] = nan;
      if (xi) xi[i].val = nan;
    }
  };
  if (sort_type < 0 || sort_type > 4) {
    fail(WDY_EINVAL, "unknown sort_type %d", sort_type);
    return set_err(WDY_EINVAL);
  }
  if (N < 1 || !xi || !x || !v || !m || nleap < 1 || (ext_force && !t0)) {
    fail(WDY_EINVAL, "invalid arguments");
    return set_err(WDY_EINVAL);
  }
  const size_t n = (size_t)N;
  std::vector<double> xs(n);
  std::vector<i32> ids(n);
  std::vector<char> seen(n, 0);
  bool eq = true;
  for (size_t i = 0; i < n; ++i) {
    const int j = xi[i].idx;
    if (j < 0 || (size_t)j >= n || seen[j]) {
      fail(WDY_EINVAL, "xi[].idx is not a permutation of 0..N-1 (at %zu)", i);
      return set_err(WDY_EINVAL);
    }
    seen[j] = 1;
    ids[i] = j;
    xs[i] = xi[i].val;
    eq = eq && (m[i] == m[0]);
  }
  wdy_sim *s = nullptr;
  // the reference hands xi back in sorted order, which the radix path keeps;
  // an external force is the caller's host function, called before every
  // step as the reference calls it (radix path)
  int rc = create_impl(N, xs.data(), ids.data(), v, m, totmass, omega2,
                       (eq ? WDY_EQUAL_MASS : 0) | WDY_NO_WARM |
                           (ext_force ? WDY_FORCE_LARGE : 0),
                       &s);
  if (rc) return set_err(rc);
  if (ext_force) {
    s->host_force = ext_force;
    s->t0 = *t0;
  }
  double *acc = nullptr, *acc_o = nullptr, *cum_o = nullptr;
  rc = [&]() -> int {
    WDY_HIP(hipMalloc(&acc, n * 8));
    WDY_HIP(hipMalloc(&acc_o, n * 8));
    WDY_HIP(hipMalloc(&cum_o, n * 8));
    // small mode writes the last acceleration per sheet, large mode per slot
    int r = advance_impl(s, dt, nleap, s->small ? acc_o : acc, cum_o);
    if (r) return r;
    if ((r = settle(s))) return r;
    r = launch_desort(s, s->out_x, s->out_v, s->small ? nullptr : acc, acc_o,
                      nullptr);
    if (r) return r;
    const int c = s->small ? 0 : s->cur;
    std::vector<u64> keys(n);
    WDY_HIP(hipMemcpyAsync(keys.data(), s->key[c], n * 8,
                           hipMemcpyDeviceToHost, s->stream));
    WDY_HIP(hipMemcpyAsync(ids.data(), s->idx[c], n * 4,
                           hipMemcpyDeviceToHost, s->stream));
    WDY_HIP(hipMemcpyAsync(x, s->out_x, n * 8, hipMemcpyDeviceToHost, s->stream));
    WDY_HIP(hipMemcpyAsync(v, s->small ? s->vel[0] : s->out_v, n * 8,
                           hipMemcpyDeviceToHost, s->stream));
    if (a)
      WDY_HIP(hipMemcpyAsync(a, acc_o, n * 8, hipMemcpyDeviceToHost, s->stream));
    if (cumulmass)
      WDY_HIP(hipMemcpyAsync(cumulmass, cum_o, n * 8, hipMemcpyDeviceToHost,
                             s->stream));
    WDY_HIP(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < n; ++i) {
      xi[i].idx = ids[i];
      xi[i].val = decode_pos(keys[i]);
    }
    float ms_ = 0.f;
    if (hipEventElapsedTime(&ms_, s->ev0, s->ev1) != hipSuccess) ms_ = 0.f;
    if (time_elapsed) *time_elapsed = ms_ * 1e-3;
    if (ext_force) *t0 = s->t0;
    return WDY_OK;
  }();
  if (acc) (void)hipFree(acc);
  if (acc_o) (void)hipFree(acc_o);
  if (cum_o) (void)hipFree(cum_o);
  wdy_destroy(s);
  set_err(rc);
}

}  // extern "C"
