// Exact event-driven solver on the CPU: the reference's approx=False path
// (wendy/wendy.c:33-315 for the steps, wendy/wendy.py:158-279 for the set-up),
// kept as a machine-precision check of the GPU leapfrog.
//
// Between collisions every sheet moves under a constant force (parabola) or,
// with an external harmonic potential, on a harmonic orbit about its
// equilibrium offset; the only events are crossings of neighbours in
// position order.  Each crossing swaps the pair, changes their forces by
// the swapped mass, and re-solves the crossing times of the pair and of its
// two neighbour pairs.  Pending crossing times live in an indexed binary
// min-heap keyed by (time, pair index).
//
// Arithmetic follows the reference expression by expression (the set-up's
// sequential mass sums, the flight updates, the crossing-time formulas,
// the force update on a swap), so trajectories agree with the reference to
// the last bit whenever events happen in the same order.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <ctime>
#include <cstdint>
#include <new>
#include <numeric>
#include <vector>

#include "wendy_amd.h"

namespace wdy {
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
}

namespace {

// Root of c0 + c1 t + c2 t^2 = 0 reached by an approaching pair
// (wendy/wendy.c:33-38).
double cross_quad(double c0, double c1, double c2) {
  const double r = -c1 / c2;
  return 0.5 * (r + std::sqrt(r * r - 4. * c0 / c2));
}

// First t > 0 with A cos(w t) + B sin(w t) = c2, A = c0 - c2, B = c1 / w
// (wendy/wendy.c:39-52).
double cross_harm(double c0, double c1, double c2, double w) {
  const double A = c0 - c2, B = c1 / w;
  const double amp = std::sqrt(B * B + A * A);
  double out = -std::asin(c2 / amp) - std::atan2(A, B);
  if (out < 0) out = M_PI + std::asin(c2 / amp) - std::atan2(A, B);
  return out / w;
}

class PairHeap {
 public:
  void build(const std::vector<double> *key) {
    key_ = key;
    const int n = (int)key->size();
    h_.resize(n);
    at_.resize(n);
    std::iota(h_.begin(), h_.end(), 0);
    std::iota(at_.begin(), at_.end(), 0);
    for (int i = n / 2 - 1; i >= 0; --i) down(i);
  }
  bool empty() const { return h_.empty(); }
  int top() const { return h_[0]; }
  void changed(int k) {  // key of pair k was modified
    const int i = at_[k];
    if (i > 0 && before(h_[i], h_[(i - 1) / 2]))
      up(i);
    else
      down(i);
  }

 private:
  bool before(int a, int b) const {
    const double ka = (*key_)[a], kb = (*key_)[b];
    return ka < kb || (ka == kb && a < b);
  }
  void place(int i, int k) {
    h_[i] = k;
    at_[k] = i;
  }
  void up(int i) {
    const int k = h_[i];
    while (i > 0) {
      const int p = (i - 1) / 2;
      if (!before(k, h_[p])) break;
      place(i, h_[p]);
      i = p;
    }
    place(i, k);
  }
  void down(int i) {
    const int n = (int)h_.size(), k = h_[i];
    while (true) {
      int c = 2 * i + 1;
      if (c >= n) break;
      if (c + 1 < n && before(h_[c + 1], h_[c])) ++c;
      if (!before(h_[c], k)) break;
      place(i, h_[c]);
      i = c;
    }
    place(i, k);
  }
  const std::vector<double> *key_ = nullptr;
  std::vector<int> h_, at_;
};

}  // namespace

struct wdy_exact {
  int64_t n = 0;
  double omega = 0.0;            // > 0: external harmonic potential
  std::vector<double> x, v, a, m, t;  // per sheet
  std::vector<int> ord;          // sheet at each position rank
  std::vector<double> tc;        // crossing time of ranks (k, k+1)
  PairHeap heap;

  // Flight of sheet i from its own time t[i] to time tn.
  void fly(int i, double tn) {
    const double dt = tn - t[i];
    if (omega > 0.0) {  // wendy/wendy.c:204-215
      const double s = std::sin(omega * dt);
      const double c = std::sqrt(1. - s * s);
      const double d = x[i] - a[i];
      x[i] = d * c + v[i] / omega * s + a[i];
      v[i] = -d * omega * s + v[i] * c;
    } else {  // wendy/wendy.c:80-84
      const double dv = a[i] * dt;
      x[i] += dv * dt / 2. + v[i] * dt;
      v[i] += dv;
    }
    t[i] = tn;
  }

  // Crossing time of ranks (k, k+1) from scratch, both sheets at time 0
  // (wendy/wendy.py:257-263).
  double pair_time0(int k) const {
    const int l = ord[k], u = ord[k + 1];
    if (omega > 0.0) return cross_harm(x[l] - x[u], v[l] - v[u], a[l] - a[u], omega);
    return cross_quad(x[l] - x[u], v[l] - v[u], (a[l] - a[u]) / 2.);
  }

  // Order, forces and crossing times from the current x, v
  // (wendy/wendy.py:243-264).  Ties in x keep the lower sheet index first.
  void setup() {
    ord.resize(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int p, int q) { return x[p] < x[q]; });
    // mass below: running sum from the bottom; mass above: from the top
    std::vector<double> below(n), above(n);
    double run = 0.0;
    for (int64_t k = 0; k < n; ++k) {
      below[k] = run;
      run += m[ord[k]];
    }
    run = 0.0;
    for (int64_t k = n - 1; k >= 0; --k) {
      above[k] = run;
      run += m[ord[k]];
    }
    for (int64_t k = 0; k < n; ++k) a[ord[k]] = above[k] - below[k];
    tc.assign(n > 1 ? n - 1 : 0, 0.0);
    for (int64_t k = 0; k + 1 < n; ++k) tc[k] = pair_time0((int)k);
    heap.build(&tc);
    std::fill(t.begin(), t.end(), 0.0);
  }

  // One output interval of length dt; returns the number of crossings.
  // first >= 0 names the pair that crosses first (a caller that keeps its
  // own record of the earliest crossing, as the reference's interface does);
  // later crossings come from the heap.
  int64_t step(double dt, int64_t maxcoll, bool *hit_max, int first = -1) {
    std::fill(t.begin(), t.end(), 0.0);
    int64_t cnt = 0;
    while (!heap.empty() && cnt < maxcoll) {
      const int k = (cnt == 0 && first >= 0) ? first : heap.top();
      if (!(tc[k] < dt)) break;
      ++cnt;
      const double tn = tc[k];
      const int i = ord[k], j = ord[k + 1];  // i below j until now
      fly(i, tn);
      fly(j, tn);
      std::swap(ord[k], ord[k + 1]);
      // each loses the other's mass on one side and gains it on the other
      // (wendy/wendy.c:98-103)
      const double dm = m[j] - m[i];
      const double ai = a[j] - dm, aj = a[i] - dm;
      a[i] = ai;
      a[j] = aj;
      // the pair separates and meets again (wendy/wendy.c:107-113, 232-236)
      double dtn;
      if (omega > 0.0) {
        double q = (v[i] - v[j]) / (a[j] - a[i]) / omega;
        q *= q;
        q = (1. - q) / (1. + q);
        dtn = std::acos(q) / omega;
      } else {
        dtn = 2. * (v[i] - v[j]) / (a[j] - a[i]);
      }
      tc[k] = tn + dtn;
      heap.changed(k);
      if (k > 0) update_neighbour(k - 1, ord[k - 1], j, tn);
      if (k + 2 < n) update_neighbour(k + 1, ord[k + 2], i, tn);
    }
    *hit_max = cnt == maxcoll;
    for (int64_t i = 0; i < n; ++i) fly((int)i, dt);
    for (auto &c : tc) c -= dt;
    return cnt;
  }

  // Re-solve the crossing of pair p between sheet o (the outer neighbour,
  // last moved at t[o]) and sheet s (just moved to time tn).
  // (wendy/wendy.c:114-141 and, with the harmonic potential, :240-287.)
  void update_neighbour(int p, int o, int s, double tn) {
    if (omega > 0.0) {
      fly(o, tn);
      tc[p] = tn + cross_harm(x[o] - x[s], v[o] - v[s], a[o] - a[s], omega);
    } else {
      const double d = t[o] - tn;  // <= 0: o's state is older
      tc[p] = tn + cross_quad(x[o] + a[o] * d * d / 2. - v[o] * d - x[s],
                              v[o] - a[o] * d - v[s], 0.5 * (a[o] - a[s]));
    }
    heap.changed(p);
  }
};

using wdy::fail;

extern "C" {

int wdy_exact_create(int64_t n, const double *x, const double *v, const double *m,
                     double omega, wdy_exact **out) {
  if (!out) return fail(WDY_EINVAL, "wdy_exact_create: out is NULL");
  *out = nullptr;
  if (n < 1 || !x || !v || !m) return fail(WDY_EINVAL, "wdy_exact_create: bad arguments");
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(v[i]) || !std::isfinite(m[i]) ||
        !(m[i] > 0.0))
      return fail(WDY_EINVAL, "wdy_exact_create: non-finite input or mass <= 0 "
                  "at sheet %lld", (long long)i);
  if (!std::isfinite(omega)) return fail(WDY_EINVAL, "wdy_exact_create: omega");
  wdy_exact *e = new (std::nothrow) wdy_exact();
  if (!e) return fail(WDY_ENOMEM, "wdy_exact_create: allocation failed");
  e->n = n;
  e->omega = omega > 0.0 ? omega : 0.0;
  e->x.assign(x, x + n);
  e->v.assign(v, v + n);
  e->m.assign(m, m + n);
  e->a.assign(n, 0.0);
  e->t.assign(n, 0.0);
  e->setup();
  *out = e;
  return WDY_OK;
}

int wdy_exact_advance(wdy_exact *e, double dt, int64_t maxcoll, int64_t *ncoll,
                      int *hit_max, double *seconds) {
  if (!e) return fail(WDY_EINVAL, "wdy_exact_advance: handle is NULL");
  if (!(dt > 0.0) || !std::isfinite(dt) || maxcoll < 0)
    return fail(WDY_EINVAL, "wdy_exact_advance: bad dt or maxcoll");
  const auto t0 = std::chrono::steady_clock::now();
  bool hit = false;
  const int64_t c = e->step(dt, maxcoll, &hit);
  const auto t1 = std::chrono::steady_clock::now();
  if (ncoll) *ncoll = c;
  if (hit_max) *hit_max = hit ? 1 : 0;
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  return WDY_OK;
}

int wdy_exact_reset(wdy_exact *e) {
  if (!e) return fail(WDY_EINVAL, "wdy_exact_reset: handle is NULL");
  e->setup();
  return WDY_OK;
}

int wdy_exact_read(wdy_exact *e, double *x_out, double *v_out) {
  if (!e) return fail(WDY_EINVAL, "wdy_exact_read: handle is NULL");
  if (x_out) std::copy(e->x.begin(), e->x.end(), x_out);
  if (v_out) std::copy(e->v.begin(), e->v.end(), v_out);
  return WDY_OK;
}

int wdy_exact_destroy(wdy_exact *e) {
  delete e;
  return WDY_OK;
}

// ---- reference-signature entry points --------------------------------------
// The caller owns the whole solver state between calls, as in the reference:
// x, v, a per sheet, sindx (sheet at each position rank), tcoll (crossing
// time of ranks k, k+1, relative to the start of the call), cindx and
// next_tcoll (the earliest crossing).  One call runs one output interval dt
// and hands the state back shifted to the interval's end.

double _solve_coll_quad(double c0, double c1, double c2) {
  return cross_quad(c0, c1, c2);
}

double _solve_coll_harm(double c0, double c1, double c2, double omega) {
  return cross_harm(c0, c1, c2, omega);
}

}  // extern "C"

namespace {

void exact_compat_step(int N, double *x, double *v, double *a, double *m,
                       int *sindx, int *cindx, double *next_tcoll,
                       double *tcoll, double dt, int maxcoll, int *err,
                       int *ncoll, double *time_elapsed, double omega) {
  if (N < 1 || !x || !v || !a || !m || !sindx || !cindx || !next_tcoll ||
      (N > 1 && !tcoll)) {
    wdy::fail(WDY_EINVAL, "exact step: invalid arguments");
    if (err) *err = WDY_EINVAL;
    return;
  }
  wdy_exact e;
  e.n = N;
  e.omega = omega;
  e.x.assign(x, x + N);
  e.v.assign(v, v + N);
  e.a.assign(a, a + N);
  e.m.assign(m, m + N);
  e.t.assign(N, 0.0);
  e.ord.assign(sindx, sindx + N);
  e.tc.assign(tcoll, tcoll + (N - 1));
  e.heap.build(&e.tc);
  const int first = (N > 1 && *cindx >= 0 && *cindx < N - 1) ? *cindx : -1;
  bool hit = false;
  // the reference times the crossing loop with the process CPU clock
  const std::clock_t c0 = std::clock();
  const int64_t cnt = e.step(dt, maxcoll, &hit, first);
  const std::clock_t c1 = std::clock();
  std::copy(e.x.begin(), e.x.end(), x);
  std::copy(e.v.begin(), e.v.end(), v);
  std::copy(e.a.begin(), e.a.end(), a);
  std::copy(e.ord.begin(), e.ord.end(), sindx);
  std::copy(e.tc.begin(), e.tc.end(), tcoll);
  // the earliest pending crossing, shifted like tcoll; unchanged (but
  // shifted) when nothing crossed
  if (cnt > 0 && !e.heap.empty()) {
    *cindx = e.heap.top();
    *next_tcoll = e.tc[*cindx];
  } else {
    *next_tcoll -= dt;
  }
  if (ncoll) *ncoll = (int)cnt;
  if (time_elapsed) *time_elapsed = (double)(c1 - c0) / CLOCKS_PER_SEC;
  if (hit && err) *err = -2;
}

}  // namespace

extern "C" {

void _wendy_nbody_onestep(int N, double *x, double *v, double *a, double *m,
                          int *sindx, int *cindx, double *next_tcoll,
                          double *tcoll, double dt, int maxcoll, int *err,
                          int *ncoll, double *time_elapsed) {
  exact_compat_step(N, x, v, a, m, sindx, cindx, next_tcoll, tcoll, dt,
                    maxcoll, err, ncoll, time_elapsed, 0.0);
}

void _wendy_nbody_harm_onestep(int N, double *x, double *v, double *a,
                               double *m, int *sindx, int *cindx,
                               double *next_tcoll, double *tcoll, double dt,
                               int maxcoll, int *err, int *ncoll,
                               double *time_elapsed, double omega) {
  if (!(omega > 0.0)) {
    wdy::fail(WDY_EINVAL, "_wendy_nbody_harm_onestep: omega must be > 0");
    if (err) *err = WDY_EINVAL;
    return;
  }
  exact_compat_step(N, x, v, a, m, sindx, cindx, next_tcoll, tcoll, dt,
                    maxcoll, err, ncoll, time_elapsed, omega);
}

}  // extern "C"
