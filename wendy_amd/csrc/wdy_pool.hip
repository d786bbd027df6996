// Serving pool: simulations arrive and leave while others run.  A pool owns
// one device arena (the batched layout of wdy_batch.h, sized by the caller's
// HBM budget) and a table of slots; each wdy_pool_step advances any subset of
// the live slots, each by its own h and step count, in ONE launch of the
// batched kernel.  Admission (wdy_pool_add) places a simulation in a free
// arena range (best fit) and uploads it on a second stream, so uploads overlap
// the running step; the next step waits for them with an event.
//
// Ordering rules that keep the arena race-free:
//  * a step waits (stream wait) for every upload enqueued before it;
//  * a removal makes the upload stream wait for all compute work enqueued so
//    far, so a range freed by a removal is not overwritten while a step that
//    still reads it runs;
//  * descriptor staging is double-buffered in pinned memory, each buffer
//    reused only after the step that read it has completed.
#include <algorithm>
#include <cmath>
#include <map>
#include <new>
#include <vector>

#include "wdy_batch.h"

namespace wdy {
namespace {

__global__ void k_pool_init(u32 n, const double *__restrict__ x,
                            u64 *__restrict__ key, i32 *__restrict__ idx,
                            u32 *__restrict__ err) {
  const u32 i = blockIdx.x * 256 + threadIdx.x;
  if (i == 0) *err = 0;  // the slot's error word, for its new occupant
  if (i >= n) return;
  key[i] = encode_pos(x[i]);
  idx[i] = (i32)i;  // sheet index within its simulation
}

__global__ void k_pool_desort(BatchDesc d, BatchArrays A, double *__restrict__ xo,
                              double *__restrict__ vo) {
  const u32 i = blockIdx.x * 256 + threadIdx.x;
  if (i >= d.n) return;
  const u64 p = d.off + i;
  const i32 j = A.idx[d.cur][p];
  xo[j] = decode_pos(A.key[d.cur][p]);
  vo[j] = A.vel[d.cur][p];
}

struct Slot {
  bool used = false;
  BatchDesc d{};
};

}  // namespace
}  // namespace wdy

using namespace wdy;

struct wdy_pool {
  int device = 0;
  hipStream_t stream = nullptr, up = nullptr;
  u64 cap = 0;
  u32 max_n = 0;
  BatchArrays arr{};
  BatchDesc *d_desc[2] = {nullptr, nullptr};
  BatchDesc *h_desc[2] = {nullptr, nullptr};  // pinned staging
  hipEvent_t desc_done[2] = {nullptr, nullptr};
  bool desc_pending[2] = {false, false};
  int which = 0;
  hipEvent_t up_done = nullptr, compute_mark = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool timed = false;
  double *stage_x = nullptr;            // upload staging, max_n
  double *read_x = nullptr, *read_v = nullptr;  // read staging, max_n each
  u32 *err = nullptr;
  std::vector<Slot> slots;
  std::map<u64, u64> free_ranges;       // offset -> length
  u64 used_sheets = 0;
  int live = 0;
};

static void pool_free(wdy_pool *p) {
  if (!p) return;
  for (int s = 0; s < 2; ++s) {
    if (p->arr.key[s]) (void)hipFree(p->arr.key[s]);
    if (p->arr.vel[s]) (void)hipFree(p->arr.vel[s]);
    if (p->arr.idx[s]) (void)hipFree(p->arr.idx[s]);
    if (p->arr.mass[s]) (void)hipFree(p->arr.mass[s]);
    if (p->d_desc[s]) (void)hipFree(p->d_desc[s]);
    if (p->h_desc[s]) (void)hipHostFree(p->h_desc[s]);
    if (p->desc_done[s]) (void)hipEventDestroy(p->desc_done[s]);
  }
  if (p->stage_x) (void)hipFree(p->stage_x);
  if (p->read_x) (void)hipFree(p->read_x);
  if (p->read_v) (void)hipFree(p->read_v);
  if (p->err) (void)hipFree(p->err);
  for (hipEvent_t e : {p->up_done, p->compute_mark, p->ev0, p->ev1})
    if (e) (void)hipEventDestroy(e);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  if (p->up) (void)hipStreamDestroy(p->up);
  delete p;
}

// Best-fit range of n sheets; returns false when none is large enough.
static bool pool_alloc(wdy_pool *p, u64 n, u64 *off) {
  auto best = p->free_ranges.end();
  for (auto it = p->free_ranges.begin(); it != p->free_ranges.end(); ++it)
    if (it->second >= n && (best == p->free_ranges.end() || it->second < best->second))
      best = it;
  if (best == p->free_ranges.end()) return false;
  *off = best->first;
  const u64 rest = best->second - n, roff = best->first + n;
  p->free_ranges.erase(best);
  if (rest) p->free_ranges[roff] = rest;
  return true;
}

static void pool_release(wdy_pool *p, u64 off, u64 n) {
  auto it = p->free_ranges.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != p->free_ranges.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    p->free_ranges.erase(nx);
  }
  if (it != p->free_ranges.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      p->free_ranges.erase(it);
    }
  }
}

static int slot_check(wdy_pool *p, int slot, const char *who) {
  if (!p) return fail(WDY_EINVAL, "%s: pool is NULL", who);
  if (slot < 0 || slot >= (int)p->slots.size() || !p->slots[(size_t)slot].used)
    return fail(WDY_EINVAL, "%s: slot %d is not live", who, slot);
  return WDY_OK;
}

extern "C" {

int wdy_pool_create(int64_t capacity_sheets, int max_sims, int64_t max_n,
                    wdy_pool **out) {
  if (!out) return fail(WDY_EINVAL, "wdy_pool_create: out is NULL");
  *out = nullptr;
  if (capacity_sheets < 1 || max_sims < 1 || max_n < 1 || max_n > capacity_sheets ||
      max_n >= (int64_t(1) << 31))
    return fail(WDY_EINVAL, "wdy_pool_create: need 1 <= max_n <= capacity_sheets, "
                "max_n < 2^31, max_sims >= 1");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(WDY_ENODEV, "wdy_pool_create: no GPU visible");
  wdy_pool *p = new (std::nothrow) wdy_pool();
  if (!p) return fail(WDY_ENOMEM, "wdy_pool_create: host allocation failed");
  int rc = [&]() -> int {
    WDY_HIP(hipGetDevice(&p->device));
    WDY_HIP(hipStreamCreate(&p->stream));
    WDY_HIP(hipStreamCreate(&p->up));
    for (hipEvent_t *e : {&p->up_done, &p->compute_mark, &p->ev0, &p->ev1})
      WDY_HIP(hipEventCreate(e));
    p->cap = (u64)capacity_sheets;
    p->max_n = (u32)max_n;
    const size_t N = (size_t)capacity_sheets;
    for (int s = 0; s < 2; ++s) {
      WDY_HIP(hipMalloc(&p->arr.key[s], N * 8));
      WDY_HIP(hipMalloc(&p->arr.vel[s], N * 8));
      WDY_HIP(hipMalloc(&p->arr.idx[s], N * 4));
      WDY_HIP(hipMalloc(&p->arr.mass[s], N * 8));
      WDY_HIP(hipMalloc(&p->d_desc[s], (size_t)max_sims * sizeof(BatchDesc)));
      WDY_HIP(hipHostMalloc((void **)&p->h_desc[s], (size_t)max_sims * sizeof(BatchDesc),
                            hipHostMallocDefault));
      WDY_HIP(hipEventCreate(&p->desc_done[s]));
    }
    WDY_HIP(hipMalloc(&p->stage_x, (size_t)max_n * 8));
    WDY_HIP(hipMalloc(&p->read_x, (size_t)max_n * 8));
    WDY_HIP(hipMalloc(&p->read_v, (size_t)max_n * 8));
    // one error word per slot, cleared by each admission's upload
    WDY_HIP(hipMalloc(&p->err, (size_t)max_sims * 4));
    WDY_HIP(hipMemsetAsync(p->err, 0, (size_t)max_sims * 4, p->stream));
    WDY_HIP(hipStreamSynchronize(p->stream));
    p->slots.resize((size_t)max_sims);
    p->free_ranges[0] = p->cap;
    return WDY_OK;
  }();
  if (rc) {
    pool_free(p);
    return rc;
  }
  *out = p;
  return WDY_OK;
}

int wdy_pool_add(wdy_pool *p, int64_t n, const double *x, const double *v,
                 const double *m, double totmass, double omega2, int *slot_out) {
  if (!p || !slot_out || !x || !v || !m)
    return fail(WDY_EINVAL, "wdy_pool_add: NULL argument");
  *slot_out = -1;
  if (n < 1 || n > (int64_t)p->max_n)
    return fail(WDY_EINVAL, "wdy_pool_add: n = %lld outside 1 .. %u", (long long)n,
                p->max_n);
  if (!std::isfinite(totmass) || !std::isfinite(omega2))
    return fail(WDY_EINVAL, "wdy_pool_add: non-finite totmass/omega2");
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(v[i]) || !std::isfinite(m[i]) || m[i] < 0.0)
      return fail(WDY_EINVAL, "wdy_pool_add: non-finite or negative input at sheet %lld",
                  (long long)i);
  int slot = -1;
  for (size_t k = 0; k < p->slots.size(); ++k)
    if (!p->slots[k].used) {
      slot = (int)k;
      break;
    }
  if (slot < 0) return fail(WDY_ENOMEM, "wdy_pool_add: every slot is live");
  u64 off = 0;
  if (!pool_alloc(p, (u64)n, &off))
    return fail(WDY_ENOMEM, "wdy_pool_add: no free arena range of %lld sheets "
                "(budget)", (long long)n);
  WDY_HIP(hipSetDevice(p->device));
  Slot &s = p->slots[(size_t)slot];
  batch_desc_fill(s.d, m, (u32)n, totmass, omega2);
  s.d.off = off;
  s.d.eslot = (u32)slot;
  const size_t bytes = (size_t)n * 8;
  int rc = [&]() -> int {
    WDY_HIP(hipMemcpyAsync(p->stage_x, x, bytes, hipMemcpyHostToDevice, p->up));
    WDY_HIP(hipMemcpyAsync(p->arr.vel[0] + off, v, bytes, hipMemcpyHostToDevice, p->up));
    WDY_HIP(hipMemcpyAsync(p->arr.mass[0] + off, m, bytes, hipMemcpyHostToDevice, p->up));
    k_pool_init<<<(unsigned)((n + 255) / 256), 256, 0, p->up>>>(
        (u32)n, p->stage_x, p->arr.key[0] + off, p->arr.idx[0] + off,
        p->err + slot);
    WDY_CHECK_LAUNCH();
    WDY_HIP(hipEventRecord(p->up_done, p->up));
    return WDY_OK;
  }();
  if (rc) {
    pool_release(p, off, (u64)n);
    return rc;
  }
  s.used = true;
  p->used_sheets += (u64)n;
  ++p->live;
  *slot_out = slot;
  return WDY_OK;
}

int wdy_pool_step(wdy_pool *p, int count, const int *slots, const double *h,
                  const int *nleap) {
  if (!p) return fail(WDY_EINVAL, "wdy_pool_step: pool is NULL");
  if (count < 0 || count > (int)p->slots.size() || (count && (!slots || !h || !nleap)))
    return fail(WDY_EINVAL, "wdy_pool_step: bad count or NULL arrays");
  std::vector<char> seen(p->slots.size(), 0);
  for (int k = 0; k < count; ++k) {
    if (int rc = slot_check(p, slots[k], "wdy_pool_step")) return rc;
    if (nleap[k] < 0 || !std::isfinite(h[k]))
      return fail(WDY_EINVAL, "wdy_pool_step: bad h or nleap for slot %d", slots[k]);
    if (seen[(size_t)slots[k]]++)
      return fail(WDY_EINVAL, "wdy_pool_step: slot %d listed twice", slots[k]);
  }
  WDY_HIP(hipSetDevice(p->device));
  const int w = p->which;
  if (p->desc_pending[w]) {
    WDY_HIP(hipEventSynchronize(p->desc_done[w]));
    p->desc_pending[w] = false;
  }
  BatchDesc *hd = p->h_desc[w];
  int c = 0;
  for (int k = 0; k < count; ++k) {
    if (nleap[k] == 0) continue;
    Slot &s = p->slots[(size_t)slots[k]];
    hd[c] = s.d;
    hd[c].h = h[k];
    hd[c].nleap = nleap[k];
    s.d.cur ^= nleap[k] & 1;  // where the state will be after this step
    ++c;
  }
  std::stable_sort(hd, hd + c, [](const BatchDesc &a, const BatchDesc &b) {
    return a.n > b.n;
  });
  WDY_HIP(hipStreamWaitEvent(p->stream, p->up_done, 0));
  WDY_HIP(hipMemcpyAsync(p->d_desc[w], hd, (size_t)c * sizeof(BatchDesc),
                         hipMemcpyHostToDevice, p->stream));
  WDY_HIP(hipEventRecord(p->ev0, p->stream));
  if (int rc = batch_launch(p->stream, p->d_desc[w], c, p->arr, p->err)) return rc;
  WDY_HIP(hipEventRecord(p->ev1, p->stream));
  WDY_HIP(hipEventRecord(p->desc_done[w], p->stream));
  p->desc_pending[w] = true;
  p->timed = true;
  p->which ^= 1;
  return WDY_OK;
}

int wdy_pool_read(wdy_pool *p, int slot, double *x_out, double *v_out) {
  if (int rc = slot_check(p, slot, "wdy_pool_read")) return rc;
  WDY_HIP(hipSetDevice(p->device));
  const BatchDesc &d = p->slots[(size_t)slot].d;
  WDY_HIP(hipStreamWaitEvent(p->stream, p->up_done, 0));
  k_pool_desort<<<(d.n + 255) / 256, 256, 0, p->stream>>>(d, p->arr, p->read_x,
                                                          p->read_v);
  WDY_CHECK_LAUNCH();
  if (x_out)
    WDY_HIP(hipMemcpyAsync(x_out, p->read_x, (size_t)d.n * 8, hipMemcpyDeviceToHost,
                           p->stream));
  if (v_out)
    WDY_HIP(hipMemcpyAsync(v_out, p->read_v, (size_t)d.n * 8, hipMemcpyDeviceToHost,
                           p->stream));
  u32 err = 0;
  WDY_HIP(hipMemcpyAsync(&err, p->err + slot, 4, hipMemcpyDeviceToHost, p->stream));
  WDY_HIP(hipStreamSynchronize(p->stream));
  if (err)
    return fail(WDY_EVERIFY, "wdy_pool_read: slot %d did not sort within its "
                "bound", slot);
  return WDY_OK;
}

int wdy_pool_remove(wdy_pool *p, int slot) {
  if (int rc = slot_check(p, slot, "wdy_pool_remove")) return rc;
  WDY_HIP(hipSetDevice(p->device));
  // later uploads into this range wait for every step enqueued so far
  WDY_HIP(hipEventRecord(p->compute_mark, p->stream));
  WDY_HIP(hipStreamWaitEvent(p->up, p->compute_mark, 0));
  Slot &s = p->slots[(size_t)slot];
  pool_release(p, s.d.off, s.d.n);
  p->used_sheets -= s.d.n;
  --p->live;
  s.used = false;
  return WDY_OK;
}

int wdy_pool_sync(wdy_pool *p, double *seconds_out) {
  if (!p) return fail(WDY_EINVAL, "wdy_pool_sync: pool is NULL");
  WDY_HIP(hipSetDevice(p->device));
  WDY_HIP(hipStreamSynchronize(p->up));
  WDY_HIP(hipStreamSynchronize(p->stream));
  if (seconds_out) {
    float ms = 0.f;
    if (!p->timed || hipEventElapsedTime(&ms, p->ev0, p->ev1) != hipSuccess) ms = 0.f;
    *seconds_out = ms * 1e-3;
  }
  return WDY_OK;
}

int wdy_pool_stats(wdy_pool *p, int *live, int64_t *used_sheets,
                   int64_t *largest_free) {
  if (!p) return fail(WDY_EINVAL, "wdy_pool_stats: pool is NULL");
  u64 big = 0;
  for (const auto &r : p->free_ranges) big = std::max(big, r.second);
  if (live) *live = p->live;
  if (used_sheets) *used_sheets = (int64_t)p->used_sheets;
  if (largest_free) *largest_free = (int64_t)big;
  return WDY_OK;
}

int wdy_pool_destroy(wdy_pool *p) {
  if (!p) return WDY_OK;
  (void)hipSetDevice(p->device);
  (void)hipStreamSynchronize(p->up);
  (void)hipStreamSynchronize(p->stream);
  pool_free(p);
  return WDY_OK;
}

}  // extern "C"
