// Batched ensemble: thousands of independent simulations of uneven N advance
// together, ONE launch per output interval.
//
// Each simulation is handled by one 512-thread workgroup that runs all of
// its leapfrog steps of the interval without returning to the host (a
// persistent workgroup per simulation; descriptors are ordered largest
// first so the longest workgroups start first).  State lives in one arena:
// simulation k occupies [off_k, off_k + n_k) of two buffer sets of SoA arrays
// (key = order-preserving position encoding, velocity, global sheet index,
// mass * twopiG); `cur` in the descriptor says which set is current.
//
// * n <= BLOCK (2048): the whole simulation stays in LDS for the interval;
//   each step sorts it with the LDS bucket sort (a full sort: any input
//   order), then computes force, kick and drift in sorted order.
// * n > BLOCK: each step is one pass of a window sliding along the previous
//   order (slide_step): sorted in LDS, its lower part final and written out
//   with force, kick and drift, its upper half kept for the next window.
//   When a sheet has fallen more than BLOCK/4 ranks behind its previous place
//   (the first step from an arbitrary input order, for one) the pass detects
//   it and the step is redone in full: the arena range is sorted in place
//   with a block odd-even merge sort -- sort aligned blocks of BLOCK sheets
//   in LDS, then blocks offset by BLOCK/2 -- repeated until every block
//   boundary is in order (each block sort only permutes its own block, so
//   ordered boundaries mean the range is sorted; any input order finishes
//   within ceil(2n/BLOCK) + 1 rounds), then a last pass computes force, kick
//   and drift block by block, carrying the rank and the fixed-point mass
//   prefix, and writes the other buffer set.
//
// Arithmetic is the reference's (wendy/wendy.c:318-333, 371-383); ties break
// by sheet index; the unequal-mass prefix is exact fixed point (as in the
// warm-start step), equal masses use rank * m.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "wdy_batch.h"
#include "wdy_lds.h"

namespace wdy {

namespace {

// 512 threads, windows of 2048 sheets: 70 KB of LDS, two workgroups per
// CU, so that one's barriers overlap the other's work (one 1024-thread
// workgroup with 4096-sheet windows per CU measured 23% slower)
using BSH = Shape<512, 2048>;
constexpr int BWT = BSH::WT, BLOCK = BSH::CAP, BPER = BSH::PER;
constexpr int BNWAVE = BSH::NWAVE;

// Sorts A[lo, lo + c) (c <= BLOCK) in place by (key, index), via LDS.
__device__ void block_sort_inplace(Lds<BSH, false> &s, u64 *key, double *vel,
                                   i32 *idx, double *mass, u64 lo, u32 c,
                                   u64 *s_red, u64 *s_mm) {
  const u32 t = threadIdx.x;
#pragma unroll
  for (int i = 0; i < BPER; ++i) {
    const u32 e = i * BWT + t;
    if (e < c) {
      s.key[e] = key[lo + e];
      s.v[e] = vel[lo + e];
      s.id[e] = idx[lo + e];
      s.m[e] = mass ? mass[lo + e] : 0.0;
    }
  }
  __syncthreads();
  lds_bucket_sort<BSH, false>(s, c, s_red, s_mm, nullptr);
#pragma unroll
  for (int i = 0; i < BPER; ++i) {
    const u32 r = i * BWT + t;
    if (r < c) {
      const u32 e = s.ord[r];
      key[lo + r] = s.key[e];
      vel[lo + r] = s.v[e];
      idx[lo + r] = s.id[e];
      if (mass) mass[lo + r] = s.m[e];
    }
  }
  __syncthreads();
}

// Force, kick and drift of the sorted sheets s.ord[0..c) whose ranks start
// at rank0; `run` is the fixed-point mass below the first of them.  New
// keys and velocities are written back into LDS; returns the updated run.
__device__ u64 block_leap(Lds<BSH, false> &s, const BatchDesc &d, u32 c,
                          u64 rank0, u64 run, double hd, u64 *s_red) {
  const u32 t = threadIdx.x;
  const bool eq = d.flags & 1u, harm = d.flags & 2u;
  const u32 r0 = t * BPER;  // ranks r0 .. r0 + BPER - 1 in order
  u64 local = 0, total = 0;
  if (!eq) {
    u64 part = 0;
#pragma unroll
    for (int i = 0; i < BPER; ++i)
      if (r0 + i < c) part += mass_fix(s.m[s.ord[r0 + i]], d.fscale);
    local = block_excl_u64<BNWAVE>(part, s_red, &total);
  }
  u64 pre = run + local;
#pragma unroll
  for (int i = 0; i < BPER; ++i) {
    const u32 r = r0 + i;
    if (r < c) {
      const u32 e = s.ord[r];
      double x = decode_pos(s.key[e]);
      double F;
      if (eq) {
        F = (d.totmass - 2.0 * ((double)(rank0 + r) * d.mq)) - d.mq;
      } else {
        const double me = s.m[e];
        F = (d.totmass - 2.0 * ((double)pre * d.finv)) - me;
        pre += mass_fix(me, d.fscale);
      }
      if (harm) F = F - d.omega2 * x;
      const double a = 0.0 + F;
      const double v = s.v[e] + d.h * a;
      x = x + hd * v;
      s.v[e] = v;
      s.key[e] = encode_pos(x);
    }
  }
  __syncthreads();
  return run + total;
}

// One leapfrog step of a simulation with n > BLOCK sheets, from the arrays
// key..mass (positions already drifted to the force time) into ko..mo, in
// ONE pass: a window of BLOCK sheets slides along the previous order by
// STRIDE = BLOCK - KEEP sheets; each window is sorted in LDS, all but its
// upper KEEP sheets are final (force, kick, drift, written out) and those
// stay, in their LDS slots, for the next window, whose new sheets fill the
// slots just written out.  The first window holds KEEP + ((n - BLOCK) mod
// STRIDE) sheets (BLOCK when that remainder is 0) so that every later one
// is full.  That holds while no sheet has fallen more than KEEP ranks behind
// its place in the previous order; a sheet arriving below one already
// written shows that it did not,
// and the step reports failure (the inputs are untouched, so the caller
// redoes the step in full).  Returns true on success.
__device__ bool slide_step(Lds<BSH, false> &s, const BatchDesc &d, u32 n,
                           const u64 *key, const double *vel, const i32 *idx,
                           const double *mass, u64 *ko, double *vo, i32 *io,
                           double *mo, double hd, u64 *s_red, u64 *s_mm,
                           u32 *s_ok, u64 *s_last) {
  constexpr u32 KEEP = BLOCK / 4, STRIDE = BLOCK - KEEP;
  const u32 t = threadIdx.x;
  if (t == 0) *s_ok = 1u;
  const u32 rem = (n - (u32)BLOCK) % STRIDE;
  u32 c = rem ? KEEP + rem : (u32)BLOCK;  // sheets in the current window
  u32 lo = 0, done = 0, ne = 0;
  bool first = true;
  u64 run = 0;
  while (done < n) {
    // this window's new sheets: all c at first; then STRIDE of them, into
    // the slots written out last (s.ord[0 .. ne)) and the ones the first
    // window left unused ([c_prev, BLOCK))
    const u32 cprev = c, cn = first ? c : STRIDE;
    if (!first) c = BLOCK;
#pragma unroll
    for (int i = 0; i < BPER; ++i) {
      const u32 jn = i * BWT + t;
      if (jn < cn) {
        const u32 e = first ? jn : (jn < ne ? (u32)s.ord[jn] : cprev + (jn - ne));
        const u32 g = lo + jn;
        s.key[e] = key[g];
        s.v[e] = vel[g];
        s.id[e] = idx[g];
        s.m[e] = mass ? mass[g] : 0.0;
      }
    }
    lo += cn;
    __syncthreads();
    lds_bucket_sort<BSH, false>(s, c, s_red, s_mm, nullptr);
    ne = lo == n ? c : c - KEEP;  // sheets final in this window
    if (t == 0) {
      const u32 e0 = s.ord[0];
      if (!first && pair_lt(s.key[e0], s.id[e0], s_last[0], (i32)s_last[1]))
        *s_ok = 0u;  // precedes a sheet already written
      const u32 el = s.ord[ne - 1];
      s_last[0] = s.key[el];
      s_last[1] = (u64)(u32)s.id[el];
    }
    __syncthreads();
    run = block_leap(s, d, ne, done, run, hd, s_red);
#pragma unroll
    for (int i = 0; i < BPER; ++i) {
      const u32 r = i * BWT + t;
      if (r < ne) {
        const u32 e = s.ord[r];
        ko[done + r] = s.key[e];
        vo[done + r] = s.v[e];
        io[done + r] = s.id[e];
        if (mo) mo[done + r] = s.m[e];
      }
    }
    __syncthreads();
    done += ne;
    first = false;
  }
  const bool ok = *s_ok != 0u;
  __syncthreads();
  return ok;
}

__device__ __forceinline__ bool ordered(const u64 *key, const i32 *idx,
                                        u64 a, u64 b) {
  const u64 ka = key[a], kb = key[b];
  return ka < kb || (ka == kb && idx[a] <= idx[b]);
}

__global__ __launch_bounds__(BWT) void k_batch_advance(
    const BatchDesc *__restrict__ descs, BatchArrays A, u32 *err) {
  extern __shared__ u64 lds_dyn[];
  Lds<BSH, false> s(lds_dyn);
  __shared__ u64 s_red[BNWAVE];
  __shared__ u64 s_mm[2 * BNWAVE];
  __shared__ u32 s_flag;
  __shared__ u64 s_last[2];
  const BatchDesc d = descs[blockIdx.x];
  const u32 t = threadIdx.x;
  const u32 n = d.n;
  if (d.nleap <= 0 || n == 0) return;
  const bool eq = d.flags & 1u;
  const double hh = d.h / 2.;
  int a = d.cur;

  if (n <= (u32)BLOCK) {
    // ---- the whole simulation in LDS for the interval ------------------------
    u64 *key = A.key[a] + d.off;
    double *vel = A.vel[a] + d.off;
    i32 *idx = A.idx[a] + d.off;
    double *mass = eq ? nullptr : A.mass[a] + d.off;
#pragma unroll
    for (int i = 0; i < BPER; ++i) {
      const u32 e = i * BWT + t;
      if (e < n) {
        const double v = vel[e];
        s.key[e] = encode_pos(decode_pos(key[e]) + hh * v);  // first half drift
        s.v[e] = v;
        s.id[e] = idx[e];
        s.m[e] = mass ? mass[e] : 0.0;
      }
    }
    __syncthreads();
    for (int step = 0; step < d.nleap; ++step) {
      lds_bucket_sort<BSH, false>(s, n, s_red, s_mm, nullptr);
      (void)block_leap(s, d, n, 0, 0, step == d.nleap - 1 ? hh : d.h, s_red);
    }
    const int b = a ^ (d.nleap & 1);
    u64 *ko = A.key[b] + d.off;
    double *vo = A.vel[b] + d.off;
    i32 *io = A.idx[b] + d.off;
    double *mo = eq ? nullptr : A.mass[b] + d.off;
#pragma unroll
    for (int i = 0; i < BPER; ++i) {
      const u32 r = i * BWT + t;
      if (r < n) {
        const u32 e = s.ord[r];
        ko[r] = s.key[e];
        vo[r] = s.v[e];
        io[r] = s.id[e];
        if (mo) mo[r] = s.m[e];
      }
    }
    return;
  }

  // ---- larger simulations: arena-resident, block odd-even merge sort ---------
  {  // first half drift, in place
    u64 *key = A.key[a] + d.off;
    const double *vel = A.vel[a] + d.off;
    for (u32 i = t; i < n; i += BWT) key[i] = encode_pos(decode_pos(key[i]) + hh * vel[i]);
    __syncthreads();
  }
  const u32 half = BLOCK / 2;
  const u32 max_rounds = (2 * n + BLOCK - 1) / BLOCK + 1;
  for (int step = 0; step < d.nleap; ++step) {
    u64 *key = A.key[a] + d.off;
    double *vel = A.vel[a] + d.off;
    i32 *idx = A.idx[a] + d.off;
    double *mass = eq ? nullptr : A.mass[a] + d.off;
    {  // one sliding pass; else the full sort below
      const int b = a ^ 1;
      if (slide_step(s, d, n, key, vel, idx, mass, A.key[b] + d.off,
                     A.vel[b] + d.off, A.idx[b] + d.off,
                     eq ? nullptr : A.mass[b] + d.off,
                     step == d.nleap - 1 ? hh : d.h, s_red, s_mm, &s_flag,
                     s_last)) {
        a = b;
        continue;
      }
    }
    bool sorted = false;
    for (u32 round = 0; round < max_rounds && !sorted; ++round) {
      for (u32 lo = 0; lo < n; lo += BLOCK)
        block_sort_inplace(s, key, vel, idx, mass, lo, min((u32)BLOCK, n - lo),
                           s_red, s_mm);
      for (u32 lo = half; lo < n; lo += BLOCK)
        block_sort_inplace(s, key, vel, idx, mass, lo, min((u32)BLOCK, n - lo),
                           s_red, s_mm);
      // every boundary of the offset blocks in order -> the range is sorted
      if (t == 0) s_flag = 1;
      __syncthreads();
      for (u32 b = half + (u32)t * BLOCK; b < n; b += (u32)BWT * BLOCK)
        if (!ordered(key, idx, b - 1, b)) s_flag = 0;
      __syncthreads();
      sorted = s_flag != 0;
      __syncthreads();
    }
    if (!sorted && t == 0) atomicOr(err + d.eslot, 1u);  // not expected: bound above
    // force, kick, drift, block by block, into the other buffer set
    const int b = a ^ 1;
    u64 *ko = A.key[b] + d.off;
    double *vo = A.vel[b] + d.off;
    i32 *io = A.idx[b] + d.off;
    double *mo = eq ? nullptr : A.mass[b] + d.off;
    const double hd = step == d.nleap - 1 ? hh : d.h;
    u64 run = 0;
    for (u32 lo = 0; lo < n; lo += BLOCK) {
      const u32 c = min((u32)BLOCK, n - lo);
#pragma unroll
      for (int i = 0; i < BPER; ++i) {
        const u32 e = i * BWT + t;
        if (e < c) {
          s.key[e] = key[lo + e];
          s.v[e] = vel[lo + e];
          s.id[e] = idx[lo + e];
          s.m[e] = mass ? mass[lo + e] : 0.0;
          s.ord[e] = (unsigned short)e;
        }
      }
      __syncthreads();
      run = block_leap(s, d, c, lo, run, hd, s_red);
#pragma unroll
      for (int i = 0; i < BPER; ++i) {
        const u32 e = i * BWT + t;
        if (e < c) {
          ko[lo + e] = s.key[e];
          vo[lo + e] = s.v[e];
          io[lo + e] = s.id[e];
          if (mo) mo[lo + e] = s.m[e];
        }
      }
      __syncthreads();
    }
    a = b;
  }
}

}  // namespace

void batch_desc_fill(BatchDesc &d, const double *m, u32 n, double totmass,
                     double omega2) {
  bool eq = true;
  for (u32 i = 1; i < n && eq; ++i) eq = m[i] == m[0];
  d.off = 0;
  d.n = n;
  d.flags = (eq ? 1u : 0u) | (omega2 >= 0.0 ? 2u : 0u);
  d.cur = 0;
  d.nleap = 0;
  d.eslot = 0;
  d.pad_ = 0;
  d.h = 0.0;
  d.mq = m[0];
  d.totmass = totmass;
  d.omega2 = omega2 >= 0.0 ? omega2 : -1.0;
  // fixed-point scale of the exact prefix: 2^k * totmass <= 2^62
  int e = 0;
  (void)frexp(totmass > 0.0 ? totmass : 1.0, &e);
  d.fscale = ldexp(1.0, 62 - e);
  d.finv = ldexp(1.0, e - 62);
}

int batch_launch(hipStream_t st, const BatchDesc *d_descs, int count,
                 const BatchArrays &arr, u32 *err) {
  if (count <= 0) return WDY_OK;
  const size_t lds = Lds<BSH, false>::BYTES;
  WDY_HIP(hipFuncSetAttribute((const void *)k_batch_advance,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  k_batch_advance<<<count, BWT, lds, st>>>(d_descs, arr, err);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

}  // namespace wdy

// ============================================================================
// Host side: the wdy_batch handle
// ============================================================================
namespace wdy {
namespace {

__global__ void k_batch_init(u64 n, const double *__restrict__ x,
                             u64 *__restrict__ key, i32 *__restrict__ idx) {
  const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  key[i] = encode_pos(x[i]);
  idx[i] = (i32)i;
}

// One workgroup per descriptor: de-sort its current set into x_out, v_out
// (arena order, i.e. each simulation's sheets in input order).
__global__ void k_batch_desort(const BatchDesc *__restrict__ descs,
                               BatchArrays A, double *__restrict__ xo,
                               double *__restrict__ vo) {
  const BatchDesc d = descs[blockIdx.x];
  const int c = d.cur;
  for (u32 i = threadIdx.x; i < d.n; i += blockDim.x) {
    const u64 p = d.off + i;
    const i32 j = A.idx[c][p];
    xo[j] = decode_pos(A.key[c][p]);
    vo[j] = A.vel[c][p];
  }
}

__global__ void k_batch_flip(BatchDesc *descs, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count && descs[i].nleap > 0) descs[i].cur ^= (descs[i].nleap & 1);
}

}  // namespace
}  // namespace wdy

using namespace wdy;

struct wdy_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t nsims = 0;
  u64 ntot = 0;
  BatchArrays arr{};
  BatchDesc *d_desc = nullptr;
  std::vector<BatchDesc> h_desc;  // launch order: largest first
  double *out_x = nullptr, *out_v = nullptr;
  u32 *err = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double h = -1.0;
  int nleap = -1;
};

static void batch_free(wdy_batch *b) {
  if (!b) return;
  for (int s = 0; s < 2; ++s) {
    if (b->arr.key[s]) (void)hipFree(b->arr.key[s]);
    if (b->arr.vel[s]) (void)hipFree(b->arr.vel[s]);
    if (b->arr.idx[s]) (void)hipFree(b->arr.idx[s]);
    if (b->arr.mass[s]) (void)hipFree(b->arr.mass[s]);
  }
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->out_x) (void)hipFree(b->out_x);
  if (b->out_v) (void)hipFree(b->out_v);
  if (b->err) (void)hipFree(b->err);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" {

int wdy_batch_create(int64_t nsims, const int64_t *offsets, const double *x,
                     const double *v, const double *m, const double *totmass,
                     const double *omega2, wdy_batch **out) {
  if (!out) return fail(WDY_EINVAL, "wdy_batch_create: out is NULL");
  *out = nullptr;
  if (nsims < 1 || !offsets || !x || !v || !m || !totmass || !omega2)
    return fail(WDY_EINVAL, "wdy_batch_create: bad arguments");
  if (offsets[0] != 0)
    return fail(WDY_EINVAL, "wdy_batch_create: offsets[0] must be 0");
  for (int64_t k = 0; k < nsims; ++k) {
    if (offsets[k + 1] <= offsets[k])
      return fail(WDY_EINVAL, "wdy_batch_create: simulation %lld is empty or "
                  "offsets decrease", (long long)k);
    if (!std::isfinite(totmass[k]) || !std::isfinite(omega2[k]))
      return fail(WDY_EINVAL, "wdy_batch_create: non-finite totmass/omega2 of "
                  "simulation %lld", (long long)k);
  }
  const int64_t ntot = offsets[nsims];
  if (ntot >= (int64_t(1) << 31))
    return fail(WDY_EINVAL, "wdy_batch_create: %lld sheets in total, limit 2^31",
                (long long)ntot);
  for (int64_t i = 0; i < ntot; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(v[i]) || !std::isfinite(m[i]) ||
        m[i] < 0.0)
      return fail(WDY_EINVAL, "wdy_batch_create: non-finite or negative input "
                  "at sheet %lld", (long long)i);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(WDY_ENODEV, "wdy_batch_create: no GPU visible");
  wdy_batch *b = new (std::nothrow) wdy_batch();
  if (!b) return fail(WDY_ENOMEM, "wdy_batch_create: host allocation failed");
  int rc = [&]() -> int {
    WDY_HIP(hipGetDevice(&b->device));
    WDY_HIP(hipStreamCreate(&b->stream));
    WDY_HIP(hipEventCreate(&b->ev0));
    WDY_HIP(hipEventCreate(&b->ev1));
    b->nsims = nsims;
    b->ntot = (u64)ntot;
    bool any_unequal = false;
    std::vector<BatchDesc> descs((size_t)nsims);
    for (int64_t k = 0; k < nsims; ++k) {
      BatchDesc &d = descs[(size_t)k];
      batch_desc_fill(d, m + offsets[k], (u32)(offsets[k + 1] - offsets[k]),
                      totmass[k], omega2[k]);
      d.off = (u64)offsets[k];
      any_unequal = any_unequal || !(d.flags & 1u);
    }
    std::stable_sort(descs.begin(), descs.end(),
                     [](const BatchDesc &p, const BatchDesc &q) { return p.n > q.n; });
    b->h_desc = descs;
    const size_t N = (size_t)ntot;
    for (int s = 0; s < 2; ++s) {
      WDY_HIP(hipMalloc(&b->arr.key[s], N * 8));
      WDY_HIP(hipMalloc(&b->arr.vel[s], N * 8));
      WDY_HIP(hipMalloc(&b->arr.idx[s], N * 4));
      if (any_unequal) WDY_HIP(hipMalloc(&b->arr.mass[s], N * 8));
    }
    WDY_HIP(hipMalloc(&b->d_desc, descs.size() * sizeof(BatchDesc)));
    WDY_HIP(hipMalloc(&b->out_x, N * 8));
    WDY_HIP(hipMalloc(&b->out_v, N * 8));
    WDY_HIP(hipMalloc(&b->err, 8));
    WDY_HIP(hipMemsetAsync(b->err, 0, 8, b->stream));
    WDY_HIP(hipMemcpyAsync(b->d_desc, descs.data(), descs.size() * sizeof(BatchDesc),
                           hipMemcpyHostToDevice, b->stream));
    WDY_HIP(hipMemcpyAsync(b->out_x, x, N * 8, hipMemcpyHostToDevice, b->stream));
    WDY_HIP(hipMemcpyAsync(b->arr.vel[0], v, N * 8, hipMemcpyHostToDevice, b->stream));
    if (any_unequal)
      WDY_HIP(hipMemcpyAsync(b->arr.mass[0], m, N * 8, hipMemcpyHostToDevice,
                             b->stream));
    k_batch_init<<<(unsigned)((N + 255) / 256), 256, 0, b->stream>>>(
        (u64)N, b->out_x, b->arr.key[0], b->arr.idx[0]);
    WDY_CHECK_LAUNCH();
    WDY_HIP(hipStreamSynchronize(b->stream));
    return WDY_OK;
  }();
  if (rc) {
    batch_free(b);
    return rc;
  }
  *out = b;
  return WDY_OK;
}

int wdy_batch_advance(wdy_batch *b, double h, int nleap) {
  if (!b) return fail(WDY_EINVAL, "wdy_batch_advance: batch is NULL");
  if (nleap < 1 || !std::isfinite(h))
    return fail(WDY_EINVAL, "wdy_batch_advance: bad h or nleap");
  WDY_HIP(hipSetDevice(b->device));
  if (h != b->h || nleap != b->nleap) {
    // descriptors live on the device (their `cur` flips there); update the
    // step parameters in place by re-uploading after a read of `cur`
    std::vector<BatchDesc> cur(b->h_desc.size());
    WDY_HIP(hipMemcpyAsync(cur.data(), b->d_desc, cur.size() * sizeof(BatchDesc),
                           hipMemcpyDeviceToHost, b->stream));
    WDY_HIP(hipStreamSynchronize(b->stream));
    for (auto &d : cur) {
      d.h = h;
      d.nleap = nleap;
    }
    WDY_HIP(hipMemcpyAsync(b->d_desc, cur.data(), cur.size() * sizeof(BatchDesc),
                           hipMemcpyHostToDevice, b->stream));
    WDY_HIP(hipStreamSynchronize(b->stream));
    b->h = h;
    b->nleap = nleap;
  }
  WDY_HIP(hipEventRecord(b->ev0, b->stream));
  if (int rc = batch_launch(b->stream, b->d_desc, (int)b->nsims, b->arr, b->err))
    return rc;
  k_batch_flip<<<(unsigned)((b->nsims + 255) / 256), 256, 0, b->stream>>>(
      b->d_desc, (int)b->nsims);
  WDY_CHECK_LAUNCH();
  WDY_HIP(hipEventRecord(b->ev1, b->stream));
  return WDY_OK;
}

int wdy_batch_read(wdy_batch *b, double *x_out, double *v_out,
                   double *seconds_out) {
  if (!b) return fail(WDY_EINVAL, "wdy_batch_read: batch is NULL");
  WDY_HIP(hipSetDevice(b->device));
  const size_t N = (size_t)b->ntot;
  if (x_out || v_out) {
    k_batch_desort<<<(unsigned)b->nsims, 256, 0, b->stream>>>(b->d_desc, b->arr,
                                                               b->out_x, b->out_v);
    WDY_CHECK_LAUNCH();
    if (x_out)
      WDY_HIP(hipMemcpyAsync(x_out, b->out_x, N * 8, hipMemcpyDeviceToHost,
                             b->stream));
    if (v_out)
      WDY_HIP(hipMemcpyAsync(v_out, b->out_v, N * 8, hipMemcpyDeviceToHost,
                             b->stream));
  }
  u32 err = 0;
  WDY_HIP(hipMemcpyAsync(&err, b->err, 4, hipMemcpyDeviceToHost, b->stream));
  WDY_HIP(hipStreamSynchronize(b->stream));
  if (err)
    return fail(WDY_EVERIFY, "wdy_batch_read: a simulation did not sort within "
                "its bound");
  if (seconds_out) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess) ms = 0.f;
    *seconds_out = ms * 1e-3;
  }
  return WDY_OK;
}

int wdy_batch_destroy(wdy_batch *b) {
  if (!b) return WDY_OK;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  batch_free(b);
  return WDY_OK;
}

}  // extern "C"
