// LSD radix sort of (uint64 key, uint32 payload) pairs for gfx950.
//
// This is the cold-start / fallback sort of the leapfrog step (it replaces
// the comparison sorts of wendy/wendy.c:341-357).  Design:
//   * one histogram kernel reads every key once and counts all 8 digits
//     (8 x 256 bins) -- the "one key-read histogram pass" of a onesweep sort;
//   * each of the 8 passes is ONE scatter kernel.  A workgroup takes a ticket
//     (tile id in launch order, so every earlier tile is already resident),
//     ranks its 6144 keys stably with wave64 ballots (8 ballots give each lane
//     the mask of lanes holding the same digit; popcount below the lane is its
//     rank in the row), publishes its per-digit counts and resolves its global
//     per-digit offsets with a decoupled look-back over earlier tiles, stages
//     the tile in LDS in digit order and writes it out so that consecutive
//     lanes write consecutive addresses of one digit run.
// Look-back status words are 64-bit: [63:34] pass epoch, [33:32] flag
// (1 = tile aggregate, 2 = inclusive prefix), [31:0] count.  The epoch makes
// stale words from earlier passes invisible, so the status array is never
// cleared between passes.
#include <cstdlib>

#include "wdy_internal.h"

namespace wdy {
namespace {

constexpr int RADIX = 256;
constexpr int THREADS = 256;
constexpr int WAVES = THREADS / 64;
constexpr int ITEMS = (int)RADIX_TILE_KEYS / THREADS;
constexpr int TILE = THREADS * ITEMS;  // RADIX_TILE_KEYS keys per workgroup
static_assert(TILE == (int)RADIX_TILE_KEYS, "tile");
constexpr int PASSES = 8;

__device__ __forceinline__ u64 atomic_load_agent(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_agent(u64 *p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive scan of one u32 per thread over a 256-thread block.
__device__ __forceinline__ u32 block_excl_scan(u32 v, u32 *s_w) {
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  u32 x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    u32 y = __shfl_up(x, o, 64);
    if (lane >= (u32)o) x += y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  u32 wp = 0;
#pragma unroll
  for (int i = 0; i < WAVES; ++i)
    if (i < (int)w) wp += s_w[i];
  __syncthreads();
  return wp + x - v;
}

// Clears the histograms and the pass tickets; with clear_err also the error
// word counters[ERR_WORD] of a sort whose caller checks it per sort.
constexpr u32 ERR_WORD = 32;
__global__ void k_sort_init(u32 *hist, u32 *counters, u32 clear_err) {
  for (u32 i = threadIdx.x; i < PASSES * RADIX; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x < PASSES) counters[threadIdx.x] = 0;
  if (clear_err && threadIdx.x == 0) counters[ERR_WORD] = 0;
}

// Counts all 8 digits of every key.  Grid-stride over tiles.  The high
// digits of positions that share a sign and exponent range take few values,
// so most lanes of a wave would add to the same LDS word; the lanes whose
// digit equals the first active lane's are counted by that lane with one
// atomic, the others add one each.
__global__ __launch_bounds__(THREADS) void k_hist(const u64 *__restrict__ keys,
                                                  u32 n, u32 ntiles,
                                                  u32 *__restrict__ hist) {
  __shared__ u32 s_h[PASSES][RADIX];
  for (u32 i = threadIdx.x; i < PASSES * RADIX; i += THREADS)
    (&s_h[0][0])[i] = 0;
  __syncthreads();
  const u32 lane = threadIdx.x & 63;
  for (u32 tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const u32 base = tile * TILE;
#pragma unroll 4
    for (int j = 0; j < ITEMS; ++j) {
      const u32 i = base + j * THREADS + threadIdx.x;
      if (i < n) {
        const u64 k = keys[i];
        const u64 act = __ballot(1);
        const u32 first = (u32)__ffsll((long long)act) - 1u;
#pragma unroll
        for (int p = 0; p < PASSES; ++p) {
          const u32 d = (u32)(k >> (8 * p)) & 255u;
          const u32 d0 = __builtin_amdgcn_readfirstlane(d);
          const u64 same = __ballot(d == d0);
          if (lane == first) atomicAdd(&s_h[p][d0], (u32)__popcll(same));
          if (d != d0) atomicAdd(&s_h[p][d], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < PASSES * RADIX; i += THREADS) {
    const u32 c = (&s_h[0][0])[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// Turns the 8 digit histograms into exclusive offsets, in place.
__global__ __launch_bounds__(THREADS) void k_hist_scan(u32 *hist) {
  __shared__ u32 s_w[WAVES];
  for (int p = 0; p < PASSES; ++p) {
    const u32 c = hist[p * RADIX + threadIdx.x];
    const u32 e = block_excl_scan(c, s_w);
    hist[p * RADIX + threadIdx.x] = e;
  }
}

template <bool IDENT>
__global__ __launch_bounds__(THREADS) void k_scatter(
    const u64 *__restrict__ kin, const u32 *__restrict__ pin,
    u64 *__restrict__ kout, u32 *__restrict__ pout, u32 n, u32 shift,
    const u32 *__restrict__ offs, u64 *status, u32 *counter, u32 epoch,
    u32 spin_limit, u32 *err) {
  // 78,848 bytes: two workgroups fit a CU's 160 KiB (the 212 VGPRs of 24
  // keys per thread allow two waves per SIMD as well).  The tile number
  // and the scan's wave totals live in the
  // key staging area, which is written only after both are consumed; the
  // digits' global bases replace the tile-local ones once the local
  // scatter has read those.
  __shared__ u32 s_cnt[WAVES][RADIX];
  __shared__ u32 s_dbase[RADIX];
  __shared__ u64 s_key[TILE];
  __shared__ u32 s_pay[TILE];
  u32 &s_tile = *(u32 *)&s_key[0];
  u32 *const s_w = (u32 *)&s_key[1];

  const u32 t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (t == 0) s_tile = atomicAdd(counter, 1u);
  for (u32 i = t; i < WAVES * RADIX; i += THREADS) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const u32 tile = s_tile;
  const u32 base = tile * TILE;
  const u32 tile_n = min((u32)TILE, n - base);
  const u64 lt = (1ull << lane) - 1ull;

  u64 k[ITEMS];
  u32 p[ITEMS];
  u32 r[ITEMS];
  // Wave w owns tile-local items [w*1024, (w+1)*1024) as 16 rows of 64, so
  // ranking row by row, wave by wave, is stable.
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const u32 li = w * (ITEMS * 64) + j * 64 + lane;
    const bool valid = li < tile_n;
    const u32 gi = base + li;
    k[j] = valid ? kin[gi] : 0ull;
    p[j] = IDENT ? gi : (valid ? pin[gi] : 0u);
  }
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const u32 li = w * (ITEMS * 64) + j * 64 + lane;
    const bool valid = li < tile_n;
    const u32 d = (u32)(k[j] >> shift) & 255u;
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const u64 bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const u32 below = __popcll(peers & lt);
    const u32 old = s_cnt[w][d];
    if (valid && below == 0) s_cnt[w][d] = old + (u32)__popcll(peers);
    r[j] = old + below;
  }
  __syncthreads();

  // Thread t owns digit t: wave prefixes, tile-local digit base, look-back.
  u32 tot = 0;
  {
    u32 c[WAVES];
#pragma unroll
    for (int i = 0; i < WAVES; ++i) c[i] = s_cnt[i][t];
#pragma unroll
    for (int i = 0; i < WAVES; ++i) {
      s_cnt[i][t] = tot;
      tot += c[i];
    }
  }
  const u32 dex = block_excl_scan(tot, s_w);
  s_dbase[t] = dex;
  u32 gbase;
  {
    const u64 ep = (u64)(epoch & 0x3FFFFFFFu) << 34;
    u64 *mine = status + (u64)tile * RADIX + t;
    u32 excl = 0;
    if (tile == 0) {
      atomic_store_agent(mine, ep | (2ull << 32) | tot);
    } else {
      atomic_store_agent(mine, ep | (1ull << 32) | tot);
      int64_t pt = (int64_t)tile - 1;
      // every earlier tile publishes without waiting on a later one, so the
      // wait ends; the bound turns a broken invariant into a reported error
      // (a limit of 0 gives up at once).  A tile that gave up still
      // publishes, so that every later tile and the kernel finish.
      bool gave_up = spin_limit == 0;
      u32 spins = 0;
      while (!gave_up) {
        const u64 s = atomic_load_agent(status + (u64)pt * RADIX + t);
        const u32 flag = (u32)(s >> 32) & 3u;
        if ((u32)(s >> 34) != (epoch & 0x3FFFFFFFu) || flag == 0) {
          if (++spins > spin_limit) {
            gave_up = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += (u32)s;
        if (flag == 2) break;
        --pt;
      }
      if (gave_up) atomicOr(err, 1u);
      atomic_store_agent(mine, ep | (2ull << 32) | (excl + tot));
    }
    gbase = offs[t] + excl - dex;
  }
  __syncthreads();

#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const u32 li = w * (ITEMS * 64) + j * 64 + lane;
    if (li < tile_n) {
      const u32 d = (u32)(k[j] >> shift) & 255u;
      const u32 lpos = s_dbase[d] + s_cnt[w][d] + r[j];
      s_key[lpos] = k[j];
      s_pay[lpos] = p[j];
    }
  }
  __syncthreads();
  s_dbase[t] = gbase;
  __syncthreads();
  for (u32 i = t; i < tile_n; i += THREADS) {
    const u64 kk = s_key[i];
    const u32 d = (u32)(kk >> shift) & 255u;
    const u32 g = s_dbase[d] + i;
    // always true unless a look-back wait of this or an earlier pass gave
    // up: its tile's offsets, and the digit counts of the keys a later pass
    // reads, are then wrong, and a write could land past the array
    if (g < n) {
      kout[g] = kk;
      pout[g] = s_pay[i];
    }
  }
}

}  // namespace

u32 *radix_err_word(RadixWork *w) { return w->counters + ERR_WORD; }

u32 radix_tiles(u32 n) { return n == 0 ? 1u : (n + TILE - 1) / TILE; }
size_t radix_hist_bytes() { return (size_t)PASSES * RADIX * sizeof(u32); }
size_t radix_counter_bytes() { return 64 * sizeof(u32); }
size_t radix_status_bytes(u32 n) {
  return (size_t)radix_tiles(n) * RADIX * sizeof(u64);
}

int radix_alloc(RadixWork *w, u32 n, hipStream_t st) {
  w->n = n;
  w->ntiles = (n + TILE - 1) / TILE;
  if (w->ntiles == 0) w->ntiles = 1;
  WDY_HIP(hipMalloc(&w->hist, PASSES * RADIX * sizeof(u32)));
  WDY_HIP(hipMalloc(&w->counters, 64 * sizeof(u32)));
  WDY_HIP(hipMemsetAsync(w->counters, 0, 64 * sizeof(u32), st));
  WDY_HIP(hipMalloc(&w->status, (size_t)w->ntiles * RADIX * sizeof(u64)));
  WDY_HIP(hipMemsetAsync(w->status, 0, (size_t)w->ntiles * RADIX * sizeof(u64),
                         st));
  w->epoch = 1;
  return WDY_OK;
}

void radix_free(RadixWork *w) {
  if (w->hist) (void)hipFree(w->hist);
  if (w->counters) (void)hipFree(w->counters);
  if (w->status) (void)hipFree(w->status);
  w->hist = nullptr;
  w->counters = nullptr;
  w->status = nullptr;
}

int radix_sort_pairs(RadixWork *w, hipStream_t s, Timer *tm, u64 *k0, u64 *k1,
                     u32 *p0, u32 *p1, bool identity_payload) {
  const u32 n = w->n;
  if (n == 0) return WDY_OK;
  // bound of each look-back wait, in probes (WDY_RADIX_SPIN_LIMIT overrides
  // it; 0 makes every wait give up at once, so that tests can reach the
  // failure path with every kernel still finishing)
  u32 spin_limit = RADIX_SPIN_LIMIT;
  if (const char *e = getenv("WDY_RADIX_SPIN_LIMIT")) spin_limit = (u32)strtoul(e, nullptr, 10);
  k_sort_init<<<1, THREADS, 0, s>>>(w->hist, w->counters, w->clear_err ? 1u : 0u);
  WDY_CHECK_LAUNCH();
  const u32 hist_blocks = w->ntiles < 2048 ? w->ntiles : 2048;
  if (tm) timer_begin(tm, s, KC_SORT_HIST);
  k_hist<<<hist_blocks, THREADS, 0, s>>>(k0, n, w->ntiles, w->hist);
  if (tm) timer_end(tm, s, KC_SORT_HIST);
  WDY_CHECK_LAUNCH();
  k_hist_scan<<<1, THREADS, 0, s>>>(w->hist);
  WDY_CHECK_LAUNCH();
  for (int pass = 0; pass < PASSES; ++pass) {
    const bool even = (pass & 1) == 0;
    const u64 *ks = even ? k0 : k1;
    u64 *kd = even ? k1 : k0;
    const u32 *ps = even ? p0 : p1;
    u32 *pd = even ? p1 : p0;
    const u32 ep = w->epoch++;
    if (w->epoch > 0x3FFFFFFFu) w->epoch = 1;
    if (tm) timer_begin(tm, s, KC_SORT_PASS);
    if (pass == 0 && identity_payload)
      k_scatter<true><<<w->ntiles, THREADS, 0, s>>>(
          ks, ps, kd, pd, n, 8u * pass, w->hist + pass * RADIX, w->status,
          w->counters + pass, ep, spin_limit, w->counters + ERR_WORD);
    else
      k_scatter<false><<<w->ntiles, THREADS, 0, s>>>(
          ks, ps, kd, pd, n, 8u * pass, w->hist + pass * RADIX, w->status,
          w->counters + pass, ep, spin_limit, w->counters + ERR_WORD);
    if (tm) timer_end(tm, s, KC_SORT_PASS);
    WDY_CHECK_LAUNCH();
  }
  return WDY_OK;
}

}  // namespace wdy
