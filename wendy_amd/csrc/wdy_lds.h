// LDS building blocks shared by the warm-start step and the batched
// ensemble kernel: block scans, the (position, sheet index) order, and the
// in-LDS bucket sort (one counting pass over position bins, then each sheet
// ranks itself within its bin).
#pragma once

#include "wdy_internal.h"

namespace wdy {

__device__ __forceinline__ u64 mass_fix(double m, double fscale) {
  return (u64)llrint(m * fscale);
}

// Wave-wide scans and reductions on DPP lane moves (row shifts within rows
// of 16 lanes, then the broadcasts of lane 15 and lane 31 into the rows
// above): no LDS traffic, a few cycles per step.  Every lane of the wave
// must be active.  A move whose source lane is outside the row, or whose
// row is masked off, returns `old` (the operation's identity).
constexpr int DPP_SHR1 = 0x111, DPP_SHR2 = 0x112, DPP_SHR4 = 0x114,
              DPP_SHR8 = 0x118, DPP_BCAST15 = 0x142, DPP_BCAST31 = 0x143;

template <int CTRL, int ROWS>
__device__ __forceinline__ u32 dpp_mov(u32 old, u32 x) {
  return (u32)__builtin_amdgcn_update_dpp((int)old, (int)x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ u64 dpp_mov(u64 old, u64 x) {
  return ((u64)dpp_mov<CTRL, ROWS>((u32)(old >> 32), (u32)(x >> 32)) << 32) |
         dpp_mov<CTRL, ROWS>((u32)old, (u32)x);
}
// inclusive prefix sum over the lanes (lane 63: the wave's total)
template <class T>
__device__ __forceinline__ T wave_incl_add(T x) {
  x += dpp_mov<DPP_SHR1, 0xf>((T)0, x);
  x += dpp_mov<DPP_SHR2, 0xf>((T)0, x);
  x += dpp_mov<DPP_SHR4, 0xf>((T)0, x);
  x += dpp_mov<DPP_SHR8, 0xf>((T)0, x);
  x += dpp_mov<DPP_BCAST15, 0xa>((T)0, x);
  x += dpp_mov<DPP_BCAST31, 0xc>((T)0, x);
  return x;
}
// lane 63's value, in every lane
__device__ __forceinline__ u32 wave_last(u32 x) {
  return (u32)__builtin_amdgcn_readlane((int)x, 63);
}
__device__ __forceinline__ u64 wave_last(u64 x) {
  return ((u64)wave_last((u32)(x >> 32)) << 32) | wave_last((u32)x);
}
// the wave's minimum and maximum, in every lane
__device__ __forceinline__ u64 wave_min(u64 x) {
  u64 y;
  y = dpp_mov<DPP_SHR1, 0xf>(~0ull, x); x = y < x ? y : x;
  y = dpp_mov<DPP_SHR2, 0xf>(~0ull, x); x = y < x ? y : x;
  y = dpp_mov<DPP_SHR4, 0xf>(~0ull, x); x = y < x ? y : x;
  y = dpp_mov<DPP_SHR8, 0xf>(~0ull, x); x = y < x ? y : x;
  y = dpp_mov<DPP_BCAST15, 0xa>(~0ull, x); x = y < x ? y : x;
  y = dpp_mov<DPP_BCAST31, 0xc>(~0ull, x); x = y < x ? y : x;
  return wave_last(x);
}
__device__ __forceinline__ u64 wave_max(u64 x) {
  u64 y;
  y = dpp_mov<DPP_SHR1, 0xf>(0ull, x); x = y > x ? y : x;
  y = dpp_mov<DPP_SHR2, 0xf>(0ull, x); x = y > x ? y : x;
  y = dpp_mov<DPP_SHR4, 0xf>(0ull, x); x = y > x ? y : x;
  y = dpp_mov<DPP_SHR8, 0xf>(0ull, x); x = y > x ? y : x;
  y = dpp_mov<DPP_BCAST15, 0xa>(0ull, x); x = y > x ? y : x;
  y = dpp_mov<DPP_BCAST31, 0xc>(0ull, x); x = y > x ? y : x;
  return wave_last(x);
}

// Exclusive scan of one u64 per thread over the block; *total gets the sum.
template <int NWAVE>
__device__ u64 block_excl_u64(u64 v, u64 *s_red, u64 *total) {
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const u64 x = wave_incl_add(v);
  __syncthreads();
  if (lane == 63) s_red[w] = x;
  __syncthreads();
  u64 wp = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < NWAVE; ++i) {
    if (i < (int)w) wp += s_red[i];
    tot += s_red[i];
  }
  if (total) *total = tot;
  return wp + x - v;
}

// The same for u32 values whose total fits 32 bits, with one barrier: the
// caller guarantees that a barrier separates every earlier read of s_red
// from this call.
template <int NWAVE>
__device__ u32 block_excl_u32(u32 v, u64 *s_red) {
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const u32 x = wave_incl_add(v);
  if (lane == 63) s_red[w] = x;
  __syncthreads();
  u32 wp = 0;
#pragma unroll
  for (int i = 0; i < NWAVE; ++i)
    if (i < (int)w) wp += (u32)s_red[i];
  return wp + x - v;
}

// Exclusive scan of s[0..n) (n <= threads) in place; returns the total.
template <int NWAVE>
__device__ u32 block_excl_u32_inplace(u32 *s, u32 n, u64 *s_red) {
  const u32 t = threadIdx.x;
  const u32 v = t < n ? s[t] : 0u;
  u64 tot;
  const u64 e = block_excl_u64<NWAVE>(v, s_red, &tot);
  __syncthreads();
  if (t < n) s[t] = (u32)e;
  __syncthreads();
  return (u32)tot;
}

// The same for a short array: for n <= 64 one wave scans with shuffles and a
// single barrier publishes the result (and every LDS write made before the
// call).  The values must have been written by wave 0 (threads 0..63) or be
// visible to it already.
template <int NWAVE>
__device__ u32 short_excl_u32_inplace(u32 *s, u32 n, u64 *s_red, u32 *s_tot) {
  if (n > 64) return block_excl_u32_inplace<NWAVE>(s, n, s_red);
  const u32 t = threadIdx.x;
  if (t < 64) {
    const u32 v = t < n ? s[t] : 0u;
    const u32 x = wave_incl_add(v);
    if (t < n) s[t] = x - v;
    if (t == 63) *s_tot = x;
  }
  __syncthreads();
  return *s_tot;
}

__device__ __forceinline__ bool pair_lt(u64 ka, i32 ia, u64 kb, i32 ib) {
  return ka < kb || (ka == kb && ia < ib);
}

// One tile shape: threads per workgroup and the most sheets a bucket holds.
// WT threads, CAP sheets per bucket, NQ position bins for the bucket sort.
// PACK: two 16-bit bin counters per 32-bit word (counts and offsets never
// exceed CAP < 2^16), so twice the bins fit the same LDS.  The bin words
// (NQW of them) double as the per-wave segment counters of the warm step's
// partition.  OCC_EQ: waves per SIMD the equal-mass pipelined kernel is
// compiled for (its VGPR budget: 512 / OCC_EQ); 1 leaves the register
// count to the compiler.
template <int WT_, int CAP_, int NQ_ = CAP_, int OCC_EQ_ = 1, bool PACK_ = false>
struct Shape {
  static constexpr int WT = WT_, CAP = CAP_, NQ = NQ_, PER = CAP_ / WT_;
  static constexpr int NWAVE = WT_ / 64, OCC_EQ = OCC_EQ_;
  static constexpr bool PACK = PACK_;
  static constexpr int NQW = PACK_ ? NQ_ / 2 : NQ_;
  static_assert(CAP_ % WT_ == 0 && NQW % WT_ == 0 && NQ_ % WT_ == 0 &&
                    WT_ % 64 == 0 && CAP_ <= 65535, "shape");
};

template <class SH, bool EQ>
struct Lds {
  static constexpr int CAP = SH::CAP;
  static constexpr size_t BYTES =
      (size_t)CAP * (EQ ? 22 : 30) + (size_t)SH::NQW * 4;
  static_assert(BYTES % 8 == 0, "what follows the arrays stays 8-byte aligned");
  u64 *key;
  double *v;
  double *m;  // unequal masses only
  i32 *id;
  unsigned short *ord;
  u32 *bins;
  __device__ explicit Lds(u64 *base) {
    key = base;
    v = (double *)(key + CAP);
    m = v + CAP;
    id = (i32 *)(EQ ? (v + CAP) : (m + CAP));
    ord = (unsigned short *)(id + CAP);
    bins = (u32 *)(ord + CAP);
  }
};

// rk[i] := the rank, in the (key, id) order of s[0..c), of sheet i * WT + t
// (for i * WT + t < c).  s.ord holds the bins' contents on return; a barrier
// must separate the ranking's LDS reads from later writes to s.
// KREG: the caller's registers kr[i], ir[i] hold the key and index of its
// sheet i * WT + t (as stored in s), so they need not be read back from LDS.
// SCAT (with KREG; only when s.v holds nothing the caller needs, as in the
// equal-mass warm step, whose velocities stay in registers): each sheet's key
// is also stored at its slot in s.v's space, so the in-bin comparison reads
// the bin's keys contiguously and goes through s.ord to an index only on an
// exact tie with another sheet.
template <class SH, bool EQ, bool KREG, bool SCAT = false>
__device__ void lds_bucket_rank_impl(Lds<SH, EQ> &s, u32 c, u64 *s_red,
                                     u64 *s_mm, u64 *prof, u32 (&rk)[SH::PER],
                                     const u64 (&kr)[SH::PER],
                                     const i32 (&ir)[SH::PER]) {
  constexpr int WT = SH::WT, PER = SH::PER, NQ = SH::NQ, NWAVE = SH::NWAVE;
  static_assert(!SCAT || KREG, "the slot-ordered keys come from registers");
  const u32 t = threadIdx.x;
  u64 *const sk = (u64 *)s.v;  // SCAT: keys in slot order
  u64 kmin = ~0ull, kmax = 0ull;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const u32 e = i * WT + t;
    if (e < c) {
      const u64 k = KREG ? kr[i] : s.key[e];
      kmin = k < kmin ? k : kmin;
      kmax = k > kmax ? k : kmax;
    }
  }
  kmin = wave_min(kmin);
  kmax = wave_max(kmax);
  if ((t & 63) == 0) {
    s_mm[2 * (t >> 6)] = kmin;
    s_mm[2 * (t >> 6) + 1] = kmax;
  }
#pragma unroll
  for (int i = 0; i < SH::NQW / WT; ++i) s.bins[i * WT + t] = 0;
  __syncthreads();
  kmin = ~0ull;
  kmax = 0ull;
#pragma unroll
  for (int i = 0; i < NWAVE; ++i) {
    kmin = s_mm[2 * i] < kmin ? s_mm[2 * i] : kmin;
    kmax = s_mm[2 * i + 1] > kmax ? s_mm[2 * i + 1] : kmax;
  }
  // Bin of a sheet: monotone non-decreasing in its position.  Uniform in x
  // when the double range is finite (the usual case), uniform in key space
  // otherwise.
  const double xmin = decode_pos(kmin), xmax = decode_pos(kmax);
  const double range = xmax - xmin;
  const bool dbl = range > 0.0 && range <= 1.0e300;
  const double scale = dbl ? (double)NQ / range : 0.0;
  int kshift = 0;
  if (!dbl) {
    const u64 kr = kmax - kmin;
    while (kshift < 63 && (kr >> kshift) >= (u64)NQ) ++kshift;
  }
  u32 q[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const u32 e = i * WT + t;
    q[i] = 0;
    if (e < c) {
      const u64 k = KREG ? kr[i] : s.key[e];
      u32 b;
      if (dbl) {
        const double qd = (decode_pos(k) - xmin) * scale;
        b = qd < (double)NQ ? (u32)qd : (u32)(NQ - 1);
      } else {
        b = (u32)((k - kmin) >> kshift);
      }
      q[i] = b;
      if (SH::PACK)
        atomicAdd(&s.bins[b >> 1], 1u << ((b & 1u) * 16u));
      else
        atomicAdd(&s.bins[b], 1u);
    }
  }
  __syncthreads();
  if (prof && t == 0) prof[6] = wall_clock64();
  {
    constexpr int WPT = SH::NQW / WT;  // bin words t*WPT .. t*WPT + WPT - 1
    u32 loc[WPT];
    u32 sum = 0;  // counts total at most CAP
#pragma unroll
    for (int i = 0; i < WPT; ++i) {
      loc[i] = s.bins[t * WPT + i];
      sum += SH::PACK ? (loc[i] & 0xffffu) + (loc[i] >> 16) : loc[i];
    }
    // (two barriers above separate any earlier read of s_red; the scan's
    // barrier orders every thread's bin reads before the offsets replace
    // them)
    u32 run = block_excl_u32<NWAVE>(sum, s_red);
#pragma unroll
    for (int i = 0; i < WPT; ++i) {
      if (SH::PACK) {  // exclusive offsets of both bins of the word
        const u32 lo = run, hi = lo + (loc[i] & 0xffffu);
        s.bins[t * WPT + i] = lo | (hi << 16);
        run += (loc[i] & 0xffffu) + (loc[i] >> 16);
      } else {
        s.bins[t * WPT + i] = run;
        run += loc[i];
      }
    }
  }
  __syncthreads();
  u32 myslot[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const u32 e = i * WT + t;
    myslot[i] = 0;
    if (e < c) {
      u32 slot;
      if (SH::PACK) {
        const u32 sh = (q[i] & 1u) * 16u;
        slot = (atomicAdd(&s.bins[q[i] >> 1], 1u << sh) >> sh) & 0xffffu;
      } else {
        slot = atomicAdd(&s.bins[q[i]], 1u);
      }
      s.ord[slot] = (unsigned short)e;
      if (SCAT) {
        sk[slot] = kr[i];
        myslot[i] = slot;
      }
    }
  }
  __syncthreads();
  if (prof && t == 0) prof[7] = wall_clock64();
  // bin b now ends at bin_end(b).  Rank each sheet inside its bin by
  // counting the bin's smaller (key, id) pairs -- every sheet in parallel, a
  // bin of k sheets costs k comparisons per sheet (none when it is alone).
  auto bin_end = [&](u32 b) -> u32 {
    return SH::PACK ? (s.bins[b >> 1] >> ((b & 1u) * 16u)) & 0xffffu : s.bins[b];
  };
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const u32 e = i * WT + t;
    rk[i] = 0;
    if (e < c) {
      const u32 lo = q[i] == 0 ? 0u : bin_end(q[i] - 1), hi = bin_end(q[i]);
      u32 r = lo;
      if (SCAT && hi - lo > 1) {
        const u64 k = kr[i];
        const i32 id = ir[i];
        // another sheet at slot a precedes this one if its key is smaller,
        // or equal with a smaller index
        auto tie = [&](u32 a) -> u32 {
          return a != myslot[i] && s.id[s.ord[a]] < id ? 1u : 0u;
        };
        u32 a = lo;
        for (; a + 2 <= hi; a += 2) {  // two at a time: the LDS loads overlap
          const u64 k0 = sk[a], k1 = sk[a + 1];
          r += (k0 < k ? 1u : 0u) + (k1 < k ? 1u : 0u);
          if (k0 == k) r += tie(a);
          if (k1 == k) r += tie(a + 1);
        }
        if (a < hi) {
          const u64 k0 = sk[a];
          r += k0 < k ? 1u : 0u;
          if (k0 == k) r += tie(a);
        }
      } else if (hi - lo > 1) {
        const u64 k = KREG ? kr[i] : s.key[e];
        const i32 id = KREG ? ir[i] : s.id[e];
        u32 a = lo;
        for (; a + 2 <= hi; a += 2) {  // two at a time: the LDS loads overlap
          const unsigned short f0 = s.ord[a], f1 = s.ord[a + 1];
          const u64 k0 = s.key[f0], k1 = s.key[f1];
          const i32 i0 = s.id[f0], i1 = s.id[f1];
          r += (pair_lt(k0, i0, k, id) ? 1u : 0u) + (pair_lt(k1, i1, k, id) ? 1u : 0u);
        }
        if (a < hi) {
          const unsigned short f = s.ord[a];
          r += pair_lt(s.key[f], s.id[f], k, id) ? 1u : 0u;
        }
      }
      rk[i] = r;
    }
  }
}

template <class SH, bool EQ>
__device__ __forceinline__ void lds_bucket_rank(Lds<SH, EQ> &s, u32 c, u64 *s_red,
                                                u64 *s_mm, u64 *prof,
                                                u32 (&rk)[SH::PER]) {
  u64 kr[SH::PER];
  i32 ir[SH::PER];
  lds_bucket_rank_impl<SH, EQ, false>(s, c, s_red, s_mm, prof, rk, kr, ir);
}

template <class SH, bool EQ>
__device__ __forceinline__ void lds_bucket_rank_regs(
    Lds<SH, EQ> &s, u32 c, u64 *s_red, u64 *s_mm, u64 *prof,
    u32 (&rk)[SH::PER], const u64 (&kr)[SH::PER], const i32 (&ir)[SH::PER]) {
  lds_bucket_rank_impl<SH, EQ, true>(s, c, s_red, s_mm, prof, rk, kr, ir);
}

// The same, storing the keys in slot order in s.v's space (whose contents
// it overwrites) for the in-bin comparison.
template <class SH, bool EQ>
__device__ __forceinline__ void lds_bucket_rank_regs_scat(
    Lds<SH, EQ> &s, u32 c, u64 *s_red, u64 *s_mm, u64 *prof,
    u32 (&rk)[SH::PER], const u64 (&kr)[SH::PER], const i32 (&ir)[SH::PER]) {
  lds_bucket_rank_impl<SH, EQ, true, true>(s, c, s_red, s_mm, prof, rk, kr, ir);
}

// s.ord[0..c) := the permutation of 0..c-1 sorting (key, id) ascending.
template <class SH, bool EQ, bool KREG>
__device__ void lds_bucket_sort_impl(Lds<SH, EQ> &s, u32 c, u64 *s_red,
                                     u64 *s_mm, u64 *prof,
                                     const u64 (&kr)[SH::PER],
                                     const i32 (&ir)[SH::PER]) {
  constexpr int WT = SH::WT, PER = SH::PER;
  const u32 t = threadIdx.x;
  u32 rk[PER];
  lds_bucket_rank_impl<SH, EQ, KREG>(s, c, s_red, s_mm, prof, rk, kr, ir);
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const u32 e = i * WT + t;
    if (e < c) s.ord[rk[i]] = (unsigned short)e;
  }
  __syncthreads();
}

template <class SH, bool EQ>
__device__ __forceinline__ void lds_bucket_sort(Lds<SH, EQ> &s, u32 c, u64 *s_red,
                                                u64 *s_mm, u64 *prof) {
  u64 kr[SH::PER];
  i32 ir[SH::PER];
  lds_bucket_sort_impl<SH, EQ, false>(s, c, s_red, s_mm, prof, kr, ir);
}

// The same, with the caller's sheets' keys and indices in registers (see
// lds_bucket_rank_impl).
template <class SH, bool EQ>
__device__ __forceinline__ void lds_bucket_sort_regs(
    Lds<SH, EQ> &s, u32 c, u64 *s_red, u64 *s_mm, u64 *prof,
    const u64 (&kr)[SH::PER], const i32 (&ir)[SH::PER]) {
  lds_bucket_sort_impl<SH, EQ, true>(s, c, s_red, s_mm, prof, kr, ir);
}


}  // namespace wdy
