// Warm-start leapfrog step of the large path: ONE kernel per leapfrog step.
//
// Sheets move only a few thousand ranks per step, so each step re-sorts by
// value buckets whose boundaries ("splitters") are the sorted positions at
// ranks k*T of two steps earlier.  The previous step left, for every source
// bucket t, its sheets partitioned into contiguous segments by destination
// bucket t + d - W (0 <= d <= 2W), plus one segment of far movers that are
// also copied into an overflow list (see WarmMeta).
//
// Workgroup j (one per bucket, ticketed in launch order) runs a short chain
// of dependent global-memory round trips, the rest being LDS work:
//   1. takes its ticket (and reads the error word and the overflow count);
//   2. reads its incoming segment table and the window of splitters it will
//      partition with, together;
//   3. gathers the bucket into LDS, all loads of a thread issued together,
//      applying the interval's first half drift when asked
//      (wendy/wendy.c:318-323,398), and publishes its sheet count and
//      fixed-point mass total;
//   4. sorts the bucket in LDS by (position, sheet index): one counting pass
//      over position bins, then every sheet ranks itself inside its bin;
//   5. resolves its global rank offset and mass prefix with a decoupled
//      look-back (one wave probing 64 predecessors at a time) over integer
//      values, so nothing depends on timing;
//   6. computes force, kick and drift in sorted order (wendy/wendy.c:371-383,
//      324-333) and emits the splitters for two steps ahead;
//   7. partitions the drifted sheets by next-step bucket (binary search in
//      the prefetched window) and writes them to its output region
//      [offset, offset + count) with the segment table.
// Mode 1 of the same kernel only does step 7, for an array already sorted
// by the radix path (tile j = ranks [j*T, (j+1)*T)).
//
// Per-launch counters (tickets, overflow counts) live in rings of 4 indexed
// by the launch epoch; ticket 0 of launch e clears the slots launch e+2 will
// use, so no end-of-kernel bookkeeping is needed.
//
// Masses enter the prefix as 64-bit fixed point: m_fix = llrint(m * 2^k)
// with 2^k * totmass <= 2^62, so bucket totals and prefixes are exact
// integers whatever order they are added in; C = prefix * 2^-k rounds once.
// Equal masses need no prefix: C = rank * m.
//
// Failure (a bucket over capacity, a full overflow list, a look-back that
// exceeds its bounded wait) records the failing launch's epoch in err[0];
// later launches exit at once, and the host redoes the interval from its
// untouched start state on the radix path.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "wdy_lds.h"
#include "wdy_warm.h"

namespace wdy {
namespace {

constexpr int SMAX = 2 * WARM_WMAX + 1;
constexpr int LOVF = 256;  // overflow entries one bucket may take

__device__ __forceinline__ u64 ld_rlx(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(u64 *p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 ld_u32(const u32 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

#define WARM_STAMP(k)                                              \
  do {                                                             \
    if (A.prof && t == 0) A.prof[8ull * j + (k)] = wall_clock64(); \
  } while (0)

// Look-back status of a bucket: 5 words, each tagged with the launch epoch
// so that relaxed atomics suffice (no acquire/release, which on a multi-XCD
// part cost L2 write-backs and invalidations):
//   [0] epoch<<34 | flag<<32 | count   (flag 1 aggregate, 2 inclusive)
//   [1],[2] aggregate mass, high / low 32 bits, each epoch<<34 | 1<<32 | half
//   [3],[4] inclusive mass, likewise
// A reader that sees the count word before the mass words retries.
constexpr u64 STATUS_WORDS = 5;

__device__ __forceinline__ void publish(u64 *st, u64 ep, u32 flag, u32 cnt,
                                        u64 mass, bool with_mass) {
  if (with_mass) {
    u64 *mw = st + (flag == 2 ? 3 : 1);
    const u64 tag = ep | (1ull << 32);
    st_rlx(mw, tag | (mass >> 32));
    st_rlx(mw + 1, tag | (mass & 0xffffffffull));
  }
  st_rlx(st, ep | ((u64)flag << 32) | cnt);
}

__device__ void set_err(u32 *err, u32 epoch, u32 kind) {
  atomicCAS(err, 0u, epoch);
  atomicOr(err + 1, kind);
}

// The bucket of key k: the largest b < nb with spl[b] <= k (spl[0] = 0).
__device__ __forceinline__ u32 bucket_of(const u64 *spl, u32 nb, u64 k) {
  u32 lo = 0, hi = nb - 1;
  while (lo < hi) {
    const u32 mid = (lo + hi + 1) >> 1;
    if (spl[mid] <= k) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Position of element e (0 <= e < nseg) among the incoming segments.
__device__ __forceinline__ u32 seg_pos(const u32 *segpre, const u32 *segsrc,
                                       u32 S, u32 e) {
  u32 lo = 0, hi = S - 1;  // largest d with segpre[d] <= e
  while (lo < hi) {
    const u32 mid = (lo + hi + 1) >> 1;
    if (segpre[mid] <= e) lo = mid; else hi = mid - 1;
  }
  return segsrc[lo] + (e - segpre[lo]);
}

// Smallest power of two >= x (x >= 1).
__host__ __device__ constexpr u32 pow2_ge(u32 x) {
  return x <= 1 ? 1u : 1u << (32 - __builtin_clz(x - 1));
}

// The largest a < P with tab[a] <= v, for a table that is non-decreasing on
// [0, P) with tab[0] <= v (P a power of two <= 256; entries past the real
// ones padded with the type's maximum): a fixed sequence of halving steps,
// one load, compare and select each, the same for every lane.
template <class T>
__device__ __forceinline__ u32 pow2_search(const T *tab, u32 P, T v) {
  u32 a = 0;
#pragma unroll 1
  for (u32 st = P >> 1; st; st >>= 1)
    if (tab[a + st] <= v) a += st;
  return a;
}

// One bucket of mode 0 or 1 (steps 2-7 above) for ticket j of the step
// described by A; novf_in: the overflow count of the list it reads.  s_u
// (4 words of the caller's LDS) has s_u[3] = 0 and is published by a
// barrier on entry.  Every thread of the workgroup calls it with the same
// arguments.
template <class SH, bool EQ, bool HARM, int MODE>
__device__ __forceinline__ void warm_bucket(const WarmArgs &A, const u32 j,
                                            const u32 novf_in, u32 *s_u) {
  constexpr int WT = SH::WT, PER = SH::PER, CAP = SH::CAP, NWAVE = SH::NWAVE;
  extern __shared__ u64 lds_dyn[];
  Lds<SH, EQ> s(lds_dyn);
  __shared__ u64 s_spl[SMAX + 1];
  __shared__ u32 s_segsrc[SMAX];
  __shared__ u32 s_segpre[SMAX + 1];
  __shared__ u32 s_oidx[LOVF];
  __shared__ u64 s_red[NWAVE];
  __shared__ u64 s_mm[2 * NWAVE];
  __shared__ u64 s_u64[2];

  const u32 t = threadIdx.x, lane = t & 63, w = t >> 6;
  const u32 S = A.S, W = A.W, nb = A.nb;
  WARM_STAMP(0);
  const u32 lo = j >= W ? j - W : 0u;
  const u32 hi = j + W < nb ? j + W : nb - 1;

  u32 c = 0, off = 0;

  if (MODE == 0) {
    // ---- 2. segment table and splitter window: one round trip ---------------
    if (t < S) {
      const int src = (int)j - (int)t + (int)W;
      u32 cnt = 0, start = 0;
      if (src >= 0 && src < (int)nb) {
        cnt = A.rd.segcnt[(u32)src * (S + 1) + t];
        start = A.rd.boff[src] + A.rd.segoff[(u32)src * (S + 1) + t];
      }
      s_segsrc[t] = start;
      s_segpre[t] = cnt;
    }
    for (u32 b = t; b <= hi - lo + 1; b += WT) s_spl[b] = A.rd.spl[lo + b];
    u32 novf_all = novf_in > WARM_OVF_CAP ? WARM_OVF_CAP : novf_in;
    for (u32 q = t; q < novf_all; q += WT)
      if (A.rd.odst[q] == j) {
        const u32 slot = atomicAdd(&s_u[3], 1u);
        if (slot < (u32)LOVF) s_oidx[slot] = q;
      }
    __syncthreads();
    const u32 nseg = block_excl_u32_inplace<NWAVE>(s_segpre, S, s_red);
    const u32 novf = s_u[3];
    c = nseg + novf;
    const bool ok = c <= (u32)CAP && novf <= (u32)LOVF;

    // ---- 3. gather: positions first, then every load of the thread --------
    u64 msum = 0;
    u64 kreg[PER];  // this thread's keys and indices, for the ranking
    i32 ireg[PER];
    if (ok) {
      u32 pos[PER];
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 e = i * WT + t;
        pos[i] = e < nseg ? seg_pos(s_segpre, s_segsrc, S, e)
                          : (e < c ? s_oidx[e - nseg] : 0u);
      }
      u64 k[PER];
      double v[PER], m[PER];
      i32 id[PER];
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 e = i * WT + t;
        if (e < nseg) {  // (separate branches: see k_warm_pipe's data_issue)
          k[i] = A.kin[pos[i]];
          v[i] = A.vin[pos[i]];
          id[i] = A.iin[pos[i]];
          if (!EQ) m[i] = A.min[pos[i]];
        } else if (e < c) {
          k[i] = A.rd.okey[pos[i]];
          v[i] = A.rd.ov[pos[i]];
          id[i] = A.rd.oid[pos[i]];
          if (!EQ) m[i] = A.rd.om[pos[i]];
        }
      }
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 e = i * WT + t;
        if (e < c) {
          u64 kk = k[i];
          if (A.pre != 0.0) kk = encode_pos(decode_pos(kk) + A.pre * v[i]);
          kreg[i] = kk;
          ireg[i] = id[i];
          s.key[e] = kk;
          s.v[e] = v[i];
          s.id[e] = id[i];
          if (!EQ) {
            s.m[e] = m[i];
            msum += mass_fix(m[i], A.fscale);
          }
        }
      }
    } else if (!EQ) {
      // over capacity: still publish the exact mass total
      for (u32 e = t; e < nseg; e += WT)
        msum += mass_fix(A.min[seg_pos(s_segpre, s_segsrc, S, e)], A.fscale);
      for (u32 q = t; q < novf_all; q += WT)
        if (A.rd.odst[q] == j) msum += mass_fix(A.rd.om[q], A.fscale);
    }
    u64 mtot = 0;
    if (!EQ) (void)block_excl_u64<NWAVE>(msum, s_red, &mtot);
    const u64 ep = (u64)(A.epoch & 0x3FFFFFFFu) << 34;
    if (t == 0) {
      publish(A.status + STATUS_WORDS * j, ep, j == 0 ? 2u : 1u, c, mtot, !EQ);
      if (!ok)
        set_err(A.err, A.epoch,
                novf > (u32)LOVF ? WARM_ERR_LOVF : WARM_ERR_CAPACITY);
    }
    WARM_STAMP(1);

    // ---- 4. sort the bucket in LDS, while the predecessors publish ---------
    if (ok) lds_bucket_sort_regs<SH, EQ>(s, c, s_red, s_mm,
                                         A.prof ? A.prof + 8ull * j : nullptr,
                                         kreg, ireg);
    else __syncthreads();
    WARM_STAMP(2);

    // ---- 5. look-back, one wave probing 64 predecessors at a time ----------
    if (w == 0) {
      u64 ecnt = 0, emass = 0;
      bool stalled = false;
      if (j > 0) {
        int64_t base = (int64_t)j - 1;  // lane l probes bucket base - l
        u32 spins = 0;
        while (true) {
          const int64_t pt = base - (int64_t)lane;
          u64 wd = 0, mv = 0;
          u32 flag = 2;  // buckets below 0: an inclusive prefix of zero
          if (pt >= 0) {
            const u64 *st = A.status + STATUS_WORDS * (u64)pt;
            const u32 e30 = A.epoch & 0x3FFFFFFFu;
            wd = ld_rlx(st);
            flag = (u32)(wd >> 34) == e30 ? (u32)(wd >> 32) & 3u : 0u;
            if (!EQ && flag != 0) {
              const u64 *mw = st + (flag == 2 ? 3 : 1);
              const u64 hw = ld_rlx(mw), lw = ld_rlx(mw + 1);
              const bool good = (u32)(hw >> 32) == ((e30 << 2) | 1u) &&
                                (u32)(lw >> 32) == ((e30 << 2) | 1u);
              if (good) mv = ((hw & 0xffffffffull) << 32) | (lw & 0xffffffffull);
              else flag = 0;  // mass not visible yet: retry
            }
          }
          const u64 incm = __ballot(flag == 2);
          const u64 zerom = __ballot(flag == 0);
          const u32 first = incm ? (u32)__ffsll((long long)incm) - 1u : 64u;
          const u64 upto = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
          if (zerom & upto) {
            // every predecessor publishes without waiting on anyone, so this
            // ends; the bound turns a broken invariant into a reported
            // failure instead of a hang
            if (++spins > (1u << 20)) {
              stalled = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          u64 cc = 0, mm = 0;
          if (lane <= first && pt >= 0) {
            cc = (u32)wd;
            mm = mv;
          }
          ecnt += wave_last(wave_incl_add(cc));
          if (!EQ) emass += wave_last(wave_incl_add(mm));
          if (first < 64) break;
          base -= 64;
        }
        if (lane == 0) {
          if (stalled) {
            set_err(A.err, A.epoch, WARM_ERR_STALL);
          } else {
            publish(A.status + STATUS_WORDS * j, ep, 2u, (u32)(ecnt + c),
                    emass + mtot, !EQ);
          }
        }
      }
      if (lane == 0) {
        s_u64[0] = ecnt;
        s_u64[1] = emass;
        s_u[2] = stalled ? 1u : 0u;
      }
    }
    __syncthreads();
    WARM_STAMP(3);
    off = (u32)s_u64[0];
    const u64 bfix = s_u64[1];
    if (!ok || s_u[2]) return;  // nothing of this bucket is written

    // ---- 6. force, kick, drift in sorted order -----------------------------------
    {
      const u32 r0 = t * PER;  // ranks r0 .. r0 + PER - 1, in order
      // the one splitter rank (a multiple of T >= PER) among them, if any
      const u32 sidx = (off + r0 + A.T - 1) / A.T, srank = sidx * A.T;
      u64 run = 0;
      if (!EQ) {
        u64 part = 0;
#pragma unroll
        for (int i = 0; i < PER; ++i)
          if (r0 + i < c) part += mass_fix(s.m[s.ord[r0 + i]], A.fscale);
        run = bfix + block_excl_u64<NWAVE>(part, s_red, nullptr);
      }
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 r = r0 + i;
        if (r < c) {
          const u32 e = s.ord[r];
          const u32 R = off + r;
          const u64 kold = s.key[e];
          double x = decode_pos(kold);
          double F;
          if (EQ) {
            F = (A.totmass - 2.0 * ((double)R * A.mq)) - A.mq;
          } else {
            const double me = s.m[e];
            F = (A.totmass - 2.0 * ((double)run * A.finv)) - me;
            run += mass_fix(me, A.fscale);
          }
          if (HARM) F = F - A.omega2 * x;
          const double a = (A.ext.kind ? ext_accel(A.ext, x, A.t) : 0.0) + F;
          const double v = s.v[e] + A.h * a;
          x = x + A.hd * v;
          s.v[e] = v;
          s.key[e] = encode_pos(x);
          if (R == srank) A.wr.spl[sidx] = R == 0 ? 0ull : kold;
        }
      }
      if (j == nb - 1 && t == 0) A.wr.spl[nb] = ~0ull;
    }
    __syncthreads();
    WARM_STAMP(4);
  } else {
    // MODE 1: bucket j is the sorted tile [j*T, min(n, (j+1)*T))
    off = j * A.T;
    c = min(A.T, A.n - off);
    for (u32 b = t; b <= hi - lo + 1; b += WT) s_spl[b] = A.rd.spl[lo + b];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const u32 e = i * WT + t;
      if (e < c) {
        const u64 k = A.kin[off + e];
        const double v = A.vin[off + e];
        const i32 id = A.iin[off + e];
        s.key[e] = k;
        s.v[e] = v;
        s.id[e] = id;
        if (!EQ) s.m[e] = A.min[off + e];
        s.ord[e] = (unsigned short)e;
      }
    }
    if (t == 0) {
      A.wr.spl[j] = A.rd.spl[j];
      if (j == nb - 1) A.wr.spl[nb] = A.rd.spl[nb];
    }
    __syncthreads();
  }

  // ---- 7. partition by next-step bucket and write out ---------------------------
  // one segment counter per destination; each sheet takes its slot with one
  // LDS atomic (the order within a segment does not matter: the next step
  // ranks by (position, index) and sums fixed-point masses)
  {
    u32 *wc = s.bins;  // [S + 1] segment counters, then starts
    if (t <= S) wc[t] = 0;
    __syncthreads();
    u32 dd[PER], rr[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const u32 r = i * WT + t;
      const bool valid = r < c;
      u32 d = S;
      rr[i] = 0;
      if (valid) {
        const u32 e = s.ord[r];
        u64 kp = s.key[e];
        if (A.hp != 0.0) kp = encode_pos(decode_pos(kp) + A.hp * s.v[e]);
        const bool below = lo > 0 && kp < s_spl[0];
        const bool above = hi + 1 < nb && kp >= s_spl[hi - lo + 1];
        if (!below && !above) {
          u32 a = 0, z = hi - lo;
          while (a < z) {
            const u32 mid = (a + z + 1) >> 1;
            if (s_spl[mid] <= kp) a = mid; else z = mid - 1;
          }
          d = lo + a + W - j;
        } else {
          const u32 dst = bucket_of(A.rd.spl, nb, kp);
          const u32 q = atomicAdd(A.ovf_w, 1u);
          if (q < WARM_OVF_CAP) {
            A.wr.okey[q] = s.key[e];
            A.wr.ov[q] = s.v[e];
            A.wr.oid[q] = s.id[e];
            A.wr.om[q] = EQ ? A.mq : s.m[e];
            A.wr.odst[q] = dst;
          } else {
            set_err(A.err, A.epoch, WARM_ERR_OVF);
          }
        }
        rr[i] = atomicAdd(&wc[d], 1u);
      }
      dd[i] = d;
    }
    __syncthreads();
    block_excl_u32_inplace<NWAVE>(wc, S + 1, s_red);
    if (t <= S) {
      const u32 nxt = t < S ? wc[t + 1] : c;
      A.wr.segoff[j * (S + 1) + t] = wc[t];
      A.wr.segcnt[j * (S + 1) + t] = nxt - wc[t];
    }
    if (t == 0) A.wr.boff[j] = off;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const u32 r = i * WT + t;
      if (r < c) {
        const u32 e = s.ord[r];
        const u32 p = off + wc[dd[i]] + rr[i];
        A.kout[p] = s.key[e];
        A.vout[p] = s.v[e];
        A.iout[p] = s.id[e];
        if (!EQ) A.mout[p] = s.m[e];
      }
    }
  }
  WARM_STAMP(5);
}

template <class SH, bool EQ, bool HARM, int MODE>
__global__ __launch_bounds__(SH::WT) void k_warm_step(WarmArgs A) {
  __shared__ u32 s_u[4];
  const u32 t = threadIdx.x;
  // ---- 1. ticket, error word and overflow count: one round trip ------------
  if (t == 0) {
    s_u[0] = atomicAdd(A.tick, 1u);
    s_u[1] = ld_u32(A.err);
    s_u[3] = 0;  // overflow entries destined to this bucket
  }
  const u32 novf_in = MODE == 0 ? ld_u32(A.ovf_r) : 0u;
  __syncthreads();
  const u32 j = s_u[0];
  if (s_u[1] != 0 && s_u[1] < A.epoch) return;  // an earlier launch failed
  if (j >= A.nb) {  // tickets start at 0 each launch; never expected
    if (t == 0) set_err(A.err, A.epoch, WARM_ERR_STALL);
    return;
  }
  if (j == 0 && t == 0) {  // clear the slots launch epoch+2 will use
    *A.tick_next2 = 0;
    *A.ovf_next2 = 0;
  }
  warm_bucket<SH, EQ, HARM, MODE>(A, j, novf_in, s_u);
}

// Gang form of mode 0: one launch runs one leapfrog step of several
// simulations, args[i] describing simulation i's step (its own epoch,
// look-back words, error word and overflow rings).  Gang tickets
// bstart[i] .. bstart[i+1]-1 are simulation i's buckets in order, so a
// bucket's look-back only ever waits on lower tickets, as in k_warm_step.
// nsim <= 64: each wave finds its simulation with one ballot over bstart.
template <class SH, bool EQ, bool HARM>
__global__ __launch_bounds__(SH::WT) void k_warm_gang(
    const WarmArgs *__restrict__ args, const u32 *__restrict__ bstart,
    u32 nsim, u32 *tick, u32 *tick_next2) {
  __shared__ u32 s_u[4];
  const u32 t = threadIdx.x, lane = t & 63;
  if (t == 0) s_u[0] = atomicAdd(tick, 1u);
  __syncthreads();
  const u32 g = s_u[0];
  if (g >= bstart[nsim]) return;  // the grid is the ticket count; never expected
  if (g == 0 && t == 0) *tick_next2 = 0;
  const u64 le = __ballot(lane < nsim && bstart[lane] <= g);
  const u32 i = __builtin_amdgcn_readfirstlane((u32)__popcll(le) - 1u);
  const WarmArgs &A = args[i];
  const u32 j = g - bstart[i];
  if (t == 0) {
    s_u[1] = ld_u32(A.err);
    s_u[3] = 0;
  }
  const u32 novf_in = ld_u32(A.ovf_r);
  __syncthreads();
  if (s_u[1] != 0 && s_u[1] < A.epoch) return;  // an earlier step failed
  if (j == 0 && t == 0) *A.ovf_next2 = 0;
  warm_bucket<SH, EQ, HARM, 0>(A, j, novf_in, s_u);
}

// LDS bytes of the pipelined kernel's tables for S segments, placed after
// the bucket arrays: two splitter windows (the current bucket's, read by its
// partition, and the next one's), then one segment table (S sources and
// the segments' start offsets) and LOVF overflow indices.  The splitter windows (S+1 entries) and the start
// offsets (S) are padded to pow2_ge(S + 1) entries with the type's maximum,
// so both are searched with fixed halving steps.  Only the splitter window
// is needed by two buckets at once: the next bucket's segment table and
// overflow indices are read when its loads are issued, before the current
// bucket's partition, and again (over capacity) at its own commit, before
// the following table overwrites them.
__host__ __device__ constexpr size_t pipe_table_bytes(u32 S) {
  return 16ull * pow2_ge(S + 1) + 4ull * (S + pow2_ge(S + 1) + LOVF);
}

// Persistent, software-pipelined form of mode 0.  One workgroup per CU loops
// over tickets; while it sorts, looks back and writes bucket j, the loads of
// the next bucket are in flight: its segment table and splitter window are
// requested right after bucket j is committed to LDS, and its sheets right
// after j's look-back, so they arrive during j's force and partition.  The
// arithmetic and the outputs are those of k_warm_step.  Each workgroup holds
// at most two unpublished tickets, and the lowest unpublished ticket always
// belongs to a workgroup whose own look-back waits only on lower, published
// buckets, so the chain always advances.
// EXT: the simulation has an external force (A.ext); without one the kernel
// carries none of its code, which keeps the register count down.
// A field of a kernel's argument block, loaded where it is used: the empty
// asm statement keeps the load from being hoisted to the kernel's entry, so
// a field that only rare paths read (overflow lists, error word, profile
// buffer) does not hold scalar registers for the whole kernel.  For a kernel
// whose only argument is a WarmArgs (it starts at offset 0 of the block).
typedef __attribute__((address_space(4))) const char karg_char;
template <class T>
__device__ __forceinline__ T karg_load(size_t off) {
  karg_char *p = (karg_char *)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return *(__attribute__((address_space(4))) const T *)(p + off);
}
#define KARG(f) karg_load<decltype(A.f)>(offsetof(WarmArgs, f))
#define PIPE_STAMP(k)                                                  \
  do {                                                                 \
    if (prof_on && t == 0) KARG(prof)[8ull * j + (k)] = wall_clock64(); \
  } while (0)

template <class SH, bool EQ, bool HARM, bool EXT>
__global__ __launch_bounds__(SH::WT)
__attribute__((amdgpu_waves_per_eu(EQ ? SH::OCC_EQ : 1))) void k_warm_pipe(WarmArgs A) {
  constexpr int WT = SH::WT, PER = SH::PER, CAP = SH::CAP, NWAVE = SH::NWAVE;
  extern __shared__ u64 lds_dyn[];
  Lds<SH, EQ> s(lds_dyn);
  const u32 S = A.S, W = A.W, nb = A.nb;
  const u32 TP = pow2_ge(S + 1);  // padded table length
  u64 *const s_spl0 = lds_dyn + Lds<SH, EQ>::BYTES / 8;  // [2][TP]
  u32 *const s_segsrc = (u32 *)(s_spl0 + 2 * TP);        // [S]
  u32 *const s_segpre = s_segsrc + S;                    // [TP]
  u32 *const s_oidx = s_segpre + TP;                     // [LOVF]
  auto s_spl = [&](int bb) { return s_spl0 + (u32)bb * TP; };
  const u32 PSEG = pow2_ge(S);  // the segment search's span
  __shared__ u32 s_novf;
  __shared__ u32 s_tot;
  __shared__ u64 s_red[NWAVE];
  __shared__ u64 s_mm[2 * NWAVE];
  __shared__ u32 s_u[4];
  __shared__ u64 s_u64[2];

  const u32 t = threadIdx.x, lane = t & 63, w = t >> 6;
  const bool prof_on = A.prof != nullptr;
  // padding of the start offsets (the scans write only [0, S))
  for (u32 q = S + t; q < TP; q += WT) s_segpre[q] = 0xffffffffu;
  if (t == 0) {
    s_u[0] = atomicAdd(A.tick, 1u);
    s_u[1] = ld_u32(KARG(err));
    s_novf = 0;
  }
  const u32 novf_in = ld_u32(A.ovf_r);
  const u32 novf_all = novf_in > WARM_OVF_CAP ? WARM_OVF_CAP : novf_in;
  __syncthreads();
  if (s_u[1] != 0 && s_u[1] < A.epoch) return;  // an earlier launch failed
  u32 j = s_u[0];
  if (j == 0 && t == 0) {  // clear the slots launch epoch+2 will use
    *A.tick_next2 = 0;
    *A.ovf_next2 = 0;
  }
  const u64 ep = (u64)(A.epoch & 0x3FFFFFFFu) << 34;

  // table registers of the bucket being prefetched
  u32 tb_cnt = 0, tb_start = 0;
  u64 tb_spl = 0;
  // requests bucket q's segment table and splitter window into registers
  auto table_issue = [&](u32 q) {
    const u32 lo = q >= W ? q - W : 0u, hi = q + W < nb ? q + W : nb - 1;
    tb_cnt = tb_start = 0;
    if (t < S) {
      const int src = (int)q - (int)t + (int)W;
      if (src >= 0 && src < (int)nb) {
        tb_cnt = A.rd.segcnt[(u32)src * (S + 1) + t];
        tb_start = A.rd.boff[src] + A.rd.segoff[(u32)src * (S + 1) + t];
      }
    }
    if (t <= hi - lo + 1) tb_spl = A.rd.spl[lo + t];
  };
  // stores the registers into the segment table and splitter window bb,
  // collects q's overflow entries and scans the segment counts; returns the
  // segment sheets
  auto table_commit = [&](u32 q, int bb) -> u32 {
    const u32 lo = q >= W ? q - W : 0u, hi = q + W < nb ? q + W : nb - 1;
    if (t < S) {
      s_segsrc[t] = tb_start;
      s_segpre[t] = tb_cnt;
    }
    if (t <= hi - lo + 1) s_spl(bb)[t] = tb_spl;
    else if (t < TP) s_spl(bb)[t] = ~0ull;  // padding past the window
    for (u32 q2 = t; q2 < novf_all; q2 += WT)
      if (KARG(rd.odst)[q2] == q) {
        const u32 slot = atomicAdd(&s_novf, 1u);
        if (slot < (u32)LOVF) s_oidx[slot] = q2;
      }
    if (S > 64) __syncthreads();  // the long scan reads other waves' counts
    return short_excl_u32_inplace<NWAVE>(s_segpre, S, s_red, &s_tot);
  };

  // data registers of the bucket being prefetched.  Segment sheets and far
  // movers load in separate branches: a per-lane choice between two of the
  // argument's pointers would make the compiler fetch the pointer itself
  // from the argument block before each load, and the wait for it (vector
  // loads retire in order) would serialise every earlier load of the
  // prefetch behind it.
  u64 dk[PER];
  double dv[PER], dm[PER];
  i32 di[PER];
  auto data_issue = [&](u32 nseg, u32 c) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const u32 e = i * WT + t;
      if (e < nseg) {
        const u32 d = pow2_search(s_segpre, PSEG, e);
        const u32 pos = s_segsrc[d] + (e - s_segpre[d]);
        dk[i] = A.kin[pos];
        dv[i] = A.vin[pos];
        di[i] = A.iin[pos];
        if (!EQ) dm[i] = A.min[pos];
      } else if (e < c) {
        const u32 q = s_oidx[e - nseg];
        dk[i] = KARG(rd.okey)[q];
        dv[i] = KARG(rd.ov)[q];
        di[i] = KARG(rd.oid)[q];
        if (!EQ) dm[i] = KARG(rd.om)[q];
      }
    }
  };

  int b = 0;
  u32 nseg = 0, c = 0;
  bool ok = false;
  if (j < nb) {  // prologue: bucket j's tables and loads
    table_issue(j);
    nseg = table_commit(j, b);
    const u32 novf = s_novf;
    c = nseg + novf;
    ok = c <= (u32)CAP && novf <= (u32)LOVF;
    // equal masses: the aggregate is the count alone, known from the table
    if (EQ && t == 0) publish(A.status + STATUS_WORDS * j, ep, j == 0 ? 2u : 1u, c, 0, false);
    if (ok) data_issue(nseg, c);
  }

  while (j < nb) {
    const u32 lo = j >= W ? j - W : 0u, hi = j + W < nb ? j + W : nb - 1;
    // ---- commit bucket j to LDS; publish its aggregate ----------------------
    PIPE_STAMP(0);
    if (t == 0) s_u[0] = atomicAdd(A.tick, 1u);  // the next ticket
    u64 msum = 0;
    if (ok) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 e = i * WT + t;
        if (e < c) {
          u64 kk = dk[i];
          if (A.pre != 0.0) kk = encode_pos(decode_pos(kk) + A.pre * dv[i]);
          dk[i] = kk;  // the ranking and the force read it from here
          // (equal masses: the ranking stores the keys in slot order itself,
          // and the drifted ones stay in the owning thread's registers)
          if (!EQ) s.key[e] = kk;
          // equal masses: only this thread's force reads its velocities,
          // from registers, and it stores the kicked ones here itself
          if (!EQ) s.v[e] = dv[i];
          s.id[e] = di[i];
          if (!EQ) {
            s.m[e] = dm[i];
            msum += mass_fix(dm[i], A.fscale);
          }
        }
      }
    } else if (!EQ) {
      // over capacity: still publish the exact mass total
      for (u32 e = t; e < nseg; e += WT)
        msum += mass_fix(A.min[seg_pos(s_segpre, s_segsrc, S, e)], A.fscale);
      for (u32 q2 = t; q2 < novf_all; q2 += WT)
        if (KARG(rd.odst)[q2] == j) msum += mass_fix(KARG(rd.om)[q2], A.fscale);
    }
    u64 mtot = 0;
    if (!EQ) (void)block_excl_u64<NWAVE>(msum, s_red, &mtot);
    if (t == 0) {
      // (equal masses published the count with the table)
      if (!EQ) publish(A.status + STATUS_WORDS * j, ep, j == 0 ? 2u : 1u, c, mtot, true);
      if (!ok)
        set_err(KARG(err), A.epoch,
                c - nseg > (u32)LOVF ? WARM_ERR_LOVF : WARM_ERR_CAPACITY);
    }
    __syncthreads();
    PIPE_STAMP(1);
    const u32 jn = s_u[0];
    const int bn = b ^ 1;
    if (t == 0) s_novf = 0;
    if (jn < nb) table_issue(jn);  // in flight during the sort

    // ---- sort, look-back ---------------------------------------------------------
    // equal masses need only each sheet's rank (the force is a function of
    // it); the sheets stay where their threads put them, and the force and
    // the partition read them there.  Unequal masses need the sorted order
    // for the mass prefix.
    u32 rk[PER];
    if (ok) {
      if constexpr (EQ)
        lds_bucket_rank_regs_scat<SH, EQ>(s, c, s_red, s_mm,
                                     prof_on ? KARG(prof) + 8ull * j : nullptr, rk,
                                     dk, di);
      else
        lds_bucket_sort_regs<SH, EQ>(s, c, s_red, s_mm,
                                     prof_on ? KARG(prof) + 8ull * j : nullptr, dk, di);
    } else {
      __syncthreads();
    }
    PIPE_STAMP(2);
    if (w == 0) {
      u64 ecnt = 0, emass = 0;
      bool stalled = false;
      if (j > 0) {
        int64_t base = (int64_t)j - 1;
        u32 spins = 0;
        while (true) {
          const int64_t pt = base - (int64_t)lane;
          u64 wd = 0, mv = 0;
          u32 flag = 2;
          if (pt >= 0) {
            const u64 *st = A.status + STATUS_WORDS * (u64)pt;
            const u32 e30 = A.epoch & 0x3FFFFFFFu;
            wd = ld_rlx(st);
            flag = (u32)(wd >> 34) == e30 ? (u32)(wd >> 32) & 3u : 0u;
            if (!EQ && flag != 0) {
              const u64 *mw = st + (flag == 2 ? 3 : 1);
              const u64 hw = ld_rlx(mw), lw = ld_rlx(mw + 1);
              const bool good = (u32)(hw >> 32) == ((e30 << 2) | 1u) &&
                                (u32)(lw >> 32) == ((e30 << 2) | 1u);
              if (good) mv = ((hw & 0xffffffffull) << 32) | (lw & 0xffffffffull);
              else flag = 0;
            }
          }
          const u64 incm = __ballot(flag == 2);
          const u64 zerom = __ballot(flag == 0);
          const u32 first = incm ? (u32)__ffsll((long long)incm) - 1u : 64u;
          const u64 upto = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
          if (zerom & upto) {
            if (++spins > (1u << 20)) {
              stalled = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          u64 cc = 0, mm = 0;
          if (lane <= first && pt >= 0) {
            cc = (u32)wd;
            mm = mv;
          }
          ecnt += wave_last(wave_incl_add(cc));
          if (!EQ) emass += wave_last(wave_incl_add(mm));
          if (first < 64) break;
          base -= 64;
        }
        if (lane == 0) {
          if (stalled) {
            set_err(KARG(err), A.epoch, WARM_ERR_STALL);
          } else {
            publish(A.status + STATUS_WORDS * j, ep, 2u, (u32)(ecnt + c),
                    emass + mtot, !EQ);
          }
        }
      }
      if (lane == 0) {
        s_u64[0] = ecnt;
        s_u64[1] = emass;
        s_u[2] = stalled ? 1u : 0u;
      }
    }
    __syncthreads();
    PIPE_STAMP(3);
    const u32 off = (u32)s_u64[0];
    const u64 bfix = s_u64[1];
    const bool write = ok && !s_u[2];

    // ---- next bucket's tables to LDS --------------------------------------------
    u32 nseg_n = 0, c_n = 0;
    bool ok_n = false;
    if (jn < nb) {
      nseg_n = table_commit(jn, bn);
      const u32 novf_n = s_novf;
      c_n = nseg_n + novf_n;
      ok_n = c_n <= (u32)CAP && novf_n <= (u32)LOVF;
      // equal masses: publish the next bucket's count now, half an iteration
      // before its commit, so that later buckets' look-backs wait less
      if (EQ && t == 0)
        publish(A.status + STATUS_WORDS * jn, ep, jn == 0 ? 2u : 1u, c_n, 0, false);
    }

    // equal masses: the drifted keys and velocities stay in the registers
    // of the thread that owns the sheet until the write
    u64 kn[PER];
    double vn[PER];
    if (write && EQ) {
      // ---- force, kick, drift of each thread's own sheets, by rank ------------------
      // (the look-back's barrier separates the ranking's reads from these
      // writes); the splitters are the sheets whose ranks are multiples of T
      const u32 sfirst = (off + A.T - 1) / A.T, send = off + c;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 e = i * WT + t;
        if (e < c) {
          const u32 R = off + rk[i];
          const u64 kold = dk[i];  // this thread's sheet: still in registers
          double x = decode_pos(kold);
          double F = (A.totmass - 2.0 * ((double)R * A.mq)) - A.mq;
          if (HARM) F = F - A.omega2 * x;
          const double a = (EXT ? ext_accel(A.ext, x, A.t) : 0.0) + F;
          const double v = dv[i] + A.h * a;
          x = x + A.hd * v;
          vn[i] = v;
          kn[i] = encode_pos(x);
          for (u32 si = sfirst; si * A.T < send; ++si)
            if (R == si * A.T) KARG(wr.spl)[si] = R == 0 ? 0ull : kold;
        }
      }
      if (j == nb - 1 && t == 0) KARG(wr.spl)[nb] = ~0ull;
      // (no barrier: the partition reads only this thread's sheets)
    } else if (write) {
      // ---- force, kick, drift in sorted order ------------------------------------
      {
        const u32 r0 = t * PER;
        // the one splitter rank (a multiple of T) this thread's PER <= T
        // consecutive ranks may contain: one division per thread
        const u32 sidx = (off + r0 + A.T - 1) / A.T, srank = sidx * A.T;
        u64 run = 0;
        if (!EQ) {
          u64 part = 0;
#pragma unroll
          for (int i = 0; i < PER; ++i)
            if (r0 + i < c) part += mass_fix(s.m[s.ord[r0 + i]], A.fscale);
          run = bfix + block_excl_u64<NWAVE>(part, s_red, nullptr);
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
          const u32 r = r0 + i;
          if (r < c) {
            const u32 e = s.ord[r];
            const u32 R = off + r;
            const u64 kold = s.key[e];
            double x = decode_pos(kold);
            double F;
            if (EQ) {
              F = (A.totmass - 2.0 * ((double)R * A.mq)) - A.mq;
            } else {
              const double me = s.m[e];
              F = (A.totmass - 2.0 * ((double)run * A.finv)) - me;
              run += mass_fix(me, A.fscale);
            }
            if (HARM) F = F - A.omega2 * x;
            const double a = (EXT ? ext_accel(A.ext, x, A.t) : 0.0) + F;
            const double v = s.v[e] + A.h * a;
            x = x + A.hd * v;
            s.v[e] = v;
            s.key[e] = encode_pos(x);
            if (R == srank) KARG(wr.spl)[sidx] = R == 0 ? 0ull : kold;
          }
        }
        if (j == nb - 1 && t == 0) KARG(wr.spl)[nb] = ~0ull;
      }
      __syncthreads();
    }
    // ---- the next bucket's loads, in flight during the partition --------------------
    if (ok_n) data_issue(nseg_n, c_n);
    PIPE_STAMP(4);
    if (write) {
      // ---- partition by next-step bucket and write out ---------------------------
      // segment counters [S + 1], one per workgroup: each sheet takes its
      // slot in its segment with one LDS atomic.  The order of sheets within
      // a segment does not matter (the next step ranks them by (position,
      // index) and sums fixed-point masses), only that the slots are
      // distinct.
      u32 *wc = s.bins;
      const u64 *spl = s_spl(b);
      // the window's ends, read once: the compiler must assume that the LDS
      // atomics below may write them, and would reload them for every sheet
      const u64 spl_first = spl[0], spl_end = spl[hi - lo + 1];
      if (t <= S) wc[t] = 0;
      __syncthreads();
      u32 dd[PER], rr[PER];
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 r = i * WT + t;
        const bool valid = r < c;
        u32 d = S;
        rr[i] = 0;
        if (valid) {
          const u32 e = EQ ? r : (u32)s.ord[r];
          u64 kp = EQ ? kn[i] : s.key[e];
          if (A.hp != 0.0) kp = encode_pos(decode_pos(kp) + A.hp * (EQ ? vn[i] : s.v[e]));
          const bool below = lo > 0 && kp < spl_first;
          const bool above = hi + 1 < nb && kp >= spl_end;
          if (!below && !above) {
            // the largest a with spl[a] <= kp: spl[0] <= kp < spl[hi-lo+1],
            // and the window is padded past hi-lo+1
            const u32 a = pow2_search(spl, pow2_ge(hi - lo + 1), kp);
            d = lo + a + W - j;
          } else {
            const u32 dst = bucket_of(A.rd.spl, nb, kp);
            const u32 q = atomicAdd(KARG(ovf_w), 1u);
            if (q < WARM_OVF_CAP) {
              KARG(wr.okey)[q] = EQ ? kn[i] : s.key[e];
              KARG(wr.ov)[q] = EQ ? vn[i] : s.v[e];
              KARG(wr.oid)[q] = s.id[e];
              KARG(wr.om)[q] = EQ ? A.mq : s.m[e];
              KARG(wr.odst)[q] = dst;
            } else {
              set_err(KARG(err), A.epoch, WARM_ERR_OVF);
            }
          }
          rr[i] = atomicAdd(&wc[d], 1u);
        }
        dd[i] = d;
      }
      __syncthreads();
      // segment starts (exclusive scan of the counts), then the tables
      (void)short_excl_u32_inplace<NWAVE>(wc, S + 1, s_red, &s_tot);
      if (t <= S) {
        const u32 nxt = t < S ? wc[t + 1] : c;
        A.wr.segoff[j * (S + 1) + t] = wc[t];
        A.wr.segcnt[j * (S + 1) + t] = nxt - wc[t];
      }
      if (t == 0) A.wr.boff[j] = off;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const u32 r = i * WT + t;
        if (r < c) {
          const u32 e = EQ ? r : (u32)s.ord[r];
          const u32 p = off + wc[dd[i]] + rr[i];
          A.kout[p] = EQ ? kn[i] : s.key[e];
          A.vout[p] = EQ ? vn[i] : s.v[e];
          A.iout[p] = s.id[e];
          if (!EQ) A.mout[p] = s.m[e];
        }
      }
    }
    __syncthreads();  // LDS of bucket j free before the next commit
    PIPE_STAMP(5);
    j = jn;
    b = bn;
    nseg = nseg_n;
    c = c_n;
    ok = ok_n;
  }
}

__global__ void k_warm_reset(u32 *ring, u32 keep, u32 *err) {
  const u32 t = threadIdx.x;
  // ring[0..3] tickets, ring[4..7] overflow counts; slot `keep` of the
  // overflow counts still describes the current state's overflow list
  if (t < 4) ring[t] = 0;
  if (t >= 4 && t < 8 && t - 4 != keep) ring[t] = 0;
  if (t < 2) err[t] = 0;
}

template <class SH, bool EQ, bool HARM, int MODE>
int launch_one(hipStream_t st, const WarmArgs &a) {
  auto fn = k_warm_step<SH, EQ, HARM, MODE>;
  const size_t lds = Lds<SH, EQ>::BYTES;
  if (lds > 65536)
    WDY_HIP(hipFuncSetAttribute((const void *)fn,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  fn<<<a.nb, SH::WT, lds, st>>>(a);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

// Workgroups of the persistent mode-0 kernel: as many as are resident at
// once (more would only find the tickets exhausted).
template <class SH, bool EQ, bool HARM, bool EXT>
int launch_pipe(hipStream_t st, const WarmArgs &a) {
  auto fn = k_warm_pipe<SH, EQ, HARM, EXT>;
  const size_t lds = Lds<SH, EQ>::BYTES + pipe_table_bytes(a.S);
  // per device: (LDS bytes << 32) | resident workgroups; 0: not computed yet
  static std::atomic<u64> resident[64];
  if (lds > 65536)
    WDY_HIP(hipFuncSetAttribute((const void *)fn,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  int dev = 0;
  WDY_HIP(hipGetDevice(&dev));
  const u64 cached = dev < 64 ? resident[dev].load() : 0ull;
  int res = (cached >> 32) == lds ? (int)(u32)cached : 0;
  if (res == 0) {
    int cus = 0, per = 0;
    WDY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    WDY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)fn,
                                                         SH::WT, lds));
    res = (per > 0 ? per : 1) * (cus > 0 ? cus : 1);
    if (dev < 64) resident[dev].store(((u64)lds << 32) | (u32)res);
  }
  const u32 grid = a.nb < (u32)res ? a.nb : (u32)res;
  fn<<<grid, SH::WT, lds, st>>>(a);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

template <class SH>
int launch_shape(hipStream_t st, const WarmArgs &a, bool eq, bool harm,
                 int mode) {
  if (mode == 2) {  // mode 0, persistent pipelined kernel
    if (eq)
      return a.ext.kind ? (harm ? launch_pipe<SH, true, true, true>(st, a)
                                : launch_pipe<SH, true, false, true>(st, a))
                        : (harm ? launch_pipe<SH, true, true, false>(st, a)
                                : launch_pipe<SH, true, false, false>(st, a));
    return a.ext.kind ? (harm ? launch_pipe<SH, false, true, true>(st, a)
                              : launch_pipe<SH, false, false, true>(st, a))
                      : (harm ? launch_pipe<SH, false, true, false>(st, a)
                              : launch_pipe<SH, false, false, false>(st, a));
  }
  if (mode == 0) {
    if (eq)
      return harm ? launch_one<SH, true, true, 0>(st, a)
                  : launch_one<SH, true, false, 0>(st, a);
    return harm ? launch_one<SH, false, true, 0>(st, a)
                : launch_one<SH, false, false, 0>(st, a);
  }
  if (eq)
    return harm ? launch_one<SH, true, true, 1>(st, a)
                : launch_one<SH, true, false, 1>(st, a);
  return harm ? launch_one<SH, false, true, 1>(st, a)
              : launch_one<SH, false, false, 1>(st, a);
}

template <class SH, bool EQ, bool HARM>
int launch_gang(hipStream_t st, const WarmArgs *args, const u32 *bstart,
                u32 nsim, u32 total, u32 *tick, u32 *tick_next2) {
  auto fn = k_warm_gang<SH, EQ, HARM>;
  const size_t lds = Lds<SH, EQ>::BYTES;
  if (lds > 65536)
    WDY_HIP(hipFuncSetAttribute((const void *)fn,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
  fn<<<total, SH::WT, lds, st>>>(args, bstart, nsim, tick, tick_next2);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

template <class SH>
int launch_gang_shape(hipStream_t st, bool eq, bool harm, const WarmArgs *args,
                      const u32 *bstart, u32 nsim, u32 total, u32 *tick,
                      u32 *tick_next2) {
  if (eq)
    return harm ? launch_gang<SH, true, true>(st, args, bstart, nsim, total, tick, tick_next2)
                : launch_gang<SH, true, false>(st, args, bstart, nsim, total, tick, tick_next2);
  return harm ? launch_gang<SH, false, true>(st, args, bstart, nsim, total, tick, tick_next2)
              : launch_gang<SH, false, false>(st, args, bstart, nsim, total, tick, tick_next2);
}

// Per-member reset of a gang interval (k_warm_reset's work for each member)
// and the gang's own ticket ring.
struct GangReset {
  u32 *ring;
  u32 *err;
  u32 keep;
  u32 pad;
};

// out[2 i], out[2 i + 1] := member i's error words (behind its interval)
__global__ void k_gang_errs(const GangReset *r, u32 count, u32 *out) {
  for (u32 i = threadIdx.x; i < count; i += blockDim.x) {
    out[2 * i] = r[i].err[0];
    out[2 * i + 1] = r[i].err[1];
  }
}

__global__ void k_gang_reset(const GangReset *r, u32 count, u32 *gring) {
  const u32 t = threadIdx.x;
  if (t < 4) gring[t] = 0;
  for (u32 i = t; i < count; i += blockDim.x) {
    u32 *ring = r[i].ring;
    for (u32 q = 0; q < 4; ++q) ring[q] = 0;
    for (u32 q = 0; q < 4; ++q)
      if (q != r[i].keep) ring[4 + q] = 0;
    r[i].err[0] = 0;
    r[i].err[1] = 0;
  }
}

}  // namespace

// Tile shapes (threads per workgroup, bucket capacity in sheets, 32-bit
// words of position bins).  Shapes 4-6 keep the LDS footprint small enough
// for two or three workgroups per CU, so one workgroup's memory phases
// overlap another's sort; shape 8 packs two 16-bit bins per word (2048 bins).
static const WarmShape kShapes[WARM_NSHAPES] = {
    {1024, 4096, 4096}, {512, 2048, 2048}, {256, 1024, 1024}, {512, 1024, 1024},
    {512, 2048, 1024}, {512, 3072, 1024}, {768, 3072, 768},
    {256, 2048, 1024}, {256, 1536, 1024}, {256, 1792, 1024}};

const WarmShape &warm_shape(int v) { return kShapes[v]; }

int warm_initial_shape(int64_t n, bool equal_mass) {
  const double lg = std::log2((double)(n > 1 ? n : 1));
  if (lg < 16.5) return 3;
  if (lg < 18.5) return 1;
  if (lg < 20.5) return equal_mass ? 5 : 6;
  if (lg < 21.5) return equal_mass ? 0 : 6;
  if (lg < 22.5) return equal_mass ? 0 : WARM_SHAPE_WIDE;
  if (lg < 23.5) return WARM_SHAPE_WIDE;
  return WARM_SHAPE_DEFAULT;
}

int warm_step_launch(hipStream_t st, const WarmArgs &a, bool equal_mass,
                     bool harm, int mode, int shape) {
  switch (shape) {
    case 0: return launch_shape<Shape<1024, 4096>>(st, a, equal_mass, harm, mode);
    case 1: return launch_shape<Shape<512, 2048>>(st, a, equal_mass, harm, mode);
    case 2: return launch_shape<Shape<256, 1024>>(st, a, equal_mass, harm, mode);
    case 3: return launch_shape<Shape<512, 1024>>(st, a, equal_mass, harm, mode);
    case 4: return launch_shape<Shape<512, 2048, 1024>>(st, a, equal_mass, harm, mode);
    case 5: return launch_shape<Shape<512, 3072, 1024>>(st, a, equal_mass, harm, mode);
    case 6: return launch_shape<Shape<768, 3072, 768>>(st, a, equal_mass, harm, mode);
    case 7: return launch_shape<Shape<256, 2048, 1024>>(st, a, equal_mass, harm, mode);
    case 8: return launch_shape<Shape<256, 1536, 2048, 4, true>>(st, a, equal_mass, harm, mode);
    default: return launch_shape<Shape<256, 1792, 1024>>(st, a, equal_mass, harm, mode);
  }
}

int warm_reset_launch(hipStream_t st, u32 *ring, u32 keep, u32 *err) {
  k_warm_reset<<<1, 64, 0, st>>>(ring, keep, err);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

bool warm_gang_shape(int shape) {
  return shape == WARM_SHAPE_DEFAULT || shape == WARM_SHAPE_WIDE || shape == 0;
}

static int warm_gang_launch(hipStream_t st, int shape, bool eq, bool harm,
                            const WarmArgs *args, const u32 *bstart, u32 nsim,
                            u32 total, u32 *tick, u32 *tick_next2) {
  switch (shape) {
    case WARM_SHAPE_DEFAULT:
      return launch_gang_shape<Shape<256, 1536, 2048, 4, true>>(
          st, eq, harm, args, bstart, nsim, total, tick, tick_next2);
    case WARM_SHAPE_WIDE:
      return launch_gang_shape<Shape<256, 2048, 1024>>(
          st, eq, harm, args, bstart, nsim, total, tick, tick_next2);
    case 0:
      return launch_gang_shape<Shape<1024, 4096>>(st, eq, harm, args, bstart,
                                                  nsim, total, tick, tick_next2);
    default:
      return fail(WDY_EUNSUPPORTED, "gang launch: shape %d not compiled", shape);
  }
}

}  // namespace wdy

// ============================================================================
// Host side: metadata allocation and interval orchestration
// ============================================================================
namespace wdy {

// Completion of one gang interval, shared by its members until each has
// checked its error words: `start` is recorded before the interval's first
// launch; the members' error words are gathered on the device, copied to
// h_err[2 i], h_err[2 i + 1] for member i, and `done` recorded behind the
// copy.  One event and one copy per interval, whatever the member count.
struct GangRecord {
  int device = 0;
  hipEvent_t start = nullptr, done = nullptr;
  u32 *h_err = nullptr;  // pinned [2 * WARM_GANG_MAX]
  u32 *d_err = nullptr;  // [2 * WARM_GANG_MAX]
  ~GangRecord() {
    (void)hipSetDevice(device);
    if (done) (void)hipEventSynchronize(done);
    if (start) (void)hipEventDestroy(start);
    if (done) (void)hipEventDestroy(done);
    if (h_err) (void)hipHostFree(h_err);
    if (d_err) (void)hipFree(d_err);
  }
};

struct WarmCtx {
  u32 T = 0, W = 0, S = 0, nb = 0, cap = 0;
  bool pipe = true;  // mode 0 by the persistent pipelined kernel
  bool fitted = false;  // window sized from the displacement estimate
  bool tuned = false;   // set by wdy_tune: no automatic sizing
  int shape = WARM_SHAPE_DEFAULT;
  u32 *ring = nullptr;   // [8] ticket ring, overflow-count ring
  size_t seg_alloc = 0, nb_alloc = 0;  // allocated entries
  WarmMeta meta[3];
  u64 *status = nullptr;
  u32 *err = nullptr;
  u32 *h_err = nullptr;  // pinned copy of err
  // the last warm interval, enqueued but not yet checked: its error word is
  // copied to h_err and ev_err marks the copy's completion; until checked,
  // the interval's start set and metadata (p_start, p_mstart) stay intact
  hipEvent_t ev_err = nullptr;
  // when the last interval ran in a gang: the gang interval's record (its
  // events and this member's error words at rec_idx) instead of ev_err/h_err
  std::shared_ptr<GangRecord> rec;
  int rec_idx = 0;
  bool pending = false;
  int p_start = 0, p_mstart = 0, p_nleap = 0;
  double p_h = 0.0, p_t0 = 0.0;
  u32 epoch = 0;
  bool ready = false;    // meta[rd] partitions the current state set
  double hp = 0.0;       // half drift the partition anticipates
  int rd = 0;
  double fscale = 1.0, finv = 1.0;
  // optional per-phase profile (environment variable WDY_PROFILE_WARM=1)
  u64 *prof = nullptr;
  size_t prof_alloc = 0;
  double prof_us[9] = {};  // gather, sort, look-back, force, write, span,
                           // sort: min/max+histogram, scan+scatter, fix-ups
  int64_t prof_n = 0;
  double clk_khz = 100000.0;
};

static void meta_free(WarmMeta &m) {
  void *ps[] = {m.segcnt, m.segoff, m.boff, m.spl,
                m.okey, m.ov, m.oid, m.om, m.odst};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  m = WarmMeta();
}

static int meta_alloc(WarmMeta &m, size_t nseg, size_t nbp1) {
  WDY_HIP(hipMalloc(&m.segcnt, nseg * 4));
  WDY_HIP(hipMalloc(&m.segoff, nseg * 4));
  WDY_HIP(hipMalloc(&m.boff, nbp1 * 4));
  WDY_HIP(hipMalloc(&m.spl, nbp1 * 8));
  WDY_HIP(hipMalloc(&m.okey, (size_t)WARM_OVF_CAP * 8));
  WDY_HIP(hipMalloc(&m.ov, (size_t)WARM_OVF_CAP * 8));
  WDY_HIP(hipMalloc(&m.oid, (size_t)WARM_OVF_CAP * 4));
  WDY_HIP(hipMalloc(&m.om, (size_t)WARM_OVF_CAP * 8));
  WDY_HIP(hipMalloc(&m.odst, (size_t)WARM_OVF_CAP * 4));
  return WDY_OK;
}

void warm_free(wdy_sim *s) {
  WarmCtx *w = s->warm;
  if (!w) return;
  if (w->prof) {
    if (w->prof_n > 0) {
      const double k = 1.0 / (double)w->prof_n;
      fprintf(stderr,
              "wendy_amd warm-step profile (n=%lld, %lld launches, mean us "
              "per bucket): %s %.2f sort %.2f look-back %.2f %s %.2f "
              "partition+write %.2f | launch span %.1f | sort = histogram "
              "%.2f + scan/scatter %.2f + ranks %.2f\n",
              (long long)s->n, (long long)w->prof_n,
              w->pipe ? "commit" : "gather", w->prof_us[0] * k,
              w->prof_us[1] * k, w->prof_us[2] * k,
              w->pipe ? "next-tables+force" : "force", w->prof_us[3] * k,
              w->prof_us[4] * k, w->prof_us[5] * k, w->prof_us[6] * k,
              w->prof_us[7] * k, w->prof_us[8] * k);
    }
    (void)hipFree(w->prof);
  }
  for (auto &m : w->meta) meta_free(m);
  if (w->status) (void)hipFree(w->status);
  if (w->err) (void)hipFree(w->err);
  if (w->ring) (void)hipFree(w->ring);
  if (w->h_err) (void)hipHostFree(w->h_err);
  if (w->ev_err) (void)hipEventDestroy(w->ev_err);
  delete w;
  s->warm = nullptr;
}

// (Re)sizes the metadata for tile shape `shape` (-1: keep), bucket target T
// (0: 3/4 of the shape's capacity) and window W.
int warm_configure(wdy_sim *s, int shape, u32 T, u32 W) {
  if (!s->warm) {
    s->warm = new (std::nothrow) WarmCtx();
    if (!s->warm) return fail(WDY_ENOMEM, "warm context allocation failed");
    WDY_HIP(hipMalloc(&s->warm->err, 8));
    WDY_HIP(hipMemsetAsync(s->warm->err, 0, 8, s->stream));
    WDY_HIP(hipMalloc(&s->warm->ring, 32));
    WDY_HIP(hipMemsetAsync(s->warm->ring, 0, 32, s->stream));
    WDY_HIP(hipHostMalloc((void **)&s->warm->h_err, 8, hipHostMallocDefault));
    WDY_HIP(hipEventCreateWithFlags(&s->warm->ev_err, hipEventDisableTiming));
    // fixed-point scale: 2^k * totmass <= 2^62
    int e = 0;
    (void)frexp(s->totmass > 0.0 ? s->totmass : 1.0, &e);
    s->warm->fscale = ldexp(1.0, 62 - e);
    s->warm->finv = ldexp(1.0, e - 62);
    // WDY_WARM_PIPE=0 selects the one-workgroup-per-bucket kernel instead
    const char *pp = getenv("WDY_WARM_PIPE");
    if (pp && pp[0] == '0') s->warm->pipe = false;
  }
  WarmCtx &w = *s->warm;
  const u32 n = (u32)s->n;
  if (shape >= 0) w.shape = shape;
  w.cap = (u32)warm_shape(w.shape).cap;
  if (T == 0) T = warm_default_target(w.cap);
  if (T > w.cap) T = w.cap;
  w.T = T;
  w.W = W;
  w.S = 2 * W + 1;
  w.nb = (n + T - 1) / T;
  const size_t nseg = (size_t)w.nb * (w.S + 1), nbp1 = (size_t)w.nb + 1;
  if (nseg > w.seg_alloc || nbp1 > w.nb_alloc) {
    for (auto &m : w.meta) meta_free(m);
    if (w.status) (void)hipFree(w.status);
    w.status = nullptr;
    for (auto &m : w.meta)
      if (int rc = meta_alloc(m, nseg, nbp1)) return rc;
    WDY_HIP(hipMalloc(&w.status, nbp1 * 5 * 8));
    WDY_HIP(hipMemsetAsync(w.status, 0, nbp1 * 5 * 8, s->stream));
    w.seg_alloc = nseg;
    w.nb_alloc = nbp1;
  }
  const char *pe = getenv("WDY_PROFILE_WARM");
  if (pe && pe[0] == '1' && nbp1 * 8 > w.prof_alloc) {
    if (w.prof) (void)hipFree(w.prof);
    WDY_HIP(hipMalloc(&w.prof, nbp1 * 8 * sizeof(u64)));
    w.prof_alloc = nbp1 * 8;
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                              s->device) == hipSuccess && khz > 0)
      w.clk_khz = khz;
  }
  w.ready = false;
  return WDY_OK;
}

// Folds the phase stamps of the last warm launch into the running means.
static void prof_fold(wdy_sim *s) {
  WarmCtx &w = *s->warm;
  std::vector<u64> p((size_t)w.nb * 8);
  if (hipMemcpy(p.data(), w.prof, p.size() * 8, hipMemcpyDeviceToHost) !=
      hipSuccess)
    return;
  const double us = 1000.0 / w.clk_khz;
  u64 t0 = ~0ull, t5 = 0;
  double d[8] = {};
  for (u32 j = 0; j < w.nb; ++j) {
    const u64 *q = &p[8ull * j];
    for (int k = 0; k < 5; ++k) d[k] += (double)(q[k + 1] - q[k]) * us;
    d[5] += (double)(q[6] - q[1]) * us;
    d[6] += (double)(q[7] - q[6]) * us;
    d[7] += (double)(q[2] - q[7]) * us;
    t0 = q[0] < t0 ? q[0] : t0;
    t5 = q[5] > t5 ? q[5] : t5;
  }
  for (int k = 0; k < 5; ++k) w.prof_us[k] += d[k] / w.nb;
  for (int k = 5; k < 8; ++k) w.prof_us[k + 1] += d[k] / w.nb;
  w.prof_us[5] += (double)(t5 - t0) * us;
  w.prof_n += 1;
}

u64 *warm_cold_splitters(wdy_sim *s, u32 *T, u32 *nb) {
  WarmCtx &w = *s->warm;
  *T = w.T;
  *nb = w.nb;
  return w.meta[(w.rd + 1) % 3].spl;
}

static WarmArgs base_args(wdy_sim *s) {
  WarmCtx &w = *s->warm;
  WarmArgs a;
  a.n = (u32)s->n;
  a.nb = w.nb;
  a.T = w.T;
  a.W = w.W;
  a.S = w.S;
  a.status = w.status;
  a.err = w.err;
  a.mq = s->m_equal;
  a.totmass = s->totmass;
  a.omega2 = s->omega2;
  a.fscale = w.fscale;
  a.finv = w.finv;
  a.h = a.pre = a.hd = a.hp = 0.0;
  a.ext = s->ext;
  a.t = s->t0;
  a.prof = nullptr;
  return a;
}

static void set_epoch(WarmArgs &a, WarmCtx &w, u32 e) {
  a.epoch = e;
  a.tick = w.ring + (e & 3);
  a.tick_next2 = w.ring + ((e + 2) & 3);
  a.ovf_r = w.ring + 4 + ((e + 3) & 3);
  a.ovf_w = w.ring + 4 + (e & 3);
  a.ovf_next2 = w.ring + 4 + ((e + 2) & 3);
}

static void set_state(WarmArgs &a, wdy_sim *s, int in, int out) {
  a.kin = s->key[in];
  a.vin = s->vel[in];
  a.iin = s->idx[in];
  a.min = s->mass[in];
  a.kout = s->key[out];
  a.vout = s->vel[out];
  a.iout = s->idx[out];
  a.mout = s->mass[out];
}

// After a cold (radix) step whose leap wrote the splitters into
// meta[(rd + 1) % 3]: partition the sorted current set into the segment
// layout the next warm step reads, anticipating a half drift hp.
int warm_partition_after_cold(wdy_sim *s, double hp) {
  WarmCtx &w = *s->warm;
  if (w.T > w.cap) return fail(WDY_EINVAL, "bucket target above capacity");
  const int x = (w.rd + 1) % 3, y = (w.rd + 2) % 3;
  if (int rc = warm_reset_launch(s->stream, w.ring, 4u, w.err)) return rc;
  const int in = s->cur, out = (s->cur + 1) % 3;
  WarmArgs a = base_args(s);
  set_state(a, s, in, out);
  a.rd = w.meta[x];
  a.wr = w.meta[y];
  set_epoch(a, w, ++w.epoch);
  a.hp = hp;
  if (s->timing) timer_begin(s->timer, s->stream, KC_WARM);
  int rc = warm_step_launch(s->stream, a, s->equal_mass, s->omega2 >= 0.0, 1,
                            w.shape);
  if (s->timing) timer_end(s->timer, s->stream, KC_WARM);
  if (rc) return rc;
  s->cur = out;
  w.rd = y;
  w.ready = true;
  w.hp = hp;
  return WDY_OK;
}

void warm_mark_tuned(wdy_sim *s) {
  if (s->warm) s->warm->tuned = true;
}

void warm_invalidate(wdy_sim *s) {
  if (s->warm) s->warm->ready = false;
}

int warm_fit(wdy_sim *s, double h) {
  WarmCtx &w = *s->warm;
  if (w.fitted || w.tuned) return WDY_OK;
  w.fitted = true;
  const double need = s->disp_rate * std::fabs(h);  // ranks per step, fast tail
  if (!(need > 0.0) || !std::isfinite(need)) return WDY_OK;
  // the current (default) shape while its window stays moderate (<= 48
  // buckets either side), then 2048-sheet buckets (<= 60), else the largest
  // buckets with the window they need.  Measured at ~4.5e4 ranks per step
  // (N = 2^24, equal masses): 1536-sheet buckets with window 36 ran 24%
  // faster than 4096-sheet ones; at ~7.7e4 (N = 2^27, unequal masses, bulk
  // flow) 2048-sheet buckets with window 48 ran 21% faster than 4096-sheet
  // ones with window 23, and 1536-sheet ones with window 64 10% faster.
  const int order[3] = {w.shape, WARM_SHAPE_WIDE, 0};
  const u32 lim[3] = {48u, 60u, ~0u};
  for (int k = 0; k < 3; ++k) {
    const int sh = order[k];
    if (k > 0 && warm_shape(sh).cap <= warm_shape(order[k - 1]).cap) continue;
    const u32 T = warm_default_target(warm_shape(sh).cap);
    const u32 W = (u32)std::ceil(need / (double)T) + 1u;
    if (W <= lim[k] || sh == 0) {
      const u32 wmax = (u32)warm_shape(sh).wmax();
      const u32 Wf = W < w.W ? w.W : (W < wmax ? W : wmax);
      if (sh == w.shape && Wf == w.W) return WDY_OK;
      return warm_configure(s, sh, 0u, Wf);
    }
  }
  return WDY_OK;
}

int warm_current_shape(wdy_sim *s) {
  return s->warm ? s->warm->shape : WARM_SHAPE_DEFAULT;
}

bool warm_config(wdy_sim *s, int *shape, u32 *T, u32 *W) {
  if (!s->warm) return false;
  *shape = s->warm->shape;
  *T = s->warm->T;
  *W = s->warm->W;
  return true;
}

bool warm_ready_for(wdy_sim *s, double h) {
  return s->warm && s->warm->ready && s->warm->hp == h / 2.;
}

// The buffer sets of one warm interval: it starts from state set `start`
// and metadata set `mstart` and writes only the two other sets of each
// (three rotate), so a failure found later can still rewind to its start.
struct IntervalPlan {
  int start = 0, mstart = 0, so[2] = {0, 0}, mo[2] = {0, 0};
  int in = 0, rdm = 0;  // the sets the next step reads
  double t = 0.0;       // force time of the next step (wendy/wendy.c:401-408)
};

static IntervalPlan plan_begin(wdy_sim *s) {
  IntervalPlan p;
  p.start = s->cur;
  p.mstart = s->warm->rd;
  int ns = 0, nm = 0;
  for (int i = 0; i < 3; ++i) {
    if (i != p.start) p.so[ns++] = i;
    if (i != p.mstart) p.mo[nm++] = i;
  }
  p.in = p.start;
  p.rdm = p.mstart;
  p.t = s->t0;
  return p;
}

// Arguments of step k (0 <= k < nleap) of the interval; advances the plan.
static WarmArgs plan_step(wdy_sim *s, IntervalPlan &p, int k, int nleap,
                          double h) {
  WarmCtx &w = *s->warm;
  const int out = p.so[k & 1], wrm = p.mo[k & 1];
  WarmArgs a = base_args(s);
  set_state(a, s, p.in, out);
  a.rd = w.meta[p.rdm];
  a.wr = w.meta[wrm];
  set_epoch(a, w, ++w.epoch);
  a.h = h;
  a.pre = k == 0 ? h / 2. : 0.0;
  a.hd = k == nleap - 1 ? h / 2. : h;
  a.hp = k == nleap - 1 ? h / 2. : 0.0;
  a.t = p.t;
  p.t = p.t + h;
  a.prof = w.prof;
  p.in = out;
  p.rdm = wrm;
  return a;
}

// After the interval's launches (and the copy of its error word to h_err,
// marked by ev_err, on the stream they ran on): the interval is pending and
// the sim's state is the one it leaves when it succeeds.
static void plan_end(wdy_sim *s, const IntervalPlan &p, double h, int nleap) {
  WarmCtx &w = *s->warm;
  w.pending = true;
  w.p_start = p.start;
  w.p_mstart = p.mstart;
  w.p_h = h;
  w.p_nleap = nleap;
  w.p_t0 = s->t0;
  s->cur = p.in;
  w.rd = p.rdm;
  if (s->ext.kind) s->t0 = p.t;
}

// Enqueues one output interval of nleap warm steps and returns without
// waiting for it: the device's error word is copied to the host behind the
// interval and checked by warm_settle.
int warm_interval(wdy_sim *s, double h, int nleap) {
  WarmCtx &w = *s->warm;
  if (w.pending) return fail(WDY_ERUNTIME, "warm interval enqueued before "
                             "the previous one was checked");
  IntervalPlan p = plan_begin(s);
  w.rec.reset();  // checked through ev_err / h_err
  // keep the overflow count of the list the first launch reads (written by
  // the launch with the current epoch)
  if (int rc = warm_reset_launch(s->stream, w.ring, w.epoch & 3u, w.err))
    return rc;
  for (int k = 0; k < nleap; ++k) {
    const WarmArgs a = plan_step(s, p, k, nleap, h);
    if (s->timing) timer_begin(s->timer, s->stream, KC_WARM);
    int rc = warm_step_launch(s->stream, a, s->equal_mass, s->omega2 >= 0.0,
                              w.pipe ? 2 : 0, w.shape);
    if (s->timing) timer_end(s->timer, s->stream, KC_WARM);
    if (rc) return rc;
  }
  WDY_HIP(hipMemcpyAsync(w.h_err, w.err, 8, hipMemcpyDeviceToHost, s->stream));
  WDY_HIP(hipEventRecord(w.ev_err, s->stream));
  plan_end(s, p, h, nleap);
  return WDY_OK;
}

// Waits for the warm interval in flight, if any, and checks its error word.
// On a device-side failure the state is rewound to the interval's start, the
// bucket target or window is adapted (or the warm path dropped), and
// *failed is set with the interval's h and nleap, which the caller must redo
// on the radix path.
int warm_settle(wdy_sim *s, bool *failed, double *h, int *nleap) {
  *failed = false;
  WarmCtx *wp = s->warm;
  if (!wp || !wp->pending) return WDY_OK;
  WarmCtx &w = *wp;
  u32 e0, kind;
  if (w.rec) {
    WDY_HIP(hipEventSynchronize(w.rec->done));
    e0 = w.rec->h_err[2 * w.rec_idx];
    kind = w.rec->h_err[2 * w.rec_idx + 1];
  } else {
    WDY_HIP(hipEventSynchronize(w.ev_err));
    e0 = w.h_err[0];
    kind = w.h_err[1];
  }
  w.pending = false;
  if (e0 == 0) {
    if (w.prof && !w.rec) prof_fold(s);
    s->warm_steps += w.p_nleap;
    return WDY_OK;
  }
  *failed = true;
  *h = w.p_h;
  *nleap = w.p_nleap;
  if (kind & WARM_ERR_CAPACITY) s->fail_capacity += 1;
  if (kind & (WARM_ERR_OVF | WARM_ERR_LOVF)) s->fail_overflow += 1;
  if (kind & WARM_ERR_STALL) s->fail_stall += 1;
  s->cur = w.p_start;
  s->t0 = w.p_t0;
  w.rd = w.p_mstart;
  w.ready = false;
  if (kind & WARM_ERR_STALL) {  // not a sizing problem: stop using it
    warm_free(s);
    return WDY_OK;
  }
  u32 T = w.T, W = w.W;
  int shape = w.shape;
  if (kind & WARM_ERR_CAPACITY) T = T * 15 / 16 < WARM_T_MIN ? WARM_T_MIN : T * 15 / 16;
  if (kind & (WARM_ERR_OVF | WARM_ERR_LOVF)) {
    // sheets cross more buckets per step than the window holds: widen it,
    // and once it is at its limit move to larger buckets (2048 sheets,
    // then 4096), as the window fitting does
    const u32 wmax = (u32)warm_shape(shape).wmax();
    if (W * 2 + 1 <= wmax) {
      W = W * 2 + 1;
    } else if (warm_shape(shape).cap < warm_shape(WARM_SHAPE_WIDE).cap) {
      shape = WARM_SHAPE_WIDE;
      T = 0;
      const u32 wm = (u32)warm_shape(shape).wmax();
      W = W < wm ? W : wm;
    } else if (warm_shape(shape).cap < warm_shape(0).cap) {
      shape = 0;
      T = 0;
      const u32 wm = (u32)warm_shape(shape).wmax();
      W = W < wm ? W : wm;
    } else {
      W = wmax;
    }
  }
  if (shape == w.shape && T == w.T && W == w.W) {  // limits reached: radix
    warm_free(s);                                  // path from now on
    return WDY_OK;
  }
  return warm_configure(s, shape, T, W);
}


// ---- gangs: one launch per leapfrog step for several simulations ------------
struct WarmGang {
  int device = 0;
  hipStream_t stream = nullptr;
  u32 *ring = nullptr;  // [4] ticket counters of the gang's launches
  u32 epoch = 0;
  unsigned char *h_buf = nullptr;  // pinned staging of one interval's tables
  unsigned char *d_buf = nullptr;
  size_t cap = 0;
  hipEvent_t ev_up = nullptr;  // the last upload from h_buf has completed
  // interval records; one is free again once no member refers to it
  std::vector<std::shared_ptr<GangRecord>> recs;
};

static int gang_record(WarmGang *g, std::shared_ptr<GangRecord> *out) {
  for (auto &r : g->recs)
    if (r.use_count() == 1) {
      *out = r;
      return WDY_OK;
    }
  auto r = std::make_shared<GangRecord>();
  r->device = g->device;
  WDY_HIP(hipEventCreate(&r->start));
  WDY_HIP(hipEventCreate(&r->done));
  WDY_HIP(hipHostMalloc((void **)&r->h_err, 8 * WARM_GANG_MAX, hipHostMallocDefault));
  WDY_HIP(hipMalloc(&r->d_err, 8 * WARM_GANG_MAX));
  g->recs.push_back(r);
  *out = r;
  return WDY_OK;
}

int warm_gang_create(int device, WarmGang **out) {
  WarmGang *g = new (std::nothrow) WarmGang();
  if (!g) return fail(WDY_ENOMEM, "gang allocation failed");
  g->device = device;
  int rc = [&]() -> int {
    WDY_HIP(hipSetDevice(device));
    // not synchronised with the null stream: allocations and uploads of
    // other simulations do not stall the gang's steps
    WDY_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    WDY_HIP(hipMalloc(&g->ring, 16));
    WDY_HIP(hipMemsetAsync(g->ring, 0, 16, g->stream));
    WDY_HIP(hipEventCreateWithFlags(&g->ev_up, hipEventDisableTiming));
    WDY_HIP(hipStreamSynchronize(g->stream));
    return WDY_OK;
  }();
  if (rc) {
    warm_gang_destroy(g);
    return rc;
  }
  *out = g;
  return WDY_OK;
}

void warm_gang_destroy(WarmGang *g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->ring) (void)hipFree(g->ring);
  if (g->d_buf) (void)hipFree(g->d_buf);
  if (g->h_buf) (void)hipHostFree(g->h_buf);
  if (g->ev_up) (void)hipEventDestroy(g->ev_up);
  g->recs.clear();  // records members still refer to live on with them
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
}

int warm_gang_device(const WarmGang *g) { return g->device; }

bool warm_gang_eligible(wdy_sim *s, double h) {
  return !s->small && !s->host_force && warm_ready_for(s, h) &&
         !s->warm->pending && warm_gang_shape(s->warm->shape);
}

void warm_wait(wdy_sim *s) {
  if (!s->warm || !s->warm->pending) return;
  if (s->warm->rec) (void)hipEventSynchronize(s->warm->rec->done);
  else (void)hipEventSynchronize(s->warm->ev_err);
}

void warm_gang_release(wdy_sim *s) {
  if (s->warm && !s->warm->pending) s->warm->rec.reset();
}

bool warm_gang_events(wdy_sim *s, hipEvent_t *start, hipEvent_t *done) {
  if (!s->warm || !s->warm->rec) return false;
  *start = s->warm->rec->start;
  *done = s->warm->rec->done;
  return true;
}

// Enqueues one output interval of every member (nleap[i] steps of h[i]) on
// the gang's stream: step k is one launch per (shape, mass kind, harmonic)
// group of the members that still have a step k, each bucket of each
// member one gang ticket.  Each member's arguments are those its own
// warm_interval would use (same epochs, buffer sets and clock), so its
// results are the same bit for bit; each member's error word is copied to
// its host word behind the interval and checked by its own warm_settle.
// Every member must satisfy warm_gang_eligible(h[i]) and be on the gang's
// device; the members' own streams must be idle.
int warm_gang_interval(WarmGang *g, wdy_sim *const *sims, int count,
                       const double *h, const int *nleap) {
  struct Member {
    wdy_sim *s;
    IntervalPlan p;
    int key;  // group: shape, mass kind, harmonic term
  };
  struct Launch {
    int key;
    size_t args_off, bst_off;
    u32 nsim, total;
  };
  std::vector<Member> m((size_t)count);
  int kmax = 0;
  for (int i = 0; i < count; ++i) {
    wdy_sim *s = sims[i];
    m[i].s = s;
    m[i].key = s->warm->shape * 4 + (s->equal_mass ? 2 : 0) + (s->omega2 >= 0.0 ? 1 : 0);
    kmax = nleap[i] > kmax ? nleap[i] : kmax;
  }
  std::vector<int> keys;
  for (const Member &q : m)
    if (std::find(keys.begin(), keys.end(), q.key) == keys.end()) keys.push_back(q.key);
  // resets first: they keep the overflow slot of each member's current epoch
  std::vector<GangReset> resets((size_t)count);
  for (int i = 0; i < count; ++i) {
    WarmCtx &w = *m[i].s->warm;
    resets[i].ring = w.ring;
    resets[i].err = w.err;
    resets[i].keep = w.epoch & 3u;
    resets[i].pad = 0;
    m[i].p = plan_begin(m[i].s);
  }
  std::vector<WarmArgs> args;
  std::vector<u32> bst;
  std::vector<Launch> launches;
  for (int k = 0; k < kmax; ++k) {
    for (int key : keys) {
      Launch L;
      L.key = key;
      L.args_off = args.size();
      L.bst_off = bst.size();
      L.nsim = 0;
      bst.push_back(0u);
      for (int i = 0; i < count; ++i) {
        if (m[i].key != key || nleap[i] <= k) continue;
        WarmArgs a = plan_step(m[i].s, m[i].p, k, nleap[i], h[i]);
        a.prof = nullptr;
        args.push_back(a);
        bst.push_back(bst.back() + a.nb);
        L.nsim += 1;
      }
      if (L.nsim == 0) {
        bst.pop_back();
        continue;
      }
      L.total = bst.back();
      launches.push_back(L);
    }
  }
  // one upload: resets, then every launch's arguments and ticket starts
  const size_t r_bytes = (resets.size() * sizeof(GangReset) + 255) & ~(size_t)255;
  const size_t a_bytes = (args.size() * sizeof(WarmArgs) + 255) & ~(size_t)255;
  const size_t bytes = r_bytes + a_bytes + bst.size() * sizeof(u32);
  WDY_HIP(hipSetDevice(g->device));
  WDY_HIP(hipEventSynchronize(g->ev_up));  // h_buf free again
  if (bytes > g->cap) {
    if (g->d_buf) (void)hipFree(g->d_buf);
    if (g->h_buf) (void)hipHostFree(g->h_buf);
    g->d_buf = g->h_buf = nullptr;
    g->cap = 0;
    const size_t cap = bytes * 2;
    WDY_HIP(hipMalloc(&g->d_buf, cap));
    WDY_HIP(hipHostMalloc((void **)&g->h_buf, cap, hipHostMallocDefault));
    g->cap = cap;
  }
  memcpy(g->h_buf, resets.data(), resets.size() * sizeof(GangReset));
  memcpy(g->h_buf + r_bytes, args.data(), args.size() * sizeof(WarmArgs));
  memcpy(g->h_buf + r_bytes + a_bytes, bst.data(), bst.size() * sizeof(u32));
  WDY_HIP(hipMemcpyAsync(g->d_buf, g->h_buf, bytes, hipMemcpyHostToDevice,
                         g->stream));
  WDY_HIP(hipEventRecord(g->ev_up, g->stream));
  const WarmArgs *d_args = (const WarmArgs *)(g->d_buf + r_bytes);
  const u32 *d_bst = (const u32 *)(g->d_buf + r_bytes + a_bytes);
  std::shared_ptr<GangRecord> rec;
  if (int rc = gang_record(g, &rec)) return rc;
  WDY_HIP(hipEventRecord(rec->start, g->stream));
  k_gang_reset<<<1, 256, 0, g->stream>>>((const GangReset *)g->d_buf,
                                         (u32)count, g->ring);
  WDY_CHECK_LAUNCH();
  for (const Launch &L : launches) {
    const u32 e = ++g->epoch;
    const int shape = L.key / 4;
    if (int rc = warm_gang_launch(g->stream, shape, (L.key & 2) != 0,
                                  (L.key & 1) != 0, d_args + L.args_off,
                                  d_bst + L.bst_off, L.nsim, L.total,
                                  g->ring + (e & 3u), g->ring + ((e + 2) & 3u)))
      return rc;
  }
  k_gang_errs<<<1, 64, 0, g->stream>>>((const GangReset *)g->d_buf, (u32)count,
                                       rec->d_err);
  WDY_CHECK_LAUNCH();
  WDY_HIP(hipMemcpyAsync(rec->h_err, rec->d_err, 8 * (size_t)count,
                         hipMemcpyDeviceToHost, g->stream));
  WDY_HIP(hipEventRecord(rec->done, g->stream));
  for (int i = 0; i < count; ++i) {
    wdy_sim *s = m[i].s;
    s->warm->rec = rec;
    s->warm->rec_idx = i;
    plan_end(s, m[i].p, h[i], nleap[i]);
  }
  return WDY_OK;
}

}  // namespace wdy
