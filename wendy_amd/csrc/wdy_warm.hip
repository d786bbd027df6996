// Warm-start leapfrog step of the large path: ONE kernel per leapfrog step.
//
// Sheets move only a few thousand ranks per step, so each step re-sorts by
// value buckets whose boundaries ("splitters") are the sorted positions at
// ranks k*T of two steps earlier.  The previous step left, for every source
// bucket t, its sheets partitioned into contiguous segments by destination
// bucket t + d - W (|d - W| <= W), plus one segment of far movers that are
// also copied into an overflow list (see WarmMeta).
//
// Workgroup j (one per bucket, ticketed in launch order):
//   1. gathers its incoming segments and overflow entries into LDS, applying
//      the interval's first half drift when asked (wendy/wendy.c:318-323,398);
//   2. publishes its sheet count and fixed-point mass total, then sorts the
//      bucket in LDS by (position, sheet index): one counting pass over 4096
//      position bins, then insertion fix-ups of the (tiny) bins;
//   3. resolves its global rank offset and mass prefix with a decoupled
//      look-back over integer values, so nothing depends on timing;
//   4. computes force, kick and drift in sorted order (wendy/wendy.c:371-383,
//      324-333) and emits the splitters for two steps ahead;
//   5. partitions the drifted sheets by next-step bucket (binary search in
//      the window of splitters) and writes them to its output region
//      [offset, offset + count) with the segment table.
// Mode 1 of the same kernel only does step 5, for an array already sorted
// by the radix path (tile j = ranks [j*T, (j+1)*T)).
//
// Masses enter the prefix as 64-bit fixed point: m_fix = llrint(m * 2^k)
// with 2^k * totmass <= 2^62, so bucket totals and prefixes are exact
// integers whatever order they are added in; C = prefix * 2^-k rounds once.
// Equal masses need no prefix: C = rank * m.
//
// Failure (a bucket over the LDS capacity, a full overflow list) records the
// failing launch's epoch in err[0]; later launches exit before taking a
// ticket, and the host redoes the interval from its untouched start state on
// the radix path.
#include "wdy_warm.h"

namespace wdy {
namespace {

constexpr int WT = WARM_THREADS;
constexpr int NWAVE = WT / 64;
constexpr int NQ = WARM_NQ;
constexpr int SMAX = 2 * WARM_WMAX + 1;
constexpr int LOVF = 1024;  // overflow entries one bucket may take

__device__ __forceinline__ u64 ld_acq(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 ld_rlx(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(u64 *p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(u64 *p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 ld_u32(const u32 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ void set_err(u32 *err, u32 epoch, u32 kind) {
  atomicCAS(err, 0u, epoch);
  atomicOr(err + 1, kind);
}

__device__ __forceinline__ u64 mass_fix(double m, double fscale) {
  return (u64)llrint(m * fscale);
}

// Exclusive scan of one u64 per thread over the block; *total gets the sum.
__device__ u64 block_excl_u64(u64 v, u64 *s_red, u64 *total) {
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  u64 x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const u64 y = __shfl_up(x, o, 64);
    if (lane >= (u32)o) x += y;
  }
  __syncthreads();
  if (lane == 63) s_red[w] = x;
  __syncthreads();
  u64 wp = 0, tot = 0;
  for (int i = 0; i < NWAVE; ++i) {
    if (i < (int)w) wp += s_red[i];
    tot += s_red[i];
  }
  if (total) *total = tot;
  return wp + x - v;
}

// Exclusive scan of s[0..n) (n <= WT) in place; returns the total.
__device__ u32 block_excl_u32_inplace(u32 *s, u32 n, u64 *s_red) {
  const u32 t = threadIdx.x;
  const u32 v = t < n ? s[t] : 0u;
  u64 tot;
  const u64 e = block_excl_u64(v, s_red, &tot);
  __syncthreads();
  if (t < n) s[t] = (u32)e;
  __syncthreads();
  return (u32)tot;
}

__device__ __forceinline__ bool pair_lt(u64 ka, i32 ia, u64 kb, i32 ib) {
  return ka < kb || (ka == kb && ia < ib);
}

template <bool EQ>
struct Lds {
  static constexpr int CAP = EQ ? WARM_CAP_EQ : WARM_CAP_NEQ;
  static constexpr size_t BYTES =
      (size_t)CAP * (EQ ? 22 : 30) + (size_t)NQ * 4;
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

// s.ord[0..c) := the permutation of 0..c-1 sorting (key, id) ascending.
template <bool EQ>
__device__ void lds_bucket_sort(Lds<EQ> &s, u32 c, u64 *s_red, u64 *s_mm) {
  constexpr int PER = (Lds<EQ>::CAP + WT - 1) / WT;
  const u32 t = threadIdx.x;
  u64 kmin = ~0ull, kmax = 0ull;
  for (u32 e = t; e < c; e += WT) {
    const u64 k = s.key[e];
    kmin = k < kmin ? k : kmin;
    kmax = k > kmax ? k : kmax;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const u64 a = __shfl_down(kmin, o, 64), b = __shfl_down(kmax, o, 64);
    kmin = a < kmin ? a : kmin;
    kmax = b > kmax ? b : kmax;
  }
  if ((t & 63) == 0) {
    s_mm[2 * (t >> 6)] = kmin;
    s_mm[2 * (t >> 6) + 1] = kmax;
  }
  for (u32 b = t; b < NQ; b += WT) s.bins[b] = 0;
  __syncthreads();
  kmin = ~0ull;
  kmax = 0ull;
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
    const u32 e = t + i * WT;
    q[i] = 0;
    if (e < c) {
      u32 b;
      if (dbl) {
        const double qd = (decode_pos(s.key[e]) - xmin) * scale;
        b = qd < (double)NQ ? (u32)qd : (u32)(NQ - 1);
      } else {
        b = (u32)((s.key[e] - kmin) >> kshift);
      }
      q[i] = b;
      atomicAdd(&s.bins[b], 1u);
    }
  }
  __syncthreads();
  {
    constexpr int BPT = NQ / WT;
    u32 loc[BPT];
    u64 sum = 0;
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      loc[i] = s.bins[t * BPT + i];
      sum += loc[i];
    }
    u64 run = block_excl_u64(sum, s_red, nullptr);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < BPT; ++i) {
      s.bins[t * BPT + i] = (u32)run;
      run += loc[i];
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const u32 e = t + i * WT;
    if (e < c) s.ord[atomicAdd(&s.bins[q[i]], 1u)] = (unsigned short)e;
  }
  __syncthreads();
  // bins[b] is now the end of bin b; insertion-sort bins holding >= 2 sheets
  for (u32 b = t; b < NQ; b += WT) {
    const u32 lo = b == 0 ? 0u : s.bins[b - 1];
    const u32 hi = s.bins[b];
    for (u32 a = lo + 1; a < hi; ++a) {
      const unsigned short ea = s.ord[a];
      const u64 ka = s.key[ea];
      const i32 ia = s.id[ea];
      u32 p = a;
      while (p > lo) {
        const unsigned short eb = s.ord[p - 1];
        if (!pair_lt(ka, ia, s.key[eb], s.id[eb])) break;
        s.ord[p] = eb;
        --p;
      }
      s.ord[p] = ea;
    }
  }
  __syncthreads();
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

template <bool EQ, bool HARM, int MODE>
__global__ __launch_bounds__(WT) void k_warm_step(WarmArgs A) {
  constexpr int CAP = Lds<EQ>::CAP;
  constexpr int PER = (CAP + WT - 1) / WT;
  extern __shared__ u64 lds_dyn[];
  Lds<EQ> s(lds_dyn);
  __shared__ u64 s_spl[SMAX + 1];
  __shared__ u32 s_segsrc[SMAX];
  __shared__ u32 s_segpre[SMAX + 1];
  __shared__ u32 s_oidx[LOVF];
  __shared__ u64 s_red[NWAVE];
  __shared__ u64 s_mm[2 * NWAVE];
  __shared__ u32 s_u[4];
  __shared__ u64 s_u64[2];

  const u32 t = threadIdx.x, lane = t & 63, w = t >> 6;
  const u32 S = A.S, W = A.W, nb = A.nb;
  {
    const u32 e0 = ld_u32(A.err);
    if (e0 != 0 && e0 < A.epoch) return;  // an earlier launch failed
  }
  if (t == 0) s_u[0] = atomicAdd(&A.wr.ctr[0], 1u);
  __syncthreads();
  const u32 j = s_u[0];
  if (j >= nb) {  // tickets start at 0 each launch; never expected
    if (t == 0) set_err(A.err, A.epoch, WARM_ERR_STALL);
    return;
  }

  u32 c = 0, off = 0;
  bool ok = true;

  if (MODE == 0) {
    // ---- 1. incoming segments and overflow entries --------------------------
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
    if (t == 0) s_u[1] = 0;
    __syncthreads();
    const u32 nseg = block_excl_u32_inplace(s_segpre, S, s_red);
    u32 novf_all = ld_u32(A.rd.ctr + 2);
    if (novf_all > WARM_OVF_CAP) novf_all = WARM_OVF_CAP;
    for (u32 q = t; q < novf_all; q += WT)
      if (A.rd.odst[q] == j) {
        const u32 slot = atomicAdd(&s_u[1], 1u);
        if (slot < (u32)LOVF) s_oidx[slot] = q;
      }
    __syncthreads();
    const u32 novf = s_u[1];
    c = nseg + novf;
    ok = c <= (u32)CAP && novf <= (u32)LOVF;

    // ---- 2. gather, with the optional first half drift -----------------------
    u64 msum = 0;
    if (ok) {
      for (u32 e = t; e < c; e += WT) {
        u64 k;
        double v, m = 0.0;
        i32 id;
        if (e < nseg) {
          const u32 p = seg_pos(s_segpre, s_segsrc, S, e);
          k = A.kin[p];
          v = A.vin[p];
          id = A.iin[p];
          if (!EQ) m = A.min[p];
        } else {
          const u32 q = s_oidx[e - nseg];
          k = A.rd.okey[q];
          v = A.rd.ov[q];
          id = A.rd.oid[q];
          if (!EQ) m = A.rd.om[q];
        }
        if (A.pre != 0.0) k = encode_pos(decode_pos(k) + A.pre * v);
        s.key[e] = k;
        s.v[e] = v;
        s.id[e] = id;
        if (!EQ) {
          s.m[e] = m;
          msum += mass_fix(m, A.fscale);
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
    if (!EQ) (void)block_excl_u64(msum, s_red, &mtot);

    // ---- 3. publish the aggregate --------------------------------------------
    const u64 ep = (u64)(A.epoch & 0x3FFFFFFFu) << 34;
    if (t == 0) {
      u64 *st = A.status + 3ull * j;
      if (j == 0) {
        st_rlx(st + 2, mtot);
        st_rel(st, ep | (2ull << 32) | c);
      } else {
        st_rlx(st + 1, mtot);
        st_rel(st, ep | (1ull << 32) | c);
      }
      if (!ok)
        set_err(A.err, A.epoch,
                novf > (u32)LOVF ? WARM_ERR_LOVF : WARM_ERR_CAPACITY);
    }

    // ---- 4. local sort, while the predecessors publish ------------------------
    if (ok) lds_bucket_sort<EQ>(s, c, s_red, s_mm);

    // ---- 5. look-back ------------------------------------------------------------
    if (t == 0) {
      u64 ecnt = 0, emass = 0;
      s_u[2] = 0;
      if (j > 0) {
        int64_t pt = (int64_t)j - 1;
        u32 spins = 0;
        while (true) {
          const u64 *st = A.status + 3ull * (u64)pt;
          const u64 wd = ld_acq(st);
          const u32 flag = (u32)(wd >> 32) & 3u;
          if ((u32)(wd >> 34) != (A.epoch & 0x3FFFFFFFu) || flag == 0) {
            // every predecessor publishes without waiting on anyone, so this
            // ends; the bound turns a broken invariant into a reported
            // failure instead of a hang
            if (++spins > (1u << 20)) {
              set_err(A.err, A.epoch, WARM_ERR_STALL);
              s_u[2] = 1;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          ecnt += (u32)wd;
          if (!EQ) emass += ld_rlx(st + (flag == 2 ? 2 : 1));
          if (flag == 2) break;
          --pt;
        }
        u64 *mine = A.status + 3ull * j;
        st_rlx(mine + 2, emass + mtot);
        st_rel(mine, ep | (2ull << 32) | (ecnt + c));
      }
      s_u64[0] = ecnt;
      s_u64[1] = emass;
    }
    __syncthreads();
    off = (u32)s_u64[0];
    const u64 bfix = s_u64[1];
    if (!ok || s_u[2]) goto done;  // nothing of this bucket is written

    // ---- 6. force, kick, drift in sorted order -----------------------------------
    {
      const u32 r0 = t * PER;
      u64 run = 0;
      if (!EQ) {
        u64 part = 0;
        for (int i = 0; i < PER; ++i)
          if (r0 + i < c) part += mass_fix(s.m[s.ord[r0 + i]], A.fscale);
        run = bfix + block_excl_u64(part, s_red, nullptr);
      }
      for (int i = 0; i < PER; ++i) {
        const u32 r = r0 + i;
        if (r >= c) break;
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
        const double a = 0.0 + F;
        const double v = s.v[e] + A.h * a;
        x = x + A.hd * v;
        s.v[e] = v;
        s.key[e] = encode_pos(x);
        if (R % A.T == 0) A.wr.spl[R / A.T] = R == 0 ? 0ull : kold;
      }
      if (j == nb - 1 && t == 0) A.wr.spl[nb] = ~0ull;
    }
    __syncthreads();
  } else {
    // MODE 1: bucket j is the sorted tile [j*T, min(n, (j+1)*T))
    off = j * A.T;
    c = min(A.T, A.n - off);
    for (u32 e = t; e < c; e += WT) {
      s.key[e] = A.kin[off + e];
      s.v[e] = A.vin[off + e];
      s.id[e] = A.iin[off + e];
      if (!EQ) s.m[e] = A.min[off + e];
      s.ord[e] = (unsigned short)e;
    }
    if (t == 0) {
      A.wr.spl[j] = A.rd.spl[j];
      if (j == nb - 1) A.wr.spl[nb] = A.rd.spl[nb];
    }
    __syncthreads();
  }

  // ---- 7. partition by next-step bucket and write out ---------------------------
  {
    const u32 lo = j >= W ? j - W : 0u;
    const u32 hi = j + W < nb ? j + W : nb - 1;
    for (u32 b = t; b <= hi - lo + 1; b += WT) s_spl[b] = A.rd.spl[lo + b];
    u32 *wc = s.bins;  // [NWAVE][S + 1] per-wave segment counters
    for (u32 b = t; b < NWAVE * (S + 1); b += WT) wc[b] = 0;
    __syncthreads();
    u32 dd[PER], rk[PER];
#pragma unroll
    for (int row = 0; row < PER; ++row) {
      const u32 r = row * WT + t;
      const bool valid = r < c;
      u32 d = S;
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
          const u32 q = atomicAdd(&A.wr.ctr[2], 1u);
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
      }
      // wave-level ranking on d (8 ballots cover d <= 255)
      u64 peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const u64 bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      const u32 below_l = (u32)__popcll(peers & ((1ull << lane) - 1ull));
      const u32 old = valid ? wc[w * (S + 1) + d] : 0u;
      if (valid && below_l == 0) wc[w * (S + 1) + d] = old + (u32)__popcll(peers);
      dd[row] = d;
      rk[row] = old + below_l;
    }
    __syncthreads();
    if (t <= S) {  // thread t owns segment t: per-wave prefixes and total
      u32 tot = 0;
      for (int i = 0; i < NWAVE; ++i) {
        const u32 x = wc[i * (S + 1) + t];
        wc[i * (S + 1) + t] = tot;
        tot += x;
      }
      s_segpre[t] = tot;
    }
    __syncthreads();
    block_excl_u32_inplace(s_segpre, S + 1, s_red);
    if (t <= S) {
      const u32 nxt = t < S ? s_segpre[t + 1] : c;
      A.wr.segoff[j * (S + 1) + t] = s_segpre[t];
      A.wr.segcnt[j * (S + 1) + t] = nxt - s_segpre[t];
    }
    if (t == 0) A.wr.boff[j] = off;
#pragma unroll
    for (int row = 0; row < PER; ++row) {
      const u32 r = row * WT + t;
      if (r < c) {
        const u32 e = s.ord[r];
        const u32 p = off + s_segpre[dd[row]] + wc[w * (S + 1) + dd[row]] + rk[row];
        A.kout[p] = s.key[e];
        A.vout[p] = s.v[e];
        A.iout[p] = s.id[e];
        if (!EQ) A.mout[p] = s.m[e];
      }
    }
  }

done:
  // ---- 8. the last workgroup resets the counters the next launches reuse ------
  __syncthreads();
  if (t == 0) {
    __threadfence();
    const u32 d = atomicAdd(&A.wr.ctr[1], 1u);
    if (d == nb - 1) {
      if (MODE == 0) {
        A.rd.ctr[0] = 0;
        A.rd.ctr[2] = 0;
      }
      A.wr.ctr[1] = 0;
    }
  }
}

__global__ void k_warm_reset(u32 *ctr_a, u32 *ctr_b, u32 *err) {
  const u32 t = threadIdx.x;
  if (t < 4) {
    ctr_a[t] = 0;
    ctr_b[t] = 0;
  }
  if (t < 2) err[t] = 0;
}

template <bool EQ, bool HARM, int MODE>
int launch_one(hipStream_t st, const WarmArgs &a) {
  auto fn = k_warm_step<EQ, HARM, MODE>;
  const size_t lds = Lds<EQ>::BYTES;
  static bool attr_set = false;
  if (!attr_set) {
    WDY_HIP(hipFuncSetAttribute((const void *)fn,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds));
    attr_set = true;
  }
  fn<<<a.nb, WT, lds, st>>>(a);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

}  // namespace

int warm_step_launch(hipStream_t st, const WarmArgs &a, bool equal_mass,
                     bool harm, int mode) {
  if (mode == 0) {
    if (equal_mass)
      return harm ? launch_one<true, true, 0>(st, a) : launch_one<true, false, 0>(st, a);
    return harm ? launch_one<false, true, 0>(st, a) : launch_one<false, false, 0>(st, a);
  }
  if (equal_mass)
    return harm ? launch_one<true, true, 1>(st, a) : launch_one<true, false, 1>(st, a);
  return harm ? launch_one<false, true, 1>(st, a) : launch_one<false, false, 1>(st, a);
}

int warm_reset_launch(hipStream_t st, u32 *ctr_a, u32 *ctr_b, u32 *err) {
  k_warm_reset<<<1, 64, 0, st>>>(ctr_a, ctr_b, err);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

u32 warm_capacity(bool equal_mass) {
  return equal_mass ? (u32)WARM_CAP_EQ : (u32)WARM_CAP_NEQ;
}

}  // namespace wdy

// ============================================================================
// Host side: metadata allocation and interval orchestration
// ============================================================================
namespace wdy {

struct WarmCtx {
  u32 T = 0, W = 0, S = 0, nb = 0, cap = 0;
  size_t seg_alloc = 0, nb_alloc = 0;  // allocated entries
  WarmMeta meta[3];
  u64 *status = nullptr;
  u32 *err = nullptr;
  u32 *h_err = nullptr;  // pinned copy of err
  u32 epoch = 0;
  bool ready = false;    // meta[rd] partitions the current state set
  double hp = 0.0;       // half drift the partition anticipates
  int rd = 0;
  double fscale = 1.0, finv = 1.0;
};

static void meta_free(WarmMeta &m) {
  void *ps[] = {m.segcnt, m.segoff, m.boff, m.spl, m.ctr,
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
  WDY_HIP(hipMalloc(&m.ctr, 16));
  WDY_HIP(hipMemset(m.ctr, 0, 16));
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
  for (auto &m : w->meta) meta_free(m);
  if (w->status) (void)hipFree(w->status);
  if (w->err) (void)hipFree(w->err);
  if (w->h_err) (void)hipHostFree(w->h_err);
  delete w;
  s->warm = nullptr;
}

// (Re)sizes the metadata for bucket target T and window W.
int warm_configure(wdy_sim *s, u32 T, u32 W) {
  if (!s->warm) {
    s->warm = new (std::nothrow) WarmCtx();
    if (!s->warm) return fail(WDY_ENOMEM, "warm context allocation failed");
    WDY_HIP(hipMalloc(&s->warm->err, 8));
    WDY_HIP(hipMemset(s->warm->err, 0, 8));
    WDY_HIP(hipHostMalloc((void **)&s->warm->h_err, 8, hipHostMallocDefault));
    // fixed-point scale: 2^k * totmass <= 2^62
    int e = 0;
    (void)frexp(s->totmass > 0.0 ? s->totmass : 1.0, &e);
    s->warm->fscale = ldexp(1.0, 62 - e);
    s->warm->finv = ldexp(1.0, e - 62);
  }
  WarmCtx &w = *s->warm;
  const u32 n = (u32)s->n;
  w.cap = s->equal_mass ? (u32)WARM_CAP_EQ : (u32)WARM_CAP_NEQ;
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
    WDY_HIP(hipMalloc(&w.status, nbp1 * 3 * 8));
    WDY_HIP(hipMemset(w.status, 0, nbp1 * 3 * 8));
    w.seg_alloc = nseg;
    w.nb_alloc = nbp1;
  }
  w.ready = false;
  return WDY_OK;
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
  return a;
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
  if (int rc = warm_reset_launch(s->stream, w.meta[x].ctr, w.meta[y].ctr, w.err))
    return rc;
  const int in = s->cur, out = (s->cur + 1) % 3;
  WarmArgs a = base_args(s);
  set_state(a, s, in, out);
  a.rd = w.meta[x];
  a.wr = w.meta[y];
  a.epoch = ++w.epoch;
  a.hp = hp;
  if (s->timing) timer_begin(s->timer, s->stream, KC_WARM);
  int rc = warm_step_launch(s->stream, a, s->equal_mass, s->omega2 >= 0.0, 1);
  if (s->timing) timer_end(s->timer, s->stream, KC_WARM);
  if (rc) return rc;
  s->cur = out;
  w.rd = y;
  w.ready = true;
  w.hp = hp;
  return WDY_OK;
}

bool warm_ready_for(wdy_sim *s, double h) {
  return s->warm && s->warm->ready && s->warm->hp == h / 2.;
}

// One output interval of nleap warm steps.  On a device-side failure the
// state is rewound to the interval's start (never written by the interval),
// the bucket target or window is adapted, and *failed is set.
int warm_interval(wdy_sim *s, double h, int nleap, bool *failed) {
  WarmCtx &w = *s->warm;
  *failed = false;
  const int start = s->cur, mstart = w.rd;
  int so[2], mo[2], ns = 0, nm = 0;
  for (int i = 0; i < 3; ++i) {
    if (i != start) so[ns++] = i;
    if (i != mstart) mo[nm++] = i;
  }
  if (int rc = warm_reset_launch(s->stream, w.meta[mo[0]].ctr,
                                 w.meta[mo[1]].ctr, w.err))
    return rc;
  int in = start, rdm = mstart;
  for (int k = 0; k < nleap; ++k) {
    const int out = so[k & 1], wrm = mo[k & 1];
    WarmArgs a = base_args(s);
    set_state(a, s, in, out);
    a.rd = w.meta[rdm];
    a.wr = w.meta[wrm];
    a.epoch = ++w.epoch;
    a.h = h;
    a.pre = k == 0 ? h / 2. : 0.0;
    a.hd = k == nleap - 1 ? h / 2. : h;
    a.hp = k == nleap - 1 ? h / 2. : 0.0;
    if (s->timing) timer_begin(s->timer, s->stream, KC_WARM);
    int rc = warm_step_launch(s->stream, a, s->equal_mass, s->omega2 >= 0.0, 0);
    if (s->timing) timer_end(s->timer, s->stream, KC_WARM);
    if (rc) return rc;
    in = out;
    rdm = wrm;
  }
  WDY_HIP(hipMemcpyAsync(w.h_err, w.err, 8, hipMemcpyDeviceToHost, s->stream));
  WDY_HIP(hipStreamSynchronize(s->stream));
  if (w.h_err[0] != 0) {
    const u32 kind = w.h_err[1];
    *failed = true;
    s->cur = start;
    w.rd = mstart;
    w.ready = false;
    if (kind & WARM_ERR_STALL) {  // not a sizing problem: stop using it
      warm_free(s);
      return WDY_OK;
    }
    u32 T = w.T, W = w.W;
    if (kind & WARM_ERR_CAPACITY) T = T * 3 / 4 < WARM_T_MIN ? WARM_T_MIN : T * 3 / 4;
    if (kind & (WARM_ERR_OVF | WARM_ERR_LOVF))
      W = W * 2 + 1 > (u32)WARM_WMAX ? (u32)WARM_WMAX : W * 2 + 1;
    if (T == w.T && W == w.W) {  // limits reached: radix path from now on
      warm_free(s);
      return WDY_OK;
    }
    if (int rc = warm_configure(s, T, W)) return rc;
    return WDY_OK;
  }
  s->cur = in;
  w.rd = rdm;
  return WDY_OK;
}

}  // namespace wdy
