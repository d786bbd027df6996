// Leapfrog step kernels for gfx950: drift, fused force + kick + drift, mass
// prefix, de-sort, and the energy/momentum/potential diagnostics.
//
// Arithmetic follows the reference's operation order exactly (this file is
// compiled with -ffp-contract=off, so a*b+c stays a rounded product followed
// by a rounded sum, as in the reference's x86-64 build without FMA):
//   drift   x = x + h_d * v                       (wendy/wendy.c:322,331)
//   force   F = ((totmass - 2*C) - m) [- omega2*x] (wendy/wendy.c:378,382-383)
//           a = 0.0 + F                           (a[] zeroed, then +=)
//           a = E(x, t) + F with an external force (a[] = E, then +=)
//   kick    v = v + h_k * a                       (wendy/wendy.c:330)
// where C is the exclusive prefix of masses in sorted order
// (wendy/wendy.c:359-360).  Equal-mass runs use C = rank * m, the exact
// value of the reference's running sum whenever m is a power of two.
#include <climits>
#include <cmath>

#include "wdy_internal.h"

namespace wdy {
namespace {

constexpr int BLOCK = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

inline unsigned grid_for(int64_t n, int block = BLOCK) {
  return (unsigned)((n + block - 1) / block);
}

__global__ __launch_bounds__(BLOCK) void k_drift(u32 n, u64 *__restrict__ key,
                                                 const double *__restrict__ vel,
                                                 double hd) {
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const double x = decode_pos(key[i]);
  key[i] = encode_pos(x + hd * vel[i]);
}

// Canonical exclusive mass prefix of the sheet at sorted slot i when its
// position is exactly tied with a neighbour: tied sheets are ordered by
// original index.  Runs are found by walking the sorted keys.
template <bool EQ>
__device__ double tied_prefix(const u64 *__restrict__ ks,
                              const u32 *__restrict__ perm,
                              const i32 *__restrict__ idx_old,
                              const double *__restrict__ mass_old,
                              const double *__restrict__ cum, u32 n, u32 i,
                              u64 k, i32 myid, double mq) {
  u32 s = i;
  while (s > 0 && ks[s - 1] == k) --s;
  u32 e = i + 1;
  while (e < n && ks[e] == k) ++e;
  if (EQ) {
    u32 c = 0;
    for (u32 j = s; j < e; ++j)
      if (j != i && idx_old[sorted_slot(perm, j, n)] < myid) ++c;
    return (double)(s + c) * mq;
  } else {
    double C = cum[s];
    if (e - s <= 64) {
      // add tied predecessors in increasing index order, like the
      // sequential prefix over the canonical order
      i32 last = INT_MIN;
      while (true) {
        i32 best = INT_MAX;
        u32 bj = s;
        for (u32 j = s; j < e; ++j) {
          const i32 id = idx_old[sorted_slot(perm, j, n)];
          if (id > last && id < myid && id < best) {
            best = id;
            bj = j;
          }
        }
        if (best == INT_MAX) break;
        C = C + mass_old[sorted_slot(perm, bj, n)];
        last = best;
      }
    } else {
      for (u32 j = s; j < e; ++j)
        if (idx_old[sorted_slot(perm, j, n)] < myid)
          C = C + mass_old[sorted_slot(perm, j, n)];
    }
    return C;
  }
}

// One leapfrog force + kick + drift over the freshly sorted state.
//   ks[i]     sorted key of slot i;  perm[i] its slot before the sort
//   *_old     payload in pre-sort order;  *_new payload written in sorted order
//   knew[i]   drifted key (ks is not modified: neighbours read it for ties)
template <bool EQ, bool HARM>
__global__ __launch_bounds__(BLOCK) void k_leap(
    u32 n, const u64 *__restrict__ ks, const u32 *__restrict__ perm,
    const double *__restrict__ vel_old, const i32 *__restrict__ idx_old,
    const double *__restrict__ mass_old, const double *__restrict__ cum,
    u64 *__restrict__ knew, double *__restrict__ vel_new,
    i32 *__restrict__ idx_new, double *__restrict__ mass_new,
    double *__restrict__ acc_out, double *__restrict__ cum_out,
    u64 *__restrict__ spl_out, u32 T, u32 nb, double mq, double totmass,
    double omega2, double hk, double hd, ExtForce ext, double t) {
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const u64 k = ks[i];
  const u32 src = sorted_slot(perm, i, n);
  double x = decode_pos(k);
  double v = vel_old[src];
  const i32 id = idx_old[src];
  const double m = EQ ? mq : mass_old[src];
  const bool tied = (i > 0 && ks[i - 1] == k) || (i + 1 < n && ks[i + 1] == k);
  double C;
  if (!tied)
    C = EQ ? (double)i * mq : cum[i];
  else
    C = tied_prefix<EQ>(ks, perm, idx_old, mass_old, cum, n, i, k, id, mq);
  double F = (totmass - 2.0 * C) - m;
  if (HARM) F = F - omega2 * x;
  const double a = (ext.kind ? ext_accel_id(ext, x, t, id) : 0.0) + F;
  v = v + hk * a;
  x = x + hd * v;
  knew[i] = encode_pos(x);
  vel_new[i] = v;
  idx_new[i] = id;
  if (!EQ) mass_new[i] = m;
  if (acc_out) acc_out[i] = a;
  if (cum_out) cum_out[i] = C;
  if (spl_out) {  // sorted (pre-drift) keys at ranks k*T: warm-step splitters
    if (i % T == 0) spl_out[i / T] = i == 0 ? 0ull : k;
    if (i == n - 1) spl_out[nb] = ~0ull;
  }
}

// Cold-sort layout builder.  ks[i] sorted keys, perm[i] their staged slot,
// id_slot[] staged sheet indices; v_id/m_id staged per sheet index.
// Large mode gathers velocity and mass into sorted order; small mode keeps
// them in sheet order (copied by the host).
__global__ __launch_bounds__(BLOCK) void k_build_state(
    u32 n, const u64 *__restrict__ ks, const u32 *__restrict__ perm,
    const i32 *__restrict__ id_slot, const double *__restrict__ v_id,
    const double *__restrict__ m_id, u64 *__restrict__ key,
    i32 *__restrict__ idx, double *__restrict__ vel,
    double *__restrict__ mass) {
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const i32 id = id_slot[sorted_slot(perm, i, n)];
  key[i] = ks[i];
  idx[i] = id;
  if (vel) vel[i] = v_id[id];
  if (mass) mass[i] = m_id[id];
}

// ---- exclusive FP64 mass prefix: reduce, scan tile sums, apply ------------
// Element i of the sequence is mass[perm[i]] (or mass[i] when perm is null).
__device__ __forceinline__ double seq_elem(const double *mass, const u32 *perm,
                                           u32 i, u32 n) {
  return perm ? mass[sorted_slot(perm, i, n)] : mass[i];
}

__device__ double block_excl_scan_f64(double v, double *s_w, double *total) {
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double y = __shfl_up(x, o, 64);
    if (lane >= (u32)o) x = x + y;
  }
  if (lane == 63) s_w[w] = x;
  __syncthreads();
  double wp = 0.0, tot = 0.0;
  for (int q = 0; q < BLOCK / 64; ++q) {
    if (q < (int)w) wp = wp + s_w[q];
    tot = tot + s_w[q];
  }
  __syncthreads();
  *total = tot;
  return wp + (x - v);
}

__global__ __launch_bounds__(BLOCK) void k_scan_reduce(
    u32 n, const double *__restrict__ mass, const u32 *__restrict__ perm,
    double *__restrict__ tile_sums) {
  __shared__ double s_w[BLOCK / 64];
  const u32 base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  double acc = 0.0;
  for (int j = 0; j < SCAN_ITEMS; ++j)
    if (base + j < n) acc = acc + seq_elem(mass, perm, base + j, n);
  double tot;
  (void)block_excl_scan_f64(acc, s_w, &tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Exclusive scan of the tile sums, in place, by one block.
__global__ __launch_bounds__(BLOCK) void k_scan_top(u32 ntiles,
                                                    double *__restrict__ ts) {
  __shared__ double s_w[BLOCK / 64];
  const u32 per = (ntiles + BLOCK - 1) / BLOCK;
  const u32 b = threadIdx.x * per;
  double acc = 0.0;
  for (u32 j = 0; j < per; ++j)
    if (b + j < ntiles) acc = acc + ts[b + j];
  double tot;
  double run = block_excl_scan_f64(acc, s_w, &tot);
  for (u32 j = 0; j < per; ++j)
    if (b + j < ntiles) {
      const double t = ts[b + j];
      ts[b + j] = run;
      run = run + t;
    }
}

__global__ __launch_bounds__(BLOCK) void k_scan_apply(
    u32 n, const double *__restrict__ mass, const u32 *__restrict__ perm,
    const double *__restrict__ tile_off, double *__restrict__ cum) {
  __shared__ double s_w[BLOCK / 64];
  const u32 base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  double vals[SCAN_ITEMS];
  double acc = 0.0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    vals[j] = (base + j < n) ? seq_elem(mass, perm, base + j, n) : 0.0;
    acc = acc + vals[j];
  }
  double tot;
  double run = block_excl_scan_f64(acc, s_w, &tot);
  run = tile_off[blockIdx.x] + run;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j)
    if (base + j < n) {
      cum[base + j] = run;
      run = run + vals[j];
    }
}

__global__ __launch_bounds__(BLOCK) void k_desort(
    u32 n, const u64 *__restrict__ key, const double *__restrict__ vel,
    const i32 *__restrict__ idx, const double *__restrict__ acc,
    const double *__restrict__ mass, double *__restrict__ xo,
    double *__restrict__ vo, double *__restrict__ ao,
    double *__restrict__ mo) {
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const i32 j = idx[i];
  xo[j] = decode_pos(key[i]);
  if (vo) vo[j] = vel[i];
  if (ao) ao[j] = acc[i];
  if (mo) mo[j] = mass[i];
}

// ---- diagnostics -----------------------------------------------------------

__device__ double block_sum_f64(double v, double *s_w) {
  const u32 lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double x = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x = x + __shfl_down(x, o, 64);
  if (lane == 0) s_w[w] = x;
  __syncthreads();
  double tot = 0.0;
  if (threadIdx.x == 0)
    for (int q = 0; q < BLOCK / 64; ++q) tot = tot + s_w[q];
  __syncthreads();
  return tot;  // valid in thread 0
}

// Per-sheet energy terms on a sorted state, one partial sum per block:
//   W_i = twopiG * m_i * x_i * (2 C_i + m_i - M)   (sum_i W_i equals the
//         reference's twopiG * sum m_(i) (M_below x_(i) - XM_below))
//   K_i = m_i v_i^2 / 2,  H_i = m_i omega^2 x_i^2 / 2
// mass[] holds mscale * m (mscale = twopiG for sim state, 1 for raw arrays).
__global__ __launch_bounds__(BLOCK) void k_energy_terms(
    u32 n, const u64 *__restrict__ key, const double *__restrict__ vel,
    const double *__restrict__ mass, const double *__restrict__ cum,
    double mq, double mtot, double mscale, double twopiG, double omega2,
    double *__restrict__ partial) {
  __shared__ double s_w[BLOCK / 64];
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  double e = 0.0;
  if (i < n) {
    const double x = decode_pos(key[i]);
    const double v = vel[i];
    const double mr = mass ? mass[i] / mscale : mq / mscale;
    const double C = (cum ? cum[i] : (double)i * mq) / mscale;
    // unequal masses: the total is the last prefix plus the last mass
    const double M = (cum ? cum[n - 1] + mass[n - 1] : mtot) / mscale;
    e = twopiG * mr * x * ((2.0 * C + mr) - M);
    e = e + mr * (v * v) / 2.0;
    if (omega2 >= 0.0) e = e + mr * omega2 * (x * x) / 2.0;
  }
  const double t = block_sum_f64(e, s_w);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(BLOCK) void k_momentum_terms(
    u32 n, const double *__restrict__ v, const double *__restrict__ m,
    double *__restrict__ partial) {
  __shared__ double s_w[BLOCK / 64];
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  const double e = i < n ? m[i] * v[i] : 0.0;
  const double t = block_sum_f64(e, s_w);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// Deterministic final sum of block partials (fixed order).
// out[0] = sum of the partials; NaN when *err (a sort the sum depends on
// gave up) is set.
__global__ __launch_bounds__(BLOCK) void k_sum_partials(
    u32 np, const double *__restrict__ partial, double *__restrict__ out,
    const u32 *__restrict__ err) {
  __shared__ double s_w[BLOCK / 64];
  double acc = 0.0;
  for (u32 j = threadIdx.x; j < np; j += BLOCK) acc = acc + partial[j];
  const double t = block_sum_f64(acc, s_w);
  if (threadIdx.x == 0) out[0] = (err && *err) ? __builtin_nan("") : t;
}

// out[k] = [omega2 * y_k^2 / 2] + twopiG * sum_j m_j |x_j - y_k|
// (wendy/wendy.py:512-517).  One block per query point.
__global__ __launch_bounds__(BLOCK) void k_potential(
    u32 n, const double *__restrict__ x, const double *__restrict__ m,
    const double *__restrict__ y, double twopiG, double omega2,
    double *__restrict__ out) {
  __shared__ double s_w[BLOCK / 64];
  const double yk = y[blockIdx.x];
  double acc = 0.0;
  for (u32 j = threadIdx.x; j < n; j += BLOCK) acc = acc + m[j] * fabs(x[j] - yk);
  const double t = block_sum_f64(acc, s_w);
  if (threadIdx.x == 0) {
    const double o = omega2 >= 0.0 ? omega2 * (yk * yk) / 2.0 : 0.0;
    out[blockIdx.x] = o + twopiG * t;
  }
}

// E_i = [m_i omega^2 x_i^2 / 2] + twopiG m_i sum_j m_j |x_i - x_j|
//       + m_i v_i^2 / 2   (wendy/wendy.py:458-466).  One block per sheet.
__global__ __launch_bounds__(BLOCK) void k_energy_individual(
    u32 n, const double *__restrict__ x, const double *__restrict__ v,
    const double *__restrict__ m, double twopiG, double omega2,
    double *__restrict__ out) {
  __shared__ double s_w[BLOCK / 64];
  const u32 i = blockIdx.x;
  const double xi = x[i];
  double acc = 0.0;
  for (u32 j = threadIdx.x; j < n; j += BLOCK) acc = acc + m[j] * fabs(x[j] - xi);
  const double t = block_sum_f64(acc, s_w);
  if (threadIdx.x == 0) {
    const double o = omega2 >= 0.0 ? m[i] * omega2 * (xi * xi) / 2.0 : 0.0;
    out[i] = (o + twopiG * m[i] * t) + m[i] * (v[i] * v[i]) / 2.0;
  }
}

__global__ __launch_bounds__(BLOCK) void k_encode(u32 n,
                                                  const double *__restrict__ x,
                                                  u64 *__restrict__ key) {
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) key[i] = encode_pos(x[i]);
}

__global__ __launch_bounds__(BLOCK) void k_gather2(
    u32 n, const u32 *__restrict__ perm, const double *__restrict__ a,
    const double *__restrict__ b, double *__restrict__ ao,
    double *__restrict__ bo) {
  const u32 i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const u32 s = sorted_slot(perm, i, n);
  ao[i] = a[s];
  if (b) bo[i] = b[s];
}

int mass_prefix(wdy_sim *s, const double *mass, const u32 *perm, double *cum) {
  const u32 n = (u32)s->n;
  const u32 nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (s->timing) timer_begin(s->timer, s->stream, KC_SCAN);
  k_scan_reduce<<<nt, BLOCK, 0, s->stream>>>(n, mass, perm, s->tile_sums);
  k_scan_top<<<1, BLOCK, 0, s->stream>>>(nt, s->tile_sums);
  k_scan_apply<<<nt, BLOCK, 0, s->stream>>>(n, mass, perm, s->tile_sums, cum);
  if (s->timing) timer_end(s->timer, s->stream, KC_SCAN);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

struct DevBufs {
  void *p[8] = {};
  int n = 0;
  ~DevBufs() {
    for (int i = 0; i < n; ++i)
      if (p[i]) (void)hipFree(p[i]);
  }
  template <class T>
  hipError_t alloc(T **out, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T) + 16);
    if (e == hipSuccess) {
      p[n++] = q;
      *out = (T *)q;
    }
    return e;
  }
};

}  // namespace

int launch_drift(wdy_sim *s, double hd) {
  const u32 n = (u32)s->n;
  if (s->timing) timer_begin(s->timer, s->stream, KC_OTHER);
  k_drift<<<grid_for(n), BLOCK, 0, s->stream>>>(n, s->key[s->cur],
                                                 s->vel[s->cur], hd);
  if (s->timing) timer_end(s->timer, s->stream, KC_OTHER);
  WDY_CHECK_LAUNCH();
  s->sorted_now = false;
  return WDY_OK;
}

int launch_sort_state(wdy_sim *s) {
  return radix_sort_pairs(&s->radix, s->stream, s->timing ? s->timer : nullptr,
                          s->key[s->cur], s->skey, s->perm[0], s->perm[1],
                          true);
}

int launch_leap(wdy_sim *s, double hk, double hd, double *acc_out,
                double *cum_out, u64 *spl_out, u32 T, u32 nb, double t) {
  const u32 n = (u32)s->n;
  int rc = launch_sort_state(s);
  if (rc) return rc;
  const int c = s->cur, o = (s->cur + 1) % 3;
  if (!s->equal_mass) {
    rc = mass_prefix(s, s->mass[c], s->perm[0], s->cum);
    if (rc) return rc;
  }
  const bool harm = s->omega2 >= 0.0;
  if (s->timing) timer_begin(s->timer, s->stream, KC_LEAP);
#define WDY_LEAP(EQ, HARM)                                                      \
  k_leap<EQ, HARM><<<grid_for(n), BLOCK, 0, s->stream>>>(                       \
      n, s->key[c], s->perm[0], s->vel[c], s->idx[c], s->mass[c], s->cum,       \
      s->key[o], s->vel[o], s->idx[o], s->mass[o], acc_out, cum_out, spl_out,   \
      T ? T : 1u, nb, s->m_equal, s->totmass, s->omega2, hk, hd, s->ext, t)
  if (s->equal_mass) {
    if (harm) WDY_LEAP(true, true); else WDY_LEAP(true, false);
  } else {
    if (harm) WDY_LEAP(false, true); else WDY_LEAP(false, false);
  }
#undef WDY_LEAP
  if (s->timing) timer_end(s->timer, s->stream, KC_LEAP);
  WDY_CHECK_LAUNCH();
  s->cur = o;
  s->sorted_now = false;
  return WDY_OK;
}

int launch_build_state(wdy_sim *s, const double *x_slot, const i32 *id_slot,
                       const double *v_id, const double *m_id) {
  const u32 n = (u32)s->n;
  // sort the staged positions: keys in key[1] (skey as partner), slots in perm[0]
  k_encode<<<grid_for(n), BLOCK, 0, s->stream>>>(n, x_slot, s->key[1]);
  WDY_CHECK_LAUNCH();
  int rc = radix_sort_pairs(&s->radix, s->stream, nullptr, s->key[1], s->skey,
                            s->perm[0], s->perm[1], true);
  if (rc) return rc;
  k_build_state<<<grid_for(n), BLOCK, 0, s->stream>>>(
      n, s->key[1], s->perm[0], id_slot, v_id, m_id, s->key[0], s->idx[0],
      s->small ? nullptr : s->vel[0],
      (s->small || s->equal_mass) ? nullptr : s->mass[0]);
  WDY_CHECK_LAUNCH();
  if (s->small) {
    WDY_HIP(hipMemcpyAsync(s->vel[0], v_id, (size_t)n * 8,
                           hipMemcpyDeviceToDevice, s->stream));
    if (!s->equal_mass)
      WDY_HIP(hipMemcpyAsync(s->mass[0], m_id, (size_t)n * 8,
                             hipMemcpyDeviceToDevice, s->stream));
  }
  s->cur = 0;
  s->sorted_now = true;
  return WDY_OK;
}

int launch_desort(wdy_sim *s, double *x_out, double *v_out, const double *acc,
                  double *acc_out, double *m_out) {
  const u32 n = (u32)s->n;
  const int c = s->small ? 0 : s->cur;
  if (s->timing) timer_begin(s->timer, s->stream, KC_OTHER);
  k_desort<<<grid_for(n), BLOCK, 0, s->stream>>>(
      n, s->key[c], s->vel[c], s->idx[c], acc, s->mass[c], x_out,
      s->small ? nullptr : v_out, acc ? acc_out : nullptr,
      (s->small || s->equal_mass) ? nullptr : m_out);
  if (s->timing) timer_end(s->timer, s->stream, KC_OTHER);
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

// Scratch layout of energy_enqueue for n sheets (unequal masses when
// `masses`): sorted keys (two buffers), permutation (two), gathered
// velocities, gathered masses and their prefix, tile sums and block
// partials, then the radix sort's histograms and tickets.  (The sort's
// look-back words are passed separately: they must stay where earlier
// calls left them.)
struct EnergyLayout {
  size_t k0, k1, p0, p1, sv, sm, cum, ts, hist, counters, bytes;
  u32 nt, nb, ntiles;
};

static EnergyLayout energy_layout(u32 n, bool masses) {
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  EnergyLayout L{};
  L.nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  L.nb = grid_for(n);
  L.ntiles = radix_tiles(n);
  size_t o = 0;
  L.k0 = o; o += up((size_t)n * 8);
  L.k1 = o; o += up((size_t)n * 8);
  L.p0 = o; o += up((size_t)n * 4);
  L.p1 = o; o += up((size_t)n * 4);
  L.sv = o; o += up((size_t)n * 8);
  L.sm = o; if (masses) o += up((size_t)n * 8);
  L.cum = o; if (masses) o += up((size_t)n * 8);
  L.ts = o; o += up(((size_t)L.nt + L.nb + 1) * 8);
  L.hist = o; o += up(radix_hist_bytes());
  L.counters = o; o += up(radix_counter_bytes());
  L.bytes = o;
  return L;
}

size_t energy_scratch_bytes(u32 n, bool masses) {
  return energy_layout(n, masses).bytes;
}

int energy_enqueue(hipStream_t st, u32 n, const double *x, const double *v,
                   const double *m, double mq, double mscale, double twopiG,
                   double omega2, double *e_dev, u64 *status, void *ws,
                   u32 *epoch, const u32 **err_word) {
  const EnergyLayout L = energy_layout(n, m != nullptr);
  char *w = (char *)ws;
  u64 *k0 = (u64 *)(w + L.k0), *k1 = (u64 *)(w + L.k1);
  u32 *p0 = (u32 *)(w + L.p0), *p1 = (u32 *)(w + L.p1);
  double *sv = (double *)(w + L.sv);
  double *sm = m ? (double *)(w + L.sm) : nullptr;
  double *cum = m ? (double *)(w + L.cum) : nullptr;
  double *ts = (double *)(w + L.ts), *red = ts;
  RadixWork rw;
  rw.n = n;
  rw.ntiles = L.ntiles;
  rw.hist = (u32 *)(w + L.hist);
  rw.counters = (u32 *)(w + L.counters);
  rw.status = status;
  rw.epoch = *epoch;
  rw.clear_err = true;
  if (err_word) *err_word = radix_err_word(&rw);
  k_encode<<<L.nb, BLOCK, 0, st>>>(n, x, k0);
  WDY_CHECK_LAUNCH();
  const int rc = radix_sort_pairs(&rw, st, nullptr, k0, k1, p0, p1, true);
  *epoch = rw.epoch;
  if (rc) return rc;
  k_gather2<<<L.nb, BLOCK, 0, st>>>(n, p0, v, m, sv, sm);
  if (m) {
    // exclusive prefix of the sorted masses; total = C[n-1] + m[n-1]
    k_scan_reduce<<<L.nt, BLOCK, 0, st>>>(n, sm, nullptr, ts);
    k_scan_top<<<1, BLOCK, 0, st>>>(L.nt, ts);
    k_scan_apply<<<L.nt, BLOCK, 0, st>>>(n, sm, nullptr, ts, cum);
  }
  k_energy_terms<<<L.nb, BLOCK, 0, st>>>(n, k0, sv, sm, cum, mq,
                                         (double)n * mq, mscale, twopiG, omega2,
                                         red);
  k_sum_partials<<<1, BLOCK, 0, st>>>(L.nb, red, e_dev, radix_err_word(&rw));
  WDY_CHECK_LAUNCH();
  return WDY_OK;
}

int energy_core(hipStream_t st, u32 n, const double *x, const double *v,
                const double *m, double mq, double mscale, double twopiG,
                double omega2, double *e_out) {
  DevBufs b;
  char *ws;
  double *e_dev;
  u64 *status;
  WDY_HIP(b.alloc(&ws, energy_scratch_bytes(n, m != nullptr)));
  WDY_HIP(b.alloc(&e_dev, 1));
  WDY_HIP(b.alloc(&status, radix_status_bytes(n) / 8));
  // fresh look-back words: zero, then tagged by epochs from 1
  WDY_HIP(hipMemsetAsync(status, 0, radix_status_bytes(n), st));
  u32 epoch = 1;
  const u32 *err_word = nullptr;
  if (int rc = energy_enqueue(st, n, x, v, m, mq, mscale, twopiG, omega2, e_dev,
                              status, ws, &epoch, &err_word))
    return rc;
  u32 gave_up = 0;
  WDY_HIP(hipMemcpyAsync(e_out, e_dev, 8, hipMemcpyDeviceToHost, st));
  WDY_HIP(hipMemcpyAsync(&gave_up, err_word, 4, hipMemcpyDeviceToHost, st));
  WDY_HIP(hipStreamSynchronize(st));
  if (gave_up)
    return fail(WDY_ERUNTIME, "energy: a look-back wait of the sort gave up");
  return WDY_OK;
}

int launch_energy(wdy_sim *s, double twopiG, double omega2, double *e_out) {
  const u32 n = (u32)s->n;
  double *m_id = nullptr;
  if (!s->equal_mass) m_id = s->small ? s->mass[0] : s->cum;  // cum: scratch
  int rc = launch_desort(s, s->out_x, s->out_v, nullptr, nullptr,
                         s->small ? nullptr : m_id);
  if (rc) return rc;
  return energy_core(s->stream, n, s->out_x, s->small ? s->vel[0] : s->out_v,
                     m_id, s->m_equal, twopiG, twopiG, omega2, e_out);
}

// ---- stateless diagnostics on host arrays ---------------------------------

int launch_energy_arrays(int64_t n64, const double *x, const double *v,
                         const double *m, double twopiG, double omega2,
                         double *e_out) {
  const u32 n = (u32)n64;
  DevBufs b;
  double *dx, *dv, *dm;
  WDY_HIP(b.alloc(&dx, n));
  WDY_HIP(b.alloc(&dv, n));
  WDY_HIP(b.alloc(&dm, n));
  WDY_HIP(hipMemcpy(dx, x, 8 * (size_t)n, hipMemcpyHostToDevice));
  WDY_HIP(hipMemcpy(dv, v, 8 * (size_t)n, hipMemcpyHostToDevice));
  WDY_HIP(hipMemcpy(dm, m, 8 * (size_t)n, hipMemcpyHostToDevice));
  return energy_core(0, n, dx, dv, dm, 0.0, 1.0, twopiG, omega2, e_out);
}

int launch_momentum_arrays(int64_t n64, const double *v, const double *m,
                           double *p_out) {
  const u32 n = (u32)n64;
  DevBufs b;
  double *dv, *dm, *red;
  const u32 nb = grid_for(n);
  WDY_HIP(b.alloc(&dv, n));
  WDY_HIP(b.alloc(&dm, n));
  WDY_HIP(b.alloc(&red, nb + 1));
  WDY_HIP(hipMemcpy(dv, v, 8 * (size_t)n, hipMemcpyHostToDevice));
  WDY_HIP(hipMemcpy(dm, m, 8 * (size_t)n, hipMemcpyHostToDevice));
  k_momentum_terms<<<nb, BLOCK>>>(n, dv, dm, red);
  k_sum_partials<<<1, BLOCK>>>(nb, red, red + nb, nullptr);
  WDY_CHECK_LAUNCH();
  WDY_HIP(hipMemcpy(p_out, red + nb, 8, hipMemcpyDeviceToHost));
  return WDY_OK;
}

int launch_potential_arrays(int64_t ny, const double *y, int64_t n64,
                            const double *x, const double *m, double twopiG,
                            double omega2, double *out) {
  const u32 n = (u32)n64;
  DevBufs b;
  double *dy, *dx, *dm, *dout;
  WDY_HIP(b.alloc(&dy, ny));
  WDY_HIP(b.alloc(&dx, n));
  WDY_HIP(b.alloc(&dm, n));
  WDY_HIP(b.alloc(&dout, ny));
  WDY_HIP(hipMemcpy(dy, y, 8 * (size_t)ny, hipMemcpyHostToDevice));
  WDY_HIP(hipMemcpy(dx, x, 8 * (size_t)n, hipMemcpyHostToDevice));
  WDY_HIP(hipMemcpy(dm, m, 8 * (size_t)n, hipMemcpyHostToDevice));
  k_potential<<<(unsigned)ny, BLOCK>>>(n, dx, dm, dy, twopiG, omega2, dout);
  WDY_CHECK_LAUNCH();
  WDY_HIP(hipMemcpy(out, dout, 8 * (size_t)ny, hipMemcpyDeviceToHost));
  return WDY_OK;
}

int launch_energy_individual_arrays(int64_t n64, const double *x,
                                    const double *v, const double *m,
                                    double twopiG, double omega2, double *out) {
  const u32 n = (u32)n64;
  DevBufs b;
  double *dx, *dv, *dm, *dout;
  WDY_HIP(b.alloc(&dx, n));
  WDY_HIP(b.alloc(&dv, n));
  WDY_HIP(b.alloc(&dm, n));
  WDY_HIP(b.alloc(&dout, n));
  WDY_HIP(hipMemcpy(dx, x, 8 * (size_t)n, hipMemcpyHostToDevice));
  WDY_HIP(hipMemcpy(dv, v, 8 * (size_t)n, hipMemcpyHostToDevice));
  WDY_HIP(hipMemcpy(dm, m, 8 * (size_t)n, hipMemcpyHostToDevice));
  k_energy_individual<<<n, BLOCK>>>(n, dx, dv, dm, twopiG, omega2, dout);
  WDY_CHECK_LAUNCH();
  WDY_HIP(hipMemcpy(out, dout, 8 * (size_t)n, hipMemcpyDeviceToHost));
  return WDY_OK;
}

}  // namespace wdy

