// Canonical-order float reductions: reduce_add_f64 / reduce_mul_f64 (CanonicalAcc,
// src/kernels/scalar/reduce.rs:249-323), dot_f64, sumsq_f64 and condsum_<cmp>_f64
// (fused.rs:72-210). Accumulator j folds rows j, j+16, ... in index order with plain IEEE
// adds (or multiplies); no FMA anywhere (built with -ffp-contract=off).
//
// Two forms. The sequential form is one wave with lane j < 16 owning accumulator j; its bound
// is the dependent-add latency, n/16 adds in a row. The fast form (sums only) evaluates the
// same 16 chains in parallel with binade run summaries (canon_binade.h) and returns the same
// bits; it is HBM-bound on data whose running sums stay inside a binade for long stretches.
#include <type_traits>

#include "canon_binade.h"
#include "chain.h"
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

// Row i's value in a canonical reduction of kind K: x (0: sums, and the factors of products),
// x * y (1: dot), x * x (2: sumsq), x where its comparison with c holds, else e (3: condsum).
// K is a template argument so that row loads are plain, independent loads.
template <int K>
struct Row {
  // the summary kernel's passes per thread and the waves per SIMD it is built for: two-input
  // rows (dot) take one pass at four waves, one-input rows two passes at six (spilling no more
  // than a pair of values)
  static constexpr int kSummPasses = K == 1 ? 1 : 2;
  static constexpr int kSummWaves = K == 1 ? 1 : 6;
  double c, e;
  int keep;  // condsum: the outcomes of comparing x with c that keep x (bit 0 <, 1 ==, 2 >, 3 unordered)
  __device__ __forceinline__ double operator()(const double *x, const double *y, uint64_t i) const {
    if constexpr (K == 1) {
      return x[i] * y[i];
    } else if constexpr (K == 2) {
      const double v = x[i];
      return v * v;
    } else if constexpr (K == 3) {
      const double v = x[i];
      const int o = v < c ? 1 : v == c ? 2 : v > c ? 4 : 8;
      return (keep & o) ? v : e;
    } else {
      return x[i];
    }
  }
};

constexpr int kGroup = 8;  // 64-row chunks loaded ahead of the dependent chain

// acc[16] <- fold of rows [0, 16*blocks) in the canonical order.
template <class RV>
__global__ __launch_bounds__(64) void k_canon_chain(RV rv, bool mul, const double *__restrict__ x,
                                                     const double *__restrict__ y, uint64_t blocks,
                                                     double *acc) {
  const int l = lane_id();
  const uint64_t rows = blocks * 16;
  double a = l < 16 ? acc[l] : 0.0;
  const uint64_t chunks = (rows + 63) / 64;
  double cur[kGroup], nxt[kGroup];
  auto load = [&](double *v, uint64_t g) {
#pragma unroll
    for (int k = 0; k < kGroup; ++k) {
      uint64_t r = (g * kGroup + k) * 64 + l;
      v[k] = r < rows ? rv(x, y, r) : 0.0;
    }
  };
  const uint64_t groups = (chunks + kGroup - 1) / kGroup;
  if (groups == 0) return;
  load(cur, 0);
  for (uint64_t g = 0; g < groups; ++g) {
    if (g + 1 < groups) load(nxt, g + 1);
#pragma unroll
    for (int k = 0; k < kGroup; ++k) {
      const uint64_t base = (g * kGroup + k) * 64;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double v = __shfl(cur[k], q * 16 + (l & 15));
        if (l < 16 && base + q * 16 < rows) a = mul ? a * v : a + v;
      }
    }
#pragma unroll
    for (int k = 0; k < kGroup; ++k) cur[k] = nxt[k];
  }
  if (l < 16) acc[l] = a;
}

// Batched items of fewer than kChainItemRows rows, one wave each: acc[item * 16 ..] <- the
// item's whole blocks folded in order from acc (as k_canon_chain). Many items run at once, so
// the dependent chains overlap; the tile form's per-tile machinery would cost more than the
// chain of a small item. Larger items are left to the tile form.
constexpr uint64_t kChainItemRows = uint64_t(1) << 16;

template <class RV>
__global__ __launch_bounds__(64) void k_canon_chain_batch(RV rv, const BatchItem *items, uint64_t nitems,
                                                           double *acc) {
  const int l = lane_id();
  for (uint64_t it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem &I = items[it];
    if (I.n >= kChainItemRows) continue;
    const double *x = static_cast<const double *>(I.a);
    const double *y = I.b ? static_cast<const double *>(I.b) : x;
    const uint64_t rows = I.n / 16 * 16;
    double a = l < 16 ? acc[it * 16 + l] : 0.0;
    const uint64_t chunks = (rows + 63) / 64;
    const uint64_t groups = (chunks + kGroup - 1) / kGroup;
    double cur[kGroup], nxt[kGroup];
    auto load = [&](double *v, uint64_t g) {
#pragma unroll
      for (int k = 0; k < kGroup; ++k) {
        const uint64_t r = (g * kGroup + k) * 64 + l;
        v[k] = r < rows ? rv(x, y, r) : 0.0;
      }
    };
    if (groups > 0) load(cur, 0);
    for (uint64_t g = 0; g < groups; ++g) {
      if (g + 1 < groups) load(nxt, g + 1);
#pragma unroll
      for (int k = 0; k < kGroup; ++k) {
        const uint64_t base = (g * kGroup + k) * 64;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double v = __shfl(cur[k], q * 16 + (l & 15));
          if (l < 16 && base + q * 16 < rows) a = a + v;
        }
      }
#pragma unroll
      for (int k = 0; k < kGroup; ++k) cur[k] = nxt[k];
    }
    if (l < 16) acc[it * 16 + l] = a;
  }
}

// ------------------------------------------------------------------ fast form
// The 16 chains over tiles of 1024 blocks (16384 rows), one pass (canon_binade.h); a block
// of 1024 threads, thread t owning chain j = t % 16 and run r = t / 16 (64 runs of 16 blocks):
//  1. each thread loads the 16 rows of its chain in its run (a wave load covers four
//     128-byte blocks) and keeps them in registers;
//  2. per-chain approximate sums, and a decoupled look-back over them, give each chain an
//     approximate start and so a binade hypothesis;
//  3. every thread summarises its run under the hypothesis; runs join in order (two
//     shuffles per wave, then a tree over the 16 waves) into the tile summary;
//  4. a second decoupled look-back joins the summaries of preceding tiles with the same
//     hypothesis back to a tile whose exact end values are published, applies the joined
//     summary to that exact value, and publishes this tile's exact end values. A failed
//     check waits for the previous tile's exact values and folds the tile's rows of the
//     chain sequentially, re-read from memory.
// Tiles are large so that the tile ticket (one same-address atomic per tile) never limits
// the row rate; look-back status is published with relaxed agent-scope stores ordered by
// publish_fence, never with release/acquire (which would write back and invalidate L2).
constexpr int kFastThreads = 1024;
constexpr int kFastWaves = kFastThreads / kWave;
constexpr int kRunBlocks = 16;
constexpr int kTileBlocks = kFastThreads / 16 * kRunBlocks;  // 1024 blocks
constexpr uint64_t kFastMinBlocks = kTileBlocks;
constexpr uint64_t kStreamMinTiles = 64;  // from 1M rows the streamed form (below) is used
// Status words per tile: [0] flags; approximate aggregate and inclusive sums per chain; the
// tile summary per chain (6 words, the binade word at +7); exact end value per chain, and per
// chain a flag word (epoch << 8 | 1) once that value is published. Exact values are published
// chain by chain, so one chain's slow evaluation holds up no other chain.
constexpr int kStride = 200;
constexpr int W_AAGG = 1, W_AINC = 17, W_F = 33, W_EXACT = 161, W_XF = 177;
constexpr uint64_t F_AAGG = 1, F_AINC = 2, F_FN = 4, F_EXACT = 8;
static_assert(kFastWaves == 16, "the tile summary tree joins 16 waves as 4 x 4");

// Row functors with a `map` (fused regions) are loaded raw, all rows first, then mapped, so the
// loads issue together ahead of the per-row work.
template <class RV, class = void>
struct Mapped : std::false_type {};
template <class RV>
struct Mapped<RV, decltype(void(RV::kMapped))> : std::true_type {};

// Chain j's rows of blocks b .. b + min(cnt, 16) - 1 into v (zeros past cnt).
template <class RV>
__device__ __forceinline__ void load_run(const RV &rv, const double *x, const double *y, uint64_t b, int j,
                                         int cnt, double (&v)[kRunBlocks]) {
  if constexpr (Mapped<RV>::value) {
    double u[kRunBlocks];
#pragma unroll
    for (int k = 0; k < kRunBlocks; ++k) {
      const uint64_t i = (b + k) * 16 + j;
      v[k] = k < cnt ? x[i] : 0.0;
      u[k] = RV::kTwo && k < cnt ? y[i] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < kRunBlocks; ++k) v[k] = k < cnt ? rv.map(v[k], u[k]) : 0.0;
    return;
  }
  if (cnt >= kRunBlocks) {
#pragma unroll
    for (int k = 0; k < kRunBlocks; ++k) v[k] = rv(x, y, (b + k) * 16 + j);
  } else {
#pragma unroll
    for (int k = 0; k < kRunBlocks; ++k) v[k] = k < cnt ? rv(x, y, (b + k) * 16 + j) : 0.0;
  }
}

__device__ __forceinline__ uint64_t flag_word(uint32_t epoch, uint64_t bits) {
  return (uint64_t(epoch) << 8) | bits;
}
__device__ __forceinline__ uint64_t wait_flag(const uint64_t *st, uint32_t epoch, uint64_t need) {
  while (true) {
    const uint64_t w = ld_rlx(st);
    if ((w >> 8) == epoch && (w & need)) {
      observe_fence();
      return w;
    }
    backoff();
  }
}
// Waits until chain j of the tile at sk has a published exact value (returns true) or the tile
// has published its summaries (returns false).
__device__ __forceinline__ bool wait_exact_or_fn(const uint64_t *sk, uint32_t epoch, int j) {
  while (true) {
    // both words in one round trip
    const uint64_t xf = ld_rlx(sk + W_XF + j);
    const uint64_t w = ld_rlx(sk);
    if ((xf >> 8) == epoch) {
      observe_fence();
      return true;
    }
    if ((w >> 8) == epoch && (w & F_FN)) {
      observe_fence();
      return false;
    }
    backoff();
  }
}
// Waits until chain j of the tile at sk has a published exact value; returns it.
__device__ __forceinline__ double wait_exact(const uint64_t *sk, uint32_t epoch, int j);
__device__ __forceinline__ double ld_d(const uint64_t *p) {
  return __longlong_as_double(static_cast<long long>(ld_rlx(p)));
}
__device__ __forceinline__ void st_d(uint64_t *p, double v) {
  st_rlx(p, static_cast<uint64_t>(__double_as_longlong(v)));
}
__device__ __forceinline__ double wait_exact(const uint64_t *sk, uint32_t epoch, int j) {
  while ((ld_rlx(sk + W_XF + j) >> 8) != epoch) backoff();
  observe_fence();
  return ld_d(sk + W_EXACT + j);
}
__device__ __forceinline__ uint64_t pack_binade(const Binade &h) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(h.e + 2048)) << 2) | (uint64_t(h.neg) << 1) |
         uint64_t(h.ok);
}
__device__ __forceinline__ Binade unpack_binade(uint64_t w) {
  return Binade{static_cast<int>(w >> 2) - 2048, static_cast<int>((w >> 1) & 1), static_cast<int>(w & 1), 0};
}

// The exact fold of chain c's rows of blocks b0 .. b0 + nb - 1 from the exact value s, by
// one wave: lane r holds run r (blocks 16r .. 16r + 15). Two ways, picked by how often the
// binade of the approximate running value (s plus the approximate sums of the runs before)
// changes between runs:
//  - few changes (the usual single crossing): rounds. Runs are summarised under the binade of
//    the current exact value and scanned in order; the runs before the first one whose check
//    fails are applied at once, that run is folded element by element, and the next round
//    starts after it. Each round passes at least one run.
//  - many changes (a running value near zero): the sequential fold itself, one element after
//    another, each read from its run's lane (about a thousand dependent adds, fewer cycles
//    than a summary check per run where most runs would fail theirs).
// A check fails only where the value leaves its binade or meets a huge, infinite or NaN
// element, and a failed check is always followed by the element-by-element fold, so the
// result is the sequential one either way.
template <class RV>
__device__ double eval_chain(const RV &rv, const double *x, const double *y, uint64_t b0, int nb,
                             int c, double s) {
  const int l = lane_id();
  double v[kRunBlocks];
  double asum = 0.0;
#pragma unroll
  for (int k = 0; k < kRunBlocks; ++k) {
    const int lb = l * kRunBlocks + k;
    v[k] = lb < nb ? rv(x, y, (b0 + lb) * 16 + c) : 0.0;
    asum = asum + v[k];
  }
  const int nruns = (nb + kRunBlocks - 1) / kRunBlocks;
  // approximate start of each run: an inclusive scan of the run sums, less the own sum
  double inc = asum;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double o = __shfl_up(inc, d);
    if (l >= d) inc = inc + o;
  }
  const Binade ha = binade_of(s + (inc - asum));
  const Binade hp{__shfl_up(ha.e, 1), __shfl_up(ha.neg, 1), __shfl_up(ha.ok, 1), 0};
  const uint64_t changes = __ballot(l > 0 && l < nruns && !same_binade(ha, hp));
  double cur = s;
  if (__popcll(changes) <= 2) {
    for (int r0 = 0; r0 < nruns;) {
      const Binade h = binade_of(cur);
      RunFn f = run_identity();
      if (h.ok && l >= r0) f = run_of(v, nb - l * kRunBlocks, binade_ref(h));
#pragma unroll
      for (int d = 1; d < kWave; d <<= 1) {
        const RunFn o = shfl_up_pod(f, d);
        if (l >= d) f = run_join(o, f);
      }
      double val = cur;
      const bool ok = l < r0 || run_apply(f, h, cur, &val);
      const uint64_t fails = __ballot(!ok);
      if (fails == 0) {
        cur = __shfl(val, kWave - 1);
        break;
      }
      const int m = __builtin_ctzll(fails);
      double t = m > r0 ? __shfl(val, m - 1) : cur;
      if (l == m) {
#pragma unroll
        for (int k = 0; k < kRunBlocks; ++k)
          if (m * kRunBlocks + k < nb) t = t + v[k];
      }
      cur = __shfl(t, m);
      r0 = m + 1;
    }
    return cur;
  }
  // many changes (a running value near zero): the sequential fold itself, element by element,
  // each element read from its run's lane, stopping at block nb
  for (int r = 0; r < nruns; ++r) {
    const int nk = nb - r * kRunBlocks < kRunBlocks ? nb - r * kRunBlocks : kRunBlocks;
    if (nk == kRunBlocks) {
#pragma unroll
      for (int k = 0; k < kRunBlocks; ++k) cur = cur + bcast_pod(v[k], r);
    } else {
#pragma unroll
      for (int k = 0; k < kRunBlocks; ++k)
        if (k < nk) cur = cur + bcast_pod(v[k], r);
    }
  }
  return cur;
}

// Joins the summaries of one wave's four look-back slots (lane l holds slot l / 16 of chain
// l % 16; a higher slot is an older tile) so that every lane ends with older-to-newer order.
__device__ __forceinline__ RunFn join_slots(RunFn f, int l) {
  RunFn o = shfl_xor_pod(f, 16);
  f = (l & 16) ? run_join(f, o) : run_join(o, f);
  o = shfl_xor_pod(f, 32);
  f = (l & 32) ? run_join(f, o) : run_join(o, f);
  return f;
}

// BATCH: the tile's item gives the rows (a, b), the block count (n / 16), the first tile, and
// the item's 16 start and end accumulators at acc_in / acc_out + 16 * item.
//
// Both look-backs are windowed: thread t reads tile `end - 1 - i` for slot i = 4 * (t / 64) +
// (t % 64) / 16 and chain t % 16, so one round reads 64 predecessors in parallel; a
// block-wide minimum finds the nearest tile that ends the look-back (per chain for the exact
// one), and the slots before it are summed (approximate) or joined in order (exact).
template <class RV, bool BATCH>
__global__ __launch_bounds__(kFastThreads) void k_canon_fast(RV rv, const double *__restrict__ x,
                                                             const double *__restrict__ y, uint64_t blocks,
                                                             const double *acc_in, double *acc_out,
                                                             uint64_t *status, uint32_t epoch,
                                                             uint32_t *ticket, uint32_t ntiles,
                                                             const BatchItem *items,
                                                             const uint32_t *tile_item,
                                                             const uint32_t *dev_ntiles = nullptr,
                                                             const uint32_t *dev_epoch = nullptr) {
  __shared__ double s_a[kFastWaves][16];
  __shared__ RunFn s_f[kFastWaves][16];
  __shared__ Binade s_h[16];
  __shared__ int s_stop[16];
  __shared__ int s_more;
  __shared__ double s_base[16];
  __shared__ double s_exact[16];
  __shared__ int s_fail[16];
  const uint32_t tile = take_ticket(ticket, ntiles);
  // tiles planned on the device: blocks past the planned count exit, the epoch is the plan's
  if (dev_ntiles && tile >= *dev_ntiles) return;
  if (dev_epoch) epoch = *dev_epoch;
  uint32_t first = 0, last = ntiles - 1;
  if (BATCH) {
    const uint32_t it = tile_item[tile];
    const BatchItem &I = items[it];
    x = static_cast<const double *>(I.a);
    y = I.b ? static_cast<const double *>(I.b) : x;
    blocks = I.n / 16;
    first = static_cast<uint32_t>(I.tile0);
    last = first + static_cast<uint32_t>((blocks + kTileBlocks - 1) / kTileBlocks) - 1;
    acc_in += 16 * uint64_t(it);
    acc_out += 16 * uint64_t(it);
  }
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const int j = t & 15, r = t >> 4;
  const int slot = 4 * w + (l >> 4);  // look-back slot of this thread (chain j)
  const uint64_t b0 = uint64_t(tile - first) * kTileBlocks;
  const int nb = static_cast<int>(blocks - b0 < uint64_t(kTileBlocks) ? blocks - b0 : kTileBlocks);
  uint64_t *st = status + uint64_t(tile) * kStride;

  // 1. load the run
  double v[kRunBlocks];
  double a = 0.0;
  if constexpr (Mapped<RV>::value) {
    load_run(rv, x, y, b0 + uint64_t(r) * kRunBlocks, j, nb - r * kRunBlocks, v);
#pragma unroll
    for (int k = 0; k < kRunBlocks; ++k) a = a + v[k];
  } else {
#pragma unroll
    for (int k = 0; k < kRunBlocks; ++k) {
      const int lb = r * kRunBlocks + k;
      v[k] = lb < nb ? rv(x, y, (b0 + lb) * 16 + j) : 0.0;
      a = a + v[k];
    }
  }
  // 2. approximate chain sums of the tile, published as an aggregate (or inclusive at the
  //    item's first tile)
  a = a + __shfl_xor(a, 16);
  a = a + __shfl_xor(a, 32);
  if (l < 16) s_a[w][j] = a;
  if (t == 0) s_stop[0] = 64;
  __syncthreads();
  double atile = 0.0;  // lanes 0..15 of wave 0
  if (w == 0) {
    if (l < 16) {
#pragma unroll
      for (int k = 0; k < kFastWaves; ++k) atile = atile + s_a[k][l];
      if (tile == first) st_d(st + W_AINC + l, acc_in[l] + atile);
      else st_d(st + W_AAGG + l, atile);
    }
    publish_fence();
    if (l == 0) st_rlx(st, flag_word(epoch, tile == first ? (F_AAGG | F_AINC) : F_AAGG));
  }
  // 2b. approximate look-back: the sum of aggregates back to an inclusive or exact tile
  double pre = 0.0;  // lanes 0..15 of wave 0
  if (tile == first) {
    if (w == 0 && l < 16) pre = acc_in[l];
  } else {
    for (int64_t end = tile;; end -= 64) {
      __syncthreads();  // s_a and s_stop[0] are free
      const int64_t T = end - 1 - slot;
      bool stop = true;
      double val = 0.0;
      if (T < int64_t(first)) {
        val = acc_in[j];
      } else {
        const uint64_t *sk = status + uint64_t(T) * kStride;
        const uint64_t fw = wait_flag(sk, epoch, F_AAGG | F_AINC | F_EXACT);
        if (fw & F_EXACT) val = ld_d(sk + W_EXACT + j);
        else if (fw & F_AINC) val = ld_d(sk + W_AINC + j);
        else {
          val = ld_d(sk + W_AAGG + j);
          stop = false;
        }
      }
      if (stop) atomicMin(&s_stop[0], slot);
      __syncthreads();
      const int sidx = s_stop[0];
      double cv = slot <= sidx ? val : 0.0;
      cv = cv + __shfl_xor(cv, 16);
      cv = cv + __shfl_xor(cv, 32);
      if (l < 16) s_a[w][j] = cv;
      __syncthreads();
      if (w == 0 && l < 16) {
#pragma unroll
        for (int k = 0; k < kFastWaves; ++k) pre = pre + s_a[k][l];
      }
      if (sidx < 64) break;
      if (t == 0) s_stop[0] = 64;
    }
  }
  if (w == 0) {
    if (tile > first) {
      if (l < 16) st_d(st + W_AINC + l, pre + atile);
      publish_fence();
      if (l == 0) st_rlx(st, flag_word(epoch, F_AAGG | F_AINC));
    }
    if (l < 16) s_h[l] = binade_of(pre);
  }
  __syncthreads();
  // 3. summarise the run under the chain's hypothesis, join the runs of the wave in order
  const Binade h = s_h[j];
  RunFn f = run_identity();
  if (h.ok) f = run_of(v, nb - r * kRunBlocks, binade_ref(h));
  {
    RunFn o = shfl_xor_pod(f, 16);
    f = (l & 16) ? run_join(o, f) : run_join(f, o);
    o = shfl_xor_pod(f, 32);
    f = (l & 32) ? run_join(o, f) : run_join(f, o);
  }
  if (l < 16) s_f[w][j] = f;
  __syncthreads();
  // 4. the tile summary (wave 0: lane l joins waves 4q..4q+3 of chain l % 16, q = l / 16,
  //    then the four quarters by shuffles), published
  RunFn ft = run_identity();
  const Binade hj = s_h[j];
  if (w == 0) {
    const int q = l >> 4;
    ft = run_join(run_join(s_f[4 * q][j], s_f[4 * q + 1][j]), run_join(s_f[4 * q + 2][j], s_f[4 * q + 3][j]));
    RunFn o = shfl_xor_pod(ft, 16);
    ft = (l & 16) ? run_join(o, ft) : run_join(ft, o);
    o = shfl_xor_pod(ft, 32);
    ft = (l & 32) ? run_join(o, ft) : run_join(ft, o);
    if (l < 16) {
      uint64_t *fp = st + W_F + 8 * l;
      st_d(fp + 0, ft.d[0]);
      st_d(fp + 1, ft.d[1]);
      st_d(fp + 2, ft.lo[0]);
      st_d(fp + 3, ft.lo[1]);
      st_d(fp + 4, ft.hi[0]);
      st_d(fp + 5, ft.hi[1]);
      // A summary that fails even at the approximate start (a binade crossing or a huge,
      // infinite or NaN element inside the tile) is published as opaque: later tiles stop at
      // it and wait for its exact value instead of joining a summary that cannot apply.
      double probe;
      const bool usable = run_apply(ft, hj, pre, &probe);
      st_rlx(fp + 7, pack_binade(usable ? hj : Binade{0, 0, 0, 0}));
    }
    publish_fence();
    if (l == 0) st_rlx(st, flag_word(epoch, F_AAGG | F_AINC | F_FN));
  }
  // 5. exact look-back: per chain, join the summaries of same-hypothesis tiles back to a tile
  //    with published exact values (or the item's start); a tile with another hypothesis
  //    ends it too, and its exact values are awaited
  RunFn g = run_identity();  // wave 0 lanes 0..15: joined summaries after `base`
  double base = 0.0;
  int64_t base_tile = int64_t(tile) - 1;  // the tile whose end `base` is
  if (tile == first) {
    if (w == 0 && l < 16) base = acc_in[l];
  } else {
    if (t < 16) s_stop[t] = 64;
    bool open = true;  // wave 0 lanes 0..15: chain still looking back
    for (int64_t end = tile;; end -= 64) {
      __syncthreads();  // s_f, s_stop and s_more are free
      const int64_t T = end - 1 - slot;
      const bool chain_open = s_stop[j] >= 0;  // -1 marks a finished chain
      RunFn fk = run_identity();
      bool stop = false;
      if (chain_open) {
        if (T < int64_t(first)) {
          stop = true;
        } else {
          const uint64_t *sk = status + uint64_t(T) * kStride;
          const bool exact_there = wait_exact_or_fn(sk, epoch, j);
          const uint64_t *fp = sk + W_F + 8 * j;
          if (exact_there) {
            stop = true;
          } else if (!same_binade(unpack_binade(ld_rlx(fp + 7)), hj)) {
            stop = true;
          } else {
            fk.d[0] = ld_d(fp + 0);
            fk.d[1] = ld_d(fp + 1);
            fk.lo[0] = ld_d(fp + 2);
            fk.lo[1] = ld_d(fp + 3);
            fk.hi[0] = ld_d(fp + 4);
            fk.hi[1] = ld_d(fp + 5);
          }
        }
      }
      __syncthreads();  // every thread has read s_stop[j] for chain_open
      if (stop) atomicMin(&s_stop[j], slot);
      __syncthreads();
      const int sidx = s_stop[j];
      if (!(chain_open && slot < sidx)) fk = run_identity();
      fk = join_slots(fk, l);
      if (l < 16) s_f[w][j] = fk;
      __syncthreads();
      if (w == 0) {
        const int q = l >> 4;
        // waves 4q+3 (oldest) .. 4q, then the quarters, older first
        RunFn gw = run_join(run_join(s_f[4 * q + 3][j], s_f[4 * q + 2][j]),
                            run_join(s_f[4 * q + 1][j], s_f[4 * q][j]));
        gw = join_slots(gw, l);
        if (l < 16 && open) {
          g = run_join(gw, g);
          if (sidx < 64) {
            const int64_t Ts = end - 1 - sidx;
            base_tile = Ts;
            if (Ts < int64_t(first)) {
              base = acc_in[l];
            } else {
              base = wait_exact(status + uint64_t(Ts) * kStride, epoch, l);
            }
            open = false;
          }
        }
        const bool more = __ballot(l < 16 && open) != 0;
        if (l < 16) s_stop[l] = open ? 64 : -1;
        if (l == 0) s_more = more;
      }
      __syncthreads();
      if (!s_more) break;
    }
  }
  // 6. apply; a chain whose check fails is evaluated exactly by wave (chain), from the exact
  //    value at the tile's start: `base` moved over the joined predecessors, or, where that
  //    check fails too, the previous tile's published end
  //    Chains that apply are published at once; the others after their evaluation.
  if (w == 0) {
    double exact = 0.0;
    bool done = true;
    if (l < 16) {
      done = run_apply(run_join(g, ft), hj, base, &exact);
      s_exact[l] = exact;
      s_fail[l] = !done;
      if (done) st_d(st + W_EXACT + l, exact);
    }
    publish_fence();
    if (l < 16 && done) st_rlx(st + W_XF + l, flag_word(epoch, 1));
    if (l < 16 && !done && base_tile != int64_t(tile) - 1) {
      double start = 0.0;
      if (run_apply(g, hj, base, &start)) base = start;
      else base = wait_exact(status + uint64_t(tile - 1) * kStride, epoch, l);
    }
    if (l < 16) s_base[l] = base;
  }
  __syncthreads();
  static_assert(kFastWaves == 16, "one wave per chain");
  if (s_fail[w]) {
    const double e = eval_chain(rv, x, y, b0, nb, w, s_base[w]);
    if (l == 0) {
      s_exact[w] = e;
      st_d(st + W_EXACT + w, e);
    }
    publish_fence();
    if (l == 0) st_rlx(st + W_XF + w, flag_word(epoch, 1));
  }
  __syncthreads();
  if (w != 0) return;
  if (l < 16 && tile == last) acc_out[l] = s_exact[l];
  if (l == 0) st_rlx(st, flag_word(epoch, F_AAGG | F_AINC | F_FN | F_EXACT));
}

// ---------------------------------------------------------------- streamed form
// For long inputs the same exact evaluation in plain launches, with no look-back:
//  1. k_canon_sample, k_canon_prefix: predicted approximate starts of every tile per chain,
//     from every 16th block (a sixteenth of the rows) scaled by 16 and prefixed from the start
//     state; their binades are the first hypotheses;
//  2. k_canon_summ: every tile's run summaries under its predicted hypothesis, joined
//     into a tile summary, and the tile's approximate chain sums (one streaming read);
//  3. k_canon_prefix: the exclusive prefix of those sums from the start state, the approximate
//     start of every tile (any order of these approximate sums serves: the applies below prove
//     every result);
//  4. k_canon_check: every chain's summary is checked at its approximate start; a chain whose
//     summary is not usable there takes a slot for a wide table; tiles with a chain whose
//     prediction missed the binade of its approximate start are listed;
//  5. k_canon_fix: the listed tiles are re-read and their missed chains summarised again (a
//     chain still not usable takes a slot); k_canon_wide: the wide table (canon_binade.h) of
//     every slot's chain, simulated from its approximate start;
//  6. k_canon_groups: the ordered join of each group of 64 tiles whose summaries are all usable
//     under one hypothesis ("clean" groups);
//  7. k_canon_walk: one wave per chain walks the groups in order from the exact start value:
//     consecutive clean groups with one hypothesis join, and their joined summary is applied
//     once; other groups are split into segments of usable same-hypothesis tiles, each applied
//     once; a summary that fails is retried tile by tile; a tile that fails applies its wide
//     table, or, where that does not apply or it has none, is evaluated exactly by the wave
//     (eval_chain).
// The applies prove each result equal to the sequential fold, as in the single-pass form. A
// wrong prediction costs only the re-read of its tile in step 5. On zero-mean data, where the
// running sums keep returning near zero, most tiles that cross a binade take one table apply.
constexpr int kGroupTiles = 64;
constexpr uint64_t kUsableBit = uint64_t(1) << 16;

struct TileSum {
  double d0, d1, lo0, lo1, hi0, hi1;
  uint64_t hword;  // pack_binade(h) | kUsableBit when the summary applies at the approximate start
  uint64_t pad;
};
struct GroupSum {
  double d0, d1, lo0, lo1, hi0, hi1;
  uint64_t hword;  // the group's hypothesis
  uint64_t clean;  // every tile usable under that hypothesis
};


__device__ __forceinline__ RunFn fn_of(double d0, double d1, double lo0, double lo1, double hi0, double hi1) {
  RunFn f;
  f.d[0] = d0;
  f.d[1] = d1;
  f.lo[0] = lo0;
  f.lo[1] = lo1;
  f.hi[0] = hi0;
  f.hi[1] = hi1;
  return f;
}


// The tile summary of chain j from the 16 waves' joined runs: wave 0's lanes join four waves
// each (lane group l / 16), then the four groups (xor 16, 32); every lane of chain j gets it.
__device__ __forceinline__ RunFn tile_tree(const RunFn (&s_f)[kFastWaves][16], int j, int l) {
  const int q = l >> 4;
  RunFn ft = run_join(run_join(s_f[4 * q][j], s_f[4 * q + 1][j]), run_join(s_f[4 * q + 2][j], s_f[4 * q + 3][j]));
  RunFn o = shfl_xor_pod(ft, 16);
  ft = (l & 16) ? run_join(o, ft) : run_join(ft, o);
  o = shfl_xor_pod(ft, 32);
  ft = (l & 32) ? run_join(o, ft) : run_join(ft, o);
  return ft;
}

// 1. Predicted chain sums of a tile, asum[j * tiles + tile]: 16 times the sum of the tile's
//    blocks 0, 16, 32, ...; thread t reads chain t % 16 of blocks 16 (t / 16 + 16 m).
template <class RV>
__global__ __launch_bounds__(kBlock) void k_canon_sample(RV rv, const double *__restrict__ x,
                                                         const double *__restrict__ y, uint64_t blocks,
                                                         uint64_t tiles, double *asum) {
  __shared__ double s_a[kBlock / kWave][16];
  const uint64_t tile = blockIdx.x;
  const uint64_t b0 = tile * kTileBlocks;
  const uint64_t nb = blocks - b0 < uint64_t(kTileBlocks) ? blocks - b0 : kTileBlocks;
  const int t = threadIdx.x, l = lane_id(), j = t & 15;
  double u[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const uint64_t b = 16 * uint64_t((t >> 4) + 16 * m);
    u[m] = b < nb ? rv(x, y, (b0 + b) * 16 + j) : 0.0;
  }
  double a = (u[0] + u[1]) + (u[2] + u[3]);
  a = a + __shfl_xor(a, 16);
  a = a + __shfl_xor(a, 32);
  if (l < 16) s_a[t / kWave][l] = a;
  __syncthreads();
  if (t < 16) asum[t * tiles + tile] = 16.0 * ((s_a[0][t] + s_a[1][t]) + (s_a[2][t] + s_a[3][t]));
}

// Prefix: pre[tile * 16 + j] = the start state of chain j plus the approximate sums of chain j
//    over the tiles before `tile` (any order of these sums serves: the applies prove every
//    result). Blocks take equal tile ranges in ticket order, one wave per chain: a block sums its
//    range per chain, publishes the 16 sums (the canonical family's status layout: aggregates at
//    W_AAGG, flag F_AAGG in word 0), adds the published sums of every block before it, then scans
//    its range 64 tiles at a time and writes each round of 64 x 16 starts as consecutive words
//    through LDS. A block waits only on blocks that took earlier tickets, which publish before
//    they wait on anyone.
constexpr int kPrefixThreads = 1024;
constexpr uint32_t kPrefixMaxBlocks = 64;
__global__ __launch_bounds__(kPrefixThreads) void k_canon_prefix(const double *asum, uint64_t tiles,
                                                                 const double *acc_in, double *pre,
                                                                 uint64_t *status, uint32_t epoch, uint32_t *ticket,
                                                                 uint32_t nblk) {
  static_assert(kPrefixThreads == 16 * kWave, "one wave per chain");
  __shared__ double s_agg[16];
  __shared__ double s_lb[kPrefixMaxBlocks][17];
  __shared__ double s_out[kWave * 16];
  const uint32_t p = take_ticket(ticket, nblk);
  const uint64_t per = (tiles + nblk - 1) / nblk;
  const uint64_t t0 = uint64_t(p) * per < tiles ? uint64_t(p) * per : tiles;
  const uint64_t t1 = t0 + per < tiles ? t0 + per : tiles;
  const int t = threadIdx.x, j = t / kWave, l = lane_id();
  const double *a = asum + uint64_t(j) * tiles;
  double part = 0.0;
  for (uint64_t r0 = t0; r0 < t1; r0 += kWave) part = part + (r0 + l < t1 ? a[r0 + l] : 0.0);
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) part = part + __shfl_xor(part, d);
  if (l == 0) s_agg[j] = part;
  __syncthreads();
  uint64_t *st = status + uint64_t(p) * kStride;
  if (t < 16) st_d(st + W_AAGG + t, s_agg[t]);
  if (t < kWave) {
    publish_fence();
    if (t == 0) st_rlx(st, flag_word(epoch, F_AAGG));
    // lane q reads block q's sums (q < p)
    if (static_cast<uint32_t>(t) < p) {
      const uint64_t *sq = status + uint64_t(t) * kStride;
      wait_flag(sq, epoch, F_AAGG);
#pragma unroll
      for (int k = 0; k < 16; ++k) s_lb[t][k] = ld_d(sq + W_AAGG + k);
    }
  }
  __syncthreads();
  double run = acc_in[j];
  for (uint32_t q = 0; q < p; ++q) run = run + s_lb[q][j];
  // rounds of 64 tiles: lane l's exclusive prefix within the round, then the round's starts
  // written as 64 x 16 consecutive words
  for (uint64_t r0 = t0; r0 < t1; r0 += kWave) {
    const double v = r0 + l < t1 ? a[r0 + l] : 0.0;
    double inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const double o = __shfl_up(inc, d);
      if (l >= d) inc = o + inc;
    }
    s_out[l * 16 + j] = run + (inc - v);
    run = run + __shfl(inc, kWave - 1);
    __syncthreads();
    const uint64_t rows = (t1 - r0 < uint64_t(kWave) ? t1 - r0 : uint64_t(kWave)) * 16;
    if (static_cast<uint64_t>(t) < rows) pre[r0 * 16 + t] = s_out[t];
    __syncthreads();
  }
}

// Launches the prefix over `tiles` tiles: up to kPrefixMaxBlocks blocks of at least 256 tiles.
void canon_prefix(Ctx &c, const double *asum, uint64_t tiles, const double *acc_in, double *pre) {
  uint64_t nb = tiles / 256;
  nb = nb < 1 ? 1 : nb > kPrefixMaxBlocks ? kPrefixMaxBlocks : nb;
  uint64_t *status = status_words(c, ST_CANON, nb * kStride);
  k_canon_prefix<<<uint32_t(nb), kPrefixThreads, 0, c.stream>>>(asum, tiles, acc_in, pre, status, next_epoch(c),
                                                                c.scratch.counters, uint32_t(nb));
}

// 2. Like steps 1, 3 and 4 of the single-pass form, with the hypothesis known at the start;
// also writes the tile's approximate chain sums: each run's sum in order, then runs joined in
// the order of the summary tree. Block 0 clears the two counters of steps 4 and 5.
// With P = RV::kSummPasses passes, a block has 1024 / P threads and thread t summarises runs
// t / 16 + k * 64 / P (k < P) of chain t % 16 in turn. One-input rows take two passes: a block
// needs half the registers of one run per thread and three blocks fit a compute unit. Two-input
// rows (dot) hold twice the loads per run and take one pass over 1024 threads. The runs of wave
// w in pass k are run group w + k * 16 / P, so the 16 groups keep the tile's run order for the
// summary tree.
template <class RV>
constexpr int summ_threads() {
  return 1024 / RV::kSummPasses;
}

template <class RV>
__global__ __launch_bounds__(summ_threads<RV>()) __attribute__((amdgpu_waves_per_eu(RV::kSummWaves, 8))) void k_canon_summ(RV rv, const double *__restrict__ x,
                                                             const double *__restrict__ y, uint64_t blocks,
                                                             const double *pre, uint64_t tiles,
                                                             TileSum *sums, double *asum, uint32_t *counts) {
  __shared__ RunFn s_f[kFastWaves][16];
  __shared__ double s_a[kFastWaves][16];
  const uint64_t tile = blockIdx.x;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const int j = t & 15;
  if (tile == 0 && t < 2) counts[t] = 0;
  const uint64_t b0 = tile * kTileBlocks;
  const int nb = static_cast<int>(blocks - b0 < uint64_t(kTileBlocks) ? blocks - b0 : kTileBlocks);
  const double start = pre[tile * 16 + j];
  const Binade h = binade_of(start);
  constexpr int kThreads = summ_threads<RV>();
#pragma unroll 1
  for (int pass = 0; pass < RV::kSummPasses; ++pass) {
    const int r = (t >> 4) + pass * (kThreads / 16);
    const int cnt = nb - r * kRunBlocks;
    RunFn f = run_identity();
    double a = 0.0;
    {
      double v[kRunBlocks];
      load_run(rv, x, y, b0 + r * kRunBlocks, j, cnt, v);
      // the one-track pass also sums the run; a tie (rare) reloads the run for the two-track
      // form, so the loaded rows die as the pass consumes them
      RunFn f1;
      const bool one = run_of_one(v, cnt, binade_ref(h), &f1, &a);
      if (h.ok) {
        if (__builtin_expect(one, 1)) {
          f = f1;
        } else {
          double w2[kRunBlocks];
          load_run(rv, x, y, b0 + r * kRunBlocks, j, cnt, w2);
          f = run_of_two(w2, cnt, binade_ref(h));
        }
      }
    }
    a = a + __shfl_xor(a, 16);
    a = a + __shfl_xor(a, 32);
    const int grp = w + pass * (kThreads / kWave);
    if (l < 16) s_a[grp][j] = a;
    {
      RunFn o = shfl_xor_pod(f, 16);
      f = (l & 16) ? run_join(o, f) : run_join(f, o);
      o = shfl_xor_pod(f, 32);
      f = (l & 32) ? run_join(o, f) : run_join(f, o);
    }
    if (l < 16) s_f[grp][j] = f;
  }
  __syncthreads();
  if (w != 0) return;
  const RunFn ft = tile_tree(s_f, j, l);
  double a = (s_a[4 * (l >> 4)][j] + s_a[4 * (l >> 4) + 1][j]) + (s_a[4 * (l >> 4) + 2][j] + s_a[4 * (l >> 4) + 3][j]);
  a = a + __shfl_xor(a, 16);
  a = a + __shfl_xor(a, 32);
  if (l < 16) {
    asum[uint64_t(l) * tiles + tile] = a;
    double probe;
    const bool usable = run_apply(ft, h, start, &probe);
    TileSum ts{ft.d[0], ft.d[1], ft.lo[0], ft.lo[1], ft.hi[0], ft.hi[1],
               pack_binade(h) | (usable ? kUsableBit : 0), 0};
    sums[uint64_t(l) * tiles + tile] = ts;
  }
}

// Slot k + 1 for each lane of the wave where `need` holds (0 where it does not, or where the
// rcap slots have run out): one atomic per wave on the slot counter; the slot's (tile, chain)
// key goes to keys[k].
__device__ __forceinline__ uint32_t take_slot(bool need, uint32_t key, uint32_t *counter, uint32_t rcap,
                                              uint32_t *keys) {
  const uint64_t m = __ballot(need);
  if (!m) return 0;
  const int leader = __builtin_ctzll(m);
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(m)));
  base = __shfl(base, leader);
  const uint32_t k = base + static_cast<uint32_t>(__builtin_amdgcn_mbcnt_hi(
                                static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0)));
  if (!need || k >= rcap) return 0;
  keys[k] = key;
  return k + 1;
}

// 4. One thread per tile and chain: a chain whose stored hypothesis is the binade of its
// approximate start has its usability re-checked there, and takes a wide-table slot when it is
// not usable (its TileSum.pad is the slot + 1); a tile with a chain whose hypothesis differs
// is listed for step 5.
__global__ __launch_bounds__(kBlock) void k_canon_check(const double *pre, uint64_t tiles, TileSum *sums,
                                                        uint32_t *list, uint32_t *counts, uint32_t rcap,
                                                        uint32_t *keys) {
  // thread (tile, chain j): j = t % 16, so a tile's chains are 16 lanes of one wave
  const uint64_t g = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
  const uint64_t tile = g / 16;
  const int j = static_cast<int>(g % 16), l = lane_id();
  bool listed = false, need = false;
  TileSum *ts = sums + uint64_t(j) * tiles + tile;
  if (tile < tiles) {
    const double start = pre[tile * 16 + j];
    const Binade h = binade_of(start);
    const uint64_t hw = ts->hword;
    if ((hw & ~kUsableBit) == pack_binade(h)) {
      double probe;
      const bool usable = run_apply(fn_of(ts->d0, ts->d1, ts->lo0, ts->lo1, ts->hi0, ts->hi1), h, start, &probe);
      const uint64_t nw = pack_binade(h) | (usable ? kUsableBit : 0);
      if (nw != hw) ts->hword = nw;
      need = !usable;
    } else {
      listed = true;
    }
  }
  const uint32_t slot = take_slot(need, static_cast<uint32_t>(tile * 16 + j), counts + 1, rcap, keys);
  if (slot) ts->pad = slot;
  const uint64_t any = __ballot(listed);
  if (tile < tiles && j == 0 && ((any >> (l & ~15)) & 0xFFFF)) list[atomicAdd(counts, 1u)] = static_cast<uint32_t>(tile);
}

// 5. The listed tiles, one block each (a grid-stride loop over the list): chains whose
// hypothesis differs are summarised again under the binade of their approximate start, and
// those whose summary is then not usable take a wide-table slot.
template <class RV>
__global__ __launch_bounds__(kFastThreads) void k_canon_fix(RV rv, const double *__restrict__ x,
                                                            const double *__restrict__ y, uint64_t blocks,
                                                            const double *pre, uint64_t tiles, TileSum *sums,
                                                            const uint32_t *list, uint32_t *counts, uint32_t rcap,
                                                            uint32_t *keys) {
  __shared__ RunFn s_f[kFastWaves][16];
  __shared__ int s_mis[16];
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const int j = t & 15, r = t >> 4;
  const uint32_t n = counts[0];
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint64_t tile = list[i];
    const uint64_t b0 = tile * kTileBlocks;
    const int nb = static_cast<int>(blocks - b0 < uint64_t(kTileBlocks) ? blocks - b0 : kTileBlocks);
    const double start = pre[tile * 16 + j];
    const Binade h = binade_of(start);
    if (t < 16) s_mis[t] = (sums[uint64_t(t) * tiles + tile].hword & ~kUsableBit) != pack_binade(h);
    __syncthreads();
    const bool mis = s_mis[j] != 0;
    RunFn f = run_identity();
    if (mis && h.ok) {
      double v[kRunBlocks];
      const int cnt = nb - r * kRunBlocks;
      load_run(rv, x, y, b0 + r * kRunBlocks, j, cnt, v);
      f = run_of(v, cnt, binade_ref(h));
    }
    {
      RunFn o = shfl_xor_pod(f, 16);
      f = (l & 16) ? run_join(o, f) : run_join(f, o);
      o = shfl_xor_pod(f, 32);
      f = (l & 32) ? run_join(o, f) : run_join(f, o);
    }
    if (l < 16) s_f[w][j] = f;
    __syncthreads();
    if (w == 0) {
      const RunFn ft = tile_tree(s_f, j, l);
      TileSum *ts = sums + uint64_t(j) * tiles + tile;
      bool need = false;
      if (l < 16 && s_mis[l]) {
        double probe;
        const bool usable = run_apply(ft, h, start, &probe);
        *ts = TileSum{ft.d[0], ft.d[1], ft.lo[0], ft.lo[1], ft.hi[0], ft.hi[1],
                      pack_binade(h) | (usable ? kUsableBit : 0), 0};
        need = !usable;
      }
      const uint32_t slot = take_slot(need, static_cast<uint32_t>(tile * 16 + j), counts + 1, rcap, keys);
      if (slot) ts->pad = slot;
    }
    __syncthreads();
  }
}

// 5b. The wide tables (canon_binade.h) of the chains that took a slot, four chains per wave
// (one wave per block): the four chains' elements are first staged in LDS, every lane loading
// 64 of them in one round trip, then lane 16 q + r simulates residue r of slot s0 + q over the
// tile's blocks from Z_r = B + r u0, reading the elements from LDS (past the tile's end they
// are zeros, which leave a nonzero running value as it is, so the loop bound is uniform). The
// slot's 16 lanes then agree on M and keep 2^(M + 1) entries; the other entries never apply.
constexpr int kWideSlots = 4;
template <class RV>
__global__ __launch_bounds__(kWave) void k_canon_wide(RV rv, const double *__restrict__ x,
                                                      const double *__restrict__ y, uint64_t blocks,
                                                      const double *pre, const uint32_t *keys,
                                                      const uint32_t *counts, uint32_t rcap, WideHead *heads,
                                                      WideEnt *ents) {
  __shared__ double s_v[kWideSlots][kTileBlocks];
  const uint32_t nslots = counts[1] < rcap ? counts[1] : rcap;
  const int l = lane_id(), q = l >> 4, r = l & 15;
  for (uint32_t s0 = blockIdx.x * kWideSlots; s0 < nslots; s0 += gridDim.x * kWideSlots) {
    const uint32_t s = s0 + q;
    const bool live = s < nslots;
    const uint32_t key = live ? keys[s] : 0;
    const uint64_t tile = key >> 4;
    const int j = static_cast<int>(key & 15);
    const double a = live ? pre[tile * 16 + j] : 0.0;
    const Binade h = binade_of(a);
    const bool go = live && h.ok;
    const uint64_t b0 = tile * kTileBlocks;
    const int nb = go ? static_cast<int>(blocks - b0 < uint64_t(kTileBlocks) ? blocks - b0 : kTileBlocks) : 0;
    // stage: lane l loads elements l + 64 m of each of the four chains
    {
      double st[kWideSlots][kTileBlocks / kWave];
#pragma unroll
      for (int qq = 0; qq < kWideSlots; ++qq) {
        const uint64_t bq = __shfl(b0, 16 * qq);
        const int jq = __shfl(j, 16 * qq), nq = __shfl(nb, 16 * qq);
#pragma unroll
        for (int m = 0; m < kTileBlocks / kWave; ++m) {
          const int k = l + kWave * m;
          st[qq][m] = k < nq ? rv(x, y, (bq + k) * 16 + jq) : 0.0;
        }
      }
#pragma unroll
      for (int qq = 0; qq < kWideSlots; ++qq)
#pragma unroll
        for (int m = 0; m < kTileBlocks / kWave; ++m) s_v[qq][l + kWave * m] = st[qq][m];
    }
    __syncthreads();
    const double base = go ? wide_base(a, h) : 1.0;
    WideAcc acc = wide_start(base + r * pow2(go ? h.e - 52 : -52));
#pragma unroll 2
    for (int k0 = 0; k0 < kTileBlocks; k0 += 16) {
      double v[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = s_v[q][k0 + m];
      wide_push_n(acc, v);
    }
    __syncthreads();
    const WideEnt e = wide_end(acc);
    int mm = e.lo <= e.hi ? acc.emax - h.e : -2000;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const int o = __shfl_xor(mm, d);
      mm = o > mm ? o : mm;
    }
    const int nres = go ? wide_nres(mm) : 0;
    if (live) {
      // every entry is written: the walk stages all 16
      const double inf = __builtin_inf();
      if (r == 0) heads[s] = WideHead{base, h.e, nres};
      ents[uint64_t(s) * kWideRes + r] = r < nres ? e : WideEnt{0.0, inf, -inf};
    }
  }
}

// Ordered join of the 64 lanes' summaries (lane order is tile order); every lane gets it.
__device__ __forceinline__ RunFn join_lanes(RunFn f, int l) {
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const RunFn o = shfl_xor_pod(f, m);
    f = (l & m) ? run_join(o, f) : run_join(f, o);
  }
  return f;
}

// 6. One wave per (chain, group of 64 tiles).
__global__ __launch_bounds__(kBlock) void k_canon_groups(const TileSum *sums, uint64_t tiles, uint64_t groups,
                                                         GroupSum *gsum) {
  const uint64_t wid = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / kWave;
  if (wid >= 16 * groups) return;
  const int l = lane_id();
  const uint64_t j = wid / groups, g = wid % groups;
  const uint64_t tile = g * kGroupTiles + l;
  const bool valid = tile < tiles;
  RunFn f = run_identity();
  uint64_t hw = 0;
  if (valid) {
    const TileSum ts = sums[j * tiles + tile];
    f = fn_of(ts.d0, ts.d1, ts.lo0, ts.lo1, ts.hi0, ts.hi1);
    hw = ts.hword;
  }
  const uint64_t h0 = __shfl(hw, 0);
  const bool ok = !valid || (hw == h0 && (hw & kUsableBit));
  const bool clean = __ballot(!ok) == 0;
  f = join_lanes(f, l);
  if (l == 0) {
    GroupSum gs{f.d[0], f.d[1], f.lo[0], f.lo[1], f.hi[0], f.hi[1], h0 & ~kUsableBit, clean ? 1ull : 0ull};
    gsum[j * groups + g] = gs;
  }
}

// No record of the exact values before each lane.
struct NoRec {
  __device__ void operator()(int, int, double) const {}
};

// Lanes 0 .. n - 1 of the wave in order from the exact value cur: lane i holds a summary f
// under the hypothesis word hw (kUsableBit: it applied at its approximate start). Segments of
// consecutive usable lanes with one hypothesis are applied at once (with `singles`, every lane
// is its own segment); a segment that fails goes lane by lane, and a lane whose summary does
// not apply goes through one(i, cur), which returns the exact value after lane i.
// rec(a, b, v) receives the exact value before every lane: lanes a .. b each hold their own v.
// Inside a segment applied at once, lane i's value is the segment's prefix before it (the
// exclusive scan) applied to the segment's start; that apply always holds, because the
// prefix's running range lies inside the whole segment's, which held.
template <class One, class Rec = NoRec>
__device__ double walk_lanes(const RunFn &f, uint64_t hw, int n, bool singles, double cur, One &&one,
                             Rec &&rec = Rec{}) {
  const int l = lane_id();
  const bool valid = l < n;
  const bool usable = valid && (hw & kUsableBit);
  const uint64_t hprev = __shfl_up(hw, 1);
  const bool uprev = __shfl_up(static_cast<int>(usable), 1) != 0;
  // a segment starts at lane 0, at an unusable lane, after one, and where the hypothesis changes
  const bool start = l == 0 || !usable || !uprev || hw != hprev || singles;
  // segmented inclusive scan: (f, s) then (f', s') = s' ? (f', 1) : (f . f', s)
  RunFn sc = f;
  int ss = start ? 1 : 0;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const RunFn o = shfl_up_pod(sc, d);
    const int os = __shfl_up(ss, d);
    if (l >= d && !ss) {
      sc = run_join(o, sc);
      ss = os;
    }
  }
  const uint64_t starts = __ballot(start && valid);
  // the exclusive prefix within the segment (for the records): the previous lane's scan
  constexpr bool kRec = !std::is_same<std::decay_t<Rec>, NoRec>::value;
  RunFn scx = run_identity();
  if (kRec) scx = shfl_up_pod(sc, 1);
  // segment ends: the lane before each start, and lane n - 1
  uint64_t ends = ((starts >> 1) | (uint64_t(1) << (n - 1))) & (n == kWave ? ~uint64_t(0) : (uint64_t(1) << n) - 1);
  int first = 0;
  while (ends) {
    const int e = __builtin_ctzll(ends);
    ends &= ends - 1;
    const uint64_t hw_e = bcast_pod(hw, e);
    double out;
    if ((hw_e & kUsableBit) && run_apply(bcast_pod(sc, e), unpack_binade(hw_e & ~kUsableBit), cur, &out)) {
      if (kRec) {
        double v = cur;
        if (l > first && l <= e) run_apply(scx, unpack_binade(hw_e & ~kUsableBit), cur, &v);
        rec(first, e, v);
      }
      cur = out;
    } else {
      for (int i = first; i <= e; ++i) {
        rec(i, i, cur);
        const uint64_t hw_i = bcast_pod(hw, i);
        if (!((hw_i & kUsableBit) && run_apply(bcast_pod(f, i), unpack_binade(hw_i & ~kUsableBit), cur, &out)))
          out = one(i, cur);
        cur = out;
      }
    }
    first = e + 1;
  }
  return cur;
}

// The tiles of group g from the exact value cur: lane = tile, through walk_lanes; a tile whose
// summary does not apply goes by its wide table, staged in LDS (s_ent[lane][r]) when the group is
// loaded, or, where that does not apply or it has none, through eval_chain.
template <class RV, class Rec = NoRec>
__device__ double walk_group(const RV &rv, const double *x, const double *y, uint64_t blocks,
                             const TileSum *sums, uint64_t tiles, const WideHead *heads, const WideEnt *ents,
                             WideEnt (*s_ent)[kWideRes], int j, uint64_t g, double cur, bool singles,
                             Rec &&rec = Rec{}) {
  const int l = lane_id();
  const uint64_t tile = g * kGroupTiles + l;
  const int n = tiles - g * kGroupTiles < uint64_t(kWave) ? static_cast<int>(tiles - g * kGroupTiles) : kWave;
  RunFn f = run_identity();
  uint64_t hw = 0, slot = 0;
  WideHead wh{0.0, 0, 0};
  if (l < n) {
    const TileSum ts = sums[uint64_t(j) * tiles + tile];
    f = fn_of(ts.d0, ts.d1, ts.lo0, ts.lo1, ts.hi0, ts.hi1);
    hw = ts.hword;
    slot = ts.pad;
  }
  if (slot) {
    wh = heads[slot - 1];
    WideEnt e[kWideRes];
#pragma unroll
    for (int r = 0; r < kWideRes; ++r) e[r] = ents[(slot - 1) * kWideRes + r];
#pragma unroll
    for (int r = 0; r < kWideRes; ++r) s_ent[l][r] = e[r];
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const double out = walk_lanes(f, hw, n, singles, cur, [&](int i, double c) {
    const uint64_t tb = (g * kGroupTiles + i) * kTileBlocks;
    const int nb = static_cast<int>(blocks - tb < uint64_t(kTileBlocks) ? blocks - tb : kTileBlocks);
    double o;
    if (bcast_pod(slot, i) && wide_apply(bcast_pod(wh, i), s_ent[i], c, &o)) return o;
    return eval_chain(rv, x, y, tb, nb, j, c);
  }, rec);
  __builtin_amdgcn_wave_barrier();
  return out;
}

// 7. One wave per chain: groups 64 at a time through walk_lanes (lane = group; a clean group's
// joined summary is usable under its hypothesis), an unclean group, or a clean one whose
// summary fails, through walk_group.
template <class RV>
__global__ __launch_bounds__(kWave) void k_canon_walk(RV rv, const double *__restrict__ x,
                                                      const double *__restrict__ y, uint64_t blocks,
                                                      const TileSum *sums, uint64_t tiles, const WideHead *heads,
                                                      const WideEnt *ents, const GroupSum *gsum, uint64_t groups,
                                                      double *acc, double *gstart) {
  __shared__ WideEnt s_ent[kWave][kWideRes];
  const int j = blockIdx.x, l = lane_id();
  double cur = acc[j];
  for (uint64_t g0 = 0; g0 < groups; g0 += kWave) {
    const int n = groups - g0 < uint64_t(kWave) ? static_cast<int>(groups - g0) : kWave;
    RunFn f = run_identity();
    uint64_t hw = 0;
    if (l < n) {
      const GroupSum gs = gsum[uint64_t(j) * groups + g0 + l];
      f = fn_of(gs.d0, gs.d1, gs.lo0, gs.lo1, gs.hi0, gs.hi1);
      hw = gs.hword | (gs.clean ? kUsableBit : 0);
    }
    auto one = [&](int i, double c) {
      return walk_group(rv, x, y, blocks, sums, tiles, heads, ents, s_ent, j, g0 + i, c, false);
    };
    if (gstart) {
      cur = walk_lanes(f, hw, n, false, cur, one, [&](int a, int b, double v) {
        if (l >= a && l <= b) gstart[(g0 + l) * 16 + j] = v;
      });
    } else {
      cur = walk_lanes(f, hw, n, false, cur, one);
    }
  }
  if (l == 0) acc[j] = cur;
}

// For scans: one wave per (group, chain) from the group's exact start (recorded by the walk)
// through walk_group, recording the exact start of every tile, tstart[tile * 16 + j].
template <class RV>
__global__ __launch_bounds__(kWave) void k_canon_expand(RV rv, const double *__restrict__ x,
                                                        const double *__restrict__ y, uint64_t blocks,
                                                        const TileSum *sums, uint64_t tiles, const WideHead *heads,
                                                        const WideEnt *ents, uint64_t groups, const double *gstart,
                                                        double *tstart) {
  __shared__ WideEnt s_ent[kWave][kWideRes];
  const uint64_t g = blockIdx.x / 16;
  const int j = static_cast<int>(blockIdx.x % 16), l = lane_id();
  if (g >= groups) return;
  walk_group(rv, x, y, blocks, sums, tiles, heads, ents, s_ent, j, g, gstart[g * 16 + j], false,
             [&](int a, int b, double v) {
               if (l >= a && l <= b) tstart[(g * kGroupTiles + l) * 16 + j] = v;
             });
}

template <class RV>
__global__ void k_canon_rest(RV rv, const double *x, const double *y, uint64_t rest,
                             double *tail) {
  const uint64_t i = threadIdx.x;
  if (i < rest) tail[i] = rv(x, y, i);
}

// finish_lanes (reduce.rs:75-91): pairwise +8, +4, +2, +1, then the tail into acc 0.
__global__ void k_canon_finish(bool mul, const double *acc_in, const double *tail, int tail_len,
                               double *out) {
  if (threadIdx.x != 0) return;
  double acc[16];
  for (int j = 0; j < 16; ++j) acc[j] = acc_in[j];
  for (int step = 8; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) acc[j] = mul ? acc[j] * acc[j + step] : acc[j] + acc[j + step];
  double a = acc[0];
  for (int j = 0; j < tail_len; ++j) a = mul ? a * tail[j] : a + tail[j];
  *out = a;
}

__global__ void k_fill_d(double *p, uint64_t n, double v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// One thread per item: finish_lanes (reduce.rs:75-91) over the item's accumulators and its
// tail rows 16 * (n / 16) .. n.
template <class RV>
__global__ void k_canon_batch_finish(RV rv, bool mul, const BatchItem *items, uint64_t nitems,
                                     const double *acc, double *out) {
  const uint64_t it = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (it >= nitems) return;
  const BatchItem &I = items[it];
  const double *x = static_cast<const double *>(I.a);
  const double *y = I.b ? static_cast<const double *>(I.b) : x;
  double a[16];
  for (int j = 0; j < 16; ++j) a[j] = acc[it * 16 + j];
  for (int step = 8; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) a[j] = mul ? a[j] * a[j + step] : a[j] + a[j + step];
  double s = a[0];
  for (uint64_t r = I.n / 16 * 16; r < I.n; ++r) s = mul ? s * rv(x, y, r) : s + rv(x, y, r);
  out[it] = s;
}

__global__ void k_canon_init(double *acc, double id) {
  if (threadIdx.x < 16) acc[threadIdx.x] = id;
}

// Calls f with the Row of the spec's kind (products read plain rows).
template <class F>
void with_row(const CanonSpec &s, F &&f) {
  // comparison codes 0 eq, 1 ne, 2 lt, 3 le, 4 gt, 5 ge as the outcome sets they accept
  static const int keep[6] = {2, 1 | 4 | 8, 1, 1 | 2, 4, 4 | 2};
  switch (s.kind) {
    case 1: f(Row<1>{s.c, s.e, 0}); break;
    case 2: f(Row<2>{s.c, s.e, 0}); break;
    case 3: f(Row<3>{s.c, s.e, keep[s.cmp >= 0 && s.cmp < 6 ? s.cmp : 5]}); break;
    default: f(Row<0>{s.c, s.e, 0}); break;
  }
}

}  // namespace

namespace {

// The streamed form (below kStreamMinTiles tiles only when tstart is asked for). With tstart,
// the exact start of every tile and chain is written there (tstart[tile * 16 + j]).
template <class RV>
void canon_streamed(Ctx &c, RV rv, const double *x, const double *y, uint64_t blocks, double *acc,
                    double *tstart);

template <class RV>
void canon_blocks_as(Ctx &c, RV rv, bool mul, const double *x, const double *y, uint64_t blocks,
                     double *acc) {
  if (mul || blocks < kFastMinBlocks) {
    // products have no binade form; short inputs are cheaper as one sequential wave
    k_canon_chain<<<1, 64, 0, c.stream>>>(rv, mul, x, y, blocks, acc);
    check_hip(hipGetLastError(), "canon chain");
    return;
  }
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  if (tiles >= kStreamMinTiles) {
    canon_streamed(c, rv, x, y, blocks, acc, nullptr);
    return;
  }
  uint64_t *status = status_words(c, ST_CANON, tiles * kStride);
  const uint32_t epoch = next_epoch(c);
  double *acc_in = reinterpret_cast<double *>(c.scratch.slots + 32);  // 16 words: the start state
  copy_d2d(c, acc_in, acc, 16 * 8);
  k_canon_fast<RV, false><<<uint32_t(tiles), kFastThreads, 0, c.stream>>>(rv, x, y, blocks, acc_in, acc,
                                                               status, epoch, c.scratch.counters,
                                                               uint32_t(tiles), nullptr, nullptr);
  check_hip(hipGetLastError(), "canon fast");
}

template <class RV>
void canon_streamed(Ctx &c, RV rv, const double *x, const double *y, uint64_t blocks, double *acc,
                    double *tstart) {
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  {
    const uint64_t groups = (tiles + kGroupTiles - 1) / kGroupTiles;
    const uint64_t pairs = tiles * 16;
    const uint32_t rcap = static_cast<uint32_t>(pairs < 8192 + pairs / 4 ? pairs : 8192 + pairs / 4);
    double *asum = static_cast<double *>(tmp_alloc(c, tiles * 16 * 8));
    double *pre = static_cast<double *>(tmp_alloc(c, tiles * 16 * 8));
    auto *sums = static_cast<TileSum *>(tmp_alloc(c, tiles * 16 * sizeof(TileSum)));
    auto *gsum = static_cast<GroupSum *>(tmp_alloc(c, groups * 16 * sizeof(GroupSum)));
    auto *keys = static_cast<uint32_t *>(tmp_alloc(c, uint64_t(rcap) * 4));
    auto *heads = static_cast<WideHead *>(tmp_alloc(c, uint64_t(rcap) * sizeof(WideHead)));
    auto *ents = static_cast<WideEnt *>(tmp_alloc(c, uint64_t(rcap) * kWideRes * sizeof(WideEnt)));
    auto *list = static_cast<uint32_t *>(tmp_alloc(c, tiles * 4 + 8));
    uint32_t *counts = list + tiles;  // [0] listed tiles, [1] record slots taken
    const uint32_t fix_grid = static_cast<uint32_t>(tiles < 512 ? tiles : 512);
    k_canon_sample<RV><<<uint32_t(tiles), kBlock, 0, c.stream>>>(rv, x, y, blocks, tiles, asum);
    canon_prefix(c, asum, tiles, acc, pre);
    k_canon_summ<RV><<<uint32_t(tiles), summ_threads<RV>(), 0, c.stream>>>(rv, x, y, blocks, pre, tiles, sums, asum,
                                                                    counts);
    canon_prefix(c, asum, tiles, acc, pre);
    k_canon_check<<<uint32_t((tiles * 16 + kBlock - 1) / kBlock), kBlock, 0, c.stream>>>(pre, tiles, sums, list,
                                                                                       counts, rcap, keys);
    k_canon_fix<RV><<<fix_grid, kFastThreads, 0, c.stream>>>(rv, x, y, blocks, pre, tiles, sums, list, counts, rcap,
                                                            keys);
    const uint32_t wide_grid = static_cast<uint32_t>((rcap + kWideSlots - 1) / kWideSlots < 4096
                                                          ? (rcap + kWideSlots - 1) / kWideSlots
                                                          : 4096);
    k_canon_wide<RV><<<wide_grid, kWave, 0, c.stream>>>(rv, x, y, blocks, pre, keys, counts, rcap, heads, ents);
    k_canon_groups<<<uint32_t((16 * groups * kWave + kBlock - 1) / kBlock), kBlock, 0, c.stream>>>(sums, tiles, groups,
                                                                                              gsum);
    double *gstart = tstart ? static_cast<double *>(tmp_alloc(c, groups * 16 * 8)) : nullptr;
    k_canon_walk<RV><<<16, kWave, 0, c.stream>>>(rv, x, y, blocks, sums, tiles, heads, ents, gsum, groups, acc,
                                                 gstart);
    if (tstart)
      k_canon_expand<RV><<<uint32_t(groups * 16), kWave, 0, c.stream>>>(rv, x, y, blocks, sums, tiles, heads, ents,
                                                                       groups, gstart, tstart);
    check_hip(hipGetLastError(), "canon streamed");
    if (gstart) tmp_free(c, gstart);
    tmp_free(c, asum);
    tmp_free(c, pre);
    tmp_free(c, sums);
    tmp_free(c, gsum);
    tmp_free(c, keys);
    tmp_free(c, heads);
    tmp_free(c, ents);
    tmp_free(c, list);
  }
}

}  // namespace

void canon_init(Ctx &c, bool mul, double *acc) {
  k_canon_init<<<1, 64, 0, c.stream>>>(acc, mul ? 1.0 : 0.0);
  check_hip(hipGetLastError(), "canon init");
}

void canon_blocks(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t blocks,
                  double *acc) {
  if (blocks == 0) return;
  with_row(s, [&](auto rv) { canon_blocks_as(c, rv, s.kind == 4, x, y ? y : x, blocks, acc); });
}

// Batched canonical sums: every item's 16 chains start at the identity; a finish kernel
// combines each item's accumulators pairwise and folds its tail rows.
void canon_batch(Ctx &c, const CanonSpec &s, std::vector<BatchItem> &items, double *d_out) {
  const uint64_t nitems = items.size();
  if (nitems == 0) return;
  double *acc = static_cast<double *>(tmp_alloc(c, nitems * 16 * 8));
  double *zero = static_cast<double *>(tmp_alloc(c, nitems * 16 * 8));
  const double id = s.kind == 4 ? 1.0 : 0.0;
  k_fill_d<<<stream_grid(nitems * 16, kBlock, 4096), kBlock, 0, c.stream>>>(acc, nitems * 16, id);
  k_fill_d<<<stream_grid(nitems * 16, kBlock, 4096), kBlock, 0, c.stream>>>(zero, nitems * 16, id);
  if (s.kind == 4) usage_error("batched products are not supported");
  std::vector<BatchItem> plan_items(items);
  bool small = false;
  for (auto &it : plan_items) {  // tiles cover whole blocks of the larger items only
    small = small || (it.n < kChainItemRows && it.n >= 16);
    it.n = it.n < kChainItemRows ? 0 : it.n / 16 * 16;
  }
  BatchPlan p = batch_plan(c, plan_items, uint64_t(kTileBlocks) * 16);
  BatchPlan q = batch_plan(c, items, ~uint64_t(0) >> 1);  // the item table (full rows)
  if (small) {
    with_row(s, [&](auto rv) {
      k_canon_chain_batch<<<uint32_t(nitems < 65535 ? nitems : 65535), 64, 0, c.stream>>>(rv, q.items, nitems, acc);
    });
    check_hip(hipGetLastError(), "canon batch chains");
  }
  if (p.ntiles > 0) {
    uint64_t *status = status_words(c, ST_CANON, uint64_t(p.ntiles) * kStride);
    const uint32_t epoch = next_epoch(c);
    with_row(s, [&](auto rv) {
      k_canon_fast<decltype(rv), true><<<p.ntiles, kFastThreads, 0, c.stream>>>(
          rv, nullptr, nullptr, 0, zero, acc, status, epoch, c.scratch.counters, p.ntiles, p.items, p.tile_item);
    });
    check_hip(hipGetLastError(), "canon batch");
  }
  with_row(s, [&](auto rv) {
    k_canon_batch_finish<<<stream_grid(nitems, kBlock, 4096), kBlock, 0, c.stream>>>(rv, s.kind == 4, q.items,
                                                                                    nitems, acc, d_out);
  });
  check_hip(hipGetLastError(), "canon batch finish");
  batch_plan_free(c, p);
  batch_plan_free(c, q);
  tmp_free(c, acc);
  tmp_free(c, zero);
}

// The exact start of every 1024-block tile of chain j (tstart[tile * 16 + j]) for the scan's
// fold of x's blocks from acc, which ends holding the accumulators after the last block.
void canon_tile_starts(Ctx &c, const double *x, uint64_t blocks, double *acc, double *tstart) {
  if (blocks == 0) return;
  canon_streamed(c, Row<0>{0.0, 0.0, 0}, x, x, blocks, acc, tstart);
}

uint64_t canon_tile_blocks() { return kTileBlocks; }

uint64_t canon_tile_min_rows() { return kChainItemRows; }

// Batched canonical sums over device-planned tiles (plan_tiles_dev with unit 16 and
// canon_tile_blocks() blocks per tile): acc and zero are nitems x 16 doubles of scratch.
void canon_batch_dev(Ctx &c, const CanonSpec &s, const DevPlan &p, uint32_t nitems, double *acc, double *zero,
                     double *d_out) {
  if (nitems == 0) return;
  if (s.kind == 4) usage_error("batched products are not supported");
  k_fill_d<<<stream_grid(uint64_t(nitems) * 16, kBlock, 4096), kBlock, 0, c.stream>>>(acc, uint64_t(nitems) * 16, 0.0);
  k_fill_d<<<stream_grid(uint64_t(nitems) * 16, kBlock, 4096), kBlock, 0, c.stream>>>(zero, uint64_t(nitems) * 16, 0.0);
  with_row(s, [&](auto rv) {
    k_canon_chain_batch<<<nitems < 65535 ? nitems : 65535, 64, 0, c.stream>>>(rv, p.items, nitems, acc);
    if (p.max_tiles > 0)
      k_canon_fast<decltype(rv), true><<<p.max_tiles, kFastThreads, 0, c.stream>>>(
          rv, nullptr, nullptr, 0, zero, acc, p.status, 0, p.ticket, p.max_tiles, p.items, p.tile_item, p.ntiles,
          p.epoch);
    k_canon_batch_finish<<<stream_grid(nitems, kBlock, 4096), kBlock, 0, c.stream>>>(rv, false, p.items, nitems, acc,
                                                                                    d_out);
  });
  check_hip(hipGetLastError(), "canon batch (device plan)");
}

uint64_t canon_status_words_per_tile() { return kStride; }

void canon_rest(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t rest,
                double *tail) {
  if (rest == 0) return;
  with_row(s, [&](auto rv) { k_canon_rest<<<1, 64, 0, c.stream>>>(rv, x, y ? y : x, rest, tail); });
  check_hip(hipGetLastError(), "canon rest");
}

namespace {

// The row value of a fused region: the chain over rows x (and y when TWO), in registers.
template <bool TWO>
struct RowChain {
  static constexpr int kSummPasses = TWO ? 1 : 2;
  static constexpr int kSummWaves = TWO ? 1 : 6;
  static constexpr bool kTwo = TWO;
  static constexpr bool kMapped = true;
  ChainSpec ch;
  __device__ __forceinline__ double operator()(const double *x, const double *y, uint64_t i) const {
    return chain_eval<double>(ch, x[i], TWO ? y[i] : 0.0);
  }
  __device__ __forceinline__ double map(double x, double y) const { return chain_eval<double>(ch, x, y); }
};

}  // namespace

void canon_blocks_chain(Ctx &c, const ChainSpec &ch, const double *x, const double *y, uint64_t blocks,
                        double *acc) {
  if (blocks == 0) return;
  if (y) canon_blocks_as(c, RowChain<true>{ch}, false, x, y, blocks, acc);
  else canon_blocks_as(c, RowChain<false>{ch}, false, x, x, blocks, acc);
}

void canon_rest_chain(Ctx &c, const ChainSpec &ch, const double *x, const double *y, uint64_t rest,
                      double *tail) {
  if (rest == 0) return;
  if (y) k_canon_rest<<<1, 64, 0, c.stream>>>(RowChain<true>{ch}, x, y, rest, tail);
  else k_canon_rest<<<1, 64, 0, c.stream>>>(RowChain<false>{ch}, x, x, rest, tail);
  check_hip(hipGetLastError(), "canon rest (chain)");
}

void canon_finish(Ctx &c, bool mul, const double *acc, const double *tail, int tail_len,
                  double *out) {
  k_canon_finish<<<1, 64, 0, c.stream>>>(mul, acc, tail, tail_len, out);
  check_hip(hipGetLastError(), "canon finish");
}

}  // namespace vk
