// Canonical-order float reductions: reduce_add_f64 / reduce_mul_f64 (CanonicalAcc,
// src/kernels/scalar/reduce.rs:249-323), dot_f64, sumsq_f64 and condsum_<cmp>_f64
// (fused.rs:72-210). Accumulator j folds rows j, j+16, ... in index order with plain IEEE
// adds (or multiplies); no FMA anywhere (built with -ffp-contract=off).
//
// Two forms. The sequential form is one wave with lane j < 16 owning accumulator j; its bound
// is the dependent-add latency, n/16 adds in a row. The fast form (sums only) evaluates the
// same 16 chains in parallel with binade run summaries (canon_binade.h) and returns the same
// bits; it is HBM-bound on data whose running sums stay inside a binade for long stretches.
#include "canon_binade.h"
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

struct RowVal {
  int kind;  // 0 sum, 1 dot, 2 sumsq, 3 condsum, 4 mul
  int cmp;
  double c, e;
  __device__ __forceinline__ double operator()(const double *x, const double *y, uint64_t i) const {
    switch (kind) {
      case 1: return x[i] * y[i];
      case 2: return x[i] * x[i];
      case 3: { double v = x[i]; return cmp_test(cmp, v, c) ? v : e; }
      default: return x[i];
    }
  }
};

constexpr int kGroup = 8;  // 64-row chunks loaded ahead of the dependent chain

// acc[16] <- fold of rows [0, 16*blocks) in the canonical order.
__global__ __launch_bounds__(64) void k_canon_chain(RowVal rv, bool mul, const double *__restrict__ x,
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

// ------------------------------------------------------------------ fast form
// The 16 chains over tiles of 256 blocks (4096 rows), one pass (canon_binade.h):
//  1. thread (chain j, run r) loads the 16 blocks 16r..16r+15 of its chain (4 x 128-byte
//     segments per wave load) and keeps the row values in LDS for a sequential fallback;
//  2. per-chain approximate sums, and a decoupled look-back over them, give each chain an
//     approximate start and so a binade hypothesis;
//  3. every thread summarises its run under the hypothesis, runs join in order (two
//     shuffles, then across the four waves) into the tile summary;
//  4. a second decoupled look-back joins the summaries of preceding tiles with the same
//     hypothesis back to a tile whose exact end values are published, applies the joined
//     summary to that exact value, and publishes this tile's exact end values. A failed
//     check waits for the previous tile's exact values and folds the tile's rows of the
//     chain sequentially from LDS.
constexpr int kTileBlocks = 256;
constexpr uint64_t kFastMinBlocks = 512;
constexpr int kStride = 192;  // status words per tile
constexpr int W_AAGG = 1, W_AINC = 17, W_F = 33, W_EXACT = 161;
constexpr uint64_t F_AAGG = 1, F_AINC = 2, F_FN = 4, F_EXACT = 8;

__device__ __forceinline__ uint64_t flag_word(uint32_t epoch, uint64_t bits) {
  return (uint64_t(epoch) << 8) | bits;
}
__device__ __forceinline__ uint64_t wait_flag(const uint64_t *st, uint32_t epoch, uint64_t need) {
  while (true) {
    const uint64_t w = ld_acq(st);
    if ((w >> 8) == epoch && (w & need)) return w;
    backoff();
  }
}
__device__ __forceinline__ double ld_d(const uint64_t *p) {
  return __longlong_as_double(static_cast<long long>(ld_rlx(p)));
}
__device__ __forceinline__ void st_d(uint64_t *p, double v) {
  st_rlx(p, static_cast<uint64_t>(__double_as_longlong(v)));
}
__device__ __forceinline__ uint64_t pack_binade(const Binade &h) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(h.e + 2048)) << 2) | (uint64_t(h.neg) << 1) |
         uint64_t(h.ok);
}
__device__ __forceinline__ Binade unpack_binade(uint64_t w) {
  return Binade{static_cast<int>(w >> 2) - 2048, static_cast<int>((w >> 1) & 1), static_cast<int>(w & 1), 0};
}

// BATCH: the tile's item gives the rows (a, b), the block count (n / 16), the first tile, and
// the item's 16 start and end accumulators at acc_in / acc_out + 16 * item.
template <bool BATCH>
__global__ __launch_bounds__(kBlock) void k_canon_fast(RowVal rv, const double *__restrict__ x,
                                                       const double *__restrict__ y, uint64_t blocks,
                                                       const double *acc_in, double *acc_out,
                                                       uint64_t *status, uint32_t epoch,
                                                       uint32_t *ticket, uint32_t ntiles,
                                                       const BatchItem *items,
                                                       const uint32_t *tile_item) {
  __shared__ double s_val[kTileBlocks * 16];
  __shared__ double s_a[kBlock / kWave][16];
  __shared__ RunFn s_f[kBlock / kWave][16];
  __shared__ Binade s_h[16];
  const uint32_t tile = take_ticket(ticket, ntiles);
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
  const uint64_t b0 = uint64_t(tile - first) * kTileBlocks;
  const int nb = static_cast<int>(blocks - b0 < uint64_t(kTileBlocks) ? blocks - b0 : kTileBlocks);
  uint64_t *st = status + uint64_t(tile) * kStride;

  // 1. load the run
  double v[16];
  double a = 0.0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int lb = r * 16 + k;
    if (lb < nb) {
      v[k] = rv(x, y, (b0 + lb) * 16 + j);
      s_val[lb * 16 + j] = v[k];
      a = a + v[k];
    } else {
      v[k] = 0.0;
    }
  }
  // 2. approximate chain sums of the tile
  a = a + __shfl_xor(a, 16);
  a = a + __shfl_xor(a, 32);
  if (l < 16) s_a[w][j] = a;
  __syncthreads();
  if (w == 0) {
    double atile = 0.0, pre = 0.0;
    if (l < 16) {
      atile = s_a[0][l] + s_a[1][l] + s_a[2][l] + s_a[3][l];
      if (tile == first) {
        pre = acc_in[l];
        st_d(st + W_AINC + l, pre + atile);
      } else {
        st_d(st + W_AAGG + l, atile);
      }
    }
    __threadfence();
    if (l == 0) st_rel(st, flag_word(epoch, tile == first ? (F_AAGG | F_AINC) : F_AAGG));
    if (tile > first && l < 16) {
      double sum = 0.0;
      for (int64_t k = int64_t(tile) - 1;; --k) {
        const uint64_t *sk = status + uint64_t(k) * kStride;
        const uint64_t fw = wait_flag(sk, epoch, F_AAGG | F_AINC | F_EXACT);
        if (fw & F_EXACT) {
          sum = sum + ld_d(sk + W_EXACT + l);
          break;
        }
        if (fw & F_AINC) {
          sum = sum + ld_d(sk + W_AINC + l);
          break;
        }
        sum = sum + ld_d(sk + W_AAGG + l);
      }
      pre = sum;
      st_d(st + W_AINC + l, pre + atile);
    }
    if (tile > first) {
      __threadfence();
      if (l == 0) st_rel(st, flag_word(epoch, F_AAGG | F_AINC));
    }
    if (l < 16) s_h[l] = binade_of(pre);
  }
  __syncthreads();
  // 3. summarise the run under the chain's hypothesis, join the runs in order
  const Binade h = s_h[j];
  RunFn f = run_identity();
  if (h.ok) {
    const BinadeRef ref = binade_ref(h);
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (r * 16 + k < nb) run_push(f, ref, v[k]);
  }
  {
    RunFn o = shfl_xor_pod(f, 16);
    f = (l & 16) ? run_join(o, f) : run_join(f, o);
    o = shfl_xor_pod(f, 32);
    f = (l & 32) ? run_join(o, f) : run_join(f, o);
  }
  if (l < 16) s_f[w][j] = f;
  __syncthreads();
  if (w != 0) return;
  // 4. publish the tile summary, then the exact look-back (lane = chain)
  RunFn ft = run_identity();
  Binade hj{0, 0, 0, 0};
  if (l < 16) {
    hj = s_h[l];
    ft = run_join(run_join(s_f[0][l], s_f[1][l]), run_join(s_f[2][l], s_f[3][l]));
    uint64_t *fp = st + W_F + 8 * l;
    st_d(fp + 0, ft.d[0]);
    st_d(fp + 1, ft.d[1]);
    st_d(fp + 2, ft.lo[0]);
    st_d(fp + 3, ft.lo[1]);
    st_d(fp + 4, ft.hi[0]);
    st_d(fp + 5, ft.hi[1]);
    st_d(fp + 6, ft.bad);
    st_rlx(fp + 7, pack_binade(hj));
  }
  __threadfence();
  if (l == 0) st_rel(st, flag_word(epoch, F_AAGG | F_AINC | F_FN));
  double exact = 0.0;
  if (l < 16) {
    RunFn g = run_identity();  // joined summaries of the tiles between `base` and this one
    double base = 0.0;
    int64_t k = int64_t(tile) - 1;
    for (;; --k) {
      if (k < int64_t(first)) {
        base = acc_in[l];
        break;
      }
      const uint64_t *sk = status + uint64_t(k) * kStride;
      const uint64_t fw = wait_flag(sk, epoch, F_FN | F_EXACT);
      if (fw & F_EXACT) {
        base = ld_d(sk + W_EXACT + l);
        break;
      }
      const uint64_t *fp = sk + W_F + 8 * l;
      const Binade hk = unpack_binade(ld_rlx(fp + 7));
      if (!same_binade(hk, hj)) {
        wait_flag(sk, epoch, F_EXACT);
        base = ld_d(sk + W_EXACT + l);
        break;
      }
      RunFn fk;
      fk.d[0] = ld_d(fp + 0);
      fk.d[1] = ld_d(fp + 1);
      fk.lo[0] = ld_d(fp + 2);
      fk.lo[1] = ld_d(fp + 3);
      fk.hi[0] = ld_d(fp + 4);
      fk.hi[1] = ld_d(fp + 5);
      fk.bad = ld_d(fp + 6);
      fk.pad = 0.0;
      g = run_join(fk, g);
    }
    bool done = run_apply(run_join(g, ft), hj, base, &exact);
    if (!done) {
      if (k != int64_t(tile) - 1 && tile > first) {
        const uint64_t *sp = status + uint64_t(tile - 1) * kStride;
        wait_flag(sp, epoch, F_EXACT);
        base = ld_d(sp + W_EXACT + l);
        done = run_apply(ft, hj, base, &exact);
      }
      if (!done) {
        double s = base;
        for (int lb = 0; lb < nb; ++lb) s = s + s_val[lb * 16 + l];
        exact = s;
      }
    }
    st_d(st + W_EXACT + l, exact);
    if (tile == last) acc_out[l] = exact;
  }
  __threadfence();
  if (l == 0) st_rel(st, flag_word(epoch, F_AAGG | F_AINC | F_FN | F_EXACT));
}

__global__ void k_canon_rest(RowVal rv, const double *x, const double *y, uint64_t rest,
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
__global__ void k_canon_batch_finish(RowVal rv, const BatchItem *items, uint64_t nitems,
                                     const double *acc, double *out) {
  const uint64_t it = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (it >= nitems) return;
  const BatchItem &I = items[it];
  const double *x = static_cast<const double *>(I.a);
  const double *y = I.b ? static_cast<const double *>(I.b) : x;
  const bool mul = rv.kind == 4;
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

}  // namespace

void canon_init(Ctx &c, bool mul, double *acc) {
  k_canon_init<<<1, 64, 0, c.stream>>>(acc, mul ? 1.0 : 0.0);
  check_hip(hipGetLastError(), "canon init");
}

void canon_blocks(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t blocks,
                  double *acc) {
  if (blocks == 0) return;
  RowVal rv{s.kind, s.cmp, s.c, s.e};
  if (s.kind == 4 || blocks < kFastMinBlocks) {
    // products have no binade form; short inputs are cheaper as one sequential wave
    k_canon_chain<<<1, 64, 0, c.stream>>>(rv, s.kind == 4, x, y ? y : x, blocks, acc);
    check_hip(hipGetLastError(), "canon chain");
    return;
  }
  const uint64_t tiles = (blocks + kTileBlocks - 1) / kTileBlocks;
  uint64_t *status = status_words(c, ST_CANON, tiles * kStride);
  const uint32_t epoch = next_epoch(c);
  double *acc_in = reinterpret_cast<double *>(c.scratch.slots + 32);  // 16 words: the start state
  copy_d2d(c, acc_in, acc, 16 * 8);
  k_canon_fast<false><<<uint32_t(tiles), kBlock, 0, c.stream>>>(rv, x, y ? y : x, blocks, acc_in, acc,
                                                               status, epoch, c.scratch.counters,
                                                               uint32_t(tiles), nullptr, nullptr);
  check_hip(hipGetLastError(), "canon fast");
}

// Batched canonical sums: every item's 16 chains start at the identity; a finish kernel
// combines each item's accumulators pairwise and folds its tail rows.
void canon_batch(Ctx &c, const CanonSpec &s, std::vector<BatchItem> &items, double *d_out) {
  const uint64_t nitems = items.size();
  if (nitems == 0) return;
  RowVal rv{s.kind, s.cmp, s.c, s.e};
  double *acc = static_cast<double *>(tmp_alloc(c, nitems * 16 * 8));
  double *zero = static_cast<double *>(tmp_alloc(c, nitems * 16 * 8));
  const double id = s.kind == 4 ? 1.0 : 0.0;
  k_fill_d<<<stream_grid(nitems * 16, kBlock, 4096), kBlock, 0, c.stream>>>(acc, nitems * 16, id);
  k_fill_d<<<stream_grid(nitems * 16, kBlock, 4096), kBlock, 0, c.stream>>>(zero, nitems * 16, id);
  std::vector<BatchItem> plan_items(items);
  for (auto &it : plan_items) it.n = it.n / 16 * 16;  // tiles cover whole blocks only
  BatchPlan p = batch_plan(c, plan_items, uint64_t(kTileBlocks) * 16);
  if (p.ntiles > 0) {
    if (s.kind == 4) usage_error("batched products are not supported");
    uint64_t *status = status_words(c, ST_CANON, uint64_t(p.ntiles) * kStride);
    const uint32_t epoch = next_epoch(c);
    k_canon_fast<true><<<p.ntiles, kBlock, 0, c.stream>>>(rv, nullptr, nullptr, 0, zero, acc, status, epoch,
                                                         c.scratch.counters, p.ntiles, p.items, p.tile_item);
    check_hip(hipGetLastError(), "canon batch");
  }
  BatchPlan q = batch_plan(c, items, 1);  // the full rows, for the tails
  k_canon_batch_finish<<<stream_grid(nitems, kBlock, 4096), kBlock, 0, c.stream>>>(rv, q.items, nitems, acc,
                                                                                  d_out);
  check_hip(hipGetLastError(), "canon batch finish");
  batch_plan_free(c, p);
  batch_plan_free(c, q);
  tmp_free(c, acc);
  tmp_free(c, zero);
}

void canon_rest(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t rest,
                double *tail) {
  if (rest == 0) return;
  RowVal rv{s.kind, s.cmp, s.c, s.e};
  k_canon_rest<<<1, 64, 0, c.stream>>>(rv, x, y ? y : x, rest, tail);
  check_hip(hipGetLastError(), "canon rest");
}

void canon_finish(Ctx &c, bool mul, const double *acc, const double *tail, int tail_len,
                  double *out) {
  k_canon_finish<<<1, 64, 0, c.stream>>>(mul, acc, tail, tail_len, out);
  check_hip(hipGetLastError(), "canon finish");
}

}  // namespace vk
