// Canonical-order f32 reductions: reduce_add_f32 / reduce_mul_f32 (CanonicalAcc instantiated
// at f32, src/kernels/scalar/reduce.rs:249-323 and 567-575), dot_f32, sumsq_f32 and
// condsum_<cmp>_f32 (fused.rs:72-210 with T = f32). Accumulator j folds rows j, j + 16, ... in
// index order; every add (or multiply) rounds to f32, and there is no FMA (built with
// -ffp-contract=off).
//
// Values are carried in doubles that hold f32 values exactly; a chain step rounds through
// PrecF32::rnd, which equals the f32 operation (canon_binade.h). Products of two f32 values are
// exact in a double, so rounding them once to f32 is the f32 product.
//
// Two forms:
//  - sequential: one wave, lane j < 16 owning accumulator j (products, and inputs of fewer
//    than kC32MinBlocks blocks of 16 rows);
//  - streamed (sums): tiles of 16384 rows (1024 blocks). (1) the approximate sum of every tile
//    chain, (2) their prefix per chain: the approximate start of every tile chain, (3) every
//    tile chain summarised under the binade of its approximate start (run summaries of
//    canon_binade.h at f32 precision, joined in order), (4) one wave per chain walks the tiles
//    from the exact start, 64 tiles at a time: the summaries of the tiles whose hypothesis is the
//    running value's binade are joined by a lane scan and applied; a tile whose summary does not
//    apply goes by its wide table (canon_binade.h, built in parallel for every tile chain whose
//    summary fails at its approximate start: the tiles where the running value crosses a binade),
//    or, where that does not apply either, is folded from its rows in f32 arithmetic (the rows of
//    the first kC32RowSlots slots are kept chain-major by the table kernel, so such a tile chain
//    is read with 16-byte loads instead of one 64-byte line per row). Every applied summary or
//    table proves its result, so the result is the sequential one.
// HBM traffic: two reads of the rows (4 bytes each, 8 for dot) plus 64 bytes per tile chain.
#include "canon_binade.h"
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

// Row i's value in a canonical f32 reduction of kind K: x (0: sums, and the factors of
// products), x * y (1: dot), x * x (2: sumsq), x where its comparison with c holds, else e
// (3: condsum); the f32 result held in a double.
template <int K>
struct Row32 {
  float c, e;
  uint32_t keep;  // condsum: the comparison outcomes (outcome()) that keep x
  __device__ __forceinline__ double operator()(const float *x, const float *y, uint64_t i) const {
    if constexpr (K == 1) {
      const float p = x[i] * y[i];
      return p;
    } else if constexpr (K == 2) {
      const float v = x[i];
      const float p = v * v;
      return p;
    } else if constexpr (K == 3) {
      const float v = x[i];
      return (keep & outcome(v, c)) ? v : e;
    } else {
      return x[i];
    }
  }
};

__device__ __forceinline__ double r32(double v) { return PrecF32::rnd(v); }

constexpr int kC32Threads = 1024;
constexpr int kC32Run = 16;                                   // blocks per run
constexpr uint64_t kC32TileBlocks = kC32Threads / 16 * kC32Run;  // 1024 blocks, 16384 rows
constexpr uint64_t kC32MinBlocks = 2 * kC32TileBlocks;
constexpr uint32_t kC32RowSlots = 32768;  // slots whose rows k_c32_wide also keeps, chain-major (128 MB)
constexpr uint64_t kC32GroupTiles = 64;  // tiles per group of the walk

struct TileFn32 {
  double d0, d1, lo0, lo1, hi0, hi1;
  uint64_t hword;  // the hypothesis (pack below)
  uint64_t slot;   // wide-table slot + 1, or 0
};

__device__ __forceinline__ uint64_t pack_h(const Binade &h) {
  return (static_cast<uint64_t>(static_cast<uint32_t>(h.e + 2048)) << 2) | (uint64_t(h.neg) << 1) | uint64_t(h.ok);
}
__device__ __forceinline__ Binade unpack_h(uint64_t w) {
  return Binade{static_cast<int>(w >> 2) - 2048, static_cast<int>((w >> 1) & 1), static_cast<int>(w & 1), 0};
}
__device__ __forceinline__ RunFn run_bad() {
  const double inf = __builtin_inf();
  return RunFn{{0.0, 0.0}, {-inf, -inf}, {inf, inf}};
}

// acc[16] <- fold of rows [0, 16 * blocks) in the canonical order, one wave (as k_canon_chain).
constexpr int kGroup32 = 8;
template <class RV>
__global__ __launch_bounds__(64) void k_c32_chain(RV rv, bool mul, const float *__restrict__ x,
                                                   const float *__restrict__ y, uint64_t blocks, double *acc) {
  const int l = lane_id();
  const uint64_t rows = blocks * 16;
  double a = l < 16 ? acc[l] : 0.0;
  const uint64_t chunks = (rows + 63) / 64;
  const uint64_t groups = (chunks + kGroup32 - 1) / kGroup32;
  if (groups == 0) return;
  double cur[kGroup32], nxt[kGroup32];
  auto load = [&](double *v, uint64_t g) {
#pragma unroll
    for (int k = 0; k < kGroup32; ++k) {
      const uint64_t r = (g * kGroup32 + k) * 64 + l;
      v[k] = r < rows ? rv(x, y, r) : 0.0;
    }
  };
  load(cur, 0);
  for (uint64_t g = 0; g < groups; ++g) {
    if (g + 1 < groups) load(nxt, g + 1);
#pragma unroll
    for (int k = 0; k < kGroup32; ++k) {
      const uint64_t base = (g * kGroup32 + k) * 64;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double v = __shfl(cur[k], q * 16 + (l & 15));
        if (l < 16 && base + q * 16 < rows) a = mul ? r32(a * v) : r32(a + v);
      }
    }
#pragma unroll
    for (int k = 0; k < kGroup32; ++k) cur[k] = nxt[k];
  }
  if (l < 16) acc[l] = a;
}

// Thread t of a tile: chain j = t % 16, run r = t / 16 (blocks 16 r .. 16 r + 15 of the tile).
template <class RV>
__device__ __forceinline__ void load_run32(const RV &rv, const float *x, const float *y, uint64_t b0, int nb,
                                           int r, int j, double (&v)[kC32Run]) {
#pragma unroll
  for (int k = 0; k < kC32Run; ++k) {
    const int lb = r * kC32Run + k;
    v[k] = lb < nb ? rv(x, y, (b0 + lb) * 16 + j) : 0.0;
  }
}

// (1) asum[tile * 16 + j]: the approximate sum of chain j's rows in the tile.
template <class RV>
__global__ __launch_bounds__(kC32Threads) void k_c32_asum(RV rv, const float *__restrict__ x,
                                                           const float *__restrict__ y, uint64_t blocks,
                                                           double *asum) {
  const uint64_t tile = blockIdx.x;
  const uint64_t b0 = tile * kC32TileBlocks;
  const int nb = static_cast<int>(blocks - b0 < kC32TileBlocks ? blocks - b0 : kC32TileBlocks);
  const int t = threadIdx.x, j = t & 15, r = t >> 4, w = t / kWave, l = lane_id();
  double v[kC32Run];
  load_run32(rv, x, y, b0, nb, r, j, v);
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kC32Run; ++k) s = s + v[k];
  s = s + __shfl_xor(s, 16);
  s = s + __shfl_xor(s, 32);
  __shared__ double part[kC32Threads / kWave][16];
  if (l < 16) part[w][j] = s;
  __syncthreads();
  if (t < 16) {
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < kC32Threads / kWave; ++q) a = a + part[q][t];
    asum[tile * 16 + t] = a;
  }
}

// (2) pre[tile * 16 + j] = acc[j] + the approximate sums of chain j's earlier tiles; one block
// per chain, each thread a contiguous range of tiles.
__global__ __launch_bounds__(kC32Threads) void k_c32_prefix(const double *asum, uint64_t tiles, const double *acc,
                                                             double *pre) {
  const int j = blockIdx.x, t = threadIdx.x;
  const uint64_t per = (tiles + kC32Threads - 1) / kC32Threads;
  const uint64_t a = t * per < tiles ? t * per : tiles, b = a + per < tiles ? a + per : tiles;
  double s = 0.0;
  for (uint64_t q = a; q < b; ++q) s = s + asum[q * 16 + j];
  __shared__ double sc[kC32Threads];
  sc[t] = s;
  __syncthreads();
  for (int d = 1; d < kC32Threads; d <<= 1) {
    const double o = t >= d ? sc[t - d] : 0.0;
    __syncthreads();
    sc[t] = sc[t] + o;
    __syncthreads();
  }
  double run = acc[j] + (sc[t] - s);  // exclusive
  for (uint64_t q = a; q < b; ++q) {
    pre[q * 16 + j] = run;
    run = run + asum[q * 16 + j];
  }
}

// (3) fns[tile * 16 + j]: chain j's tile summary under the binade of its approximate start; a
// summary that does not apply at that start takes a wide-table slot (counter, keys: tile * 16 + j)
// while fewer than rcap are taken.
template <class RV>
__global__ __launch_bounds__(kC32Threads) void k_c32_summ(RV rv, const float *__restrict__ x,
                                                           const float *__restrict__ y, uint64_t blocks,
                                                           const double *pre, TileFn32 *fns, uint32_t *counter,
                                                           uint32_t rcap, uint32_t *keys) {
  const uint64_t tile = blockIdx.x;
  const uint64_t b0 = tile * kC32TileBlocks;
  const int nb = static_cast<int>(blocks - b0 < kC32TileBlocks ? blocks - b0 : kC32TileBlocks);
  const int t = threadIdx.x, j = t & 15, r = t >> 4, w = t / kWave, l = lane_id();
  double v[kC32Run];
  load_run32(rv, x, y, b0, nb, r, j, v);
  const Binade h = binade_of<PrecF32>(pre[tile * 16 + j]);
  RunFn f = run_bad();
  if (h.ok) {
    const int cnt = nb - r * kC32Run;
    f = run_of<kC32Run, PrecF32>(v, cnt < 0 ? 0 : cnt, binade_ref<PrecF32>(h));
  }
  // lanes j, 16 + j, 32 + j, 48 + j hold runs 4w .. 4w + 3 of chain j: join them in order
  RunFn o = shfl_xor_pod(f, 16);
  f = (l & 16) ? run_join(o, f) : run_join(f, o);
  o = shfl_xor_pod(f, 32);
  f = (l & 32) ? run_join(o, f) : run_join(f, o);
  __shared__ RunFn s_f[kC32Threads / kWave][16];
  if (l < 16) s_f[w][j] = f;
  __syncthreads();
  if (t < 16) {
    RunFn g = s_f[0][t];
    for (int q = 1; q < kC32Threads / kWave; ++q) g = run_join(g, s_f[q][t]);
    const double start = pre[tile * 16 + t];
    double probe;
    uint64_t slot = 0;
    if (h.ok && !run_apply<PrecF32>(g, h, start, &probe)) {
      const uint32_t q = atomicAdd(counter, 1u);
      if (q < rcap) {
        keys[q] = static_cast<uint32_t>(tile * 16 + t);
        slot = uint64_t(q) + 1;
      }
    }
    TileFn32 out{g.d[0], g.d[1], g.lo[0], g.lo[1], g.hi[0], g.hi[1], pack_h(h), slot};
    fns[tile * 16 + t] = out;
  }
}

// (3b) The wide tables of the tile chains that took a slot, four per wave: the four chains'
// rows staged in LDS, then lane 16 q + r simulates residue r of chain q from Z_r = B + r u0 at
// f32 precision (as k_canon_wide does at f64).
constexpr int kC32WideSlots = 4;
template <class RV>
__global__ __launch_bounds__(kWave) void k_c32_wide(RV rv, const float *__restrict__ x, const float *__restrict__ y,
                                                     uint64_t blocks, const double *pre, const uint32_t *keys,
                                                     const uint32_t *counter, uint32_t rcap, WideHead *heads,
                                                     WideEnt *ents, float *rows) {
  __shared__ double s_v[kC32WideSlots][kC32TileBlocks];
  const uint32_t nslots = *counter < rcap ? *counter : rcap;
  const int l = lane_id(), q = l >> 4, r = l & 15;
  for (uint32_t s0 = blockIdx.x * kC32WideSlots; s0 < nslots; s0 += gridDim.x * kC32WideSlots) {
    const uint32_t sl = s0 + q;
    const bool live = sl < nslots;
    const uint32_t key = live ? keys[sl] : 0;
    const uint64_t tile = key >> 4;
    const int j = static_cast<int>(key & 15);
    const double a = live ? pre[tile * 16 + j] : 0.0;
    const Binade h = binade_of<PrecF32>(a);
    const bool go = live && h.ok;
    const uint64_t b0 = tile * kC32TileBlocks;
    const int nb = live ? static_cast<int>(blocks - b0 < kC32TileBlocks ? blocks - b0 : kC32TileBlocks) : 0;
    {
      double st[kC32WideSlots][kC32TileBlocks / kWave];
#pragma unroll
      for (int qq = 0; qq < kC32WideSlots; ++qq) {
        const uint64_t bq = __shfl(b0, 16 * qq);
        const int jq = __shfl(j, 16 * qq), nq = __shfl(nb, 16 * qq);
#pragma unroll
        for (int m = 0; m < static_cast<int>(kC32TileBlocks / kWave); ++m) {
          const int k = l + kWave * m;
          st[qq][m] = k < nq ? rv(x, y, (bq + k) * 16 + jq) : 0.0;
        }
      }
#pragma unroll
      for (int qq = 0; qq < kC32WideSlots; ++qq)
#pragma unroll
        for (int m = 0; m < static_cast<int>(kC32TileBlocks / kWave); ++m) s_v[qq][l + kWave * m] = st[qq][m];
      // the first kC32RowSlots slots keep their chain's rows (zero past the tile's end) for the
      // walk, which evaluates such a tile chain from them when its table does not apply
#pragma unroll
      for (int qq = 0; qq < kC32WideSlots; ++qq) {
        const uint32_t sq = s0 + qq;
        if (sq < nslots && sq < kC32RowSlots) {
#pragma unroll
          for (int m = 0; m < static_cast<int>(kC32TileBlocks / kWave); ++m)
            rows[uint64_t(sq) * kC32TileBlocks + l + kWave * m] = static_cast<float>(st[qq][m]);
        }
      }
    }
    __syncthreads();
    const double base = go ? wide_base<PrecF32>(a, h) : 1.0;
    WideAcc acc = wide_start(base + r * pow2(go ? h.e - PrecF32::kMant : -PrecF32::kMant));
#pragma unroll 2
    for (int k0 = 0; k0 < static_cast<int>(kC32TileBlocks); k0 += 16) {
      double v[16];
#pragma unroll
      for (int m = 0; m < 16; ++m) v[m] = s_v[q][k0 + m];
      wide_push_n<16, PrecF32>(acc, v);
    }
    __syncthreads();
    const WideEnt e = wide_end<PrecF32>(acc);
    int mm = e.lo <= e.hi ? acc.emax - h.e : -2000;
#pragma unroll
    for (int d = 1; d < 16; d <<= 1) {
      const int o = __shfl_xor(mm, d);
      mm = o > mm ? o : mm;
    }
    const int nres = go ? wide_nres(mm) : 0;
    if (live) {
      const double inf = __builtin_inf();
      if (r == 0) heads[sl] = WideHead{base, h.e, nres};
      ents[uint64_t(sl) * kWideRes + r] = r < nres ? e : WideEnt{0.0, inf, -inf};
    }
  }
}

// The exact fold of chain j's rows in blocks b0 .. b0 + nb - 1 (nb <= 1024) from the exact value
// s, by one wave: the rows (kept: as k_c32_wide kept them for the tile chain's slot; else read
// from x and y) staged in LDS in block order, then the sequential fold in f32 arithmetic (the same
// operations: every value is an f32 value), the next 16 rows' LDS loads in flight while 16 are
// added. Rounds of run summaries from the exact value (the first failing run folded element by
// element, the rest summarised again) measured slower than this fold at 12 cycles a row: on
// normal(0,1) rows at 2^28, 36.5 ms with up to three rounds, 29.0 ms without.
template <class RV>
__device__ double eval_tile32(const RV &rv, const float *x, const float *y, uint64_t b0, int nb, int j,
                              double s, float *s_f, const float *kept) {
  const int l = lane_id();
  float vf[kC32Run];
  if (kept) {
#pragma unroll
    for (int k = 0; k < kC32Run; k += 4) {
      const float4 q = *reinterpret_cast<const float4 *>(kept + l * kC32Run + k);
      vf[k] = q.x;
      vf[k + 1] = q.y;
      vf[k + 2] = q.z;
      vf[k + 3] = q.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kC32Run; ++k) {
      const int lb = l * kC32Run + k;
      vf[k] = lb < nb ? static_cast<float>(rv(x, y, (b0 + lb) * 16 + j)) : 0.0f;
    }
  }
#pragma unroll
  for (int k = 0; k < kC32Run; k += 4)
    *reinterpret_cast<float4 *>(s_f + l * kC32Run + k) = make_float4(vf[k], vf[k + 1], vf[k + 2], vf[k + 3]);
  __syncthreads();
  float cf = static_cast<float>(s);
  int i = 0;
  const int full = nb / kC32Run * kC32Run;
  if (i < full) {
    const float4 *p4 = reinterpret_cast<const float4 *>(s_f);
    float4 a[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) a[u] = p4[u];
    for (; i + kC32Run < full; i += kC32Run) {
      float4 b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) b[u] = p4[(i + kC32Run) / 4 + u];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        cf = cf + a[u].x;
        cf = cf + a[u].y;
        cf = cf + a[u].z;
        cf = cf + a[u].w;
        a[u] = b[u];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      cf = cf + a[u].x;
      cf = cf + a[u].y;
      cf = cf + a[u].z;
      cf = cf + a[u].w;
    }
    i += kC32Run;
  }
  for (; i < nb; ++i) cf = cf + s_f[i];
  __syncthreads();
  return cf;
}

// Segmented walk of up to 64 lanes from the exact value s (lane i: summary f of unit i under
// the hypothesis word hw, bit 0 set when there is one): a segmented scan joins each run of
// consecutive lanes with one hypothesis word; segment by segment the joined summary is applied
// at once, and a segment whose summary does not apply goes lane by lane, a lane whose own
// summary does not apply going through one(i, s).
template <class One>
__device__ double walk_lanes32(const RunFn &fl, uint64_t hw, int n, double s, One &&one) {
  const int l = lane_id();
  const bool valid = l < n;
  const bool usable = valid && (hw & 1);
  const uint64_t hprev = __shfl_up(hw, 1);
  const bool uprev = __shfl_up(static_cast<int>(usable), 1) != 0;
  const bool start = l == 0 || !usable || !uprev || hw != hprev;
  RunFn sc = fl;
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
  uint64_t ends = ((starts >> 1) | (uint64_t(1) << (n - 1))) & (n == kWave ? ~uint64_t(0) : (uint64_t(1) << n) - 1);
  int first = 0;
  while (ends) {
    const int e = __builtin_ctzll(ends);
    ends &= ends - 1;
    const uint64_t he = bcast_pod(hw, e);
    double out;
    if ((he & 1) && run_apply<PrecF32>(bcast_pod(sc, e), unpack_h(he), s, &out)) {
      s = out;
    } else {
      for (int i = first; i <= e; ++i) {
        const uint64_t hi = bcast_pod(hw, i);
        if ((hi & 1) && run_apply<PrecF32>(bcast_pod(fl, i), unpack_h(hi), s, &out)) s = out;
        else s = one(i, s);
      }
    }
    first = e + 1;
  }
  return s;
}

// The tiles t0 .. t0 + n - 1 of chain j from the exact value s; a tile whose summary does not
// apply takes its wide table (s_e: the group's tables, kWideRes entries per tile) or, failing
// that, its rows.
template <class RV>
__device__ double walk_tiles32(const RV &rv, const float *x, const float *y, uint64_t blocks, const TileFn32 *fns,
                               const WideHead *heads, const WideEnt *ents, const float *rows, WideEnt *s_e,
                               float *s_f, int j, uint64_t t0, int n, double s) {
  const int l = lane_id();
  TileFn32 q{};
  if (l < n) q = fns[(t0 + l) * 16 + j];
  const RunFn fl{{q.d0, q.d1}, {q.lo0, q.lo1}, {q.hi0, q.hi1}};
  const uint64_t hw = l < n ? q.hword : 0;
  const uint64_t slot = l < n ? q.slot : 0;
  // every table of the group's tiles staged in LDS at once (lane = tile), so that applying one
  // costs no memory round trip
  WideHead wh{0.0, 0, 0};
  if (slot) {
    wh = heads[slot - 1];
    WideEnt e[kWideRes];
#pragma unroll
    for (int r = 0; r < kWideRes; ++r) e[r] = ents[(slot - 1) * kWideRes + r];
#pragma unroll
    for (int r = 0; r < kWideRes; ++r) s_e[l * kWideRes + r] = e[r];
  }
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_wave_barrier();
  const double res = walk_lanes32(fl, hw, n, s, [&](int i, double c) {
    const uint64_t sl = bcast_pod(slot, i);
    double out;
    if (sl && wide_apply<PrecF32>(bcast_pod(wh, i), s_e + i * kWideRes, c, &out)) return out;
    const uint64_t b0 = (t0 + i) * kC32TileBlocks;
    const int nb = static_cast<int>(blocks - b0 < kC32TileBlocks ? blocks - b0 : kC32TileBlocks);
    const float *kept = sl && sl - 1 < kC32RowSlots ? rows + (sl - 1) * kC32TileBlocks : nullptr;
    return eval_tile32(rv, x, y, b0, nb, j, c, s_f, kept);
  });
  __builtin_amdgcn_wave_barrier();
  return res;
}

// (3c) One wave per (chain j, group of 64 tiles): the group's tiles joined in order into
// gfn[j * groups + g], with the group's hypothesis word when every tile has that one (else 0).
__global__ __launch_bounds__(kBlock) void k_c32_groups(const TileFn32 *fns, uint64_t tiles, uint64_t groups,
                                                        TileFn32 *gfn) {
  const uint64_t wid = (blockIdx.x * uint64_t(blockDim.x) + threadIdx.x) / kWave;
  if (wid >= 16 * groups) return;
  const int l = lane_id();
  const uint64_t j = wid / groups, g = wid % groups;
  const uint64_t tile = g * kC32GroupTiles + l;
  const bool valid = tile < tiles;
  RunFn f = run_identity();
  uint64_t hw = 0;
  if (valid) {
    const TileFn32 q = fns[tile * 16 + j];
    f = RunFn{{q.d0, q.d1}, {q.lo0, q.lo1}, {q.hi0, q.hi1}};
    hw = q.hword;
  }
  const uint64_t h0 = __shfl(hw, 0);
  const bool clean = __ballot(valid && (hw != h0 || !(hw & 1))) == 0;
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const RunFn o = shfl_xor_pod(f, m);
    f = (l & m) ? run_join(o, f) : run_join(f, o);
  }
  if (l == 0) gfn[j * groups + g] = TileFn32{f.d[0], f.d[1], f.lo[0], f.lo[1], f.hi[0], f.hi[1], clean ? h0 : 0, 0};
}

// (4) One wave per chain: acc[j] <- the exact fold of chain j over every tile, 64 groups of 64
// tiles at a time; a group whose joined summary does not apply goes tile by tile.
template <class RV>
__global__ __launch_bounds__(kWave) void k_c32_walk(RV rv, const float *__restrict__ x,
                                                     const float *__restrict__ y, uint64_t blocks,
                                                     const TileFn32 *fns, uint64_t tiles, const TileFn32 *gfn,
                                                     uint64_t groups, const WideHead *heads, const WideEnt *ents,
                                                     const float *rows, double *acc) {
  __shared__ WideEnt s_e[kWave * kWideRes];
  __shared__ __align__(16) float s_f[kC32TileBlocks];
  const int j = blockIdx.x, l = lane_id();
  double s = acc[j];
  for (uint64_t g0 = 0; g0 < groups; g0 += kWave) {
    const int n = groups - g0 < uint64_t(kWave) ? static_cast<int>(groups - g0) : kWave;
    TileFn32 q{};
    if (l < n) q = gfn[uint64_t(j) * groups + g0 + l];
    const RunFn fl{{q.d0, q.d1}, {q.lo0, q.lo1}, {q.hi0, q.hi1}};
    const uint64_t hw = l < n ? q.hword : 0;
    s = walk_lanes32(fl, hw, n, s, [&](int i, double c) {
      const uint64_t t0 = (g0 + i) * kC32GroupTiles;
      const int nt = tiles - t0 < kC32GroupTiles ? static_cast<int>(tiles - t0) : static_cast<int>(kC32GroupTiles);
      return walk_tiles32(rv, x, y, blocks, fns, heads, ents, rows, s_e, s_f, j, t0, nt, c);
    });
  }
  if (l == 0) acc[j] = s;
}

template <class RV>
__global__ void k_c32_rest(RV rv, const float *x, const float *y, uint64_t rest, double *tail) {
  const uint64_t i = threadIdx.x;
  if (i < rest) tail[i] = rv(x, y, i);
}

// finish_lanes (reduce.rs:75-91) at f32: pairwise +8, +4, +2, +1, then the tail into acc 0;
// out receives the f32 bits.
__global__ void k_c32_finish(bool mul, const double *acc_in, const double *tail, int tail_len, uint64_t *out) {
  if (threadIdx.x != 0) return;
  double acc[16];
  for (int j = 0; j < 16; ++j) acc[j] = acc_in[j];
  for (int step = 8; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) acc[j] = mul ? r32(acc[j] * acc[j + step]) : r32(acc[j] + acc[j + step]);
  double a = acc[0];
  for (int j = 0; j < tail_len; ++j) a = mul ? r32(a * tail[j]) : r32(a + tail[j]);
  const float f = static_cast<float>(a);
  uint32_t b;
  __builtin_memcpy(&b, &f, 4);
  *out = b;
}

template <class F>
void with_row32(const CanonSpec &s, F &&f) {
  const float c = static_cast<float>(s.c), e = static_cast<float>(s.e);
  switch (s.kind) {
    case 1: f(Row32<1>{c, e, 0}); break;
    case 2: f(Row32<2>{c, e, 0}); break;
    case 3: f(Row32<3>{c, e, keepset_of(s.cmp)}); break;
    default: f(Row32<0>{c, e, 0}); break;
  }
}

template <class RV>
void c32_blocks_as(Ctx &c, RV rv, bool mul, const float *x, const float *y, uint64_t blocks, double *acc) {
  if (mul || blocks < kC32MinBlocks) {
    k_c32_chain<<<1, 64, 0, c.stream>>>(rv, mul, x, y, blocks, acc);
    check_hip(hipGetLastError(), "canon f32 chain");
    return;
  }
  const uint64_t tiles = (blocks + kC32TileBlocks - 1) / kC32TileBlocks;
  if (tiles > 0xFFFFFFFFull) usage_error("reduction too large");
  // wide-table slots: every tile chain when there are few, else 8192 and a quarter of them
  const uint64_t pairs = tiles * 16;
  const uint32_t rcap = static_cast<uint32_t>(pairs < 8192 + pairs / 4 ? pairs : 8192 + pairs / 4);
  double *asum = static_cast<double *>(tmp_alloc(c, tiles * 16 * 8));
  double *pre = static_cast<double *>(tmp_alloc(c, tiles * 16 * 8));
  auto *fns = static_cast<TileFn32 *>(tmp_alloc(c, tiles * 16 * sizeof(TileFn32)));
  auto *keys = static_cast<uint32_t *>(tmp_alloc(c, uint64_t(rcap) * 4 + 8));
  auto *heads = static_cast<WideHead *>(tmp_alloc(c, uint64_t(rcap) * sizeof(WideHead)));
  auto *ents = static_cast<WideEnt *>(tmp_alloc(c, uint64_t(rcap) * kWideRes * sizeof(WideEnt)));
  const uint64_t row_slots = rcap < kC32RowSlots ? rcap : kC32RowSlots;
  auto *rows = static_cast<float *>(tmp_alloc(c, row_slots * kC32TileBlocks * 4));
  uint32_t *counter = static_cast<uint32_t *>(tmp_alloc(c, 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(counter), 1, 0);
  const uint32_t grid = static_cast<uint32_t>(tiles);
  k_c32_asum<RV><<<grid, kC32Threads, 0, c.stream>>>(rv, x, y, blocks, asum);
  k_c32_prefix<<<16, kC32Threads, 0, c.stream>>>(asum, tiles, acc, pre);
  k_c32_summ<RV><<<grid, kC32Threads, 0, c.stream>>>(rv, x, y, blocks, pre, fns, counter, rcap, keys);
  const uint64_t wide_blocks = (uint64_t(rcap) + kC32WideSlots - 1) / kC32WideSlots;
  k_c32_wide<RV><<<uint32_t(wide_blocks < 4096 ? wide_blocks : 4096), kWave, 0, c.stream>>>(
      rv, x, y, blocks, pre, keys, counter, rcap, heads, ents, rows);
  const uint64_t groups = (tiles + kC32GroupTiles - 1) / kC32GroupTiles;
  auto *gfn = static_cast<TileFn32 *>(tmp_alloc(c, groups * 16 * sizeof(TileFn32)));
  k_c32_groups<<<uint32_t((16 * groups * kWave + kBlock - 1) / kBlock), kBlock, 0, c.stream>>>(fns, tiles, groups,
                                                                                             gfn);
  k_c32_walk<RV><<<16, kWave, 0, c.stream>>>(rv, x, y, blocks, fns, tiles, gfn, groups, heads, ents, rows,
                                                    acc);
  check_hip(hipGetLastError(), "canon f32 streamed");
  tmp_free(c, asum);
  tmp_free(c, pre);
  tmp_free(c, fns);
  tmp_free(c, keys);
  tmp_free(c, heads);
  tmp_free(c, ents);
  tmp_free(c, rows);
  tmp_free(c, counter);
  tmp_free(c, gfn);
}

}  // namespace

void canon32_blocks(Ctx &c, const CanonSpec &s, const float *x, const float *y, uint64_t blocks, double *acc) {
  if (blocks == 0) return;
  with_row32(s, [&](auto rv) { c32_blocks_as(c, rv, s.kind == 4, x, y ? y : x, blocks, acc); });
}

void canon32_rest(Ctx &c, const CanonSpec &s, const float *x, const float *y, uint64_t rest, double *tail) {
  if (rest == 0) return;
  with_row32(s, [&](auto rv) { k_c32_rest<<<1, 64, 0, c.stream>>>(rv, x, y ? y : x, rest, tail); });
  check_hip(hipGetLastError(), "canon f32 rest");
}

void canon32_finish(Ctx &c, bool mul, const double *acc, const double *tail, int tail_len, uint64_t *out) {
  k_c32_finish<<<1, 64, 0, c.stream>>>(mul, acc, tail, tail_len, out);
  check_hip(hipGetLastError(), "canon f32 finish");
}

}  // namespace vk
