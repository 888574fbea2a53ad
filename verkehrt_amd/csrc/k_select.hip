// Selections: compress_<t> (select.rs:79-92), filter_<cmp>_<t> (fused.rs:212-246) and where
// (select.rs:94-104). The kept rows keep their row order.
//
// One pass, no count pass: each 16384-row tile flags its rows, ranks the kept ones with
// wave64 ballots and mbcnt, learns the number of kept rows before it by decoupled
// look-back over the tiles' published counts, and writes its survivors. Algorithmic bytes:
// n x (width [+1 mask byte]) read, kept x width written.
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

// 1024 threads x 16 rows: large enough that the tile ticket (one same-address atomic per
// tile, about 11 ns apart at best) never limits the row rate. Batched launches may use 8 rows
// per thread (chosen by the plan autotuner).
constexpr int kSelBlock = 1024;
constexpr int kItems = 16;
constexpr int kTile = kSelBlock * kItems;  // 16384 rows
constexpr int kWaves = kSelBlock / kWave;

// Look-back word: epoch (16 bits) | flag (2 bits: 1 aggregate, 2 inclusive) | count (46 bits).
__device__ __forceinline__ uint64_t pack(uint32_t epoch, uint64_t flag, uint64_t v) {
  return (uint64_t(epoch) << 48) | (flag << 46) | v;
}
constexpr uint64_t kCountMask = (uint64_t(1) << 46) - 1;

// The number of kept rows of all tiles before `tile`, by one wave: lanes read the status of
// 64 predecessors at once, wait until all are published, add aggregates back to the nearest
// inclusive prefix.
// Tiles below `first` (another item of a batched launch, or none) count as an inclusive 0.
__device__ uint64_t look_back(uint64_t *status, uint32_t epoch, uint32_t tile, uint32_t first) {
  uint64_t sum = 0;
  int64_t end = static_cast<int64_t>(tile);  // exclusive
  const int l = lane_id();
  while (end > static_cast<int64_t>(first)) {
    const int64_t idx = end - 64 + l;
    uint64_t w = 0;
    bool ready;
    do {
      w = idx >= static_cast<int64_t>(first) ? ld_rlx(&status[idx]) : pack(epoch, 2, 0);
      ready = (w >> 48) == epoch && ((w >> 46) & 3) != 0;
      if (__all(ready)) break;
      backoff();
    } while (true);
    const uint64_t incl = __ballot(((w >> 46) & 3) == 2);
    // lanes at or above the highest inclusive lane contribute
    const int top = incl ? 63 - __builtin_clzll(incl) : -1;
    uint64_t v = (l >= top) ? (w & kCountMask) : 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    sum += v;
    if (top >= 0) break;
    end -= 64;
  }
  return sum;
}

// kind: 0 compress (keep = mask != 0), 1 filter (keep = x cmp c), 2 where (value base + row).
// `keepset` is the set of comparison outcomes a filter keeps (outcome()).
// BATCH: the tile's item (row pointers, row count, output, first tile, index base) comes
// from the batch tables, and the item's last tile writes its count to counts[item].
// All of a thread's loads are issued before the first ballot (a ballot is convergent, so
// loads interleaved with ballots are not hoisted and would complete one at a time); rows past
// the end of the last tile load the last row and are never kept, so no load is predicated.
// Two blocks per compute unit (8 waves per SIMD, at most 64 VGPRs) keep one block loading while
// the other waits on its look-back.
// ITEMS rows per thread (a tile is 1024 x ITEMS rows). A launch whose tiles were planned on
// the device passes dev_ntiles (blocks past it exit) and dev_epoch (the epoch to publish with).
// PM (how a tile learns where its survivors go), for single calls: 0 the look-back; 1 (compress
// and where over long inputs) a count pass over the mask and a scan have left the tile's
// exclusive prefix in status[tile]; 2 (filters over long inputs with few survivors) the tile
// writes its survivors to its own 16384-row slot of out (a staging buffer) and its count to
// status[tile], for a scan and a move pass. Modes 1 and 2 take tiles in block order, with no
// ticket and no look-back.
template <typename T, int KIND, bool BATCH, int ITEMS = kItems, int PM = 0>
__global__ __launch_bounds__(kSelBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_select(const T *__restrict__ x,
                                                   const uint8_t *__restrict__ mask, uint64_t n,
                                                   uint32_t keepset, T c, uint64_t base, void *out_v,
                                                   uint64_t *status, uint32_t epoch,
                                                   uint32_t *ticket, uint32_t ntiles,
                                                   uint64_t *d_count, const BatchItem *items,
                                                   const uint32_t *tile_item, uint64_t *counts,
                                                   const uint32_t *dev_ntiles, const uint32_t *dev_epoch) {
  using Out = typename std::conditional<KIND == 2, int64_t, T>::type;
  constexpr int kTile = kSelBlock * ITEMS;
  constexpr int kItems = ITEMS;
  const uint32_t tile = PM ? blockIdx.x : take_ticket(ticket, ntiles);
  if (dev_ntiles && tile >= *dev_ntiles) return;
  if (dev_epoch) epoch = *dev_epoch;
  uint32_t first = 0, last = ntiles - 1;
  if (BATCH) {
    const uint32_t it = tile_item[tile];
    const BatchItem &I = items[it];
    x = static_cast<const T *>(I.a);
    mask = KIND == 2 ? static_cast<const uint8_t *>(I.a) : static_cast<const uint8_t *>(I.b);
    n = I.n;
    out_v = I.out;
    base = I.aux;
    first = static_cast<uint32_t>(I.tile0);
    last = first + static_cast<uint32_t>((n + kTile - 1) / kTile) - 1;
    d_count = counts + it;
  }
  Out *out = static_cast<Out *>(out_v);
  const uint64_t t0 = uint64_t(tile - first) * kTile;
  const uint64_t rem = n - t0;
  const uint32_t lim = rem >= uint64_t(kTile) ? uint32_t(kTile) : uint32_t(rem);  // rows of the tile
  const int t = threadIdx.x, w = t / kWave, l = lane_id();

  Out val[kItems];
  uint8_t mv[kItems];
  uint64_t keepm[kItems];
  __shared__ uint32_t s_cnt[kItems][kWaves];
  __shared__ uint32_t s_excl[kItems][kWaves];
  __shared__ uint64_t s_prefix;
  // item i of thread t is row t0 + i * kSelBlock + t: every wave load is 64 consecutive rows.
  // The mask bytes of a wave's rows are 16 (kItems) segments of 64 bytes: in a full tile of an
  // aligned mask, lane l loads 16 bytes of segment l / 4 and passes them through LDS, one load
  // per lane instead of one byte load per item.
  __shared__ uint8_t s_mask[KIND == 1 ? 1 : kWaves][KIND == 1 ? 1 : kItems][kWave];
  {
    const T *xt = KIND == 2 ? nullptr : x + t0;
    const uint8_t *mt = KIND == 1 ? nullptr : mask + t0;
    bool wide = false;
    if constexpr (KIND != 1) {
      static_assert(kItems * 4 <= kWave, "a lane loads 16 bytes of one segment");
      wide = lim == uint32_t(kTile) && (reinterpret_cast<uintptr_t>(mt) & 15) == 0;
      if (wide && (l >> 2) < kItems) {
        struct alignas(16) Q {
          uint32_t a[4];
        };
        const int i = l >> 2, part = (l & 3) * 16;
        const Q q = *reinterpret_cast<const Q *>(mt + uint32_t(i) * kSelBlock + uint32_t(w) * kWave + part);
        *reinterpret_cast<Q *>(&s_mask[w][i][part]) = q;
      }
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint32_t lr = uint32_t(i) * kSelBlock + t;
      const uint32_t rr = lr < lim ? lr : lim - 1;
      if (KIND != 1 && !wide) mv[i] = mt[rr];
      if (KIND != 2) val[i] = static_cast<Out>(xt[rr]);
    }
    if constexpr (KIND != 1) {
      // the wave reads back what it wrote (LDS operations of one wave complete in order)
      if (wide) {
#pragma unroll
        for (int i = 0; i < kItems; ++i) mv[i] = s_mask[w][i][l];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint32_t lr = uint32_t(i) * kSelBlock + t;
    const bool keep = KIND == 1 ? (keepset & outcome(static_cast<T>(val[i]), c)) != 0 : mv[i] != 0;
    keepm[i] = __ballot(keep && lr < lim);
    if (l == 0) s_cnt[i][w] = __popcll(keepm[i]);
  }
  __syncthreads();
  // Exclusive scan of the kItems x kWaves counts in row order (item-major), by wave 0:
  // lane l owns the PER consecutive entries PER l .. PER l + PER - 1.
  constexpr int PER = kItems * kWaves / kWave;
  static_assert(PER * kWave == kItems * kWaves, "whole counts per lane");
  if (w == 0) {
    uint32_t c4[PER], v = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int k = PER * l + q;
      c4[q] = s_cnt[k / kWaves][k % kWaves];
      v += c4[q];
    }
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t o = __shfl_up(inc, d);
      if (l >= d) inc += o;
    }
    uint32_t e = inc - v;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int k = PER * l + q;
      s_excl[k / kWaves][k % kWaves] = e;
      e += c4[q];
    }
    const uint32_t total = __shfl(inc, 63);
    // publish and look back; the status word carries its payload, so relaxed stores suffice
    uint64_t prefix = 0;
    if constexpr (PM == 1) {
      prefix = status[tile];
    } else if constexpr (PM == 2) {
      prefix = t0;
      if (l == 0) status[tile] = total;  // a raw count: epoch field 0, which no launch uses
    } else if (tile == first) {
      if (l == 0) st_rlx(&status[tile], pack(epoch, 2, total));
    } else {
      if (l == 0) st_rlx(&status[tile], pack(epoch, 1, total));
      prefix = look_back(status, epoch, tile, first);
      if (l == 0) st_rlx(&status[tile], pack(epoch, 2, prefix + total));
    }
    if (l == 0) {
      s_prefix = prefix;
      if (PM != 2 && tile == last) *d_count = prefix + total;
    }
  }
  __syncthreads();
  Out *ot = out + s_prefix;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    if ((keepm[i] >> l) & 1) {
      const uint32_t at = s_excl[i][w] + rank_below(keepm[i]);
      if (KIND == 2) ot[at] = static_cast<Out>(base + t0 + uint32_t(i) * kSelBlock + t);
      else ot[at] = val[i];
    }
  }
}

// Persistent form for long inputs: one 1024-thread block per compute unit takes 16384-row tiles
// in ticket order and loads the next tile into registers before it resolves the current one,
// so the loads of one tile are in flight while the block waits on the look-back and writes the
// survivors of the other: HBM stays busy through every look-back. A tile's count is published
// as soon as its rows are in registers, and the ticket after the next tile is taken while the
// current one resolves. Ticket: a block takes no more after one finds no tile, so a launch
// takes ntiles + gridDim.x tickets and the last one returns the ticket to zero. A block only
// waits on tiles below its current one, all held by running blocks, so no wait is circular.
// BATCH: a tile's item (rows, row count, output, first tile, count slot) comes from the batch
// tables, and its look-back stops at the item's first tile; a launch whose tiles were planned
// on the device passes dev_ntiles (every block reads the same planned count) and dev_epoch.
template <typename T, int KIND, bool BATCH = false>
__global__ __launch_bounds__(kSelBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_select_p(
    const T *__restrict__ x, const uint8_t *__restrict__ mask, uint64_t n, uint32_t keepset, T c, uint64_t base,
    void *out_v, uint64_t *status, uint32_t epoch, uint32_t *ticket, uint32_t ntiles, uint64_t *d_count,
    const BatchItem *items = nullptr, const uint32_t *tile_item = nullptr, uint64_t *counts = nullptr,
    const uint32_t *dev_ntiles = nullptr, const uint32_t *dev_epoch = nullptr) {
  using Out = typename std::conditional<KIND == 2, int64_t, T>::type;
  struct alignas(16) Q {
    uint32_t a[4];
  };
  if (dev_ntiles) ntiles = *dev_ntiles;
  if (dev_epoch) epoch = *dev_epoch;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const uint32_t takes = ntiles + gridDim.x;
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_cnt[kItems][kWaves];
  __shared__ uint32_t s_excl[kItems][kWaves];
  __shared__ uint32_t s_total;
  __shared__ uint64_t s_prefix;
  __shared__ uint8_t s_mask[KIND == 1 ? 1 : kWaves][KIND == 1 ? 1 : kItems][kWave];

  auto take = [&]() -> uint32_t {
    __syncthreads();
    if (t == 0) {
      const uint32_t k = atomicAdd(ticket, 1u);
      if (k == takes - 1) atomicExch(ticket, 0u);
      s_tile = k;
    }
    __syncthreads();
    return s_tile;
  };
  // a tile's rows: its array (values x, mask m), first row t0 in the array, row count, the
  // first tile of its array, whether it is the array's last tile, output, count slot, base
  struct Ref {
    const T *x;
    const uint8_t *m;
    uint64_t t0;
    uint32_t lim, first;
    bool last;
    Out *out;
    uint64_t *cnt;
    uint64_t base;
  };
  auto ref = [&](uint32_t tile) -> Ref {
    Ref r;
    uint64_t rows;
    if constexpr (BATCH) {
      const uint32_t it = tile_item[tile];
      const BatchItem &I = items[it];
      r.x = static_cast<const T *>(I.a);
      r.m = KIND == 2 ? static_cast<const uint8_t *>(I.a) : static_cast<const uint8_t *>(I.b);
      r.first = static_cast<uint32_t>(I.tile0);
      r.out = static_cast<Out *>(I.out);
      r.cnt = counts + it;
      r.base = I.aux;
      rows = I.n;
    } else {
      r.x = x;
      r.m = mask;
      r.first = 0;
      r.out = static_cast<Out *>(out_v);
      r.cnt = d_count;
      r.base = base;
      rows = n;
    }
    r.t0 = uint64_t(tile - r.first) * kTile;
    const uint64_t rem = rows - r.t0;
    r.lim = rem >= uint64_t(kTile) ? uint32_t(kTile) : uint32_t(rem);
    r.last = rem <= uint64_t(kTile);
    return r;
  };
  // the tile's rows into registers (rows past the end load the last row), and the mask bytes
  // of an aligned full tile (16 per lane); other tiles' mask bytes are loaded when counted
  auto load = [&](uint32_t tile, Out (&val)[kItems], Q &mq, bool &wide) {
    const Ref r = ref(tile);
    const uint32_t lim = r.lim;
    const T *xt = KIND == 2 ? nullptr : r.x + r.t0;
    wide = false;
    if constexpr (KIND != 1) {
      const uint8_t *mt = r.m + r.t0;
      wide = lim == uint32_t(kTile) && (reinterpret_cast<uintptr_t>(mt) & 15) == 0;
      if (wide && (l >> 2) < kItems)
        mq = *reinterpret_cast<const Q *>(mt + uint32_t(l >> 2) * kSelBlock + uint32_t(w) * kWave + (l & 3) * 16);
    }
    if constexpr (KIND != 2) {
#pragma unroll
      for (int i = 0; i < kItems; ++i) {
        const uint32_t lr = uint32_t(i) * kSelBlock + t;
        val[i] = static_cast<Out>(xt[lr < lim ? lr : lim - 1]);
      }
    }
  };
  // ballots, the block's exclusive counts (s_excl, s_total) and the published aggregate
  auto count = [&](uint32_t tile, const Out (&val)[kItems], const Q &mq, bool wide, uint64_t (&keepm)[kItems]) {
    const Ref r = ref(tile);
    const uint32_t lim = r.lim;
    uint8_t mv[kItems];
    if constexpr (KIND != 1) {
      if (wide) {
        if ((l >> 2) < kItems) *reinterpret_cast<Q *>(&s_mask[w][l >> 2][(l & 3) * 16]) = mq;
        // the wave reads back what it wrote (LDS operations of one wave complete in order)
#pragma unroll
        for (int i = 0; i < kItems; ++i) mv[i] = s_mask[w][i][l];
      } else {
        const uint8_t *mt = r.m + r.t0;
#pragma unroll
        for (int i = 0; i < kItems; ++i) {
          const uint32_t lr = uint32_t(i) * kSelBlock + t;
          mv[i] = mt[lr < lim ? lr : lim - 1];
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint32_t lr = uint32_t(i) * kSelBlock + t;
      const bool keep = KIND == 1 ? (keepset & outcome(static_cast<T>(val[i]), c)) != 0 : mv[i] != 0;
      keepm[i] = __ballot(keep && lr < lim);
      if (l == 0) s_cnt[i][w] = __popcll(keepm[i]);
    }
    __syncthreads();
    constexpr int PER = kItems * kWaves / kWave;
    if (w == 0) {
      uint32_t c4[PER], v = 0;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int k = PER * l + q;
        c4[q] = s_cnt[k / kWaves][k % kWaves];
        v += c4[q];
      }
      uint32_t inc = v;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d);
        if (l >= d) inc += o;
      }
      uint32_t e = inc - v;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int k = PER * l + q;
        s_excl[k / kWaves][k % kWaves] = e;
        e += c4[q];
      }
      const uint32_t total = __shfl(inc, 63);
      if (l == 0) {
        s_total = total;
        st_rlx(&status[tile], pack(epoch, tile == r.first ? 2 : 1, total));
      }
    }
  };
  // the look-back (wave 0), the inclusive count, and the survivors written
  auto finish = [&](uint32_t tile, const Out (&val)[kItems], const uint64_t (&keepm)[kItems]) {
    const Ref r = ref(tile);
    if (w == 0) {
      const uint32_t total = s_total;
      uint64_t prefix = 0;
      if (tile > r.first) {
        prefix = look_back(status, epoch, tile, r.first);
        if (l == 0) st_rlx(&status[tile], pack(epoch, 2, prefix + total));
      }
      if (l == 0) {
        s_prefix = prefix;
        if (r.last) *r.cnt = prefix + total;
      }
    }
    __syncthreads();
    Out *ot = r.out + s_prefix;
    const uint64_t t0 = r.t0;
    const uint64_t base = r.base;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      if ((keepm[i] >> l) & 1) {
        const uint32_t at = s_excl[i][w] + rank_below(keepm[i]);
        if (KIND == 2) ot[at] = static_cast<Out>(base + t0 + uint32_t(i) * kSelBlock + t);
        else ot[at] = val[i];
      }
    }
  };

  uint32_t tile = take();
  if (tile >= ntiles) return;
  uint32_t next = take();
  Out va[kItems], vb[kItems];
  uint64_t ka[kItems];
  Q qa{}, qb{};
  bool wa = false, wb = false;
  load(tile, va, qa, wa);
  count(tile, va, qa, wa, ka);
  while (true) {
    if (next < ntiles) load(next, vb, qb, wb);
    // the ticket after `next`, taken while this tile resolves (no take once one finds no tile)
    uint32_t after = 0;
    if (t == 0 && next < ntiles) {
      after = atomicAdd(ticket, 1u);
      if (after == takes - 1) atomicExch(ticket, 0u);
    }
    finish(tile, va, ka);
    if (next >= ntiles) break;
    if (t == 0) s_tile = after;
    __syncthreads();  // the survivors of `tile` are written before s_excl is reused; s_tile is set
    tile = next;
    next = s_tile;
#pragma unroll
    for (int i = 0; i < kItems; ++i) va[i] = vb[i];
    qa = qb;
    wa = wb;
    count(tile, va, qa, wa, ka);
  }
}

// Look-back scratch of one selection launch: status words (one per tile), their epoch, and a
// tile ticket that is zero before the launch (the last tile returns it to zero).
struct SelScratch {
  uint64_t *status;
  uint32_t epoch;
  uint32_t *ticket;
};

// The count pass of mask selections over long inputs: cnt[tile] = the nonzero mask bytes of the
// tile's 16384 rows, by a block of 256 threads, 64 bytes per thread as four 16-byte loads (byte by
// byte in a partial or unaligned tile). The counts go into the look-back status words, whose epoch
// field they leave 0, which no launch uses.
constexpr int kCountThreads = 256;
__global__ __launch_bounds__(kCountThreads) void k_mask_count(const uint8_t *__restrict__ mask, uint64_t n,
                                                              uint64_t *cnt) {
  const uint64_t t0 = uint64_t(blockIdx.x) * kTile;
  const uint64_t rem = n - t0;
  const uint32_t lim = rem >= uint64_t(kTile) ? uint32_t(kTile) : uint32_t(rem);
  const uint8_t *mt = mask + t0;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  uint32_t c = 0;
  if (lim == uint32_t(kTile) && (reinterpret_cast<uintptr_t>(mt) & 15) == 0) {
    struct alignas(16) Q {
      uint32_t a[4];
    };
    constexpr int kLoads = kTile / 16 / kCountThreads;
    Q q[kLoads];
#pragma unroll
    for (int u = 0; u < kLoads; ++u) q[u] = reinterpret_cast<const Q *>(mt)[u * kCountThreads + t];
#pragma unroll
    for (int u = 0; u < kLoads; ++u)
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        // the high bit of every nonzero byte
        const uint32_t v = q[u].a[k];
        c += __popc((((v & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | v) & 0x80808080u);
      }
  } else {
    for (uint32_t r = uint32_t(t); r < lim; r += kCountThreads) c += mt[r] != 0;
  }
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) c += __shfl_xor(c, m);
  __shared__ uint32_t s_w[kCountThreads / kWave];
  if (l == 0) s_w[w] = c;
  __syncthreads();
  if (t == 0) {
    uint32_t a = 0;
#pragma unroll
    for (int k = 0; k < kCountThreads / kWave; ++k) a += s_w[k];
    cnt[blockIdx.x] = a;
  }
}

// In place: cnt[tile] becomes the exclusive prefix of the counts; *d_count the total. One block
// over chunks of 4096 tiles: a chunk is loaded into LDS with coalesced loads, thread t scans its
// four entries, the thread sums are scanned by wave shuffles and across the 16 waves, and the
// chunk goes back with coalesced stores. With `gate`, a launch whose gate word is 0 does nothing.
constexpr uint32_t kOffChunk = 4 * kSelBlock;
__global__ __launch_bounds__(kSelBlock) void k_tile_offsets(uint64_t *cnt, uint32_t tiles, uint64_t *d_count,
                                                            const uint32_t *gate = nullptr) {
  if (gate && *gate == 0) return;
  __shared__ uint64_t s_v[kOffChunk];
  __shared__ uint64_t s_w[kWaves];
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  uint64_t carry = 0;
  for (uint32_t c0 = 0; c0 < tiles; c0 += kOffChunk) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t i = c0 + uint32_t(u) * kSelBlock + t;
      s_v[u * kSelBlock + t] = i < tiles ? cnt[i] : 0;
    }
    __syncthreads();
    uint64_t a[4], sum = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = s_v[4 * t + u];
      sum += a[u];
    }
    uint64_t inc = sum;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint64_t o = __shfl_up(inc, d);
      if (l >= d) inc += o;
    }
    if (l == kWave - 1) s_w[w] = inc;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      before += k < w ? s_w[k] : 0;
      all += s_w[k];
    }
    uint64_t e = carry + before + inc - sum;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      s_v[4 * t + u] = e;
      e += a[u];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t i = c0 + uint32_t(u) * kSelBlock + t;
      if (i < tiles) cnt[i] = s_v[u * kSelBlock + t];
    }
    carry += all;
    __syncthreads();
  }
  if (t == 0) *d_count = carry;
}

// The kept rows among every `step`-th row of a filter's input (the choice of the staged form).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_filter_sample(const T *__restrict__ x, uint64_t n, uint32_t keepset, T c,
                                                          uint64_t step, uint32_t *kept) {
  uint32_t k = 0;
  const uint64_t samples = (n + step - 1) / step;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < samples; i += uint64_t(gridDim.x) * kBlock)
    k += (keepset & outcome(x[i * step], c)) != 0;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) k += __shfl_xor(k, m);
  if (lane_id() == 0 && k) atomicAdd(kept, k);
}

// The staged survivors of every tile (its slot of stage; status[tile] is its offset after the
// scan turned the counts into offsets) moved to their place in out, blocks taking tiles
// blockIdx.x, + gridDim.x, ...
template <typename T>
__global__ __launch_bounds__(kBlock) void k_stage_move(const T *__restrict__ stage, const uint64_t *off, uint32_t tiles,
                                                       const uint64_t *d_count, T *__restrict__ out,
                                                       const uint32_t *gate) {
  if (*gate == 0) return;
  for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    const uint64_t o = off[tile];
    const uint64_t e = tile + 1 < tiles ? off[tile + 1] : *d_count;
    const T *src = stage + uint64_t(tile) * kTile;
    for (uint64_t r = threadIdx.x; r < e - o; r += kBlock) out[o + r] = src[r];
  }
}

// Filters over at least this many rows sample their selectivity first and, below 1 in
// kStageInv kept rows, take the staged form (no look-back; the survivors cross HBM twice more).
// The choice is made on the device: k_stage_choose writes the tile counts the two forms run
// with (the other gets 0 and its launches return at once), so no read-back sits between them.
constexpr uint64_t kStageMinRows = uint64_t(1) << 24;
constexpr uint64_t kSampleStep = 4096;
constexpr uint32_t kStageInv = 16;

// gate[0] = tiles for the staged form and gate[1] = 0, or the other way round.
__global__ void k_stage_choose(const uint32_t *kept, uint64_t samples, uint32_t tiles, uint32_t *gate) {
  if (threadIdx.x != 0) return;
  const bool staged = uint64_t(*kept) * kStageInv < samples;
  gate[0] = staged ? tiles : 0;
  gate[1] = staged ? 0 : tiles;
}

template <typename T, int KIND>
void launch(Ctx &c, const void *x, const uint8_t *mask, uint64_t n, int cmp, uint64_t cbits,
            uint64_t base, void *out, uint64_t *d_count, const SelScratch *given) {
  const uint64_t tiles = (n + kTile - 1) / kTile;
  if (tiles > 0xFFFFFFFFull) usage_error("selection too large");
  SelScratch sc;
  if (given) {
    sc = *given;
  } else {
    sc.status = status_words(c, ST_SELECT, tiles);
    sc.epoch = next_epoch(c);
    sc.ticket = c.scratch.counters;
  }
  T cv;
  __builtin_memcpy(&cv, &cbits, sizeof(T));
  if constexpr (KIND != 1) {
    // long compress and where: the kept rows per tile are counted from the mask bytes alone (a
    // read of n bytes), scanned, and the selection pass takes each tile's offset from there
    if (tiles >= 2 * uint64_t(c.cus)) {
      k_mask_count<<<static_cast<uint32_t>(tiles), kCountThreads, 0, c.stream>>>(mask, n, sc.status);
      k_tile_offsets<<<1, kSelBlock, 0, c.stream>>>(sc.status, static_cast<uint32_t>(tiles), d_count);
      k_select<T, KIND, false, kItems, 1><<<static_cast<uint32_t>(tiles), kSelBlock, 0, c.stream>>>(
          static_cast<const T *>(x), mask, n, keepset_of(cmp), cv, base, out, sc.status, sc.epoch, sc.ticket,
          static_cast<uint32_t>(tiles), d_count, nullptr, nullptr, nullptr, nullptr, nullptr);
      check_hip(hipGetLastError(), "select launch");
      return;
    }
  }
  if constexpr (KIND == 1) {
    // long filters: the staged form when a sample finds few kept rows, else the persistent
    // look-back form; both are launched, gated on the device
    if (!given && n >= kStageMinRows && tiles >= 2 * uint64_t(c.cus)) {
      uint32_t *kept = static_cast<uint32_t *>(tmp_alloc(c, 16));
      uint32_t *gate = kept + 2;
      fill_u64(c, reinterpret_cast<uint64_t *>(kept), 2, 0);
      const uint32_t nt = static_cast<uint32_t>(tiles);
      k_filter_sample<T><<<64, kBlock, 0, c.stream>>>(static_cast<const T *>(x), n, keepset_of(cmp), cv, kSampleStep,
                                                    kept);
      k_stage_choose<<<1, 64, 0, c.stream>>>(kept, (n + kSampleStep - 1) / kSampleStep, nt, gate);
      T *stage = static_cast<T *>(tmp_alloc(c, tiles * kTile * sizeof(T)));
      k_select<T, 1, false, kItems, 2><<<nt, kSelBlock, 0, c.stream>>>(
          static_cast<const T *>(x), mask, n, keepset_of(cmp), cv, base, stage, sc.status, sc.epoch, sc.ticket, nt,
          d_count, nullptr, nullptr, nullptr, gate, nullptr);
      k_tile_offsets<<<1, kSelBlock, 0, c.stream>>>(sc.status, nt, d_count, gate);
      const uint32_t mg = nt < 1024 ? nt : 1024;
      k_stage_move<T><<<mg, kBlock, 0, c.stream>>>(stage, sc.status, nt, d_count, static_cast<T *>(out), gate);
      k_select_p<T, KIND><<<static_cast<uint32_t>(c.cus), kSelBlock, 0, c.stream>>>(
          static_cast<const T *>(x), mask, n, keepset_of(cmp), cv, base, out, sc.status, sc.epoch, sc.ticket, nt,
          d_count, nullptr, nullptr, nullptr, gate + 1, nullptr);
      check_hip(hipGetLastError(), "select launch");
      tmp_free(c, stage);
      tmp_free(c, kept);
      return;
    }
    // long filters: one persistent block per compute unit (compress and where, which also hold
    // mask bytes, keep one tile per block: their persistent form spilled registers)
    if (tiles >= 2 * uint64_t(c.cus)) {
      k_select_p<T, KIND><<<static_cast<uint32_t>(c.cus), kSelBlock, 0, c.stream>>>(
          static_cast<const T *>(x), mask, n, keepset_of(cmp), cv, base, out, sc.status, sc.epoch, sc.ticket,
          static_cast<uint32_t>(tiles), d_count);
      check_hip(hipGetLastError(), "select launch");
      return;
    }
  }
  k_select<T, KIND, false><<<static_cast<uint32_t>(tiles), kSelBlock, 0, c.stream>>>(
      static_cast<const T *>(x), mask, n, keepset_of(cmp), cv, base, out, sc.status, sc.epoch,
      sc.ticket, static_cast<uint32_t>(tiles), d_count, nullptr, nullptr, nullptr, nullptr, nullptr);
  check_hip(hipGetLastError(), "select launch");
}

template <typename T, int KIND>
void launch_batch(Ctx &c, const SelectSpec &s, std::vector<BatchItem> &items, uint64_t *d_counts) {
  BatchPlan p = batch_plan(c, items, kTile);
  if (p.ntiles > 0) {
    uint64_t *status = status_words(c, ST_SELECT, p.ntiles);
    const uint32_t epoch = next_epoch(c);
    T cv;
    __builtin_memcpy(&cv, &s.cbits, sizeof(T));
    if constexpr (KIND == 1) {
      // many tiles: the persistent form, every item's tiles in one ticket order
      if (p.ntiles >= 2 * uint32_t(c.cus)) {
        k_select_p<T, KIND, true><<<static_cast<uint32_t>(c.cus), kSelBlock, 0, c.stream>>>(
            nullptr, nullptr, 0, keepset_of(s.cmp), cv, 0, nullptr, status, epoch, c.scratch.counters, p.ntiles,
            nullptr, p.items, p.tile_item, d_counts);
        check_hip(hipGetLastError(), "select batch launch");
        batch_plan_free(c, p);
        return;
      }
    }
    k_select<T, KIND, true><<<p.ntiles, kSelBlock, 0, c.stream>>>(
        nullptr, nullptr, 0, keepset_of(s.cmp), cv, 0, nullptr, status, epoch, c.scratch.counters, p.ntiles,
        nullptr, p.items, p.tile_item, d_counts, nullptr, nullptr);
    check_hip(hipGetLastError(), "select batch launch");
  }
  batch_plan_free(c, p);
}

template <int KIND>
void batch_by_width(Ctx &c, const SelectSpec &s, std::vector<BatchItem> &items, uint64_t *d_counts) {
  switch (s.type) {
    case VK_F64: launch_batch<double, KIND>(c, s, items, d_counts); break;
    case VK_F32: launch_batch<float, KIND>(c, s, items, d_counts); break;
    case VK_I64: launch_batch<int64_t, KIND>(c, s, items, d_counts); break;
    case VK_U64: launch_batch<uint64_t, KIND>(c, s, items, d_counts); break;
    case VK_I32: launch_batch<int32_t, KIND>(c, s, items, d_counts); break;
    case VK_U8: case VK_BOOL: launch_batch<uint8_t, KIND>(c, s, items, d_counts); break;
    default: usage_error("unsupported selection type");
  }
}

template <int KIND>
void by_width(Ctx &c, int32_t type, const void *x, const uint8_t *mask, uint64_t n, int cmp,
              uint64_t cbits, uint64_t base, void *out, uint64_t *d_count, const SelScratch *sc) {
  switch (type) {
    case VK_F64: launch<double, KIND>(c, x, mask, n, cmp, cbits, base, out, d_count, sc); break;
    case VK_F32: launch<float, KIND>(c, x, mask, n, cmp, cbits, base, out, d_count, sc); break;
    case VK_I64: launch<int64_t, KIND>(c, x, mask, n, cmp, cbits, base, out, d_count, sc); break;
    case VK_U64: launch<uint64_t, KIND>(c, x, mask, n, cmp, cbits, base, out, d_count, sc); break;
    case VK_I32: launch<int32_t, KIND>(c, x, mask, n, cmp, cbits, base, out, d_count, sc); break;
    case VK_U8: case VK_BOOL: launch<uint8_t, KIND>(c, x, mask, n, cmp, cbits, base, out, d_count, sc); break;
    default: usage_error("unsupported selection type");
  }
}

__global__ void k_zero_count(uint64_t *d) { *d = 0; }

void select_dispatch(Ctx &c, const SelectSpec &s, const void *x, const uint8_t *mask, uint64_t n, void *out,
                     uint64_t *d_count, const SelScratch *sc) {
  if (n == 0) {
    k_zero_count<<<1, 1, 0, c.stream>>>(d_count);
    return;
  }
  switch (s.kind) {
    case 0: by_width<0>(c, s.type, x, mask, n, 0, 0, 0, out, d_count, sc); break;
    case 1: by_width<1>(c, s.type, x, nullptr, n, s.cmp, s.cbits, 0, out, d_count, sc); break;
    default: launch<uint8_t, 2>(c, nullptr, mask, n, 0, 0, s.base, out, d_count, sc); break;
  }
}

}  // namespace

void select_rows(Ctx &c, const SelectSpec &s, const void *x, const uint8_t *mask, uint64_t n,
                 void *out, uint64_t *d_count) {
  select_dispatch(c, s, x, mask, n, out, d_count, nullptr);
}

uint64_t select_rows_per_tile() { return kTile; }

void select_rows_scratch(Ctx &c, const SelectSpec &s, const void *x, const uint8_t *mask, uint64_t n,
                         void *out, uint64_t *d_count, uint64_t *status, uint32_t epoch, uint32_t *ticket) {
  const SelScratch sc{status, epoch, ticket};
  select_dispatch(c, s, x, mask, n, out, d_count, &sc);
}

namespace {

// ITEMS 0: the persistent form (filters only), one block per compute unit over 16K-row tiles.
template <typename T, int KIND, int ITEMS>
void launch_dev(Ctx &c, const SelectSpec &s, const DevPlan &p, uint64_t *d_counts) {
  T cv;
  __builtin_memcpy(&cv, &s.cbits, sizeof(T));
  if constexpr (ITEMS == 0) {
    static_assert(KIND == 1, "the persistent form is a filter's");
    k_select_p<T, KIND, true><<<static_cast<uint32_t>(c.cus), kSelBlock, 0, c.stream>>>(
        nullptr, nullptr, 0, keepset_of(s.cmp), cv, 0, nullptr, p.status, 0, p.ticket, p.max_tiles, nullptr,
        p.items, p.tile_item, d_counts, p.ntiles, p.epoch);
  } else {
    k_select<T, KIND, true, ITEMS><<<p.max_tiles, kSelBlock, 0, c.stream>>>(
        nullptr, nullptr, 0, keepset_of(s.cmp), cv, 0, nullptr, p.status, 0, p.ticket, p.max_tiles, nullptr,
        p.items, p.tile_item, d_counts, p.ntiles, p.epoch);
  }
}

template <int KIND, int ITEMS>
void dev_by_width(Ctx &c, const SelectSpec &s, const DevPlan &p, uint64_t *d_counts) {
  switch (s.type) {
    case VK_F64: launch_dev<double, KIND, ITEMS>(c, s, p, d_counts); break;
    case VK_F32: launch_dev<float, KIND, ITEMS>(c, s, p, d_counts); break;
    case VK_I64: launch_dev<int64_t, KIND, ITEMS>(c, s, p, d_counts); break;
    case VK_U64: launch_dev<uint64_t, KIND, ITEMS>(c, s, p, d_counts); break;
    case VK_I32: launch_dev<int32_t, KIND, ITEMS>(c, s, p, d_counts); break;
    case VK_U8: case VK_BOOL: launch_dev<uint8_t, KIND, ITEMS>(c, s, p, d_counts); break;
    default: usage_error("unsupported selection type");
  }
}

template <int ITEMS>
void dev_by_kind(Ctx &c, const SelectSpec &s, const DevPlan &p, uint64_t *d_counts) {
  if constexpr (ITEMS == 0) {
    if (s.kind != 1) usage_error("the persistent selection form is a filter's");
    dev_by_width<1, 0>(c, s, p, d_counts);
  } else {
    switch (s.kind) {
      case 0: dev_by_width<0, ITEMS>(c, s, p, d_counts); break;
      case 1: dev_by_width<1, ITEMS>(c, s, p, d_counts); break;
      default: launch_dev<uint8_t, 2, ITEMS>(c, s, p, d_counts); break;
    }
  }
}

}  // namespace

uint64_t select_tile_rows(int items_per_thread) { return uint64_t(kSelBlock) * (items_per_thread == 8 ? 8 : 16); }

void select_batch_dev(Ctx &c, const SelectSpec &s, int items_per_thread, const DevPlan &p, uint32_t nitems,
                      uint64_t *d_counts) {
  fill_u64(c, d_counts, nitems, 0);  // items without rows keep a count of 0
  if (p.max_tiles == 0) return;
  if (items_per_thread == 8) dev_by_kind<8>(c, s, p, d_counts);
  else if (items_per_thread == 0) dev_by_kind<0>(c, s, p, d_counts);
  else dev_by_kind<16>(c, s, p, d_counts);
  check_hip(hipGetLastError(), "select batch (device plan)");
}

void select_batch(Ctx &c, const SelectSpec &s, std::vector<BatchItem> &items, uint64_t *d_counts) {
  fill_u64(c, d_counts, items.size(), 0);  // items without rows keep a count of 0
  switch (s.kind) {
    case 0: batch_by_width<0>(c, s, items, d_counts); break;
    case 1: batch_by_width<1>(c, s, items, d_counts); break;
    default: launch_batch<uint8_t, 2>(c, s, items, d_counts); break;
  }
}

}  // namespace vk
