// Scans: scan_<op>_<t> and scan_by_<op>_<t> (Scan<T, O, BY>, src/kernels/scalar/scan.rs:
// 282-326) for op add (wrapping, widening i32->i64 and u8->u64, ScanSum scan.rs:142-161),
// min and max (IEEE minimum/maximum on floats, ScanExtreme scan.rs:164-183). A start byte
// restarts the fold; row 0 of the whole array always starts.
//
// One pass over 16384-row tiles (512 threads x 32 rows; scan_by: 8192-row tiles of 16 rows
// per thread), data in registers only: wave w of a tile owns its 2048 (1024) consecutive rows,
// so every wave load and store covers 64 (or, with 16-byte accesses, 128) consecutive rows. Each wave scans its rows item by item with DPP
// row shifts and broadcasts (no LDS), carrying the running value from item to item; the
// eight wave aggregates meet in wave 0, which publishes the tile aggregate and joins it to
// the tiles before by decoupled look-back on (aggregate, has-start) pairs: a predecessor whose
// aggregate holds a start, or that published its inclusive value, ends the look-back.
// The fold starts from the operator's identity (0, +inf, -inf, or the integer bounds), which
// is the same as starting at row 0; the operators are associative and commutative, so any
// association order gives the reference's bits. Algorithmic bytes: n x (in width [+1 start
// byte] + 8).
#include "vk_device.h"
#include "vk_internal.h"

#include <limits>

namespace vk {
namespace {

constexpr int kScanBlock = 512;
constexpr int kWaves = kScanBlock / kWave;
// Rows per thread: 32 for plain scans (16K-row tiles: half the tile tickets and look-backs of
// 8K-row tiles, measured faster although one block fits a compute unit), 16 for scan_by, whose
// start flags would need the registers of two waves per SIMD at 32 rows.
constexpr int scan_rows(bool by) { return by ? 16 : 32; }
constexpr int scan_tile(bool by) { return kScanBlock * scan_rows(by); }

template <typename U, int OP>
__device__ __forceinline__ U apply(U a, U b) {
  if constexpr (OP == 0) return a + b;
  else if constexpr (OP == 1) return min2(a, b);
  else return max2(a, b);
}

template <typename U, int OP>
__device__ __forceinline__ U identity() {
  if constexpr (OP == 0) return U{};
  else if constexpr (std::is_floating_point<U>::value)
    return OP == 1 ? std::numeric_limits<U>::infinity() : -std::numeric_limits<U>::infinity();
  else return OP == 1 ? std::numeric_limits<U>::max() : std::numeric_limits<U>::lowest();
}

// A span of rows: its fold, and whether a start lies in it (BY scans; always 0 otherwise).
template <typename U>
struct Seg {
  U v;
  uint32_t f;
};

// a then b: b's value stands alone if b holds a start.
template <typename U, int OP>
__device__ __forceinline__ Seg<U> join(Seg<U> a, Seg<U> b) {
  return Seg<U>{b.f ? b.v : apply<U, OP>(a.v, b.v), a.f | b.f};
}

// DPP move of every 32-bit word of a segment; lanes with no source (or masked rows) take `id`.
template <int CTRL, int ROWS, typename U>
__device__ __forceinline__ Seg<U> dpp_seg(Seg<U> x, Seg<U> id) {
  static_assert(sizeof(U) == 8 || sizeof(U) == 4, "4- or 8-byte values");
  constexpr int N = sizeof(U) / 4;
  int xw[N], iw[N];
  __builtin_memcpy(xw, &x.v, sizeof(U));
  __builtin_memcpy(iw, &id.v, sizeof(U));
  int ow[N];
#pragma unroll
  for (int k = 0; k < N; ++k) ow[k] = __builtin_amdgcn_update_dpp(iw[k], xw[k], CTRL, ROWS, 0xF, false);
  Seg<U> r;
  __builtin_memcpy(&r.v, ow, sizeof(U));
  r.f = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(static_cast<int>(id.f), static_cast<int>(x.f), CTRL, ROWS, 0xF, false));
  return r;
}

// Inclusive scan over the 64 lanes of a wave (lane order = row order): Kogge-Stone within
// rows of 16 lanes (row_shr 1, 2, 4, 8), then row_bcast:15 into rows 1 and 3 and
// row_bcast:31 into rows 2 and 3.
template <typename U, int OP, bool BY>
__device__ __forceinline__ Seg<U> wave_scan(Seg<U> x) {
  const Seg<U> id{identity<U, OP>(), 0};
  auto f = [&](Seg<U> o) { x = BY ? join<U, OP>(o, x) : Seg<U>{apply<U, OP>(o.v, x.v), 0}; };
  f(dpp_seg<0x111, 0xF>(x, id));
  f(dpp_seg<0x112, 0xF>(x, id));
  f(dpp_seg<0x114, 0xF>(x, id));
  f(dpp_seg<0x118, 0xF>(x, id));
  f(dpp_seg<0x142, 0xA>(x, id));
  f(dpp_seg<0x143, 0xC>(x, id));
  return x;
}

// The value of lane 63 in every lane.
template <typename U>
__device__ __forceinline__ Seg<U> lane63(Seg<U> x) {
  Seg<U> r;
  r.v = bcast_pod(x.v, 63);
  r.f = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x.f), 63));
  return r;
}

// Status of a tile: [0] aggregate, [1] inclusive, [2] flag word
// (epoch << 8 | 1 aggregate published | 2 inclusive published | 4 aggregate has a start).
__device__ __forceinline__ bool stat_ready(uint64_t w, uint32_t epoch) {
  return (w >> 8) == epoch && (w & 3) != 0;
}

// The fold of all tiles before `tile` back to the nearest inclusive value or start, starting
// from `acc` (the identity, or the carried value for tile 0's predecessors).
template <typename U, int OP>
__device__ U look_back(uint64_t *status, uint32_t epoch, uint32_t tile, U acc) {
  int64_t end = tile;
  const int l = lane_id();
  while (end > 0) {
    const int64_t idx = end - 64 + l;
    uint64_t w;
    while (true) {
      w = idx >= 0 ? ld_rlx(&status[idx * 4 + 2]) : ((uint64_t(epoch) << 8) | 2);
      if (__all(stat_ready(w, epoch))) break;
      backoff();
    }
    observe_fence();
    // A lane ends the look-back when it is inclusive or its aggregate holds a start.
    const bool stop = idx < 0 || (w & 2) || (w & 4);
    const uint64_t stops = __ballot(stop);
    const int top = stops ? 63 - __builtin_clzll(stops) : -1;
    U v = identity<U, OP>();
    if (l >= top && idx >= 0) {
      const uint64_t bits = (w & 2) ? ld_rlx(&status[idx * 4 + 1]) : ld_rlx(&status[idx * 4 + 0]);
      v = from_bits<U>(bits);
    }
    // fold the contributing lanes (any order: the operators commute; only the top lane may
    // hold a start, and it is the earliest)
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = apply<U, OP>(v, shfl_xor_pod(v, m));
    acc = apply<U, OP>(acc, v);
    if (top >= 0) break;
    end -= 64;
  }
  return acc;
}

// VEC rows per access (1, or 2 with 16-byte output accesses when every pointer is aligned).
template <typename TIn, typename U, int OP, bool BY, int VEC>
__global__ __launch_bounds__(kScanBlock) void k_scan(const TIn *__restrict__ x,
                                                 const uint8_t *__restrict__ starts, uint64_t n,
                                                 U *__restrict__ out, uint64_t *status,
                                                 uint32_t epoch, uint32_t *ticket, uint32_t ntiles,
                                                 const uint64_t *carry_in, uint64_t *carry_out) {
  constexpr int kRows = scan_rows(BY);
  constexpr int kTile = scan_tile(BY);
  constexpr int kWaveRows = kWave * kRows;  // rows per wave
  constexpr int L = kRows / VEC;            // accesses per thread
  __shared__ Seg<U> s_wave[kWaves];
  __shared__ U s_pre[kWaves];
  const uint32_t tile = take_ticket(ticket, ntiles);
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const uint64_t t0 = uint64_t(tile) * kTile;
  const uint64_t rem = n - t0;
  const uint32_t lim = rem >= uint64_t(kTile) ? uint32_t(kTile) : uint32_t(rem);
  const bool full = lim == uint32_t(kTile);
  // access k of lane l covers tile rows w * 1024 + k * 64 * VEC + l * VEC + (0 .. VEC - 1)
  auto lrow = [&](int k, int e) -> uint32_t {
    return uint32_t(w) * kWaveRows + uint32_t(k) * (kWave * VEC) + uint32_t(l) * VEC + e;
  };
  U v[kRows];
  uint8_t f[kRows];
  // scan_by in a full tile with 16-byte aligned starts: a wave's start bytes (its rows are
  // consecutive) are read with one 16-byte load per lane and passed to the row owners through
  // LDS, instead of a 2-byte load per access
  __shared__ uint8_t s_st[BY ? kWaves : 1][BY ? kWaveRows : 1];
  {
    const TIn *xt = x + t0;
    const uint8_t *st = BY ? starts + t0 : nullptr;
    if (VEC == 2 && full) {
      using In2 = struct alignas(2 * sizeof(TIn)) { TIn a, b; };
      using St2 = struct alignas(2) { uint8_t a, b; };
      bool wide = false;
      if constexpr (BY) {
        static_assert(kWaveRows == kWave * 16, "one 16-byte load per lane covers a wave's starts");
        wide = (reinterpret_cast<uintptr_t>(st) & 15) == 0;
        if (wide) {
          struct alignas(16) Q {
            uint32_t a[4];
          };
          *reinterpret_cast<Q *>(&s_st[w][l * 16]) =
              *reinterpret_cast<const Q *>(st + uint32_t(w) * kWaveRows + uint32_t(l) * 16);
        }
      }
#pragma unroll
      for (int k = 0; k < L; ++k) {
        const uint32_t r = lrow(k, 0);
        const In2 p = *reinterpret_cast<const In2 *>(xt + r);
        v[2 * k] = static_cast<U>(p.a);
        v[2 * k + 1] = static_cast<U>(p.b);
        if (BY && !wide) {
          const St2 q = *reinterpret_cast<const St2 *>(st + r);
          f[2 * k] = q.a;
          f[2 * k + 1] = q.b;
        }
      }
      if constexpr (BY) {
        // the wave reads back what it wrote (LDS operations of one wave complete in order)
        if (wide) {
#pragma unroll
          for (int k = 0; k < L; ++k) {
            const uint32_t o = lrow(k, 0) - uint32_t(w) * kWaveRows;
            const St2 q = *reinterpret_cast<const St2 *>(&s_st[w][o]);
            f[2 * k] = q.a;
            f[2 * k + 1] = q.b;
          }
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < L; ++k)
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          const uint32_t r = lrow(k, e);
          const uint32_t rr = r < lim ? r : lim - 1;
          v[k * VEC + e] = static_cast<U>(xt[rr]);
          if (BY) f[k * VEC + e] = st[rr];
        }
    }
  }
  // rows past the end are identities (they only ever follow the valid rows)
  if (!full) {
#pragma unroll
    for (int k = 0; k < L; ++k)
#pragma unroll
      for (int e = 0; e < VEC; ++e)
        if (lrow(k, e) >= lim) {
          v[k * VEC + e] = identity<U, OP>();
          if (BY) f[k * VEC + e] = 0;
        }
  }
  // wave scan, item by item, carrying the wave's running value; v[] becomes the inclusive
  // fold from the wave's first row (flag in f[]: a start since the wave's first row)
  Seg<U> carry{identity<U, OP>(), 0};
#pragma unroll
  for (int k = 0; k < L; ++k) {
    Seg<U> e0{v[k * VEC], BY ? uint32_t(f[k * VEC] != 0) : 0u};
    Seg<U> lane = e0;
    if (VEC == 2) lane = join<U, OP>(e0, Seg<U>{v[k * VEC + 1], BY ? uint32_t(f[k * VEC + 1] != 0) : 0u});
    const Seg<U> inc = wave_scan<U, OP, BY>(lane);
    if (VEC == 2) {
      // row 0 of the pair: the lanes before (wave_shr:1 of the inclusive) then e0
      const Seg<U> id{identity<U, OP>(), 0};
      const Seg<U> excl = dpp_seg<0x138, 0xF>(inc, id);
      const Seg<U> r0 = join<U, OP>(carry, join<U, OP>(excl, e0));
      v[k * VEC] = r0.v;
      if (BY) f[k * VEC] = uint8_t(r0.f);
    }
    const Seg<U> r1 = join<U, OP>(carry, inc);
    v[k * VEC + VEC - 1] = r1.v;
    if (BY) f[k * VEC + VEC - 1] = uint8_t(r1.f);
    carry = lane63(r1);
  }
  if (l == 0) s_wave[w] = carry;
  __syncthreads();
  // wave 0: the tile aggregate, publish, look back, then the prefix of every wave
  if (w == 0) {
    const Seg<U> id{identity<U, OP>(), 0};
    Seg<U> a = l < kWaves ? s_wave[l] : id;
    const Seg<U> inc = wave_scan<U, OP, BY>(a);  // lane l: waves 0..l
    const Seg<U> agg = lane63(inc);
    U pre = identity<U, OP>();
    if (tile == 0) {
      if (carry_in[1] != 0) pre = from_bits<U>(carry_in[0]);
    } else {
      if (l == 0) {
        st_rlx(&status[uint64_t(tile) * 4 + 0], to_bits(agg.v));
        publish_fence();
        st_rlx(&status[uint64_t(tile) * 4 + 2], (uint64_t(epoch) << 8) | 1 | (agg.f ? 4 : 0));
      }
      pre = look_back<U, OP>(status, epoch, tile, pre);
    }
    const Seg<U> incl = join<U, OP>(Seg<U>{pre, 0}, agg);
    if (l == 0) {
      st_rlx(&status[uint64_t(tile) * 4 + 1], to_bits(incl.v));
      publish_fence();
      st_rlx(&status[uint64_t(tile) * 4 + 2], (uint64_t(epoch) << 8) | 2 | 1 | (agg.f ? 4 : 0));
    }
    // wave l's prefix: the tile prefix joined with waves 0 .. l - 1
    const Seg<U> excl = dpp_seg<0x138, 0xF>(inc, id);
    if (l < kWaves) s_pre[l] = join<U, OP>(Seg<U>{pre, 0}, excl).v;
  }
  __syncthreads();
  const U wp = s_pre[w];
  U *ot = out + t0;
  auto finish = [&](int i) -> U { return BY && f[i] ? v[i] : apply<U, OP>(wp, v[i]); };
  if (VEC == 2 && full) {
    using Out2 = struct alignas(2 * sizeof(U)) { U a, b; };
#pragma unroll
    for (int k = 0; k < L; ++k) {
      Out2 o{finish(2 * k), finish(2 * k + 1)};
      *reinterpret_cast<Out2 *>(ot + lrow(k, 0)) = o;
    }
  } else {
#pragma unroll
    for (int k = 0; k < L; ++k)
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const uint32_t r = lrow(k, e);
        if (r < lim) ot[r] = finish(k * VEC + e);
      }
  }
  // the carry for the next update: the last row's value
  if (tile == ntiles - 1) {
    const uint32_t last = lim - 1;
#pragma unroll
    for (int k = 0; k < L; ++k)
#pragma unroll
      for (int e = 0; e < VEC; ++e)
        if (lrow(k, e) == last) {
          carry_out[0] = to_bits(finish(k * VEC + e));
          carry_out[1] = 1;
        }
  }
}

template <typename TIn, typename U, int OP, bool BY, int VEC>
void launch_as(Ctx &c, const void *x, const uint8_t *starts, uint64_t n, void *out, uint64_t *carry_in,
               uint64_t *carry, uint64_t *status, uint32_t epoch, uint32_t tiles) {
  k_scan<TIn, U, OP, BY, VEC><<<tiles, kScanBlock, 0, c.stream>>>(static_cast<const TIn *>(x), starts, n,
                                                                   static_cast<U *>(out), status, epoch,
                                                                   c.scratch.counters, tiles, carry_in, carry);
}

template <typename TIn, typename U, int OP>
void launch(Ctx &c, bool by, const void *x, const uint8_t *starts, uint64_t n, void *out,
            uint64_t *carry) {
  const uint64_t tile = scan_tile(by);
  const uint64_t tiles = (n + tile - 1) / tile;
  if (tiles > 0xFFFFFFFFull) usage_error("scan too large");
  uint64_t *status = status_words(c, ST_SCAN, tiles * 4);
  const uint32_t epoch = next_epoch(c);
  uint64_t *carry_in = c.scratch.slots + 8;
  copy_d2d(c, carry_in, carry, 16);
  // 16-byte output accesses when the input pairs, start pairs and outputs are aligned
  const bool vec = reinterpret_cast<uintptr_t>(x) % (2 * sizeof(TIn)) == 0 &&
                   reinterpret_cast<uintptr_t>(out) % (2 * sizeof(U)) == 0 &&
                   (!by || reinterpret_cast<uintptr_t>(starts) % 2 == 0);
  const uint32_t nt = static_cast<uint32_t>(tiles);
  if (by) {
    if (vec) launch_as<TIn, U, OP, true, 2>(c, x, starts, n, out, carry_in, carry, status, epoch, nt);
    else launch_as<TIn, U, OP, true, 1>(c, x, starts, n, out, carry_in, carry, status, epoch, nt);
  } else {
    if (vec) launch_as<TIn, U, OP, false, 2>(c, x, starts, n, out, carry_in, carry, status, epoch, nt);
    else launch_as<TIn, U, OP, false, 1>(c, x, starts, n, out, carry_in, carry, status, epoch, nt);
  }
  check_hip(hipGetLastError(), "scan launch");
}

// ---------------------------------------------------------------- f64 scan add
// Build-defined order (the reference has no f64 scan add, builtins.dd:50-54; tier 2 design
// :172 candidate): out[i] is the canonical reduction of x[0..=i]: the 16 accumulators after
// the whole blocks of x[0..=i], combined pairwise (+8, +4, +2, +1), then the rows of the
// partial block in index order. Every row needs the exact accumulators before its block.
//
// From 1024 whole blocks on: the exact canonical fold (k_canon.hip) gives the exact
// accumulators at the start of every 1024-block tile, then k_fscan_tiles runs each tile's 16
// chains through its blocks and finishes every row. Below that, one wave folds the
// accumulators block by block (k_fscan_acc) and k_fscan_rows finishes the rows.
constexpr int kFscanChunk = 64;  // blocks per step of k_fscan_tiles

// One wave per tile of 1024 blocks (the last may be shorter), 64 blocks at a time: the rows
// are staged in LDS (the next 64 blocks' rows are loaded into registers meanwhile); lanes
// 0..15 fold chain j from the tile's exact start, recording the accumulators before every
// block; then lane b finishes block b's 16 rows in place (the rows before the block's end from
// the pairwise combine of the accumulators before it, the last row from those after it), and
// the rows are written back coalesced. Rows of 16 doubles are padded to 17 in LDS, so that a
// lane reading its own block (lanes 128 bytes apart) does not hit one bank with every lane.
__global__ __launch_bounds__(64) void k_fscan_tiles(const double *__restrict__ x, uint64_t full,
                                                    uint64_t tile_blocks, const double *tstart,
                                                    double *__restrict__ out) {
  __shared__ double s_x[kFscanChunk * 17];
  __shared__ double s_acc[kFscanChunk + 1][17];
  auto px = [](int i) { return i + (i >> 4); };  // row i of the chunk in the padded layout
  const uint64_t tile = blockIdx.x;
  const int l = lane_id();
  const uint64_t b0 = tile * tile_blocks;
  const uint64_t nbt = full - b0 < tile_blocks ? full - b0 : tile_blocks;
  double a = l < 16 ? tstart[tile * 16 + l] : 0.0;
  double nx[16];
  auto load = [&](uint64_t c0) {
    const uint64_t r0 = (b0 + c0) * 16;
    const int rows = c0 < nbt ? static_cast<int>((nbt - c0 < uint64_t(kFscanChunk) ? nbt - c0 : kFscanChunk) * 16) : 0;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int k = l + 64 * m;
      nx[m] = k < rows ? x[r0 + k] : 0.0;
    }
  };
  load(0);
  for (uint64_t c0 = 0; c0 < nbt; c0 += kFscanChunk) {
    const int cb = static_cast<int>(nbt - c0 < uint64_t(kFscanChunk) ? nbt - c0 : kFscanChunk);
    const uint64_t r0 = (b0 + c0) * 16;
    const int rows = cb * 16;
#pragma unroll
    for (int m = 0; m < 16; ++m) s_x[px(l + 64 * m)] = nx[m];
    __syncthreads();
    load(c0 + kFscanChunk);
    if (l < 16) {
      if (cb == kFscanChunk) {
        double v[kFscanChunk];
#pragma unroll
        for (int b = 0; b < kFscanChunk; ++b) v[b] = s_x[b * 17 + l];
#pragma unroll
        for (int b = 0; b < kFscanChunk; ++b) {
          s_acc[b][l] = a;
          a = a + v[b];
        }
      } else {
        for (int b = 0; b < cb; ++b) {
          s_acc[b][l] = a;
          a = a + s_x[b * 17 + l];
        }
      }
      s_acc[cb][l] = a;
    }
    __syncthreads();
    if (l < cb) {
      double t[16], u[16], r[15];
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        t[j] = s_acc[l][j];
        u[j] = s_acc[l + 1][j];
      }
#pragma unroll
      for (int k = 0; k < 15; ++k) r[k] = s_x[l * 17 + k];
#pragma unroll
      for (int step = 8; step > 0; step /= 2)
#pragma unroll
        for (int j = 0; j < step; ++j) {
          t[j] = t[j] + t[j + step];
          u[j] = u[j] + u[j + step];
        }
      double v = t[0];
#pragma unroll
      for (int k = 0; k < 15; ++k) {
        v = v + r[k];
        s_x[l * 17 + k] = v;
      }
      s_x[l * 17 + 15] = u[0];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int k = l + 64 * m;
      if (k < rows) out[r0 + k] = s_x[px(k)];
    }
    __syncthreads();
  }
}

// The rows of a trailing partial block: the pairwise combine of the accumulators after every
// whole block, then the rows in index order.
__global__ void k_fscan_tail(const double *__restrict__ x, uint64_t r0, uint64_t n, const double *acc,
                             double *__restrict__ out) {
  if (threadIdx.x != 0) return;
  double t[16];
  for (int j = 0; j < 16; ++j) t[j] = acc[j];
  for (int step = 8; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) t[j] = t[j] + t[j + step];
  double v = t[0];
  for (uint64_t r = r0; r < n; ++r) {
    v = v + x[r];
    out[r] = v;
  }
}
__global__ __launch_bounds__(64) void k_fscan_acc(const double *__restrict__ x, uint64_t blocks,
                                                  double *acc_state, double *acc_before) {
  const int l = lane_id();
  double a = l < 16 ? acc_state[l] : 0.0;
  for (uint64_t b = 0; b < blocks; ++b) {
    if (l < 16) {
      acc_before[b * 16 + l] = a;
      a = a + x[b * 16 + l];
    }
  }
  if (l < 16) acc_state[l] = a;
}

__global__ void k_fscan_rows(const double *__restrict__ x, uint64_t n, const double *acc_before,
                             const double *acc_after_all, double *out) {
  const uint64_t b = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const uint64_t nb = (n + 15) / 16;
  if (b >= nb) return;
  const double *acc = acc_before + b * 16;
  double t[16];
  for (int j = 0; j < 16; ++j) t[j] = acc[j];
  for (int step = 8; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) t[j] = t[j] + t[j + step];
  double a = t[0];
  const uint64_t r0 = b * 16;
  const uint64_t r1 = r0 + 16 < n ? r0 + 16 : n;
  for (uint64_t r = r0; r < r1; ++r) {
    if (r == r0 + 15) {
      // the block completes: every row of it sits in the accumulators
      const double *na = (b + 1) * 16 <= n ? (b + 1 < nb ? acc_before + (b + 1) * 16 : acc_after_all)
                                           : acc_after_all;
      double u[16];
      for (int j = 0; j < 16; ++j) u[j] = na[j];
      for (int step = 8; step > 0; step /= 2)
        for (int j = 0; j < step; ++j) u[j] = u[j] + u[j + step];
      out[r] = u[0];
    } else {
      a = a + x[r];
      out[r] = a;
    }
  }
}

// prev (Prev<T, BOOL>, scan.rs:14-55): row i is row i - 1 of x; row 0 is the last row of
// the previous tile, or `fill` before the first row. The carry keeps the last row.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prev(const T *__restrict__ x, uint64_t n, T *__restrict__ out,
                                                 const uint64_t *carry_in, uint64_t fill_bits,
                                                 uint64_t *carry_out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x) {
    if (i == 0) {
      out[0] = carry_in[1] ? from_bits<T>(carry_in[0]) : from_bits<T>(fill_bits);
    } else {
      out[i] = x[i - 1];
    }
    if (i == n - 1) {
      carry_out[0] = to_bits(x[n - 1]);
      carry_out[1] = 1;
    }
  }
}

}  // namespace

void prev_rows(Ctx &c, uint32_t width, const void *x, uint64_t n, void *out, uint64_t *carry,
               uint64_t fill_bits) {
  if (n == 0) return;
  uint64_t *carry_in = c.scratch.slots + 10;
  copy_d2d(c, carry_in, carry, 16);
  const uint32_t g = stream_grid(n, kBlock * 4, 8192);
  switch (width) {
    case 8: k_prev<uint64_t><<<g, kBlock, 0, c.stream>>>(static_cast<const uint64_t *>(x), n,
                                                         static_cast<uint64_t *>(out), carry_in, fill_bits, carry);
      break;
    case 4: k_prev<uint32_t><<<g, kBlock, 0, c.stream>>>(static_cast<const uint32_t *>(x), n,
                                                         static_cast<uint32_t *>(out), carry_in, fill_bits, carry);
      break;
    default: k_prev<uint8_t><<<g, kBlock, 0, c.stream>>>(static_cast<const uint8_t *>(x), n,
                                                         static_cast<uint8_t *>(out), carry_in, fill_bits, carry);
      break;
  }
  check_hip(hipGetLastError(), "prev");
}

void scan_rows(Ctx &c, const ScanSpec &s, const void *x, const uint8_t *starts, uint64_t n,
               void *out, uint64_t *carry) {
  if (n == 0) return;
  if (s.op == 0) {
    switch (s.type) {
      case VK_I64: launch<int64_t, uint64_t, 0>(c, s.by, x, starts, n, out, carry); break;
      case VK_U64: launch<uint64_t, uint64_t, 0>(c, s.by, x, starts, n, out, carry); break;
      case VK_I32: launch<int32_t, uint64_t, 0>(c, s.by, x, starts, n, out, carry); break;
      case VK_U8: launch<uint8_t, uint64_t, 0>(c, s.by, x, starts, n, out, carry); break;
      default: usage_error("unsupported scan add type");
    }
    return;
  }
  const bool mx = s.op == 2;
  switch (s.type) {
    case VK_F64:
      if (mx) launch<double, double, 2>(c, s.by, x, starts, n, out, carry);
      else launch<double, double, 1>(c, s.by, x, starts, n, out, carry);
      break;
    case VK_I64:
      if (mx) launch<int64_t, int64_t, 2>(c, s.by, x, starts, n, out, carry);
      else launch<int64_t, int64_t, 1>(c, s.by, x, starts, n, out, carry);
      break;
    case VK_U64:
      if (mx) launch<uint64_t, uint64_t, 2>(c, s.by, x, starts, n, out, carry);
      else launch<uint64_t, uint64_t, 1>(c, s.by, x, starts, n, out, carry);
      break;
    default: usage_error("unsupported scan min/max type");
  }
}

// `state` holds 16 accumulators (whole-array use: start at +0.0); n rows from a block
// boundary.
void scan_add_f64(Ctx &c, const double *x, uint64_t n, double *out, double *state) {
  if (n == 0) return;
  const uint64_t full = n / 16, nb = (n + 15) / 16;
  const uint64_t tb = canon_tile_blocks();
  if (full >= tb) {
    const uint64_t tiles = (full + tb - 1) / tb;
    double *tstart = static_cast<double *>(tmp_alloc(c, tiles * 16 * 8));
    canon_tile_starts(c, x, full, state, tstart);
    k_fscan_tiles<<<uint32_t(tiles), 64, 0, c.stream>>>(x, full, tb, tstart, out);
    if (nb > full) k_fscan_tail<<<1, 64, 0, c.stream>>>(x, full * 16, n, state, out);
    check_hip(hipGetLastError(), "f64 scan (tiles)");
    tmp_free(c, tstart);
    return;
  }
  double *before = static_cast<double *>(tmp_alloc(c, (nb + 1) * 16 * 8));
  k_fscan_acc<<<1, 64, 0, c.stream>>>(x, full, state, before);
  // rows of a trailing partial block use the accumulators after all full blocks
  if (nb > full) copy_d2d(c, before + full * 16, state, 16 * 8);
  k_fscan_rows<<<stream_grid(nb, kBlock, 1u << 20), kBlock, 0, c.stream>>>(x, n, before, state, out);
  check_hip(hipGetLastError(), "f64 scan");
  tmp_free(c, before);
}

}  // namespace vk
