// Scans: scan_<op>_<t> and scan_by_<op>_<t> (Scan<T, O, BY>, src/kernels/scalar/scan.rs:
// 282-326) for op add (wrapping, widening i32->i64 and u8->u64, ScanSum scan.rs:142-161),
// min and max (IEEE minimum/maximum on floats, ScanExtreme scan.rs:164-183). A start byte
// restarts the fold; row 0 of the whole array always starts.
//
// One pass: an 8192-row tile (512 threads x 16 rows) is staged through LDS into 16 consecutive rows per thread,
// scanned per thread, per wave (shuffles) and per block, then joined to the tiles before
// it by decoupled look-back on (aggregate, has-start) pairs: a predecessor whose aggregate
// holds a start, or that published its inclusive value, ends the look-back. The operators
// are associative and commutative, so any association order gives the reference's bits.
// Algorithmic bytes: n x (in width [+1 start byte] + 8).
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

constexpr int kItems = 16;
constexpr int kScanBlock = 512;
constexpr int kTile = kScanBlock * kItems;  // 8192 rows: ~78 KB of LDS, two blocks per CU
constexpr int kWaves = kScanBlock / kWave;

template <typename U, int OP>
__device__ __forceinline__ U apply(U a, U b) {
  if constexpr (OP == 0) return a + b;
  else if constexpr (OP == 1) return min2(a, b);
  else return max2(a, b);
}

template <typename U>
struct Seg {
  U v;
  uint32_t f;  // a start lies in this span
  uint32_t pad;
};

template <typename U, int OP>
__device__ __forceinline__ Seg<U> join(Seg<U> a, Seg<U> b) {
  // a then b: b's value stands alone if b holds a start.
  return Seg<U>{b.f ? b.v : apply<U, OP>(a.v, b.v), a.f | b.f, 0};
}

// Status of a tile: [0] aggregate, [1] inclusive, [2] flag word
// (epoch << 8 | 1 aggregate published | 2 inclusive published | 4 aggregate has a start).
__device__ __forceinline__ bool stat_ready(uint64_t w, uint32_t epoch) {
  return (w >> 8) == epoch && (w & 3) != 0;
}

template <typename U, int OP>
__device__ U look_back(uint64_t *status, uint32_t epoch, uint32_t tile, bool *has_any) {
  U acc{};
  bool have = false;
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
    U v{};
    bool mine = l >= top && idx >= 0;
    if (mine) {
      uint64_t bits = (w & 2) ? ld_rlx(&status[idx * 4 + 1]) : ld_rlx(&status[idx * 4 + 0]);
      v = from_bits<U>(bits);
    }
    // Fold the contributing lanes (any order: the operators commute).
    const uint64_t contrib = __ballot(mine);
    if (contrib) {
      U r = v;
      bool rh = mine;
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        U o = shfl_xor_pod(r, m);
        int oh = __shfl_xor(static_cast<int>(rh), m);
        if (oh) r = rh ? apply<U, OP>(r, o) : o;
        rh = rh || oh;
      }
      acc = have ? apply<U, OP>(acc, r) : r;
      have = true;
    }
    if (top >= 0) break;
    end -= 64;
  }
  *has_any = have;
  return acc;
}

template <typename TIn, typename U, int OP, bool BY>
__global__ __launch_bounds__(kScanBlock) void k_scan(const TIn *__restrict__ x,
                                                 const uint8_t *__restrict__ starts, uint64_t n,
                                                 U *__restrict__ out, uint64_t *status,
                                                 uint32_t epoch, uint32_t *ticket, uint32_t ntiles,
                                                 const uint64_t *carry_in, uint64_t *carry_out) {
  __shared__ U s_v[kTile + kTile / 16];
  __shared__ uint8_t s_f[kTile];
  __shared__ Seg<U> s_wave[kWaves];
  __shared__ Seg<U> s_prefix;
  const uint32_t tile = take_ticket(ticket, ntiles);
  const uint64_t t0 = uint64_t(tile) * kTile;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  // carry from the previous update (tile contract, scan.rs:251-259)
  const bool has_carry = carry_in[1] != 0;

  // 1. coalesced load into LDS, padded one slot per 16 rows
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int lr = k * kScanBlock + t;
    const uint64_t r = t0 + lr;
    U v{};
    uint8_t f = 0;
    if (r < n) {
      v = static_cast<U>(x[r]);
      if (BY) f = starts[r] != 0;
      if (r == 0 && !has_carry) f = 1;
    }
    s_v[lr + (lr >> 4)] = v;
    s_f[lr] = f;
  }
  __syncthreads();
  // 2. the aggregate of the thread's 16 consecutive rows (the rows stay in LDS)
  Seg<U> run{U{}, 0, 0};
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int lr = t * kItems + j;
    const Seg<U> e{s_v[lr + (lr >> 4)], s_f[lr], 0};
    run = j == 0 ? e : join<U, OP>(run, e);
  }
  // 3. wave inclusive scan of thread aggregates
  Seg<U> inc = run;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    Seg<U> o = shfl_up_pod(inc, d);
    if (l >= d) inc = join<U, OP>(o, inc);
  }
  if (l == 63) s_wave[w] = inc;
  Seg<U> excl_in_wave = shfl_up_pod(inc, 1);
  __syncthreads();
  // 4. tile aggregate, publish, look back (wave 0)
  if (w == 0) {
    Seg<U> agg = s_wave[0];
    for (int k = 1; k < kWaves; ++k) agg = join<U, OP>(agg, s_wave[k]);
    Seg<U> pre{U{}, 0, 0};
    bool have = false;
    if (tile == 0) {
      if (has_carry) {
        pre = Seg<U>{from_bits<U>(carry_in[0]), 0, 0};
        have = true;
      }
    } else {
      if (l == 0) {
        st_rlx(&status[uint64_t(tile) * 4 + 0], to_bits(agg.v));
        publish_fence();
        st_rlx(&status[uint64_t(tile) * 4 + 2], (uint64_t(epoch) << 8) | 1 | (agg.f ? 4 : 0));
      }
      bool h;
      U pv = look_back<U, OP>(status, epoch, tile, &h);
      pre = Seg<U>{pv, 0, 0};
      have = h;
    }
    const Seg<U> incl = have ? join<U, OP>(pre, agg) : agg;
    if (l == 0) {
      st_rlx(&status[uint64_t(tile) * 4 + 1], to_bits(incl.v));
      publish_fence();
      st_rlx(&status[uint64_t(tile) * 4 + 2], (uint64_t(epoch) << 8) | 2 | 1 | (agg.f ? 4 : 0));
      s_prefix = pre;
      s_prefix.f = have ? 1 : 0;  // here f marks "a prefix exists"
    }
  }
  __syncthreads();
  // 5. this thread's exclusive prefix = tile prefix, waves before, lanes before
  Seg<U> pre{U{}, 0, 0};
  bool have = s_prefix.f != 0;
  if (have) pre = Seg<U>{s_prefix.v, 0, 0};
  for (int k = 0; k < w; ++k) {
    pre = have ? join<U, OP>(pre, s_wave[k]) : s_wave[k];
    have = true;
  }
  if (l > 0) {
    pre = have ? join<U, OP>(pre, excl_in_wave) : excl_in_wave;
    have = true;
  }
  // the thread's rows again, folded from the prefix (a start restarts the fold)
  U cur = pre.v;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int lr = t * kItems + j;
    const U v = s_v[lr + (lr >> 4)];
    cur = (s_f[lr] || (j == 0 && !have)) ? v : apply<U, OP>(cur, v);
    s_v[lr + (lr >> 4)] = cur;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int lr = k * kScanBlock + t;
    const uint64_t r = t0 + lr;
    if (r < n) {
      const U v = s_v[lr + (lr >> 4)];
      out[r] = v;
      if (r == n - 1) {
        carry_out[0] = to_bits(v);
        carry_out[1] = 1;
      }
    }
  }
}

template <typename TIn, typename U, int OP>
void launch(Ctx &c, bool by, const void *x, const uint8_t *starts, uint64_t n, void *out,
            uint64_t *carry) {
  const uint64_t tiles = (n + kTile - 1) / kTile;
  uint64_t *status = status_words(c, ST_SCAN, tiles * 4);
  const uint32_t epoch = next_epoch(c);
  uint64_t *carry_in = c.scratch.slots + 8;
  copy_d2d(c, carry_in, carry, 16);
  if (by)
    k_scan<TIn, U, OP, true><<<uint32_t(tiles), kScanBlock, 0, c.stream>>>(
        static_cast<const TIn *>(x), starts, n, static_cast<U *>(out), status, epoch,
        c.scratch.counters, uint32_t(tiles), carry_in, carry);
  else
    k_scan<TIn, U, OP, false><<<uint32_t(tiles), kScanBlock, 0, c.stream>>>(
        static_cast<const TIn *>(x), starts, n, static_cast<U *>(out), status, epoch,
        c.scratch.counters, uint32_t(tiles), carry_in, carry);
  check_hip(hipGetLastError(), "scan launch");
}

// ---------------------------------------------------------------- f64 scan add
// Build-defined order (the reference has no f64 scan add, builtins.dd:50-54; tier 2 design
// :172 candidate): out[i] is the canonical reduction of x[0..=i]. Pass 1 (one wave) folds
// the 16 accumulators block by block and records them before every block; pass 2 finishes
// every row of a block in parallel: pairwise combine of the recorded accumulators, then the
// block's rows up to the row in index order.
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
  double *before = static_cast<double *>(tmp_alloc(c, (nb + 1) * 16 * 8));
  k_fscan_acc<<<1, 64, 0, c.stream>>>(x, full, state, before);
  // rows of a trailing partial block use the accumulators after all full blocks
  if (nb > full) copy_d2d(c, before + full * 16, state, 16 * 8);
  k_fscan_rows<<<stream_grid(nb, kBlock, 1u << 20), kBlock, 0, c.stream>>>(x, n, before, state, out);
  check_hip(hipGetLastError(), "f64 scan");
  tmp_free(c, before);
}

}  // namespace vk
