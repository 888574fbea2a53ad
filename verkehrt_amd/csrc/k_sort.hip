// argsort_<t> (ArgSort<T>, src/kernels/scalar/index.rs:157-173; radix.rs:1-259): the
// positions of the rows in stable ascending order, as i64. The stable ascending permutation
// is unique, so any stable sort returns the reference's output; this one keeps the
// reference's structure of an LSD radix sort on (key image, position) pairs with byte
// digits, one counting pass for every digit, and digits that are equal in every row
// skipped (radix.rs:187-259).
//
// MI355X form: the counting pass builds all digit histograms in LDS per block. Each
// remaining digit is one pass over 16384-row tiles (1024 threads x 16 rows): the rows of a wave are ranked among
// equal digits with eight ballots (a wave64 match), per-wave digit counters in LDS give
// ranks in row order, each tile publishes its 256 digit counts and learns the counts of
// the tiles before it by decoupled look-back per digit, and the rows are staged sorted
// through LDS so the scatter writes runs of consecutive addresses. Positions are u32.
//
// Key images (radix.rs:51-101): i64/i32 flip the sign bit; u64/u8 are themselves. f64 has
// no reference kernel (builtins.dd:73-77); this project's order maps negative values (sign
// bit set) to their complement and others to a flipped sign bit, so
// -NaN < -inf < ... < -0.0 < +0.0 < ... < +inf < +NaN.
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

constexpr int kItems = 16;
constexpr int kSortBlock = 1024;             // threads of a pass block
constexpr int kTile = kSortBlock * kItems;   // 16384 rows
constexpr int kWaves = kSortBlock / kWave;
constexpr int kRadix = 256;
constexpr int32_t kRawU32 = 100;    // internal key type: u32 keys sorted as they are
constexpr int32_t kChainKey = 101;  // internal key type: i64 bucket k of row r as k * 16 + r % 16

template <typename U>
__device__ __forceinline__ U image(int32_t type, const void *keys, uint64_t r) {
  switch (type) {
    case VK_I64: return U(static_cast<const uint64_t *>(keys)[r] ^ 0x8000000000000000ull);
    case VK_F64: {
      const uint64_t b = static_cast<const uint64_t *>(keys)[r];
      return U((b >> 63) ? ~b : (b ^ 0x8000000000000000ull));
    }
    case VK_I32: return U(static_cast<const uint32_t *>(keys)[r] ^ 0x80000000u);
    case VK_U8: return U(static_cast<const uint8_t *>(keys)[r]);
    case kRawU32: return U(static_cast<const uint32_t *>(keys)[r]);
    case kChainKey: return U(static_cast<const uint64_t *>(keys)[r] * 16 + (r & 15));
    default: return U(static_cast<const uint64_t *>(keys)[r]);  // u64
  }
}

// Calls f(img) with the key-image function of `type` fixed, so that a loop inside f sees one
// load form and can issue all its loads together (a switch on the type per row would sit
// between them).
template <typename U, typename F>
__device__ __forceinline__ void with_image(int32_t type, const void *keys, F &&f) {
  switch (type) {
    case VK_I64: f([&](uint64_t r) { return image<U>(VK_I64, keys, r); }); break;
    case VK_F64: f([&](uint64_t r) { return image<U>(VK_F64, keys, r); }); break;
    case VK_I32: f([&](uint64_t r) { return image<U>(VK_I32, keys, r); }); break;
    case VK_U8: f([&](uint64_t r) { return image<U>(VK_U8, keys, r); }); break;
    case kRawU32: f([&](uint64_t r) { return image<U>(kRawU32, keys, r); }); break;
    case kChainKey: f([&](uint64_t r) { return image<U>(kChainKey, keys, r); }); break;
    default: f([&](uint64_t r) { return image<U>(VK_U64, keys, r); }); break;
  }
}

// ------------------------------------------------------------- counting pass

// Adds one row's digits to the block's LDS histograms. The lanes that share the first active
// lane's digit (all of them for the constant high digits of small keys and runs of equal keys,
// many for skewed keys) are added once, by that lane, instead of by one atomic each on one LDS
// word; the other lanes add themselves.
template <typename U>
__device__ __forceinline__ void hist_add(uint32_t (&s_h)[sizeof(U)][kRadix], U k) {
  const uint64_t act = __ballot(1);
  const int lead = __ffsll(static_cast<unsigned long long>(act)) - 1;
#pragma unroll
  for (int d = 0; d < static_cast<int>(sizeof(U)); ++d) {
    const uint32_t dig = static_cast<uint32_t>(k >> (8 * d)) & 255;
    const uint32_t first = __builtin_amdgcn_readfirstlane(dig);
    const uint64_t same = __ballot(dig == first);
    if (static_cast<int>(lane_id()) == lead) atomicAdd(&s_h[d][first], static_cast<uint32_t>(__popcll(same)));
    if (dig != first) atomicAdd(&s_h[d][dig], 1u);
  }
}

constexpr int kHistAhead = 8;

// CHECK (chain keys only): the bucket keys are also checked against 0..limit, the least row
// holding one outside lowering *bad, so the caller can refuse the input before any pass runs.
template <typename U, bool CHECK = false>
__global__ __launch_bounds__(kBlock) void k_hist(int32_t type, const void *keys, uint64_t n,
                                                 uint32_t *hist, uint64_t limit = 0,
                                                 uint64_t *bad = nullptr) {
  constexpr int P = sizeof(U);
  __shared__ uint32_t s_h[P][kRadix];
  for (int k = threadIdx.x; k < P * kRadix; k += kBlock) (&s_h[0][0])[k] = 0;
  __syncthreads();
  uint64_t my_bad = ~uint64_t(0);
  // kHistAhead rows per thread loaded before their ballots (which are convergent: a load in the
  // same iteration as a ballot would wait for the previous row's)
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t r0 = blockIdx.x * uint64_t(kBlock) + threadIdx.x; r0 < n; r0 += stride * kHistAhead) {
    U k[kHistAhead];
    if (CHECK) {
#pragma unroll
      for (int u = 0; u < kHistAhead; ++u) {
        const uint64_t r = r0 + u * stride;
        const uint64_t raw = r < n ? static_cast<const uint64_t *>(keys)[r] : 0;
        k[u] = U(raw * 16 + (r & 15));
        if (r < n && raw >= limit && r < my_bad) my_bad = r;
      }
    } else {
      with_image<U>(type, keys, [&](auto img) {
#pragma unroll
        for (int u = 0; u < kHistAhead; ++u) {
          const uint64_t r = r0 + u * stride;
          k[u] = r < n ? img(r) : U(0);
        }
      });
    }
#pragma unroll
    for (int u = 0; u < kHistAhead; ++u)
      if (r0 + u * stride < n) hist_add<U>(s_h, k[u]);
  }
  if (CHECK) {
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
      const uint64_t o = __shfl_xor(my_bad, m);
      my_bad = o < my_bad ? o : my_bad;
    }
    if (lane_id() == 0 && my_bad != ~uint64_t(0))
      atomicMin(reinterpret_cast<unsigned long long *>(bad), static_cast<unsigned long long>(my_bad));
  }
  __syncthreads();
  for (int k = threadIdx.x; k < P * kRadix; k += kBlock) {
    const uint32_t v = (&s_h[0][0])[k];
    if (v) atomicAdd(&hist[k], v);
  }
}

// ------------------------------------------------------------------ one pass

__device__ __forceinline__ uint64_t pack(uint32_t epoch, uint64_t flag, uint64_t v) {
  return (uint64_t(epoch) << 48) | (flag << 46) | v;
}
constexpr uint64_t kCountMask = (uint64_t(1) << 46) - 1;

// The arguments of one pass. Batched passes (BATCH) take each tile's item from the batch
// tables: the item's typed keys (`a`), rows (`n`), first tile (`tile0`), first row in the
// concatenated key and position buffers (`aux`) and its i64 output (`out`); its digit bases
// are at digit_base + (item * P + d) * 256.
// The payload V travels with the key: u32 row positions (argsort: FIRST takes the row number,
// LAST writes it to out_pos as i64) or u64 value bits (FIRST takes vals[row], LAST writes the
// sorted keys and values; batched items take their values from `b`).
template <typename U, typename V>
struct PassArgs {
  int32_t type;
  const void *keys_in;  // FIRST: typed keys; otherwise key images
  const V *pay_in;
  const uint64_t *vals;  // FIRST with u64 payload: the values
  uint64_t n;
  int shift;
  const uint32_t *digit_base;
  U *keys_out;
  V *pay_out;
  int64_t *out_pos;
  uint64_t *status;
  uint32_t epoch;
  uint32_t *ticket;
  uint32_t ntiles;
  const BatchItem *items;
  const uint32_t *tile_item;
  int P, d;
  // batch plans: the pass is slot `slot` of the digits listed on the device (slots[0] = their
  // count, slots[1 + k] = the k-th), and the epoch is read from epoch_ptr; a launch whose slot
  // is past the list, or is (or is not) the last one against LAST, exits before taking a ticket
  const int32_t *slots;
  int slot;
  const uint32_t *epoch_ptr;
};

// FIRST: keys are the typed input (imaged here); the payload starts as described above.
// LAST: positions go to out_pos as i64 (and keys to keys_out when given); values go with
// their keys to keys_out / pay_out.
template <typename U, typename V, bool FIRST, bool LAST, bool BATCH>
__global__ __launch_bounds__(kSortBlock) void k_pass(const PassArgs<U, V> a) {
  constexpr bool POS = std::is_same<V, uint32_t>::value;
  __shared__ uint32_t s_wcnt[kWaves][kRadix];  // per-wave counters, then per-wave offsets
  __shared__ uint32_t s_excl[kRadix];           // tile-local exclusive digit offsets
  __shared__ uint64_t s_gbase[kRadix];          // global position of the tile's first row per digit
  __shared__ uint64_t s_stage[kTile];          // one column of the sorted tile at a time
  __shared__ uint32_t s_wsum[kWaves];

  const int32_t type = a.type;
  int pd = a.d, shift = a.shift;
  uint32_t epoch = a.epoch;
  if (a.slots) {
    const int m = a.slots[0];
    if (a.slot >= m || (a.slot == m - 1) != LAST) return;  // uniform over the grid
    pd = a.slots[1 + a.slot];
    shift = 8 * pd;
    epoch = *a.epoch_ptr;
  }
  const uint32_t tile = take_ticket(a.ticket, a.ntiles);
  uint32_t first = 0;
  const void *keys_in = a.keys_in;
  const V *__restrict__ pos_in = a.pay_in;
  const uint64_t *__restrict__ vals = a.vals;
  uint64_t n = a.n;
  const uint32_t *__restrict__ digit_base = a.digit_base;
  U *__restrict__ keys_out = a.keys_out;
  V *__restrict__ pos_out = a.pay_out;
  int64_t *__restrict__ out_pos = a.out_pos;
  uint64_t *status = a.status;
  if (BATCH) {
    const uint32_t it = a.tile_item[tile];
    const BatchItem &I = a.items[it];
    n = I.n;
    first = static_cast<uint32_t>(I.tile0);
    const uint64_t row0 = I.aux;
    keys_in = FIRST ? I.a : static_cast<const void *>(static_cast<const U *>(a.keys_in) + row0);
    if (FIRST) vals = static_cast<const uint64_t *>(I.b);
    else pos_in += row0;
    digit_base += (uint64_t(it) * a.P + pd) * kRadix;
    if (LAST && POS) {
      out_pos = static_cast<int64_t *>(I.out);
      keys_out = nullptr;
    } else {
      keys_out += row0;
      pos_out += row0;
    }
  }
  const uint64_t t0 = uint64_t(tile - first) * kTile;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  for (int k = t; k < kWaves * kRadix; k += kSortBlock) (&s_wcnt[0][0])[k] = 0;
  __syncthreads();

  U key[kItems];
  V pos[kItems];
  uint32_t rnk[kItems];
  const uint64_t wbase = t0 + uint64_t(w) * (kTile / kWaves);
  const uint64_t below = (uint64_t(1) << l) - 1;
  // every load of the tile first: the ballots below are convergent, so loads written in the
  // same loop would each wait for the previous item's ballots
  auto load_tile = [&](auto img) {
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint64_t r = wbase + uint64_t(i) * kWave + l;
      if (r < n) {
        key[i] = img(r);
        if (FIRST) pos[i] = POS ? static_cast<V>(r) : static_cast<V>(vals[r]);
        else pos[i] = pos_in[r];
      } else {
        key[i] = 0;
        pos[i] = 0;
      }
    }
  };
  if (FIRST) with_image<U>(type, keys_in, load_tile);  // the key type fixed outside the loop
  else load_tile([&](uint64_t r) { return static_cast<const U *>(keys_in)[r]; });
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t r = wbase + uint64_t(i) * kWave + l;
    const bool valid = r < n;
    const uint32_t d = (key[i] >> shift) & 255;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? bal : ~bal;
    }
    uint32_t r_in = 0;
    if (valid) {
      r_in = s_wcnt[w][d] + __popcll(peers & below);
      const int leader = __builtin_ctzll(peers);
      if (l == leader) s_wcnt[w][d] += __popcll(peers);
    }
    rnk[i] = r_in;
  }
  __syncthreads();
  // per digit: wave offsets and the tile count (thread t < 256 owns digit t; the first four
  // waves)
  uint32_t cnt = 0, inc = 0;
  if (t < kRadix) {
    const int dd = t;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      const uint32_t v = s_wcnt[k][dd];
      s_wcnt[k][dd] = cnt;
      cnt += v;
    }
    // tile-local exclusive scan of the 256 digit counts
    inc = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(inc, d);
      if (l >= d) inc += o;
    }
    if (l == 63) s_wsum[w] = inc;
  }
  __syncthreads();
  // publish this tile's count of digit t and look back for the tiles before it
  if (t < kRadix) {
    uint32_t wpre = 0;
    for (int k = 0; k < w; ++k) wpre += s_wsum[k];
    s_excl[t] = wpre + inc - cnt;
    const int dd = t;
    uint64_t *st = status + uint64_t(tile) * kRadix + dd;
    uint64_t prefix = 0;
    if (tile == first) {
      st_rlx(st, pack(epoch, 2, cnt));
    } else {
      st_rlx(st, pack(epoch, 1, cnt));
      // the walk reads kLook predecessors per step (independent loads, one latency); a
      // tile before `first` counts as an inclusive zero
      constexpr int kLook = 8;
      int64_t j = int64_t(tile) - 1;
      bool done = false;
      while (!done) {
        uint64_t wd[kLook];
#pragma unroll
        for (int u = 0; u < kLook; ++u) {
          const int64_t jj = j - u;
          wd[u] = jj >= int64_t(first) ? ld_rlx(status + uint64_t(jj) * kRadix + dd) : pack(epoch, 2, 0);
        }
        int u = 0;
        for (; u < kLook; ++u) {
          const uint64_t flag = (wd[u] >> 46) & 3;
          if ((wd[u] >> 48) != epoch || flag == 0) break;  // not published yet: resume there
          prefix += wd[u] & kCountMask;
          if (flag == 2) {
            done = true;
            break;
          }
        }
        if (!done) {
          j -= u;
          if (u < kLook) backoff();
        }
      }
      st_rlx(st, pack(epoch, 2, prefix + cnt));
    }
    s_gbase[dd] = uint64_t(digit_base[dd]) + prefix;
  }
  __syncthreads();
  // Stage the tile sorted by digit one column at a time through one LDS buffer (128 KB; keys
  // and positions together would not fit the 160 KB of a compute unit): the key round
  // gives every sorted slot its destination, kept in registers for the payload round.
  const uint32_t rows = static_cast<uint32_t>(n - t0 < uint64_t(kTile) ? n - t0 : uint64_t(kTile));
  uint32_t at[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t r = wbase + uint64_t(i) * kWave + l;
    const uint32_t d = (key[i] >> shift) & 255;
    at[i] = r < n ? s_excl[d] + s_wcnt[w][d] + rnk[i] : 0u;
  }
  U *sk = reinterpret_cast<U *>(s_stage);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t r = wbase + uint64_t(i) * kWave + l;
    if (r < n) sk[at[i]] = key[i];
  }
  __syncthreads();
  constexpr int kSlots = kTile / kSortBlock;
  uint32_t gj[kSlots];  // destinations below 2^32: at most 2^32 - 1 rows per sort
#pragma unroll
  for (int q = 0; q < kSlots; ++q) {
    const uint32_t j = uint32_t(t) + uint32_t(q) * kSortBlock;
    gj[q] = 0;
    if (j < rows) {
      const U k = sk[j];
      const uint32_t d = (k >> shift) & 255;
      const uint64_t g = s_gbase[d] + (j - s_excl[d]);
      gj[q] = static_cast<uint32_t>(g);
      if (keys_out) keys_out[g] = k;
    }
  }
  __syncthreads();
  V *sp = reinterpret_cast<V *>(s_stage);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t r = wbase + uint64_t(i) * kWave + l;
    if (r < n) sp[at[i]] = pos[i];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kSlots; ++q) {
    const uint32_t j = uint32_t(t) + uint32_t(q) * kSortBlock;
    if (j < rows) {
      if (LAST && POS) out_pos[gj[q]] = static_cast<int64_t>(sp[j]);
      else pos_out[gj[q]] = sp[j];
    }
  }
}

__global__ void k_iota64(uint64_t n, int64_t *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = static_cast<int64_t>(i);
}

template <typename U>
__global__ void k_copy_images(int32_t type, const void *keys, uint64_t n, U *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = image<U>(type, keys, i);
}

// ------------------------------------------------- top-digit buckets, local sorts
// Roughly uniform keys need not go through every digit in HBM. Two passes by the two most
// significant non-constant digits order the (image, position) pairs by a 16-bit bucket; each
// bucket, a few thousand rows, is then sorted by its remaining digits inside one block's LDS
// and its positions written once. The result is the same stable permutation.

constexpr uint32_t kBuckets = 65536;

template <typename U>
__device__ __forceinline__ uint32_t bucket16(U k, int sh_hi, int sh_lo) {
  return ((static_cast<uint32_t>(k >> sh_hi) & 255) << 8) | (static_cast<uint32_t>(k >> sh_lo) & 255);
}

// start[b] = the first row of bucket b in keys sorted by bucket (start[] holds n beforehand,
// so buckets after the last row keep n): the row where the bucket changes writes the start of
// every bucket from the one after its predecessor's up to its own.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_bucket_starts(const U *__restrict__ k, uint64_t n, int sh_hi,
                                                          int sh_lo, uint32_t *__restrict__ start) {
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const int64_t id = bucket16<U>(k[i], sh_hi, sh_lo);
    const int64_t prev = i ? int64_t(bucket16<U>(k[i - 1], sh_hi, sh_lo)) : -1;
    for (int64_t b = prev + 1; b <= id; ++b) start[b] = static_cast<uint32_t>(i);
  }
}

struct LocalDigits {
  int n;
  int shift[8];
};

// The rows [s, s + m) of (keys, pos), ordered by their top digits (a bucket or a segment) and,
// inside, by row, sorted by one block and their positions written to out[0 .. m) as i64:
// - rows that all hold one key are in order as they stand;
// - otherwise at most THREADS * ITEMS rows are loaded into registers and ordered by the local
//   digits `dg` (ascending) in LDS: one stable counting pass by the most significant local
//   digit (wave64 match ranks, per-wave digit counters, rows written to LDS in digit order)
//   splits them into sub-buckets, each in row order; when none holds more than kSubMax rows,
//   every row's place in its sub-bucket is the number of rows there with a smaller key, or an
//   equal key and an earlier place; else the block runs a stable pass per local digit, least
//   significant first, from the rows as loaded.
// Returns false, writing nothing, for more rows than that with more than one key: the caller
// lists them for the batched passes.
constexpr uint32_t kSubMax = 96;

template <typename U, int THREADS, int ITEMS>
__device__ __forceinline__ bool local_sort_rows(const U *__restrict__ keys, const uint32_t *__restrict__ pos,
                                                uint64_t s, uint32_t m, const LocalDigits &dg,
                                                int64_t *__restrict__ out) {
  constexpr int C = THREADS * ITEMS;
  constexpr int W = THREADS / kWave;
  static_assert(THREADS >= kRadix, "one thread per digit in the digit scan");
  // per-wave digit counts stay below 2^16 (at most ITEMS * 64 rows per wave): 16-bit counters
  // keep the LDS of the 3072-row form small enough for three blocks per compute unit
  using Cnt = typename std::conditional<(ITEMS * kWave < 65536), uint16_t, uint32_t>::type;
  __shared__ U s_key[C];
  __shared__ uint32_t s_pos[C];
  __shared__ Cnt s_wcnt[W][kRadix];
  __shared__ uint32_t s_excl[kRadix + 1];
  __shared__ uint32_t s_wsum[kRadix / kWave];
  __shared__ uint32_t s_max;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  if (m == 0) return true;
  if (m > static_cast<uint32_t>(C)) {
    // too many rows for LDS: written as they stand if they all hold one key
    const U k0 = keys[s];
    bool differ = false;
    for (uint32_t j = t; j < m && !differ; j += THREADS) differ = keys[s + j] != k0;
    if (__syncthreads_or(differ)) return false;
    for (uint32_t j = t; j < m; j += THREADS) out[j] = static_cast<int64_t>(pos[s + j]);
    return true;
  }
  if (dg.n == 0) {  // the rows are in order already
    for (uint32_t j = t; j < m; j += THREADS) out[j] = static_cast<int64_t>(pos[s + j]);
    return true;
  }
  // rows are dealt to the waves in equal contiguous runs of ipw * 64
  const uint32_t ipw = (m + W * kWave - 1) / (W * kWave);
  const uint32_t wbase = uint32_t(w) * ipw * kWave;
  const uint64_t below = (uint64_t(1) << l) - 1;
  U key[ITEMS];
  uint32_t p[ITEMS];
  const U k0 = keys[s];
  bool differ = false;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t r = wbase + uint32_t(i) * kWave + l;
    const bool ok = uint32_t(i) < ipw && r < m;
    key[i] = ok ? keys[s + r] : U(0);
    p[i] = ok ? pos[s + r] : 0u;
    differ = differ || (ok && key[i] != k0);
  }
  // rows of one key are in order already
  if (!__syncthreads_or(differ)) {
    for (uint32_t j = t; j < m; j += THREADS) out[j] = static_cast<int64_t>(pos[s + j]);
    return true;
  }
  // One stable pass by the digit at `shift` over the rows in registers: they land in LDS in
  // digit order, s_excl holds the digit starts (s_excl[256] = m) and s_max the largest count.
  auto lds_pass = [&](int shift) {
    uint32_t rnk[ITEMS];
    for (int q = t; q < W * kRadix; q += THREADS) (&s_wcnt[0][0])[q] = 0;
    if (t == 0) s_max = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      rnk[i] = 0;
      if (uint32_t(i) < ipw) {  // uniform across the block
        const uint32_t r = wbase + uint32_t(i) * kWave + l;
        const bool valid = r < m;
        const uint32_t d = static_cast<uint32_t>(key[i] >> shift) & 255;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
          const uint64_t bal = __ballot((d >> bit) & 1);
          peers &= ((d >> bit) & 1) ? bal : ~bal;
        }
        if (valid) {
          rnk[i] = s_wcnt[w][d] + __popcll(peers & below);
          if (l == __builtin_ctzll(peers)) s_wcnt[w][d] += __popcll(peers);
        }
      }
    }
    __syncthreads();
    uint32_t cnt = 0, inc = 0;
    if (t < kRadix) {
#pragma unroll
      for (int q = 0; q < W; ++q) {
        const uint32_t v = s_wcnt[q][t];
        s_wcnt[q][t] = cnt;
        cnt += v;
      }
      inc = cnt;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o);
        if (l >= o) inc += u;
      }
      if (l == kWave - 1) s_wsum[w] = inc;
      uint32_t mx = cnt;
#pragma unroll
      for (int o = kWave / 2; o >= 1; o >>= 1) {
        const uint32_t u = __shfl_xor(mx, o);
        mx = u > mx ? u : mx;
      }
      if (l == 0) atomicMax(&s_max, mx);
    }
    __syncthreads();
    if (t < kRadix) {
      uint32_t wpre = 0;
      for (int q = 0; q < w; ++q) wpre += s_wsum[q];
      s_excl[t] = wpre + inc - cnt;
      if (t == 0) s_excl[kRadix] = m;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t r = wbase + uint32_t(i) * kWave + l;
      if (uint32_t(i) < ipw && r < m) {
        const uint32_t d = static_cast<uint32_t>(key[i] >> shift) & 255;
        const uint32_t at = s_excl[d] + s_wcnt[w][d] + rnk[i];
        s_key[at] = key[i];
        s_pos[at] = p[i];
      }
    }
    __syncthreads();
  };
  const int top = dg.shift[dg.n - 1];
  lds_pass(top);
  const bool counted = s_max <= kSubMax;
  __syncthreads();  // every wave has read s_max before a further pass clears it
  if (counted) {
    // each row's place, kept in registers until every row has been read, then the positions
    // placed in LDS and written out in order
    uint32_t dst[ITEMS], pv[ITEMS];
    constexpr int kPosBits = 24;  // LDS places (C <= 2^24)
    static_assert(C <= (1 << kPosBits), "places fit the composite");
    if (top + kPosBits <= 8 * static_cast<int>(sizeof(U))) {
      // The rows of a sub-bucket share every bit from `top` up, so inside it (key, place)
      // orders as (the key's bits below top, place): each key in LDS becomes that one word,
      // unique, and a row's place in its sub-bucket is the number of smaller words there.
      const U lowmask = (U(1) << top) - 1;
      uint32_t lo[ITEMS], hi[ITEMS];
      U cj[ITEMS];
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const uint32_t j = uint32_t(t) + uint32_t(i) * THREADS;
        lo[i] = 0;
        hi[i] = 0;
        pv[i] = 0;
        cj[i] = 0;
        if (j < m) {
          const U kj = s_key[j];
          const uint32_t d = static_cast<uint32_t>(kj >> top) & 255;
          lo[i] = s_excl[d];
          hi[i] = s_excl[d + 1];
          pv[i] = s_pos[j];
          cj[i] = ((kj & lowmask) << kPosBits) | U(j);
        }
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const uint32_t j = uint32_t(t) + uint32_t(i) * THREADS;
        if (j < m) s_key[j] = cj[i];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const uint32_t j = uint32_t(t) + uint32_t(i) * THREADS;
        dst[i] = 0;
        if (j < m) {
          // four reads in flight per step; past hi a read stands for no row
          uint32_t r = 0;
          for (uint32_t q = lo[i]; q < hi[i]; q += 4) {
            U kq[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) kq[u] = q + u < hi[i] ? s_key[q + u] : ~U(0);
#pragma unroll
            for (int u = 0; u < 4; ++u) r += kq[u] < cj[i] ? 1u : 0u;
          }
          dst[i] = lo[i] + r;
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const uint32_t j = uint32_t(t) + uint32_t(i) * THREADS;
        dst[i] = 0;
        pv[i] = 0;
        if (j < m) {
          const U kj = s_key[j];
          const uint32_t d = static_cast<uint32_t>(kj >> top) & 255;
          const uint32_t lo = s_excl[d], hi = s_excl[d + 1];
          // four reads in flight per step (the sub-bucket holds j, so hi - 1 is in it)
          uint32_t r = 0;
          for (uint32_t q = lo; q < hi; q += 4) {
            U kq[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) kq[u] = s_key[q + u < hi ? q + u : hi - 1];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              const uint32_t qq = q + u;
              r += (qq < hi && (kq[u] < kj || (kq[u] == kj && qq < j))) ? 1u : 0u;
            }
          }
          dst[i] = lo + r;
          pv[i] = s_pos[j];
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if (uint32_t(t) + uint32_t(i) * THREADS < m) s_pos[dst[i]] = pv[i];
    __syncthreads();
    for (uint32_t j = t; j < m; j += THREADS) out[j] = static_cast<int64_t>(s_pos[j]);
    return true;
  }
  // a sub-bucket too large to count: every local digit, from the rows as loaded
  for (int k = 0; k < dg.n; ++k) {
    lds_pass(dg.shift[k]);
    if (k + 1 < dg.n) {
#pragma unroll
      for (int i = 0; i < ITEMS; ++i) {
        const uint32_t r = wbase + uint32_t(i) * kWave + l;
        if (uint32_t(i) < ipw && r < m) {
          key[i] = s_key[r];
          p[i] = s_pos[r];
        }
      }
    }
  }
  for (uint32_t j = t; j < m; j += THREADS) out[j] = static_cast<int64_t>(s_pos[j]);
  return true;
}

// One block per bucket b: rows [start[b], start[b + 1]); a bucket the block cannot sort is
// listed in overflow[1..] (count in overflow[0]).
template <typename U, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_local_sort(const U *__restrict__ keys, const uint32_t *__restrict__ pos,
                                                        const uint32_t *__restrict__ start, const LocalDigits dg,
                                                        int64_t *__restrict__ out, uint32_t *__restrict__ overflow) {
  const uint32_t b = blockIdx.x;
  const uint32_t s = start[b];
  if (!local_sort_rows<U, THREADS, ITEMS>(keys, pos, s, start[b + 1] - s, dg, out + s) && threadIdx.x == 0)
    overflow[1 + atomicAdd(&overflow[0], 1u)] = b;
}

// One block per segment j of a batched sort: rows [seg[j], seg[j + 1]) (the last to `total`) of
// the concatenated rows, inside one item (an item's first row starts a segment); the item is
// found by binary search of the items' first rows (aux) and the positions go to its output. A
// segment the block cannot sort is listed in overflow[1..]. With ids, block i takes segment
// ids[i].
template <typename U, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_local_sort_seg(const U *__restrict__ keys, const uint32_t *__restrict__ pos,
                                                            const int64_t *__restrict__ seg, uint64_t nseg,
                                                            uint64_t total, const BatchItem *items, uint32_t nitems,
                                                            const LocalDigits dg, uint32_t *__restrict__ overflow,
                                                            const uint32_t *__restrict__ ids = nullptr) {
  const uint64_t j = ids ? ids[blockIdx.x] : blockIdx.x;
  const uint64_t s = static_cast<uint64_t>(seg[j]);
  const uint64_t e = j + 1 < nseg ? static_cast<uint64_t>(seg[j + 1]) : total;
  uint32_t lo = 0, hi = nitems;  // the last item whose first row is at most s
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (items[mid].aux <= s) lo = mid;
    else hi = mid;
  }
  int64_t *out = static_cast<int64_t *>(items[lo].out) + (s - items[lo].aux);
  if (!local_sort_rows<U, THREADS, ITEMS>(keys, pos, s, static_cast<uint32_t>(e - s), dg, out) && threadIdx.x == 0)
    overflow[1 + atomicAdd(&overflow[0], 1u)] = static_cast<uint32_t>(j);
}

template <typename U, typename V, bool FIRST, bool LAST>
void run_pass(Ctx &c, int32_t type, const void *kin, const V *pin, const uint64_t *vals, uint64_t n,
              int shift, const uint32_t *dbase, U *kout, V *pout, int64_t *out_pos) {
  const uint64_t tiles = (n + kTile - 1) / kTile;
  PassArgs<U, V> a{};
  a.type = type;
  a.keys_in = kin;
  a.pay_in = pin;
  a.vals = vals;
  a.n = n;
  a.shift = shift;
  a.digit_base = dbase;
  a.keys_out = kout;
  a.pay_out = pout;
  a.out_pos = out_pos;
  a.status = status_words(c, ST_SORT, tiles * kRadix);
  a.epoch = next_epoch(c);
  a.ticket = c.scratch.counters;
  a.ntiles = uint32_t(tiles);
  k_pass<U, V, FIRST, LAST, false><<<uint32_t(tiles), kSortBlock, 0, c.stream>>>(a);
  check_hip(hipGetLastError(), "radix pass");
}

// ------------------------------------------------------------------ batched form
// Every item sorts in its own rows: one launch per pass for all items, the tiles of an item
// looking back only to the item's first tile, per-item digit histograms and bases. Digits
// that are constant in every item are skipped; a digit constant in some items only is a
// stable identity pass for those.

template <typename U>
__global__ __launch_bounds__(kBlock) void k_hist_batch(int32_t type, const BatchItem *items,
                                                       const uint32_t *tile_item, uint32_t *hist) {
  constexpr int P = sizeof(U);
  __shared__ uint32_t s_h[P][kRadix];
  for (int k = threadIdx.x; k < P * kRadix; k += kBlock) (&s_h[0][0])[k] = 0;
  __syncthreads();
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const uint64_t t0 = (blockIdx.x - I.tile0) * uint64_t(kTile);
  const uint64_t t1 = t0 + kTile < I.n ? t0 + kTile : I.n;
  for (uint64_t r0 = t0 + threadIdx.x; r0 < t1; r0 += uint64_t(kBlock) * kHistAhead) {
    U k[kHistAhead];
    with_image<U>(type, I.a, [&](auto img) {
#pragma unroll
      for (int u = 0; u < kHistAhead; ++u) {
        const uint64_t r = r0 + uint64_t(u) * kBlock;
        k[u] = r < t1 ? img(r) : U(0);
      }
    });
#pragma unroll
    for (int u = 0; u < kHistAhead; ++u)
      if (r0 + uint64_t(u) * kBlock < t1) hist_add<U>(s_h, k[u]);
  }
  __syncthreads();
  uint32_t *h = hist + uint64_t(it) * P * kRadix;
  for (int k = threadIdx.x; k < P * kRadix; k += kBlock) {
    const uint32_t v = (&s_h[0][0])[k];
    if (v) atomicAdd(&h[k], v);
  }
}

// One thread per (item, digit): counts to exclusive bases in place; active[d] counts the
// items in which digit d takes more than one value.
__global__ void k_base_batch(const BatchItem *items, uint64_t nitems, int P, uint32_t *hist,
                             uint32_t *active) {
  const uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (k >= nitems * P) return;
  const uint64_t n = items[k / P].n;
  uint32_t *h = hist + k * kRadix;
  uint32_t s = 0;
  bool constant = false;
  for (int v = 0; v < kRadix; ++v) {
    const uint32_t cnt = h[v];
    h[v] = s;
    s += cnt;
    if (cnt == n) constant = true;
  }
  if (n > 0 && !constant) atomicAdd(&active[k % P], 1u);
}

__global__ void k_iota_batch(const BatchItem *items, const uint32_t *tile_item) {
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const uint64_t t0 = (blockIdx.x - I.tile0) * uint64_t(kTile);
  const uint64_t t1 = t0 + kTile < I.n ? t0 + kTile : I.n;
  int64_t *out = static_cast<int64_t *>(I.out);
  for (uint64_t r = t0 + threadIdx.x; r < t1; r += kBlock) out[r] = static_cast<int64_t>(r);
}

// Every item's rows are positions in order already: out[r] = pos[aux + r] as i64.
__global__ void k_pos_batch(const BatchItem *items, const uint32_t *tile_item, const uint32_t *pos) {
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const uint64_t t0 = (blockIdx.x - I.tile0) * uint64_t(kTile);
  const uint64_t t1 = t0 + kTile < I.n ? t0 + kTile : I.n;
  int64_t *out = static_cast<int64_t *>(I.out);
  for (uint64_t r = t0 + threadIdx.x; r < t1; r += kBlock) out[r] = static_cast<int64_t>(pos[I.aux + r]);
}

template <typename U, typename V, bool FIRST, bool LAST>
void run_pass_batch(Ctx &c, int32_t type, const BatchPlan &p, int P, int d, const void *kin,
                    const V *pin, const uint32_t *hist, U *kout, V *pout) {
  PassArgs<U, V> a{};
  a.type = type;
  a.keys_in = kin;
  a.pay_in = pin;
  a.shift = 8 * d;
  a.digit_base = hist;
  a.keys_out = kout;
  a.pay_out = pout;
  a.status = status_words(c, ST_SORT, uint64_t(p.ntiles) * kRadix);
  a.epoch = next_epoch(c);
  a.ticket = c.scratch.counters;
  a.ntiles = p.ntiles;
  a.items = p.items;
  a.tile_item = p.tile_item;
  a.P = P;
  a.d = d;
  k_pass<U, V, FIRST, LAST, true><<<p.ntiles, kSortBlock, 0, c.stream>>>(a);
  check_hip(hipGetLastError(), "radix batch pass");
}

// Batched sort: positions mode (V = u32) writes each item's i64 positions to item.out;
// values mode (V = u64) sorts (key, value bits of item.b) pairs into sorted_keys / sorted_vals
// at each item's first row (item.aux, set here) of the concatenated rows.
template <typename U, typename V>
void radix_batch(Ctx &c, int32_t type, std::vector<BatchItem> &items, U *sorted_keys = nullptr,
                 V *sorted_vals = nullptr) {
  constexpr bool POS = std::is_same<V, uint32_t>::value;
  constexpr int P = sizeof(U);
  uint64_t total = 0;
  for (auto &it : items) {
    if (it.n > 0xFFFFFFFFull) usage_error("argsort supports at most 2^32 - 1 rows");
    it.aux = total;
    total += it.n;
  }
  if (total == 0) return;
  const uint64_t nitems = items.size();
  BatchPlan p = batch_plan(c, items, kTile);
  const uint64_t hwords = nitems * P * kRadix;
  uint32_t *hist = static_cast<uint32_t *>(tmp_alloc(c, hwords * 4 + 64));
  uint32_t *active = hist + hwords;
  fill_u64(c, reinterpret_cast<uint64_t *>(hist), (hwords * 4 + 64) / 8, 0);
  k_hist_batch<U><<<p.ntiles, kBlock, 0, c.stream>>>(type, p.items, p.tile_item, hist);
  check_hip(hipGetLastError(), "radix batch histogram");
  k_base_batch<<<stream_grid(nitems * P, kBlock, 1u << 20), kBlock, 0, c.stream>>>(p.items, nitems, P, hist,
                                                                                 active);
  check_hip(hipGetLastError(), "radix batch bases");
  uint32_t act[8] = {0};
  check_hip(hipMemcpyAsync(act, active, P * 4, hipMemcpyDeviceToHost, c.stream), "active digits");
  check_hip(hipStreamSynchronize(c.stream), "sync");
  std::vector<int> digits;
  for (int d = 0; d < P; ++d)
    if (act[d]) digits.push_back(d);
  // with no digit to sort by, one stable pass on digit 0 still moves the rows as they are
  if (digits.empty() && !POS) digits.push_back(0);
  if (digits.empty()) {
    k_iota_batch<<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item);
    check_hip(hipGetLastError(), "radix batch iota");
  } else {
    const int np = static_cast<int>(digits.size());
    U *ka = nullptr, *kb = nullptr;
    V *pa = nullptr, *pb = nullptr;
    if (np > 1) {
      ka = static_cast<U *>(tmp_alloc(c, total * sizeof(U)));
      pa = static_cast<V *>(tmp_alloc(c, total * sizeof(V)));
      if (np > 2) {
        kb = static_cast<U *>(tmp_alloc(c, total * sizeof(U)));
        pb = static_cast<V *>(tmp_alloc(c, total * sizeof(V)));
      }
    }
    const U *kin = nullptr;
    const V *pin = nullptr;
    for (int k = 0; k < np; ++k) {
      const int d = digits[k];
      const bool first = k == 0, last = k == np - 1;
      U *ko = last ? sorted_keys : (k % 2 == 0) ? ka : kb;
      V *po = last ? sorted_vals : (k % 2 == 0) ? pa : pb;
      if (first && last) run_pass_batch<U, V, true, true>(c, type, p, P, d, nullptr, nullptr, hist, ko, po);
      else if (first) run_pass_batch<U, V, true, false>(c, type, p, P, d, nullptr, nullptr, hist, ko, po);
      else if (last) run_pass_batch<U, V, false, true>(c, type, p, P, d, kin, pin, hist, ko, po);
      else run_pass_batch<U, V, false, false>(c, type, p, P, d, kin, pin, hist, ko, po);
      kin = ko;
      pin = po;
    }
    if (ka) tmp_free(c, ka);
    if (pa) tmp_free(c, pa);
    if (kb) tmp_free(c, kb);
    if (pb) tmp_free(c, pb);
  }
  tmp_free(c, hist);
  batch_plan_free(c, p);
}

// ------------------------------------------------ batched sorts by top-digit buckets

constexpr uint32_t kPairWords = kBuckets / 32;  // one bit per (top digit, second digit) pair
constexpr uint32_t kDenseTile = 4096;           // a tile with more pairs marks its item dense

// Per item, the distinct pairs of the digits at sh_hi and sh_lo: a tile sets the bits of its
// rows' pairs in an LDS bitmap, then, unless it holds more than kDenseTile pairs (which marks
// the item dense), ORs its nonzero words into the item's bitmap (kPairWords words).
template <typename U>
__global__ __launch_bounds__(kBlock) void k_pairs_batch(int32_t type, const BatchItem *items, const uint32_t *tile_item,
                                                        int sh_hi, int sh_lo, uint32_t *bitmaps, uint32_t *dense) {
  __shared__ uint32_t s_bm[kPairWords];
  __shared__ uint32_t s_cnt;
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const uint64_t t0 = (blockIdx.x - I.tile0) * uint64_t(kTile);
  const uint64_t t1 = t0 + kTile < I.n ? t0 + kTile : I.n;
  for (uint32_t w = threadIdx.x; w < kPairWords; w += kBlock) s_bm[w] = 0;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  with_image<U>(type, I.a, [&](auto img) {
    for (uint64_t r = t0 + threadIdx.x; r < t1; r += kBlock) {
      const uint32_t b = bucket16<U>(img(r), sh_hi, sh_lo);
      atomicOr(&s_bm[b >> 5], 1u << (b & 31));
    }
  });
  __syncthreads();
  uint32_t cnt = 0;
  for (uint32_t w = threadIdx.x; w < kPairWords; w += kBlock) cnt += __popc(s_bm[w]);
  atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (s_cnt > kDenseTile) {
    if (threadIdx.x == 0) atomicOr(&dense[it], 1u);
    return;
  }
  uint32_t *bm = bitmaps + uint64_t(it) * kPairWords;
  for (uint32_t w = threadIdx.x; w < kPairWords; w += kBlock)
    if (s_bm[w]) atomicOr(&bm[w], s_bm[w]);
}

// count[item] = the item's distinct pairs, or ~0 when it is dense.
__global__ __launch_bounds__(kBlock) void k_pairs_count(const uint32_t *bitmaps, const uint32_t *dense, uint32_t *count) {
  __shared__ uint32_t s_cnt;
  const uint32_t it = blockIdx.x;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  uint32_t cnt = 0;
  for (uint32_t w = threadIdx.x; w < kPairWords; w += kBlock) cnt += __popc(bitmaps[uint64_t(it) * kPairWords + w]);
  atomicAdd(&s_cnt, cnt);
  __syncthreads();
  if (threadIdx.x == 0) count[it] = dense[it] ? ~0u : s_cnt;
}

// dst row i = src row idx[i] (rows of `width` words): the digit bases of a subset of items.
__global__ void k_gather_rows(const uint32_t *src, const uint32_t *idx, uint64_t rows, uint32_t width, uint32_t *dst) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < rows * width;
       i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] = src[uint64_t(idx[i / width]) * width + i % width];
}

// flags[aux + r] = 1 where row r of an item starts a segment: the item's first row, or a row
// whose pair of the digits at sh_hi and sh_lo differs from the row before (keys sorted by them).
template <typename U>
__global__ __launch_bounds__(kBlock) void k_seg_flags_batch(const BatchItem *items, const uint32_t *tile_item,
                                                            const U *keys, int sh_hi, int sh_lo, uint8_t *flags) {
  const BatchItem &I = items[tile_item[blockIdx.x]];
  const uint64_t t0 = (blockIdx.x - I.tile0) * uint64_t(kTile);
  const uint64_t t1 = t0 + kTile < I.n ? t0 + kTile : I.n;
  const U *k = keys + I.aux;
  for (uint64_t r = t0 + threadIdx.x; r < t1; r += kBlock)
    flags[I.aux + r] = r == 0 || bucket16<U>(k[r], sh_hi, sh_lo) != bucket16<U>(k[r - 1], sh_hi, sh_lo);
}

// flags[r] = 1 where row r of keys sorted by the bits of `mask` starts a segment: the first row,
// or a row whose masked bits differ from the row before.
template <typename U>
__global__ __launch_bounds__(kBlock) void k_seg_flags(const U *__restrict__ keys, uint64_t n, U mask,
                                                      uint8_t *__restrict__ flags) {
  for (uint64_t r = blockIdx.x * uint64_t(kBlock) + threadIdx.x; r < n; r += uint64_t(gridDim.x) * kBlock)
    flags[r] = r == 0 || ((keys[r] ^ keys[r - 1]) & mask) != 0;
}

// The local sort's capacity in rows for a bucket-size estimate: 3072 (512 threads x 6 rows,
// three blocks per compute unit) or 8192 (512 threads x 16 rows, one block), 0 when buckets
// would not fit.
inline int local_capacity(double est_max) {
  if (est_max <= 0.7 * 3072) return 3072;
  if (est_max <= 0.6 * 8192) return 8192;
  return 0;
}

// Row ranges of sorted (image, position) buffers kb / pb left by the local sorts (each range in
// row order, by the digits above `lower` equal): sorted by the batched passes over the digits of
// `lower` (ascending) that vary in some range, an identity pass where one does not, the last
// writing positions to each range's output (items: a = kb + first row, out, n, aux = first row);
// with none varying, the positions are copied as they stand. ka / pa are free buffers of the
// same rows.
template <typename U>
void finish_ranges(Ctx &c, std::vector<BatchItem> &items, const std::vector<int> &lower, U *ka, uint32_t *pa, U *kb,
                   uint32_t *pb) {
  using V = uint32_t;
  constexpr int P = sizeof(U);
  const int32_t img = sizeof(U) == 8 ? int32_t(VK_U64) : kRawU32;  // the rows are images already
  BatchPlan bp = batch_plan(c, items, kTile);
  const uint64_t hwords = items.size() * P * kRadix;
  uint32_t *bh = static_cast<uint32_t *>(tmp_alloc(c, hwords * 4 + 64));
  uint32_t *active = bh + hwords;
  fill_u64(c, reinterpret_cast<uint64_t *>(bh), (hwords * 4 + 64) / 8, 0);
  k_hist_batch<U><<<bp.ntiles, kBlock, 0, c.stream>>>(img, bp.items, bp.tile_item, bh);
  check_hip(hipGetLastError(), "range histogram");
  k_base_batch<<<stream_grid(items.size() * P, kBlock, 1u << 20), kBlock, 0, c.stream>>>(bp.items, items.size(), P,
                                                                                       bh, active);
  check_hip(hipGetLastError(), "range bases");
  uint32_t act[8] = {0};
  read_back(c, active, P * 4, act);
  std::vector<int> vary;
  for (int d : lower)
    if (act[d]) vary.push_back(d);
  const U *kin = kb;
  const V *pin = pb;
  if (vary.empty()) {
    k_pos_batch<<<bp.ntiles, kBlock, 0, c.stream>>>(bp.items, bp.tile_item, pb);
    check_hip(hipGetLastError(), "range positions");
  }
  for (size_t k = 0; k < vary.size(); ++k) {
    const int d = vary[k];
    if (k + 1 == vary.size()) {
      run_pass_batch<U, V, false, true>(c, img, bp, P, d, kin, pin, bh, nullptr, nullptr);
    } else {
      U *ko = (k % 2 == 0) ? ka : kb;
      V *po = (k % 2 == 0) ? pa : pb;
      run_pass_batch<U, V, false, false>(c, img, bp, P, d, kin, pin, bh, ko, po);
      kin = ko;
      pin = po;
    }
  }
  tmp_free(c, bh);
  batch_plan_free(c, bp);
}

// Argsort by top-digit buckets (see local_sort_rows). `digits` are the active digits, ascending,
// at least four; hist holds their exclusive digit bases. The two top digits go through HBM
// passes, the others through the local sorts; buckets the local sorts cannot take are finished
// by the batched passes over their row ranges.
template <typename U>
void argsort_buckets(Ctx &c, int32_t type, const void *keys, uint64_t n, const std::vector<int> &digits,
                     const uint32_t *hist, int cap, int64_t *out_pos) {
  using V = uint32_t;
  const int np = static_cast<int>(digits.size());
  const int dlo = digits[np - 2], dhi = digits[np - 1];
  U *ka = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
  V *pa = static_cast<V *>(tmp_alloc(c, n * sizeof(V)));
  U *kb = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
  V *pb = static_cast<V *>(tmp_alloc(c, n * sizeof(V)));
  run_pass<U, V, true, false>(c, type, keys, nullptr, nullptr, n, 8 * dlo, hist + dlo * kRadix, ka, pa, nullptr);
  run_pass<U, V, false, false>(c, type, ka, pa, nullptr, n, 8 * dhi, hist + dhi * kRadix, kb, pb, nullptr);
  // bucket starts (kBuckets + 1 words) and the overflow list (count + kBuckets ids)
  uint32_t *start = static_cast<uint32_t *>(tmp_alloc(c, (2 * kBuckets + 4) * 4));
  uint32_t *overflow = start + kBuckets + 2;
  fill_u64(c, reinterpret_cast<uint64_t *>(start), (kBuckets + 2) / 2, n | (n << 32));
  fill_u64(c, reinterpret_cast<uint64_t *>(overflow), 1, 0);
  k_bucket_starts<U><<<stream_grid(n, kBlock, 8192), kBlock, 0, c.stream>>>(kb, n, 8 * dhi, 8 * dlo, start);
  check_hip(hipGetLastError(), "bucket starts");
  LocalDigits dg{};
  dg.n = np - 2;
  for (int k = 0; k < np - 2; ++k) dg.shift[k] = 8 * digits[k];
  if (cap == 3072) k_local_sort<U, 512, 6><<<kBuckets, 512, 0, c.stream>>>(kb, pb, start, dg, out_pos, overflow);
  else k_local_sort<U, 512, 16><<<kBuckets, 512, 0, c.stream>>>(kb, pb, start, dg, out_pos, overflow);
  check_hip(hipGetLastError(), "local sort");
  uint32_t nover = 0;
  read_back(c, overflow, 4, &nover);
  if (nover) {
    std::vector<uint32_t> st(kBuckets + 1), ids(nover);
    check_hip(hipMemcpyAsync(st.data(), start, (kBuckets + 1) * 4, hipMemcpyDeviceToHost, c.stream), "starts");
    check_hip(hipMemcpyAsync(ids.data(), overflow + 1, nover * 4, hipMemcpyDeviceToHost, c.stream), "overflow");
    check_hip(hipStreamSynchronize(c.stream), "sync");
    std::vector<BatchItem> ranges;
    for (uint32_t id : ids) {
      BatchItem it{};
      it.a = kb + st[id];
      it.out = out_pos + st[id];
      it.n = st[id + 1] - st[id];
      it.aux = st[id];
      ranges.push_back(it);
    }
    finish_ranges<U>(c, ranges, std::vector<int>(digits.begin(), digits.end() - 2), ka, pa, kb, pb);
  }
  tmp_free(c, start);
  tmp_free(c, ka);
  tmp_free(c, pa);
  tmp_free(c, kb);
  tmp_free(c, pb);
}

// Argsort by segments of the three top digits: for keys whose two top digits are too skewed for
// 16-bit buckets (normal floats: sign and exponent bits), three HBM passes by the three top
// digits, then the segments where those digits change (flags and the ordered selection), one
// local sort per segment over the other digits, and the batched passes for segments the
// local sorts cannot take. `digits` are the active digits, ascending, at least five; hist
// holds their exclusive digit bases.
template <typename U>
void argsort_segments(Ctx &c, int32_t type, const void *keys, uint64_t n, const std::vector<int> &digits,
                      const uint32_t *hist, int cap, int64_t *out_pos) {
  using V = uint32_t;
  const int np = static_cast<int>(digits.size());
  const int d0 = digits[np - 3], d1 = digits[np - 2], d2 = digits[np - 1];
  U *ka = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
  V *pa = static_cast<V *>(tmp_alloc(c, n * sizeof(V)));
  U *kb = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
  V *pb = static_cast<V *>(tmp_alloc(c, n * sizeof(V)));
  run_pass<U, V, true, false>(c, type, keys, nullptr, nullptr, n, 8 * d0, hist + d0 * kRadix, ka, pa, nullptr);
  run_pass<U, V, false, false>(c, type, ka, pa, nullptr, n, 8 * d1, hist + d1 * kRadix, kb, pb, nullptr);
  run_pass<U, V, false, false>(c, type, kb, pb, nullptr, n, 8 * d2, hist + d2 * kRadix, ka, pa, nullptr);
  const U mask = (U(255) << (8 * d0)) | (U(255) << (8 * d1)) | (U(255) << (8 * d2));
  uint8_t *flags = static_cast<uint8_t *>(tmp_alloc(c, n + 8));
  int64_t *seg = static_cast<int64_t *>(tmp_alloc(c, n * 8 + 8));
  k_seg_flags<U><<<stream_grid(n, kBlock, 8192), kBlock, 0, c.stream>>>(ka, n, mask, flags);
  check_hip(hipGetLastError(), "segment flags");
  SelectSpec ss{2, VK_I64, 0, 0, 0};
  uint64_t *d_cnt = c.scratch.slots + 21;
  select_rows(c, ss, nullptr, flags, n, seg, d_cnt);
  uint64_t nseg = 0;
  read_back(c, d_cnt, 8, &nseg);
  if (nseg > 0x7FFFFFFFull) usage_error("argsort has too many segments");
  // one item: the whole output from row 0
  BatchItem whole{};
  whole.out = out_pos;
  whole.n = n;
  auto *d_items = static_cast<BatchItem *>(tmp_alloc(c, sizeof(BatchItem)));
  check_hip(hipMemcpyAsync(d_items, &whole, sizeof(BatchItem), hipMemcpyHostToDevice, c.stream), "segment item");
  // every segment through the 3072-row local sort (three blocks per compute unit); with the
  // larger capacity, the segments it lists go through the 5120-row one (two blocks per unit),
  // and those that one lists through the 8192-row one
  uint32_t *over = static_cast<uint32_t *>(tmp_alloc(c, 3 * (nseg * 4 + 8)));
  uint32_t *overflow = over;
  fill_u64(c, reinterpret_cast<uint64_t *>(overflow), 1, 0);
  LocalDigits dg{};
  dg.n = np - 3;
  for (int k = 0; k < np - 3; ++k) dg.shift[k] = 8 * digits[k];
  k_local_sort_seg<U, 512, 6><<<uint32_t(nseg), 512, 0, c.stream>>>(ka, pa, seg, nseg, n, d_items, 1, dg, overflow);
  check_hip(hipGetLastError(), "segment sort");
  uint32_t nover = 0;
  read_back(c, overflow, 4, &nover);
  if (nover && cap == 8192) {
    uint32_t *mid = over + nseg + 2;
    fill_u64(c, reinterpret_cast<uint64_t *>(mid), 1, 0);
    k_local_sort_seg<U, 512, 10><<<nover, 512, 0, c.stream>>>(ka, pa, seg, nseg, n, d_items, 1, dg, mid,
                                                             overflow + 1);
    check_hip(hipGetLastError(), "segment sort");
    overflow = mid;
    read_back(c, overflow, 4, &nover);
  }
  if (nover && cap == 8192) {
    uint32_t *wide = over + 2 * (nseg + 2);
    fill_u64(c, reinterpret_cast<uint64_t *>(wide), 1, 0);
    k_local_sort_seg<U, 512, 16><<<nover, 512, 0, c.stream>>>(ka, pa, seg, nseg, n, d_items, 1, dg, wide,
                                                             overflow + 1);
    check_hip(hipGetLastError(), "wide segment sort");
    overflow = wide;
    read_back(c, overflow, 4, &nover);
  }
  if (nover) {
    std::vector<uint32_t> ids(nover);
    std::vector<int64_t> starts(nseg);
    check_hip(hipMemcpyAsync(ids.data(), overflow + 1, nover * 4, hipMemcpyDeviceToHost, c.stream), "overflow");
    check_hip(hipMemcpyAsync(starts.data(), seg, nseg * 8, hipMemcpyDeviceToHost, c.stream), "segments");
    check_hip(hipStreamSynchronize(c.stream), "sync");
    std::vector<BatchItem> ranges;
    for (uint32_t j : ids) {
      const uint64_t s0 = static_cast<uint64_t>(starts[j]);
      const uint64_t e = j + 1 < nseg ? static_cast<uint64_t>(starts[j + 1]) : n;
      BatchItem it{};
      it.a = ka + s0;
      it.out = out_pos + s0;
      it.n = e - s0;
      it.aux = s0;
      ranges.push_back(it);
    }
    // the sorted rows are in ka / pa; kb / pb are free
    finish_ranges<U>(c, ranges, std::vector<int>(digits.begin(), digits.end() - 3), kb, pb, ka, pa);
  }
  tmp_free(c, over);
  tmp_free(c, d_items);
  tmp_free(c, flags);
  tmp_free(c, seg);
  tmp_free(c, ka);
  tmp_free(c, pa);
  tmp_free(c, kb);
  tmp_free(c, pb);
}

// Batched argsort (positions): one histogram of every item gives the digit bases and the active
// digits. With four or more active digits, the distinct pairs of the two top ones are counted
// per item (k_pairs_batch); items of at least kBucketItemRows rows with at least 64 rows per
// pair sort by buckets: two batched passes by the two top digits, the segments of one pair found
// by flags and a selection, one block per segment (local_sort_rows), segments it cannot take
// finished by finish_ranges. The other items take one batched pass per active digit.
constexpr uint64_t kBucketItemRows = 4096;

template <typename U>
void argsort_batch_pos(Ctx &c, int32_t type, std::vector<BatchItem> &items) {
  using V = uint32_t;
  constexpr int P = sizeof(U);
  constexpr uint32_t kRow = P * kRadix;  // histogram words per item
  uint64_t total = 0;
  for (auto &it : items) {
    if (it.n > 0xFFFFFFFFull) usage_error("argsort supports at most 2^32 - 1 rows");
    it.aux = total;
    total += it.n;
  }
  if (total == 0) return;
  const uint64_t nitems = items.size();
  BatchPlan p = batch_plan(c, items, kTile);
  const uint64_t hwords = nitems * kRow;
  uint32_t *hist = static_cast<uint32_t *>(tmp_alloc(c, hwords * 4 + 64));
  uint32_t *active = hist + hwords;
  fill_u64(c, reinterpret_cast<uint64_t *>(hist), (hwords * 4 + 64) / 8, 0);
  k_hist_batch<U><<<p.ntiles, kBlock, 0, c.stream>>>(type, p.items, p.tile_item, hist);
  check_hip(hipGetLastError(), "radix batch histogram");
  k_base_batch<<<stream_grid(nitems * P, kBlock, 1u << 20), kBlock, 0, c.stream>>>(p.items, nitems, P, hist, active);
  check_hip(hipGetLastError(), "radix batch bases");
  uint32_t act[8] = {0};
  read_back(c, active, P * 4, act);
  std::vector<int> digits;
  for (int d = 0; d < P; ++d)
    if (act[d]) digits.push_back(d);
  const int np = static_cast<int>(digits.size());
  if (np == 0) {
    k_iota_batch<<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item);
    check_hip(hipGetLastError(), "radix batch iota");
    tmp_free(c, hist);
    batch_plan_free(c, p);
    return;
  }
  // which items sort by buckets
  std::vector<uint32_t> pairs(nitems, ~0u);
  if (np >= 4) {
    const int dlo = digits[np - 2], dhi = digits[np - 1];
    uint32_t *bm = static_cast<uint32_t *>(tmp_alloc(c, (nitems * kPairWords + 2 * nitems) * 4 + 8));
    uint32_t *dense = bm + nitems * kPairWords;
    uint32_t *cnt = dense + nitems;
    fill_u64(c, reinterpret_cast<uint64_t *>(bm), ((nitems * kPairWords + 2 * nitems) * 4 + 8) / 8, 0);
    k_pairs_batch<U><<<p.ntiles, kBlock, 0, c.stream>>>(type, p.items, p.tile_item, 8 * dhi, 8 * dlo, bm, dense);
    k_pairs_count<<<uint32_t(nitems), kBlock, 0, c.stream>>>(bm, dense, cnt);
    check_hip(hipGetLastError(), "pair counts");
    check_hip(hipMemcpyAsync(pairs.data(), cnt, nitems * 4, hipMemcpyDeviceToHost, c.stream), "pair counts");
    check_hip(hipStreamSynchronize(c.stream), "sync");
    tmp_free(c, bm);
  }
  std::vector<uint32_t> bidx, lidx;
  for (uint32_t i = 0; i < nitems; ++i) {
    const BatchItem &it = items[i];
    const bool buckets = it.n >= kBucketItemRows && pairs[i] != ~0u && it.n >= 64 * uint64_t(pairs[i]);
    (buckets ? bidx : lidx).push_back(i);
  }
  // a subset of the items: its own rows (aux), plan and gathered digit bases
  struct Sub {
    std::vector<BatchItem> items;
    uint64_t rows = 0;
    BatchPlan plan;
    uint32_t *hist = nullptr;
  };
  auto subset = [&](const std::vector<uint32_t> &idx) {
    Sub s;
    for (uint32_t i : idx) {
      BatchItem it = items[i];
      it.aux = s.rows;
      s.rows += it.n;
      s.items.push_back(it);
    }
    s.plan = batch_plan(c, s.items, kTile);
    s.hist = static_cast<uint32_t *>(tmp_alloc(c, idx.size() * kRow * 4 + 8));
    uint32_t *d_idx = static_cast<uint32_t *>(tmp_alloc(c, idx.size() * 4 + 8));
    check_hip(hipMemcpyAsync(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, c.stream), "subset");
    k_gather_rows<<<stream_grid(idx.size() * kRow, kBlock, 8192), kBlock, 0, c.stream>>>(hist, d_idx, idx.size(), kRow,
                                                                                      s.hist);
    check_hip(hipGetLastError(), "subset bases");
    check_hip(hipStreamSynchronize(c.stream), "subset");
    tmp_free(c, d_idx);
    return s;
  };
  // one batched pass per active digit, the last writing positions
  auto lsd = [&](const BatchPlan &bp, const uint32_t *bh, uint64_t rows) {
    U *ka = nullptr, *kb = nullptr;
    V *pa = nullptr, *pb = nullptr;
    if (np > 1) {
      ka = static_cast<U *>(tmp_alloc(c, rows * sizeof(U)));
      pa = static_cast<V *>(tmp_alloc(c, rows * sizeof(V)));
      if (np > 2) {
        kb = static_cast<U *>(tmp_alloc(c, rows * sizeof(U)));
        pb = static_cast<V *>(tmp_alloc(c, rows * sizeof(V)));
      }
    }
    const U *kin = nullptr;
    const V *pin = nullptr;
    for (int k = 0; k < np; ++k) {
      const int d = digits[k];
      const bool first = k == 0, last = k == np - 1;
      U *ko = last ? nullptr : (k % 2 == 0) ? ka : kb;
      V *po = last ? nullptr : (k % 2 == 0) ? pa : pb;
      if (first && last) run_pass_batch<U, V, true, true>(c, type, bp, P, d, nullptr, nullptr, bh, ko, po);
      else if (first) run_pass_batch<U, V, true, false>(c, type, bp, P, d, nullptr, nullptr, bh, ko, po);
      else if (last) run_pass_batch<U, V, false, true>(c, type, bp, P, d, kin, pin, bh, ko, po);
      else run_pass_batch<U, V, false, false>(c, type, bp, P, d, kin, pin, bh, ko, po);
      kin = ko;
      pin = po;
    }
    if (ka) tmp_free(c, ka);
    if (pa) tmp_free(c, pa);
    if (kb) tmp_free(c, kb);
    if (pb) tmp_free(c, pb);
  };
  if (bidx.empty()) {
    lsd(p, hist, total);
  } else {
    if (!lidx.empty()) {
      Sub L = subset(lidx);
      lsd(L.plan, L.hist, L.rows);
      tmp_free(c, L.hist);
      batch_plan_free(c, L.plan);
    }
    Sub B = subset(bidx);
    const int dlo = digits[np - 2], dhi = digits[np - 1];
    const uint64_t rows = B.rows;
    U *ka = static_cast<U *>(tmp_alloc(c, rows * sizeof(U)));
    V *pa = static_cast<V *>(tmp_alloc(c, rows * sizeof(V)));
    U *kb = static_cast<U *>(tmp_alloc(c, rows * sizeof(U)));
    V *pb = static_cast<V *>(tmp_alloc(c, rows * sizeof(V)));
    run_pass_batch<U, V, true, false>(c, type, B.plan, P, dlo, nullptr, nullptr, B.hist, ka, pa);
    run_pass_batch<U, V, false, false>(c, type, B.plan, P, dhi, ka, pa, B.hist, kb, pb);
    // segments: an item's first row, and rows whose pair of the two top digits changes
    uint8_t *flags = static_cast<uint8_t *>(tmp_alloc(c, rows + 8));
    int64_t *seg = static_cast<int64_t *>(tmp_alloc(c, rows * 8 + 8));
    k_seg_flags_batch<U><<<B.plan.ntiles, kBlock, 0, c.stream>>>(B.plan.items, B.plan.tile_item, kb, 8 * dhi, 8 * dlo,
                                                                  flags);
    check_hip(hipGetLastError(), "segment flags");
    SelectSpec ss{2, VK_I64, 0, 0, 0};
    uint64_t *d_cnt = c.scratch.slots + 21;
    select_rows(c, ss, nullptr, flags, rows, seg, d_cnt);
    uint64_t nseg = 0;
    read_back(c, d_cnt, 8, &nseg);
    uint32_t *overflow = static_cast<uint32_t *>(tmp_alloc(c, nseg * 4 + 8));
    fill_u64(c, reinterpret_cast<uint64_t *>(overflow), 1, 0);
    LocalDigits dg{};
    dg.n = np - 2;
    for (int k = 0; k < np - 2; ++k) dg.shift[k] = 8 * digits[k];
    if (nseg > 0x7FFFFFFFull) usage_error("batched argsort has too many segments");
    k_local_sort_seg<U, 512, 6><<<uint32_t(nseg), 512, 0, c.stream>>>(kb, pb, seg, nseg, rows, B.plan.items,
                                                                     uint32_t(B.items.size()), dg, overflow);
    check_hip(hipGetLastError(), "segment sort");
    uint32_t nover = 0;
    read_back(c, overflow, 4, &nover);
    if (nover) {
      std::vector<uint32_t> ids(nover);
      std::vector<int64_t> starts(nseg);
      check_hip(hipMemcpyAsync(ids.data(), overflow + 1, nover * 4, hipMemcpyDeviceToHost, c.stream), "overflow");
      check_hip(hipMemcpyAsync(starts.data(), seg, nseg * 8, hipMemcpyDeviceToHost, c.stream), "segments");
      check_hip(hipStreamSynchronize(c.stream), "sync");
      std::vector<BatchItem> ranges;
      for (uint32_t j : ids) {
        const uint64_t s = static_cast<uint64_t>(starts[j]);
        const uint64_t e = j + 1 < nseg ? static_cast<uint64_t>(starts[j + 1]) : rows;
        // the item holding the segment: the last whose first row is at most s
        size_t lo = 0, hi = B.items.size();
        while (hi - lo > 1) {
          const size_t mid = (lo + hi) / 2;
          if (B.items[mid].aux <= s) lo = mid;
          else hi = mid;
        }
        BatchItem it{};
        it.a = kb + s;
        it.out = static_cast<int64_t *>(B.items[lo].out) + (s - B.items[lo].aux);
        it.n = e - s;
        it.aux = s;
        ranges.push_back(it);
      }
      finish_ranges<U>(c, ranges, std::vector<int>(digits.begin(), digits.end() - 2), ka, pa, kb, pb);
    }
    tmp_free(c, overflow);
    tmp_free(c, flags);
    tmp_free(c, seg);
    tmp_free(c, ka);
    tmp_free(c, pa);
    tmp_free(c, kb);
    tmp_free(c, pb);
    tmp_free(c, B.hist);
    batch_plan_free(c, B.plan);
  }
  tmp_free(c, hist);
  batch_plan_free(c, p);
}

// Sorts rows by image. Positions mode (V = u32): positions of the sorted rows to out_pos,
// sorted images to sorted_keys when not null. Values mode (V = u64): the sorted images and
// the value bits of `vals` that go with them to sorted_keys / sorted_vals.
// With class_shift, the sort leaves out the top non-constant digit T when there are two or
// more: the rows come out in stable order of the key bits below 8 T (their class), and
// *class_shift = 8 T; otherwise every digit is sorted and *class_shift is 8 above the top one.
template <typename U, typename V>
void radix(Ctx &c, int32_t type, const void *keys, const uint64_t *vals, uint64_t n, int64_t *out_pos,
           U *sorted_keys, V *sorted_vals, uint64_t limit = 0, uint64_t *bad_row = nullptr,
           int *class_shift = nullptr) {
  constexpr bool POS = std::is_same<V, uint32_t>::value;
  if (n == 0) return;
  if (n > 0xFFFFFFFFull) usage_error("argsort supports at most 2^32 - 1 rows");
  constexpr int P = sizeof(U);
  uint32_t *hist = static_cast<uint32_t *>(tmp_alloc(c, P * kRadix * 4));
  fill_u64(c, reinterpret_cast<uint64_t *>(hist), P * kRadix / 2, 0);
  // with bad_row (chain keys): the least row whose bucket key is outside 0..limit, or ~0; the
  // passes run only when there is none
  uint64_t *d_bad = c.scratch.slots + 24;
  const uint32_t hg = stream_grid(n, kBlock * 8, 4096);
  if (bad_row) {
    fill_u64(c, d_bad, 1, ~uint64_t(0));
    k_hist<U, true><<<hg, kBlock, 0, c.stream>>>(type, keys, n, hist, limit, d_bad);
  } else {
    k_hist<U><<<hg, kBlock, 0, c.stream>>>(type, keys, n, hist);
  }
  check_hip(hipGetLastError(), "radix histogram");
  std::vector<uint32_t> h(P * kRadix);
  uint64_t bad = ~uint64_t(0);
  check_hip(hipMemcpyAsync(h.data(), hist, P * kRadix * 4, hipMemcpyDeviceToHost, c.stream), "hist");
  if (bad_row) check_hip(hipMemcpyAsync(&bad, d_bad, 8, hipMemcpyDeviceToHost, c.stream), "bad row");
  check_hip(hipStreamSynchronize(c.stream), "sync");
  if (bad_row) {
    *bad_row = bad;
    if (bad != ~uint64_t(0)) {
      tmp_free(c, hist);
      return;
    }
  }
  // exclusive offsets per digit; passes whose digit is the same in every row are skipped
  std::vector<int> active;
  std::vector<uint32_t> base(P * kRadix);
  for (int d = 0; d < P; ++d) {
    uint64_t s = 0;
    bool constant = false;
    for (int v = 0; v < kRadix; ++v) {
      base[d * kRadix + v] = static_cast<uint32_t>(s);
      s += h[d * kRadix + v];
      if (h[d * kRadix + v] == n) constant = true;
    }
    if (!constant) active.push_back(d);
  }
  check_hip(hipMemcpyAsync(hist, base.data(), P * kRadix * 4, hipMemcpyHostToDevice, c.stream), "base");
  check_hip(hipStreamSynchronize(c.stream), "sync");
  const uint32_t g = stream_grid(n, kBlock * 4, 8192);
  if (active.empty() && !POS) active.push_back(0);  // one stable pass moves the pairs
  if (class_shift) {
    *class_shift = 8 * (active.back() + (active.size() >= 2 ? 0 : 1));
    if (active.size() >= 2) active.pop_back();
  }
  if (active.empty()) {
    k_iota64<<<g, kBlock, 0, c.stream>>>(n, out_pos);
    if (sorted_keys) k_copy_images<U><<<g, kBlock, 0, c.stream>>>(type, keys, n, sorted_keys);
    tmp_free(c, hist);
    return;
  }
  const int np = static_cast<int>(active.size());
  if constexpr (POS && sizeof(U) >= 4) {
    // positions only, at least four digits and buckets of the two top digits that fit a local
    // sort if the two digits were independent: bucket passes
    if (out_pos && !sorted_keys && np >= 4 && n >= (uint64_t(1) << 22)) {
      const int dlo = active[np - 2], dhi = active[np - 1];
      uint32_t mlo = 0, mhi = 0;
      for (int v = 0; v < kRadix; ++v) {
        mlo = h[dlo * kRadix + v] > mlo ? h[dlo * kRadix + v] : mlo;
        mhi = h[dhi * kRadix + v] > mhi ? h[dhi * kRadix + v] : mhi;
      }
      const int cap = local_capacity(double(mlo) * double(mhi) / double(n));
      if (cap) {
        argsort_buckets<U>(c, type, keys, n, active, hist, cap, out_pos);
        tmp_free(c, hist);
        return;
      }
      // two top digits too skewed: segments of three, when there are two or more digits left
      // for the local sorts and the three top digits would give segments that fit
      if (np >= 5) {
        const int d3 = active[np - 3];
        uint32_t m3 = 0;
        for (int v = 0; v < kRadix; ++v) m3 = h[d3 * kRadix + v] > m3 ? h[d3 * kRadix + v] : m3;
        const int cap3 = local_capacity(double(mlo) * double(mhi) * double(m3) / double(n) / double(n));
        if (cap3) {
          argsort_segments<U>(c, type, keys, n, active, hist, cap3, out_pos);
          tmp_free(c, hist);
          return;
        }
      }
    }
  }
  U *ka = nullptr, *kb = nullptr;
  V *pa = nullptr, *pb = nullptr;
  if (np > 1) {
    ka = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
    pa = static_cast<V *>(tmp_alloc(c, n * sizeof(V)));
    if (np > 2) {
      kb = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
      pb = static_cast<V *>(tmp_alloc(c, n * sizeof(V)));
    }
  }
  const U *kin = nullptr;
  const V *pin = nullptr;
  for (int k = 0; k < np; ++k) {
    const int d = active[k];
    const uint32_t *db = hist + d * kRadix;
    const bool first = k == 0, last = k == np - 1;
    U *ko = last ? sorted_keys : (k % 2 == 0) ? ka : kb;
    V *po = last ? sorted_vals : (k % 2 == 0) ? pa : pb;
    int64_t *op = last ? out_pos : nullptr;
    if (first && last) run_pass<U, V, true, true>(c, type, keys, nullptr, vals, n, 8 * d, db, ko, po, op);
    else if (first) run_pass<U, V, true, false>(c, type, keys, nullptr, vals, n, 8 * d, db, ko, po, op);
    else if (last) run_pass<U, V, false, true>(c, type, kin, pin, vals, n, 8 * d, db, ko, po, op);
    else run_pass<U, V, false, false>(c, type, kin, pin, vals, n, 8 * d, db, ko, po, op);
    kin = ko;
    pin = po;
  }
  tmp_free(c, hist);
  if (ka) tmp_free(c, ka);
  if (pa) tmp_free(c, pa);
  if (kb) tmp_free(c, kb);
  if (pb) tmp_free(c, pb);
}

}  // namespace

void argsort(Ctx &c, int32_t type, const void *keys, uint64_t n, int64_t *out) {
  switch (type) {
    case VK_U64: case VK_I64: case VK_F64:
      radix<uint64_t, uint32_t>(c, type, keys, nullptr, n, out, nullptr, nullptr);
      break;
    case VK_I32: radix<uint32_t, uint32_t>(c, type, keys, nullptr, n, out, nullptr, nullptr); break;
    case VK_U8: radix<uint8_t, uint32_t>(c, type, keys, nullptr, n, out, nullptr, nullptr); break;
    default: usage_error("unsupported argsort type");
  }
}

void argsort_batch(Ctx &c, int32_t type, std::vector<BatchItem> &items) {
  switch (type) {
    case VK_U64: case VK_I64: case VK_F64: argsort_batch_pos<uint64_t>(c, type, items); break;
    case VK_I32: argsort_batch_pos<uint32_t>(c, type, items); break;
    case VK_U8: radix_batch<uint8_t, uint32_t>(c, type, items); break;
    default: usage_error("unsupported argsort type");
  }
}

void sort_u64_pairs(Ctx &c, const uint64_t *keys, uint64_t n, uint64_t *sorted_keys, int64_t *perm) {
  radix<uint64_t, uint32_t>(c, VK_U64, keys, nullptr, n, perm, sorted_keys, nullptr);
}

void sort_key_values(Ctx &c, bool wide, const void *keys, const uint64_t *vals, uint64_t n,
                     void *sorted_keys, uint64_t *sorted_vals) {
  if (wide)
    radix<uint64_t, uint64_t>(c, VK_U64, keys, vals, n, nullptr, static_cast<uint64_t *>(sorted_keys), sorted_vals);
  else
    radix<uint32_t, uint64_t>(c, kRawU32, keys, vals, n, nullptr, static_cast<uint32_t *>(sorted_keys), sorted_vals);
}

void sort_chain_values(Ctx &c, bool wide, const int64_t *keys, const uint64_t *vals, uint64_t n,
                       void *sorted_keys, uint64_t *sorted_vals, uint64_t buckets, uint64_t *bad_row,
                       int *class_shift) {
  if (wide)
    radix<uint64_t, uint64_t>(c, kChainKey, keys, vals, n, nullptr, static_cast<uint64_t *>(sorted_keys), sorted_vals,
                              buckets, bad_row, class_shift);
  else
    radix<uint32_t, uint64_t>(c, kChainKey, keys, vals, n, nullptr, static_cast<uint32_t *>(sorted_keys), sorted_vals,
                              buckets, bad_row, class_shift);
}

void sort_chain_values_batch(Ctx &c, bool wide, std::vector<BatchItem> &items, void *sorted_keys,
                             uint64_t *sorted_vals) {
  if (wide)
    radix_batch<uint64_t, uint64_t>(c, kChainKey, items, static_cast<uint64_t *>(sorted_keys), sorted_vals);
  else
    radix_batch<uint32_t, uint64_t>(c, kChainKey, items, static_cast<uint32_t *>(sorted_keys), sorted_vals);
}

void sort_key_values_batch(Ctx &c, bool wide, std::vector<BatchItem> &items, void *sorted_keys,
                           uint64_t *sorted_vals) {
  if (wide)
    radix_batch<uint64_t, uint64_t>(c, VK_U64, items, static_cast<uint64_t *>(sorted_keys), sorted_vals);
  else
    radix_batch<uint32_t, uint64_t>(c, kRawU32, items, static_cast<uint32_t *>(sorted_keys), sorted_vals);
}


// ------------------------------------------------------------ sorts inside batch plans
// A batch plan's argsort, and the (chain key, value) sort of its reduce_into_add_f64, run with
// no host read, so that a plan captures them into its graph. The digits that vary in some item
// are listed on the device (k_plan_slots), and each of the P pass slots is launched twice, as a
// middle pass and as the last pass; k_pass exits at once, in every block, when its slot is not
// its role. Slot k reads what slot k - 1 wrote (buffers by parity), so the passes are the
// batched sort's over the active digits. Each slot has its own look-back status words, all
// stamped with the plan's device epoch.

namespace {

__global__ void k_plan_slots(const uint32_t *active, int P, int pos, int32_t *slots) {
  if (threadIdx.x != 0) return;
  int m = 0;
  for (int d = 0; d < P; ++d)
    if (active[d]) slots[1 + m++] = d;
  // with no digit to sort by, one stable pass on digit 0 still moves (key, value) rows as they are
  if (m == 0 && !pos) slots[1 + m++] = 0;
  slots[0] = m;
}

// Positions of items with no varying digit in any item: the identity.
__global__ void k_iota_plan(const BatchItem *items, const uint32_t *tile_item, const int32_t *slots) {
  if (slots[0] != 0) return;
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const uint64_t t0 = (blockIdx.x - I.tile0) * uint64_t(kTile);
  const uint64_t t1 = t0 + kTile < I.n ? t0 + kTile : I.n;
  int64_t *out = static_cast<int64_t *>(I.out);
  for (uint64_t r = t0 + threadIdx.x; r < t1; r += kBlock) out[r] = static_cast<int64_t>(r);
}

template <typename U, typename V>
void plan_sort_run_as(Ctx &c, PlanSort &s, const uint32_t *epoch) {
  constexpr int P = sizeof(U);
  constexpr bool POS = std::is_same<V, uint32_t>::value;
  if (s.bp.ntiles == 0) return;
  uint32_t *active = s.hist + s.nitems * P * kRadix;
  fill_u64(c, reinterpret_cast<uint64_t *>(s.hist), (s.nitems * P * kRadix + 16) / 2, 0);
  k_hist_batch<U><<<s.bp.ntiles, kBlock, 0, c.stream>>>(s.type, s.bp.items, s.bp.tile_item, s.hist);
  k_base_batch<<<stream_grid(s.nitems * P, kBlock, 1u << 20), kBlock, 0, c.stream>>>(s.bp.items, s.nitems, P,
                                                                                   s.hist, active);
  k_plan_slots<<<1, 64, 0, c.stream>>>(active, P, POS ? 1 : 0, s.slots);
  if (POS) k_iota_plan<<<s.bp.ntiles, kBlock, 0, c.stream>>>(s.bp.items, s.bp.tile_item, s.slots);
  for (int k = 0; k < P; ++k) {
    for (int last = 0; last < 2; ++last) {
      PassArgs<U, V> a{};
      a.type = s.type;
      a.keys_in = k ? s.kbuf[(k - 1) & 1] : nullptr;
      a.pay_in = k ? static_cast<const V *>(s.pbuf[(k - 1) & 1]) : nullptr;
      a.digit_base = s.hist;
      a.keys_out = static_cast<U *>(last ? s.sorted_keys : s.kbuf[k & 1]);
      a.pay_out = static_cast<V *>(last ? static_cast<void *>(s.sorted_vals) : s.pbuf[k & 1]);
      a.status = s.status + uint64_t(k) * s.bp.ntiles * kRadix;
      a.epoch_ptr = epoch;
      a.ticket = s.ticket;
      a.ntiles = s.bp.ntiles;
      a.items = s.bp.items;
      a.tile_item = s.bp.tile_item;
      a.P = P;
      a.slots = s.slots;
      a.slot = k;
      if (k == 0 && last) k_pass<U, V, true, true, true><<<s.bp.ntiles, kSortBlock, 0, c.stream>>>(a);
      else if (k == 0) k_pass<U, V, true, false, true><<<s.bp.ntiles, kSortBlock, 0, c.stream>>>(a);
      else if (last) k_pass<U, V, false, true, true><<<s.bp.ntiles, kSortBlock, 0, c.stream>>>(a);
      else k_pass<U, V, false, false, true><<<s.bp.ntiles, kSortBlock, 0, c.stream>>>(a);
    }
  }
  check_hip(hipGetLastError(), "plan sort");
}

int key_bytes(int32_t type, bool wide) {
  switch (type) {
    case VK_I32: return 4;
    case VK_U8: return 1;
    case kChainKey: return wide ? 8 : 4;
    default: return 8;
  }
}

}  // namespace

int32_t chain_key_type() { return kChainKey; }

PlanSort *plan_sort_new(Ctx &c, int32_t type, bool pos, bool wide, std::vector<BatchItem> &items, void *sorted_keys,
                        uint64_t *sorted_vals) {
  auto *s = new PlanSort{};
  s->type = type;
  s->pos = pos;
  s->wide = wide;
  s->nitems = items.size();
  uint64_t total = 0;
  for (auto &it : items) {
    if (it.n > 0xFFFFFFFFull) usage_error("argsort supports at most 2^32 - 1 rows");
    it.aux = total;
    total += it.n;
  }
  s->total = total;
  const int P = key_bytes(type, wide);
  s->P = P;
  s->bp = batch_plan(c, items, kTile);
  s->hist = static_cast<uint32_t *>(tmp_alloc(c, (s->nitems * P * kRadix + 16) * 4));
  s->slots = static_cast<int32_t *>(tmp_alloc(c, 64));
  const uint64_t vb = pos ? 4 : 8;
  for (int k = 0; k < 2; ++k) {
    s->kbuf[k] = tmp_alloc(c, (total ? total : 1) * uint64_t(P));
    s->pbuf[k] = tmp_alloc(c, (total ? total : 1) * vb);
  }
  s->sorted_keys = sorted_keys;
  s->sorted_vals = sorted_vals;
  s->status_words = uint64_t(P) * s->bp.ntiles * kRadix;
  s->status = static_cast<uint64_t *>(tmp_alloc(c, (s->status_words ? s->status_words : 1) * 8));
  fill_u64(c, s->status, s->status_words, 0);
  s->ticket = static_cast<uint32_t *>(tmp_alloc(c, 8));
  fill_u64(c, reinterpret_cast<uint64_t *>(s->ticket), 1, 0);
  return s;
}

void plan_sort_run(Ctx &c, PlanSort &s, const uint32_t *epoch) {
  if (s.pos) {
    switch (s.P) {
      case 8: plan_sort_run_as<uint64_t, uint32_t>(c, s, epoch); break;
      case 4: plan_sort_run_as<uint32_t, uint32_t>(c, s, epoch); break;
      default: plan_sort_run_as<uint8_t, uint32_t>(c, s, epoch); break;
    }
  } else if (s.P == 8) {
    plan_sort_run_as<uint64_t, uint64_t>(c, s, epoch);
  } else {
    plan_sort_run_as<uint32_t, uint64_t>(c, s, epoch);
  }
}

void plan_sort_free(Ctx &c, PlanSort *s) {
  if (!s) return;
  batch_plan_free(c, s->bp);
  tmp_free(c, s->hist);
  tmp_free(c, s->slots);
  for (int k = 0; k < 2; ++k) {
    tmp_free(c, s->kbuf[k]);
    tmp_free(c, s->pbuf[k]);
  }
  tmp_free(c, s->status);
  tmp_free(c, s->ticket);
  delete s;
}

}  // namespace vk
