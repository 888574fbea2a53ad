// argsort_<t> (ArgSort<T>, src/kernels/scalar/index.rs:157-173; radix.rs:1-259): the
// positions of the rows in stable ascending order, as i64. The stable ascending permutation
// is unique, so any stable sort returns the reference's output; this one keeps the
// reference's structure of an LSD radix sort on (key image, position) pairs with byte
// digits, one counting pass for every digit, and digits that are equal in every row
// skipped (radix.rs:187-259).
//
// MI355X form: the counting pass builds all digit histograms in LDS per block. Each
// remaining digit is one pass over 4096-row tiles: the rows of a wave are ranked among
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
constexpr int kTile = kBlock * kItems;  // 4096 rows
constexpr int kWaves = kBlock / kWave;
constexpr int kRadix = 256;

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
    default: return U(static_cast<const uint64_t *>(keys)[r]);  // u64
  }
}

// ------------------------------------------------------------- counting pass

template <typename U>
__global__ __launch_bounds__(kBlock) void k_hist(int32_t type, const void *keys, uint64_t n,
                                                 uint32_t *hist) {
  constexpr int P = sizeof(U);
  __shared__ uint32_t s_h[P][kRadix];
  for (int k = threadIdx.x; k < P * kRadix; k += kBlock) (&s_h[0][0])[k] = 0;
  __syncthreads();
  for (uint64_t r = blockIdx.x * uint64_t(kBlock) + threadIdx.x; r < n;
       r += uint64_t(gridDim.x) * kBlock) {
    const U k = image<U>(type, keys, r);
#pragma unroll
    for (int d = 0; d < P; ++d) atomicAdd(&s_h[d][(k >> (8 * d)) & 255], 1u);
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

// FIRST: keys are the typed input (imaged here) and positions are the row numbers.
// LAST: positions go to out_pos as i64 (and keys to keys_out when given).
template <typename U, bool FIRST, bool LAST>
__global__ __launch_bounds__(kBlock) void k_pass(int32_t type, const void *keys_in,
                                                 const uint32_t *__restrict__ pos_in, uint64_t n,
                                                 int shift, const uint32_t *__restrict__ digit_base,
                                                 U *__restrict__ keys_out, uint32_t *__restrict__ pos_out,
                                                 int64_t *__restrict__ out_pos, uint64_t *status,
                                                 uint32_t epoch, uint32_t *ticket, uint32_t ntiles) {
  __shared__ uint32_t s_wcnt[kWaves][kRadix];  // per-wave counters, then per-wave offsets
  __shared__ uint32_t s_excl[kRadix];           // tile-local exclusive digit offsets
  __shared__ uint64_t s_gbase[kRadix];          // global position of the tile's first row per digit
  __shared__ U s_keys[kTile];
  __shared__ uint32_t s_pos[kTile];
  __shared__ uint32_t s_wsum[kWaves];

  const uint32_t tile = take_ticket(ticket, ntiles);
  const uint64_t t0 = uint64_t(tile) * kTile;
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  for (int k = t; k < kWaves * kRadix; k += kBlock) (&s_wcnt[0][0])[k] = 0;
  __syncthreads();

  U key[kItems];
  uint32_t pos[kItems];
  uint32_t rnk[kItems];
  const uint64_t wbase = t0 + uint64_t(w) * (kTile / kWaves);
  const uint64_t below = (uint64_t(1) << l) - 1;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t r = wbase + uint64_t(i) * kWave + l;
    const bool valid = r < n;
    if (valid) {
      key[i] = FIRST ? image<U>(type, keys_in, r) : static_cast<const U *>(keys_in)[r];
      pos[i] = FIRST ? static_cast<uint32_t>(r) : pos_in[r];
    } else {
      key[i] = 0;
      pos[i] = 0;
    }
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
  // per digit: wave offsets and the tile count (thread t owns digit t)
  uint32_t cnt = 0;
  {
    const int dd = t;
#pragma unroll
    for (int k = 0; k < kWaves; ++k) {
      const uint32_t v = s_wcnt[k][dd];
      s_wcnt[k][dd] = cnt;
      cnt += v;
    }
  }
  // tile-local exclusive scan of the 256 digit counts
  uint32_t inc = cnt;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d);
    if (l >= d) inc += o;
  }
  if (l == 63) s_wsum[w] = inc;
  __syncthreads();
  uint32_t wpre = 0;
  for (int k = 0; k < w; ++k) wpre += s_wsum[k];
  s_excl[t] = wpre + inc - cnt;
  // publish this tile's count of digit t and look back for the tiles before it
  {
    const int dd = t;
    uint64_t *st = status + uint64_t(tile) * kRadix + dd;
    uint64_t prefix = 0;
    if (tile == 0) {
      st_rlx(st, pack(epoch, 2, cnt));
    } else {
      st_rlx(st, pack(epoch, 1, cnt));
      int64_t j = int64_t(tile) - 1;
      while (j >= 0) {
        const uint64_t wd = ld_rlx(status + uint64_t(j) * kRadix + dd);
        const uint64_t flag = (wd >> 46) & 3;
        if ((wd >> 48) != epoch || flag == 0) {
          backoff();
          continue;
        }
        prefix += wd & kCountMask;
        if (flag == 2) break;
        --j;
      }
      st_rlx(st, pack(epoch, 2, prefix + cnt));
    }
    s_gbase[dd] = uint64_t(digit_base[dd]) + prefix;
  }
  __syncthreads();
  // stage the rows sorted by digit within the tile
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t r = wbase + uint64_t(i) * kWave + l;
    if (r < n) {
      const uint32_t d = (key[i] >> shift) & 255;
      const uint32_t at = s_excl[d] + s_wcnt[w][d] + rnk[i];
      s_keys[at] = key[i];
      s_pos[at] = pos[i];
    }
  }
  __syncthreads();
  const uint64_t rows = n - t0 < uint64_t(kTile) ? n - t0 : uint64_t(kTile);
  for (uint32_t j = t; j < rows; j += kBlock) {
    const U k = s_keys[j];
    const uint32_t d = (k >> shift) & 255;
    const uint64_t g = s_gbase[d] + (j - s_excl[d]);
    if (LAST) {
      out_pos[g] = static_cast<int64_t>(s_pos[j]);
      if (keys_out) keys_out[g] = k;
    } else {
      keys_out[g] = k;
      pos_out[g] = s_pos[j];
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

template <typename U, bool FIRST, bool LAST>
void run_pass(Ctx &c, int32_t type, const void *kin, const uint32_t *pin, uint64_t n, int shift,
              const uint32_t *dbase, U *kout, uint32_t *pout, int64_t *out_pos) {
  const uint64_t tiles = (n + kTile - 1) / kTile;
  uint64_t *status = status_words(c, ST_SORT, tiles * kRadix);
  const uint32_t epoch = next_epoch(c);
  k_pass<U, FIRST, LAST><<<uint32_t(tiles), kBlock, 0, c.stream>>>(
      type, kin, pin, n, shift, dbase, kout, pout, out_pos, status, epoch, c.scratch.counters,
      uint32_t(tiles));
  check_hip(hipGetLastError(), "radix pass");
}

// Sorts rows by image; positions of the sorted rows to out_pos, sorted images to
// sorted_keys when not null.
template <typename U>
void radix(Ctx &c, int32_t type, const void *keys, uint64_t n, int64_t *out_pos, U *sorted_keys) {
  if (n == 0) return;
  if (n > 0xFFFFFFFFull) usage_error("argsort supports at most 2^32 - 1 rows");
  constexpr int P = sizeof(U);
  uint32_t *hist = static_cast<uint32_t *>(tmp_alloc(c, P * kRadix * 4));
  fill_u64(c, reinterpret_cast<uint64_t *>(hist), P * kRadix / 2, 0);
  k_hist<U><<<stream_grid(n, kBlock * 8, 1024), kBlock, 0, c.stream>>>(type, keys, n, hist);
  check_hip(hipGetLastError(), "radix histogram");
  std::vector<uint32_t> h(P * kRadix);
  check_hip(hipMemcpyAsync(h.data(), hist, P * kRadix * 4, hipMemcpyDeviceToHost, c.stream), "hist");
  check_hip(hipStreamSynchronize(c.stream), "sync");
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
  if (active.empty()) {
    k_iota64<<<g, kBlock, 0, c.stream>>>(n, out_pos);
    if (sorted_keys) k_copy_images<U><<<g, kBlock, 0, c.stream>>>(type, keys, n, sorted_keys);
    tmp_free(c, hist);
    return;
  }
  const int np = static_cast<int>(active.size());
  U *ka = nullptr, *kb = nullptr;
  uint32_t *pa = nullptr, *pb = nullptr;
  if (np > 1) {
    ka = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
    pa = static_cast<uint32_t *>(tmp_alloc(c, n * 4));
    if (np > 2) {
      kb = static_cast<U *>(tmp_alloc(c, n * sizeof(U)));
      pb = static_cast<uint32_t *>(tmp_alloc(c, n * 4));
    }
  }
  const U *kin = nullptr;
  const uint32_t *pin = nullptr;
  for (int k = 0; k < np; ++k) {
    const int d = active[k];
    const uint32_t *db = hist + d * kRadix;
    const bool first = k == 0, last = k == np - 1;
    U *ko = (k % 2 == 0) ? ka : kb;
    uint32_t *po = (k % 2 == 0) ? pa : pb;
    if (first && last)
      run_pass<U, true, true>(c, type, keys, nullptr, n, 8 * d, db, sorted_keys, nullptr, out_pos);
    else if (first)
      run_pass<U, true, false>(c, type, keys, nullptr, n, 8 * d, db, ko, po, nullptr);
    else if (last)
      run_pass<U, false, true>(c, type, kin, pin, n, 8 * d, db, sorted_keys, nullptr, out_pos);
    else
      run_pass<U, false, false>(c, type, kin, pin, n, 8 * d, db, ko, po, nullptr);
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
    case VK_U64: case VK_I64: case VK_F64: radix<uint64_t>(c, type, keys, n, out, nullptr); break;
    case VK_I32: radix<uint32_t>(c, type, keys, n, out, nullptr); break;
    case VK_U8: radix<uint8_t>(c, type, keys, n, out, nullptr); break;
    default: usage_error("unsupported argsort type");
  }
}

void sort_u64_pairs(Ctx &c, const uint64_t *keys, uint64_t n, uint64_t *sorted_keys, int64_t *perm) {
  radix<uint64_t>(c, VK_U64, keys, n, perm, sorted_keys);
}

}  // namespace vk
