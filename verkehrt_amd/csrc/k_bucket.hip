// Bucketed folds: histogram (index.rs:137-155) and reduce_into_<op>_<t> (ReduceInto,
// index.rs:192-222): the fold of the values per key in 0..n, an empty bucket holding the
// fold's identity. Keys outside 0..n are an error naming the first bad key in row order
// (checked_keys, index.rs:113-126).
//
// Integer add and count commute, so rows fold with atomics: into a per-block LDS copy of the
// buckets when they fit (one global atomic per bucket per block), else straight into HBM.
// min/max use compare-and-swap loops with the IEEE minimum/maximum rules.
//
// f64 add has no reference kernel (builtins.dd:67-71); its order here is the tier 2
// candidate (tier2-design.md:171): row i < b = n - n%16 folds into stripe i%16 of its bucket
// in index order, the stripes of a bucket combine pairwise, then rows >= b add in index
// order. Rows are stably sorted by (bucket, stripe), each stripe folds sequentially, and
// stripes combine per bucket.
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

constexpr uint64_t kLdsBuckets = 4096;  // 32 KB of u64 buckets per block

__device__ __forceinline__ void note_bad(uint64_t row, uint64_t *my_bad) {
  if (row < *my_bad) *my_bad = row;
}

__device__ __forceinline__ void flush_bad(uint64_t my_bad, uint64_t *bad) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint64_t o = __shfl_xor(my_bad, m);
    my_bad = o < my_bad ? o : my_bad;
  }
  if (lane_id() == 0 && my_bad != ~uint64_t(0))
    atomicMin(reinterpret_cast<unsigned long long *>(bad), my_bad);
}

// op 0 count, 1 add. Values are i64 (add) or unused (count).
template <int OP, bool LDS>
__global__ __launch_bounds__(kBlock) void k_bucket_add(const int64_t *__restrict__ keys,
                                                       const int64_t *__restrict__ vals, uint64_t n,
                                                       uint64_t buckets, unsigned long long *out,
                                                       uint64_t *bad) {
  __shared__ unsigned long long s_b[LDS ? kLdsBuckets : 1];
  if (LDS) {
    for (uint64_t k = threadIdx.x; k < buckets; k += blockDim.x) s_b[k] = 0;
    __syncthreads();
  }
  uint64_t my_bad = ~uint64_t(0);
  grid_rows(keys, n, [&](int64_t k, uint64_t row) {
    if (static_cast<uint64_t>(k) >= buckets) {
      note_bad(row, &my_bad);
      return;
    }
    const unsigned long long v = OP == 0 ? 1ull : static_cast<unsigned long long>(vals[row]);
    if (LDS) atomicAdd(&s_b[k], v);
    else atomicAdd(&out[k], v);
  });
  flush_bad(my_bad, bad);
  if (LDS) {
    __syncthreads();
    for (uint64_t k = threadIdx.x; k < buckets; k += blockDim.x)
      if (s_b[k]) atomicAdd(&out[k], s_b[k]);
  }
}

template <typename T, bool MAX>
__global__ __launch_bounds__(kBlock) void k_bucket_ext(const int64_t *__restrict__ keys,
                                                       const T *__restrict__ vals, uint64_t n,
                                                       uint64_t buckets, T *out, uint64_t *bad) {
  uint64_t my_bad = ~uint64_t(0);
  grid_rows(keys, n, [&](int64_t k, uint64_t row) {
    if (static_cast<uint64_t>(k) >= buckets) {
      note_bad(row, &my_bad);
      return;
    }
    const T v = vals[row];
    unsigned long long *p = reinterpret_cast<unsigned long long *>(&out[k]);
    unsigned long long old = *p;
    while (true) {
      const T cur = from_bits<T>(old);
      const T nv = MAX ? max2(cur, v) : min2(cur, v);
      const unsigned long long nb = to_bits(nv);
      if (nb == old) break;
      const unsigned long long got = atomicCAS(p, old, nb);
      if (got == old) break;
      old = got;
    }
  });
  flush_bad(my_bad, bad);
}

__global__ void k_check_keys(const int64_t *__restrict__ keys, uint64_t n, uint64_t buckets,
                             uint64_t *bad) {
  uint64_t my_bad = ~uint64_t(0);
  grid_rows(keys, n, [&](int64_t k, uint64_t row) {
    if (static_cast<uint64_t>(k) >= buckets) note_bad(row, &my_bad);
  });
  flush_bad(my_bad, bad);
}

// ----------------------------------------------------------- f64 reduce_into add

// The first row of every chain key 0..nchains - 1 in the sorted chain keys (its lower bound,
// by binary search), empty chains included: the chain starts without a pass over the rows.
template <typename K>
__global__ void k_chain_bounds(const K *__restrict__ sk, uint64_t b, uint64_t nchains, uint64_t *starts) {
  for (uint64_t ch = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; ch < nchains;
       ch += uint64_t(gridDim.x) * blockDim.x) {
    const K key = static_cast<K>(ch);
    uint64_t lo = 0, hi = b;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (sk[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    starts[ch] = lo;
  }
}

// One wave per chain for long chains (starts[ch] is chain key ch's first row; the last chain
// ends at b): the lanes load 64 values at a time, the wave folds them in order with
// constant-lane reads (readlane, no LDS round trip) while the next chunks load. Every lane
// holds the same running value. An empty chain leaves its stripe at 0.0.
constexpr int kFoldWaves = 4;  // chains per block of the wave fold

__global__ __launch_bounds__(kFoldWaves * kWave) void k_chain_fold_wave(const uint64_t *__restrict__ starts,
                                                                       uint64_t nchains,
                                                                       const double *__restrict__ sv,
                                                                       uint64_t b, double *stripes) {
  const int l = lane_id();
  const uint64_t w0 = blockIdx.x * uint64_t(kFoldWaves) + threadIdx.x / kWave;
  for (uint64_t ch = w0; ch < nchains; ch += uint64_t(gridDim.x) * kFoldWaves) {
    const uint64_t p0 = starts[ch];
    const uint64_t p1 = ch + 1 < nchains ? starts[ch + 1] : b;
    double a = 0.0;
    // kAhead chunks of 64 values in flight: chunk d of each round is folded, then refilled
    // with the chunk kAhead further on
    constexpr int kAhead = 8;
    double v[kAhead];
#pragma unroll
    for (int d = 0; d < kAhead; ++d) {
      const uint64_t r = p0 + d * 64 + l;
      v[d] = r < p1 ? sv[r] : 0.0;
    }
    for (uint64_t q = p0; q < p1; q += kAhead * 64) {
#pragma unroll
      for (int d = 0; d < kAhead; ++d) {
        const uint64_t qd = q + d * 64;
        if (qd >= p1) break;
        if (p1 - qd >= 64) {
#pragma unroll
          for (int j = 0; j < 64; ++j) a = a + bcast_pod(v[d], j);
        } else {
          const int cnt = static_cast<int>(p1 - qd);
          for (int j = 0; j < cnt; ++j) a = a + __shfl(v[d], j);
        }
        const uint64_t r = qd + kAhead * 64 + l;
        v[d] = r < p1 ? sv[r] : 0.0;
      }
    }
    if (l == 0 && p1 > p0) stripes[ch] = a;
  }
}

// Short chains: lane l of a wave folds bucket w * 64 + l (the compacted bucket starts). The
// wave's 64 buckets are consecutive rows of the sorted pairs, which it streams through LDS in
// chunks of kBucketChunk rows with coalesced loads; each lane folds its own rows of the chunk.
// A bucket's rows are its chains in stripe order, each in index order; every chain folds
// from 0.0 and the 16 stripes (0.0 where a stripe has no rows) combine pairwise into
// out[bucket], as k_combine_stripes does.
constexpr int kBucketChunk = 1024;

template <typename K>
__global__ __launch_bounds__(kWave) void k_bucket_fold(const uint64_t *__restrict__ starts, uint64_t nstarts,
                                                       const K *__restrict__ sk, const double *__restrict__ sv,
                                                       uint64_t b, double *out) {
  __shared__ K s_k[kBucketChunk];
  __shared__ double s_v[kBucketChunk];
  const int l = lane_id();
  for (uint64_t w0 = blockIdx.x * uint64_t(kWave); w0 < nstarts; w0 += uint64_t(gridDim.x) * kWave) {
    const uint64_t bi = w0 + l;
    const bool live = bi < nstarts;
    const uint64_t p0 = live ? starts[bi] : b;
    const uint64_t p1 = live ? (bi + 1 < nstarts ? starts[bi + 1] : b) : b;
    // the wave's rows: from lane 0's first to the last live lane's end
    const uint64_t r0 = bcast_pod(p0, 0);
    uint64_t r1 = live ? p1 : 0;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const uint64_t o = __shfl_xor(r1, m);
      r1 = o > r1 ? o : r1;
    }
    double a[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) a[j] = 0.0;
    int cur = -1;
    double acc = 0.0;
    auto put = [&](int s, double v) {
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j == s) a[j] = v;
    };
    for (uint64_t c = r0; c < r1; c += kBucketChunk) {
      K kk[kBucketChunk / kWave];
      double vv[kBucketChunk / kWave];
#pragma unroll
      for (int j = 0; j < kBucketChunk / kWave; ++j) {
        const uint64_t r = c + uint64_t(j) * kWave + l;
        kk[j] = r < r1 ? sk[r] : K(0);
        vv[j] = r < r1 ? sv[r] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < kBucketChunk / kWave; ++j) {
        s_k[j * kWave + l] = kk[j];
        s_v[j * kWave + l] = vv[j];
      }
      __syncthreads();
      const uint64_t e = c + kBucketChunk;
      const uint64_t q0 = p0 > c ? p0 : c;
      const uint64_t q1 = p1 < e ? p1 : e;
      for (uint64_t q = q0; q < q1; ++q) {
        const int s = static_cast<int>(s_k[q - c] & 15);
        const double v = s_v[q - c];
        if (s != cur) {
          if (cur >= 0) put(cur, acc);
          acc = 0.0;
          cur = s;
        }
        acc = acc + v;
      }
      __syncthreads();
    }
    if (live) {
      put(cur, acc);
#pragma unroll
      for (int step = 8; step > 0; step /= 2)
#pragma unroll
        for (int j = 0; j < step; ++j) a[j] = a[j] + a[j + step];
      out[static_cast<uint64_t>(sk[p0]) >> 4] = a[0];
    }
  }
}

// flag[p] = 1 where row p of the sorted chain keys starts a chain (SHIFT 0) or a bucket
// (SHIFT 4: chain key / 16 changes). A thread takes four consecutive rows and stores their
// flags as one word (flag is a fresh, aligned allocation).
template <int SHIFT, typename K>
__global__ void k_key_starts(const K *__restrict__ sk, uint64_t b, uint8_t *flag) {
  const uint64_t quads = (b + 3) / 4;
  for (uint64_t q = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; q < quads;
       q += uint64_t(gridDim.x) * blockDim.x) {
    const uint64_t p = q * 4;
    K k[5];
    k[0] = p > 0 ? sk[p - 1] : K(0);
#pragma unroll
    for (int u = 0; u < 4; ++u) k[u + 1] = p + u < b ? sk[p + u] : K(0);
    uint32_t word = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (p + u == 0 || (k[u] >> SHIFT) != (k[u + 1] >> SHIFT)) word |= 1u << (8 * u);
    if (p + 4 <= b) {
      *reinterpret_cast<uint32_t *>(flag + p) = word;
    } else {
      for (int u = 0; p + u < b; ++u) flag[p + u] = static_cast<uint8_t>((word >> (8 * u)) & 1);
    }
  }
}

// One lane per chain when there are many long chains: a wave takes 64 consecutive chains and
// moves 64 values of each through LDS per round (one coalesced 512-byte load per chain, all 64
// issued before the first is stored; rows padded to 65 values so that a lane's row walk
// spreads over the banks); lane l then folds its own chain's 64 values in index order. Two
// LDS buffers let the loads of round q + 1 fly while round q is folded. A chain shorter than
// the round's window adds nothing past its end. Chains are indexed by chain key, as in the
// wave fold; an empty chain leaves its stripe at 0.0.
constexpr int kLaneFoldPad = 65;

__global__ __launch_bounds__(kWave) void k_chain_fold_lanes(const uint64_t *__restrict__ starts,
                                                             uint64_t nchains,
                                                             const double *__restrict__ sv, uint64_t b,
                                                             double *stripes) {
  __shared__ double s_v[2][kWave * kLaneFoldPad];
  const int l = lane_id();
  for (uint64_t c0 = blockIdx.x * uint64_t(kWave); c0 < nchains; c0 += uint64_t(gridDim.x) * kWave) {
    const uint64_t ch = c0 + l;
    const bool live = ch < nchains;
    const uint64_t p0 = live ? starts[ch] : 0;
    const uint64_t p1 = live ? (ch + 1 < nchains ? starts[ch + 1] : b) : 0;
    const uint64_t len = p1 - p0;
    // the longest chain of the wave sets the rounds
    uint64_t most = len;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const uint64_t o = __shfl_xor(most, m);
      most = o > most ? o : most;
    }
    double v[kWave];
    // chain i's values q .. q + 63, for lane l the value q + l
    auto load_round = [&](uint64_t q) {
      const uint64_t r = q + l;
#pragma unroll
      for (int i = 0; i < kWave; ++i) {
        const uint64_t pi = bcast_pod(p0, i), li = bcast_pod(len, i);
        v[i] = r < li ? sv[pi + r] : 0.0;
      }
    };
    auto store_round = [&](int buf) {
#pragma unroll
      for (int i = 0; i < kWave; ++i) s_v[buf][i * kLaneFoldPad + l] = v[i];
    };
    double a = 0.0;
    if (most > 0) {
      load_round(0);
      store_round(0);
    }
    __syncthreads();
    int buf = 0;
    for (uint64_t q = 0; q < most; q += kWave) {
      const bool more = q + kWave < most;
      if (more) load_round(q + kWave);
      const double *row = &s_v[buf][l * kLaneFoldPad];
      if (len >= q + kWave) {
#pragma unroll 16
        for (int k = 0; k < kWave; ++k) a = a + row[k];
      } else if (len > q) {
        const int cnt = static_cast<int>(len - q);
        for (int k = 0; k < cnt; ++k) a = a + row[k];
      }
      if (more) store_round(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    }
    if (live && p1 > p0) stripes[ch] = a;
    __syncthreads();
  }
}

// Class folds. When the sort leaves the chain keys' top non-constant digit unsorted, the rows
// of one class (the key bits below `shift`) are contiguous and in row order, and the chains of
// a class differ only in the digit at `shift` (the bits above it are equal in every row).
// k_class_bounds finds every class's first row by binary search of the masked sorted keys
// (starts[nclasses] = b). k_class_fold takes one class per block and streams its rows in
// chunks of THREADS * kClassItems: a chunk is partitioned stably by the digit in LDS, then
// thread d adds the chunk's rows of digit d to its chain's running value, in row order, and
// keeps the value for the next chunk. At the end each thread that saw rows writes its chain's
// stripe. The fold of a chunk runs while the waves rank the next one (ranking uses registers
// and per-wave entries only), and the loads of the chunk after that go out once the next
// chunk is placed.
//
// Ranking: a wave takes 64 rows of the chunk per step. Each lane ORs its lane bit into its
// digit's mask in the wave's LDS entries (an OR does not depend on order), then reads the mask
// back: the lanes of its step with its digit. Its rank is the digit's count from the wave's
// earlier steps plus the lanes of that mask below it; the lowest lane of the mask clears the
// mask and advances the count. A wave's LDS operations complete in issue order, so the read
// sees every lane's bit of its step, and the next step's ORs follow the clear.
constexpr int kClassItems = 8;
constexpr int kClassRadix = 256;

__global__ void k_class_bounds(const uint32_t *__restrict__ sk, uint64_t b, uint64_t nclasses, uint32_t mask,
                               uint64_t *starts) {
  for (uint64_t c = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; c <= nclasses;
       c += uint64_t(gridDim.x) * blockDim.x) {
    uint64_t lo = 0, hi = b;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if ((sk[mid] & mask) < c) lo = mid + 1;
      else hi = mid;
    }
    starts[c] = lo;
  }
}

template <int THREADS>
__global__ __launch_bounds__(THREADS) void k_class_fold(const uint64_t *__restrict__ starts, int shift,
                                                        const uint32_t *__restrict__ sk,
                                                        const double *__restrict__ sv, double *stripes) {
  constexpr int W = THREADS / kWave;
  constexpr int C = THREADS * kClassItems;
  constexpr int kWaveRows = kClassItems * kWave;
  constexpr int kFoldAhead = 8;
  static_assert(THREADS >= kClassRadix, "one thread per digit");
  __shared__ double s_v[C];
  __shared__ unsigned long long s_mask[W][kClassRadix];
  __shared__ uint32_t s_cnt[W][kClassRadix];
  __shared__ uint32_t s_excl[kClassRadix + 1];
  __shared__ uint32_t s_wsum[kClassRadix / kWave];
  const int t = threadIdx.x, w = t / kWave, l = lane_id();
  const uint64_t p0 = starts[blockIdx.x], p1 = starts[blockIdx.x + 1];
  if (p0 == p1) return;
  const uint64_t below = (uint64_t(1) << l) - 1;
  uint32_t key[kClassItems];
  double val[kClassItems];
  uint32_t rnk[kClassItems];
  auto load = [&](uint64_t q) {
#pragma unroll
    for (int i = 0; i < kClassItems; ++i) {
      const uint64_t r = q + uint64_t(w) * kWaveRows + uint64_t(i) * kWave + l;
      const bool ok = r < p1;
      key[i] = ok ? sk[r] : 0u;
      val[i] = ok ? sv[r] : 0.0;
    }
  };
  auto rows_at = [&](uint64_t q) {
    return p1 - q < uint64_t(C) ? static_cast<uint32_t>(p1 - q) : uint32_t(C);
  };
  // the ranks of the m rows in registers, from the wave's own entries
  auto rank = [&](uint32_t m) {
#pragma unroll
    for (int i = 0; i < kClassItems; ++i) {
      const uint32_t r = uint32_t(w) * kWaveRows + uint32_t(i) * kWave + l;
      const uint32_t d = (key[i] >> shift) & (kClassRadix - 1);
      rnk[i] = 0;
      if (r < m) {
        atomicOr(&s_mask[w][d], 1ull << l);
        const uint64_t peers = s_mask[w][d];
        const uint32_t cnt = s_cnt[w][d];
        const uint32_t ahead = static_cast<uint32_t>(__popcll(peers & below));
        rnk[i] = cnt + ahead;
        if (ahead == 0) {
          s_mask[w][d] = 0;
          s_cnt[w][d] = cnt + static_cast<uint32_t>(__popcll(peers));
        }
      }
    }
  };
  // after the ranks: the digit starts (s_excl), and each wave's count of digit d becomes the
  // place of the wave's first row of digit d
  auto scan = [&](uint32_t m) {
    uint32_t cnt = 0, inc = 0;
    uint32_t col[W];
    if (t < kClassRadix) {
#pragma unroll
      for (int z = 0; z < W; ++z) {
        col[z] = cnt;
        cnt += s_cnt[z][t];
      }
      inc = cnt;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t u = __shfl_up(inc, o);
        if (l >= o) inc += u;
      }
      if (l == kWave - 1) s_wsum[w] = inc;
    }
    __syncthreads();
    if (t < kClassRadix) {
      uint32_t wpre = 0;
      for (int z = 0; z < w; ++z) wpre += s_wsum[z];
      const uint32_t ex = wpre + inc - cnt;
      s_excl[t] = ex;
      if (t == 0) s_excl[kClassRadix] = m;
#pragma unroll
      for (int z = 0; z < W; ++z) s_cnt[z][t] = ex + col[z];
    }
  };
  auto place = [&](uint32_t m) {
#pragma unroll
    for (int i = 0; i < kClassItems; ++i) {
      const uint32_t r = uint32_t(w) * kWaveRows + uint32_t(i) * kWave + l;
      if (r < m) {
        const uint32_t d = (key[i] >> shift) & (kClassRadix - 1);
        s_v[s_cnt[w][d] + rnk[i]] = val[i];
      }
    }
    // the wave's counts start again from zero for its next ranks (its reads above come first)
    for (int z = l; z < kClassRadix; z += kWave) s_cnt[w][z] = 0;
  };
  for (int z = t; z < W * kClassRadix; z += THREADS) {
    (&s_mask[0][0])[z] = 0;
    (&s_cnt[0][0])[z] = 0;
  }
  __syncthreads();
  uint64_t q = p0;
  load(q);
  rank(rows_at(q));
  __syncthreads();
  scan(rows_at(q));
  __syncthreads();
  place(rows_at(q));
  if (q + C < p1) load(q + C);
  __syncthreads();
  double acc = 0.0;
  bool seen = false;
  while (true) {
    // chunk q's rows of digit t are s_v[lo, hi)
    uint32_t lo = 0, hi = 0;
    if (t < kClassRadix) {
      lo = s_excl[t];
      hi = s_excl[t + 1];
    }
    const uint64_t qn = q + C;
    const bool more = qn < p1;
    if (t < kClassRadix) {
      seen = seen || hi > lo;
      uint32_t j = lo;
      for (; j + kFoldAhead <= hi; j += kFoldAhead) {
        double u[kFoldAhead];
#pragma unroll
        for (int k = 0; k < kFoldAhead; ++k) u[k] = s_v[j + k];
#pragma unroll
        for (int k = 0; k < kFoldAhead; ++k) acc = acc + u[k];
      }
      if (j < hi) {
        // past hi a read repeats the last row and its add is skipped
        double u[kFoldAhead];
#pragma unroll
        for (int k = 0; k < kFoldAhead; ++k) u[k] = s_v[j + k < hi ? j + k : hi - 1];
#pragma unroll
        for (int k = 0; k < kFoldAhead; ++k)
          if (j + k < hi) acc = acc + u[k];
      }
    }
    if (more) rank(rows_at(qn));
    __syncthreads();  // chunk q folded, chunk qn ranked
    if (!more) break;
    scan(rows_at(qn));
    __syncthreads();
    place(rows_at(qn));
    if (qn + C < p1) load(qn + C);
    __syncthreads();
    q = qn;
  }
  if (t < kClassRadix && seen) {
    // the chain of digit t: the bits above the digit as in every row, then t, then the class
    const uint32_t above = sk[p0] >> shift >> 8;
    stripes[(((above << 8) | uint32_t(t)) << shift) | uint32_t(blockIdx.x)] = acc;
  }
}

__global__ void k_combine_stripes(const double *__restrict__ stripes, uint64_t buckets, double *out) {
  for (uint64_t k = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; k < buckets;
       k += uint64_t(gridDim.x) * blockDim.x) {
    double a[16];
    for (int j = 0; j < 16; ++j) a[j] = stripes[k * 16 + j];
    for (int step = 8; step > 0; step /= 2)
      for (int j = 0; j < step; ++j) a[j] = a[j] + a[j + step];
    out[k] = a[0];
  }
}

__global__ void k_fold_tail(const int64_t *keys, const double *vals, uint64_t b, uint64_t n,
                            double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (uint64_t i = b; i < n; ++i) out[keys[i]] = out[keys[i]] + vals[i];
}

template <typename T>
__global__ void k_fill(T *p, uint64_t n, T v) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

}  // namespace

void bucket_init(Ctx &c, int op, int32_t type, void *out, uint64_t buckets) {
  if (buckets == 0) return;
  const uint32_t g = stream_grid(buckets, kBlock, 4096);
  if (op <= 1) {
    k_fill<uint64_t><<<g, kBlock, 0, c.stream>>>(static_cast<uint64_t *>(out), buckets, 0);
  } else if (type == VK_F64) {
    const double id = op == 2 ? __builtin_inf() : -__builtin_inf();
    k_fill<double><<<g, kBlock, 0, c.stream>>>(static_cast<double *>(out), buckets, id);
  } else {
    const int64_t id = op == 2 ? INT64_MAX : INT64_MIN;
    k_fill<int64_t><<<g, kBlock, 0, c.stream>>>(static_cast<int64_t *>(out), buckets, id);
  }
  check_hip(hipGetLastError(), "bucket init");
}

void bucket_rows(Ctx &c, int op, int32_t type, const int64_t *keys, const void *vals, uint64_t n,
                 uint64_t buckets, void *out, uint64_t *d_bad) {
  fill_u64(c, d_bad, 1, ~uint64_t(0));
  if (n == 0) return;
  const uint32_t g = stream_grid(n, kBlock * 8, 2048);
  auto *o = static_cast<unsigned long long *>(out);
  const auto *v = static_cast<const int64_t *>(vals);
  if (op <= 1) {
    const bool lds = buckets <= kLdsBuckets;
    if (op == 0) {
      if (lds) k_bucket_add<0, true><<<g, kBlock, 0, c.stream>>>(keys, v, n, buckets, o, d_bad);
      else k_bucket_add<0, false><<<g, kBlock, 0, c.stream>>>(keys, v, n, buckets, o, d_bad);
    } else {
      if (lds) k_bucket_add<1, true><<<g, kBlock, 0, c.stream>>>(keys, v, n, buckets, o, d_bad);
      else k_bucket_add<1, false><<<g, kBlock, 0, c.stream>>>(keys, v, n, buckets, o, d_bad);
    }
  } else if (type == VK_F64) {
    auto *vf = static_cast<const double *>(vals);
    auto *of = static_cast<double *>(out);
    if (op == 2) k_bucket_ext<double, false><<<g, kBlock, 0, c.stream>>>(keys, vf, n, buckets, of, d_bad);
    else k_bucket_ext<double, true><<<g, kBlock, 0, c.stream>>>(keys, vf, n, buckets, of, d_bad);
  } else {
    auto *oi = static_cast<int64_t *>(out);
    if (op == 2) k_bucket_ext<int64_t, false><<<g, kBlock, 0, c.stream>>>(keys, v, n, buckets, oi, d_bad);
    else k_bucket_ext<int64_t, true><<<g, kBlock, 0, c.stream>>>(keys, v, n, buckets, oi, d_bad);
  }
  check_hip(hipGetLastError(), "bucket launch");
}

void check_keys(Ctx &c, const int64_t *keys, uint64_t n, uint64_t buckets, uint64_t *d_bad) {
  fill_u64(c, d_bad, 1, ~uint64_t(0));
  if (n == 0) return;
  k_check_keys<<<stream_grid(n, kBlock * 8, 2048), kBlock, 0, c.stream>>>(keys, n, buckets, d_bad);
  check_hip(hipGetLastError(), "check keys");
}

// The sort checks the keys of rows 0..b before anything folds.
namespace {

// The chains of rows 0..b (bucket, i % 16): a stable sort of (chain key, value) pairs puts each
// chain's values together in index order, and a bucket's chains together in stripe order.
// Below kBucketFoldMin buckets every chain key's first row comes from a binary search of the
// sorted keys (k_chain_bounds, no pass over the rows), and the chains fold into the stripes,
// which then combine per bucket: one lane per chain (k_chain_fold_lanes) unless there are
// fewer than kLaneFoldChains
// chains averaging 256 rows or more, which fold one wave per chain (every lane of a wave on
// one chain, for few long chains that need all the parallel hardware they can get). From
// kBucketFoldMin buckets on, the buckets alone give the parallelism: one lane per bucket
// folds straight into out (k_bucket_fold, no stripe buffer of 16 x buckets values).
// Returns true when out holds the combined buckets.
// From kClassFoldMin to kClassFoldMax buckets (chain keys of two or three bytes) the sort
// leaves the top digit out and the classes fold (k_class_fold) into the stripes, with a
// stripe buffer of its own from kBucketFoldMin buckets on; the forms above serve the rest.
constexpr uint64_t kLaneFoldChains = 4096;
constexpr uint64_t kBucketFoldMin = uint64_t(1) << 16;
constexpr uint64_t kClassFoldMin = 17;
constexpr uint64_t kClassFoldMax = uint64_t(1) << 20;
constexpr uint64_t kClassWideRows = 16384;  // rows per class from which a class takes 1024 threads

template <typename K>
bool fold_chains(Ctx &c, const int64_t *keys, const double *vals, uint64_t b, uint64_t buckets,
                 double *stripes, double *out, uint64_t *bad_row) {
  constexpr bool wide = sizeof(K) == 8;
  K *sk = static_cast<K *>(tmp_alloc(c, b * sizeof(K)));
  double *sv = static_cast<double *>(tmp_alloc(c, b * 8));
  const bool class_form = !wide && buckets >= kClassFoldMin && buckets <= kClassFoldMax;
  int shift = 0;
  // the sort computes each row's chain key from its bucket key as it reads it, and checks the
  // bucket keys in its counting pass
  sort_chain_values(c, wide, keys, reinterpret_cast<const uint64_t *>(vals), b, sk,
                    reinterpret_cast<uint64_t *>(sv), buckets, bad_row, class_form ? &shift : nullptr);
  if (*bad_row != ~uint64_t(0)) {
    tmp_free(c, sk);
    tmp_free(c, sv);
    return false;
  }
  if constexpr (!wide) {
    if (class_form) {
      const uint64_t nclasses = uint64_t(1) << shift;  // shift is 8 or 16 below 2^20 buckets
      const uint64_t nchains = buckets * 16;
      uint64_t *starts = static_cast<uint64_t *>(tmp_alloc(c, (nclasses + 1) * 8));
      k_class_bounds<<<stream_grid(nclasses + 1, kBlock, 8192), kBlock, 0, c.stream>>>(sk, b, nclasses,
                                                                                      uint32_t(nclasses - 1), starts);
      double *st = stripes;
      if (!st) {
        st = static_cast<double *>(tmp_alloc(c, nchains * 8));
        k_fill<double><<<stream_grid(nchains, kBlock, 8192), kBlock, 0, c.stream>>>(st, nchains, 0.0);
      }
      if ((b >> shift) >= kClassWideRows)
        k_class_fold<1024><<<uint32_t(nclasses), 1024, 0, c.stream>>>(starts, shift, sk, sv, st);
      else
        k_class_fold<256><<<uint32_t(nclasses), 256, 0, c.stream>>>(starts, shift, sk, sv, st);
      if (!stripes)
        k_combine_stripes<<<stream_grid(buckets, kBlock, 8192), kBlock, 0, c.stream>>>(st, buckets, out);
      check_hip(hipGetLastError(), "class fold");
      if (!stripes) tmp_free(c, st);
      tmp_free(c, starts);
      tmp_free(c, sk);
      tmp_free(c, sv);
      return !stripes;
    }
  }
  const bool by_bucket = buckets >= kBucketFoldMin;
  if (by_bucket) {
    // the bucket starts, compacted in order
    uint8_t *flag = static_cast<uint8_t *>(tmp_alloc(c, b));
    uint64_t *starts = static_cast<uint64_t *>(tmp_alloc(c, b * 8));
    const uint32_t gq = stream_grid((b + 3) / 4, kBlock * 4, 8192);
    k_key_starts<4, K><<<gq, kBlock, 0, c.stream>>>(sk, b, flag);
    SelectSpec ss{2, VK_I64, 0, 0, 0};
    uint64_t *d_cnt = c.scratch.slots + 20;
    select_rows(c, ss, nullptr, flag, b, starts, d_cnt);
    uint64_t nstarts = 0;
    read_back(c, d_cnt, 8, &nstarts);
    k_fill<double><<<stream_grid(buckets, kBlock, 8192), kBlock, 0, c.stream>>>(out, buckets, 0.0);
    k_bucket_fold<K><<<stream_grid(nstarts, kWave, 65535), kWave, 0, c.stream>>>(starts, nstarts, sk, sv, b,
                                                                                  out);
    check_hip(hipGetLastError(), "bucket fold");
    tmp_free(c, flag);
    tmp_free(c, starts);
  } else {
    // every chain key's first row (at most 16 x 64K chains), then the chains fold into the
    // stripes
    const uint64_t nchains = buckets * 16;
    uint64_t *starts = static_cast<uint64_t *>(tmp_alloc(c, nchains * 8));
    k_chain_bounds<K><<<stream_grid(nchains, kBlock, 8192), kBlock, 0, c.stream>>>(sk, b, nchains, starts);
    if (nchains >= kLaneFoldChains || b / nchains < 256)
      k_chain_fold_lanes<<<stream_grid(nchains, kWave, 16384), kWave, 0, c.stream>>>(starts, nchains, sv, b, stripes);
    else
      k_chain_fold_wave<<<stream_grid(nchains, kFoldWaves, 16384), kFoldWaves * kWave, 0, c.stream>>>(
          starts, nchains, sv, b, stripes);
    check_hip(hipGetLastError(), "chain fold");
    tmp_free(c, starts);
  }
  tmp_free(c, sk);
  tmp_free(c, sv);
  return by_bucket;
}

}  // namespace

// ------------------------------------------------------------------ batched form
// Items: a = keys (i64), b = values (f64), n rows, out = buckets f64, aux2 = buckets. The
// chain keys of all items are sorted with their values in one batched sort (each item in its
// own rows, from aux), then every chain is folded, and each item combines its stripes and
// folds its tail rows.
namespace {

constexpr uint64_t kBTile = 4096;

__global__ __launch_bounds__(kBlock) void k_check_keys_batch(const BatchItem *items,
                                                             const uint32_t *tile_item, uint64_t *bad) {
  const uint32_t it = tile_item[blockIdx.x];
  const BatchItem &I = items[it];
  const uint64_t t0 = (blockIdx.x - I.tile0) * kBTile, t1 = t0 + kBTile < I.n ? t0 + kBTile : I.n;
  const int64_t *keys = static_cast<const int64_t *>(I.a);
  uint64_t my_bad = ~uint64_t(0);
  for (uint64_t r = t0 + threadIdx.x; r < t1; r += kBlock)
    if (static_cast<uint64_t>(keys[r]) >= I.aux2 && r < my_bad) my_bad = r;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const uint64_t o = __shfl_xor(my_bad, m);
    my_bad = o < my_bad ? o : my_bad;
  }
  if (lane_id() == 0 && my_bad != ~uint64_t(0)) atomicMin(reinterpret_cast<unsigned long long *>(&bad[it]), my_bad);
}

// One lane per chain key of an item (items here: n = the item's chains, aux = its first row in
// the sorted buffers, aux2 = its rows, aux3 = its first stripe): the lane finds its chain's rows
// by binary search of the sorted chain keys and folds them from +0.0 in row order (the sort is
// stable), eight loads in flight; an empty chain leaves its stripe at 0.0.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_chain_fold_keys_batch(const BatchItem *items, const uint32_t *tile_item,
                                                                  const K *sk, const double *sv, double *stripes) {
  const BatchItem &I = items[tile_item[blockIdx.x]];
  const uint64_t c = (blockIdx.x - I.tile0) * kBlock + threadIdx.x;  // the chain key
  if (c >= I.n) return;
  const K *k_ = sk + I.aux;
  const double *v_ = sv + I.aux;
  const uint64_t rows = I.aux2;
  auto lower = [&](uint64_t key) {
    uint64_t lo = 0, hi = rows;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) / 2;
      if (static_cast<uint64_t>(k_[mid]) < key) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const uint64_t s = lower(c), e = lower(c + 1);
  if (s == e) return;
  double a = 0.0;
  uint64_t q = s;
  for (; q + 8 <= e; q += 8) {
    double u[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) u[j] = v_[q + j];
#pragma unroll
    for (int j = 0; j < 8; ++j) a = a + u[j];
  }
  for (; q < e; ++q) a = a + v_[q];
  stripes[I.aux3 + c] = a;
}

// One block per item: the pairwise combine of every bucket's 16 stripes (finish_lanes,
// reduce.rs:75-91), then thread 0 folds the item's tail rows in order.
__global__ __launch_bounds__(kBlock) void k_combine_batch(const BatchItem *items, const double *stripes) {
  const BatchItem &I = items[blockIdx.x];
  const double *st = stripes + I.aux3;
  double *out = static_cast<double *>(I.out);
  for (uint64_t k = threadIdx.x; k < I.aux2; k += kBlock) {
    double a[16];
    for (int j = 0; j < 16; ++j) a[j] = st[k * 16 + j];
    for (int step = 8; step > 0; step /= 2)
      for (int j = 0; j < step; ++j) a[j] = a[j] + a[j + step];
    out[k] = a[0];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int64_t *keys = static_cast<const int64_t *>(I.a);
    const double *vals = static_cast<const double *>(I.b);
    for (uint64_t r = I.n - I.n % 16; r < I.n; ++r) out[keys[r]] = out[keys[r]] + vals[r];
  }
}

}  // namespace

void check_keys_batch(Ctx &c, std::vector<BatchItem> &items, uint64_t *d_bad) {
  const uint64_t nitems = items.size();
  fill_u64(c, d_bad, nitems, ~uint64_t(0));
  BatchPlan p = batch_plan(c, items, kBTile);
  if (p.ntiles) k_check_keys_batch<<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item, d_bad);
  check_hip(hipGetLastError(), "check keys batch");
  batch_plan_free(c, p);
}

void reduce_into_add_f64_batch(Ctx &c, std::vector<BatchItem> &items) {
  const uint64_t nitems = items.size();
  if (nitems == 0) return;
  // item rows and stripe offsets
  uint64_t stripes_n = 0, maxb = 0;
  std::vector<BatchItem> whole(items);  // whole blocks only, for the chains
  for (size_t i = 0; i < nitems; ++i) {
    items[i].aux3 = stripes_n;
    stripes_n += items[i].aux2 * 16;
    maxb = items[i].aux2 > maxb ? items[i].aux2 : maxb;
    whole[i] = items[i];
    whole[i].n = items[i].n - items[i].n % 16;
  }
  const bool wide = maxb * 16 > 0xFFFFFFFFull;
  double *stripes = static_cast<double *>(tmp_alloc(c, (stripes_n ? stripes_n : 1) * 8));
  k_fill<double><<<stream_grid(stripes_n, kBlock, 8192), kBlock, 0, c.stream>>>(stripes, stripes_n, 0.0);
  uint64_t total = 0;
  for (auto &it : whole) {
    it.aux = total;
    total += it.n;
  }
  if (total > 0) {
    const size_t kw = wide ? 8 : 4;
    void *sk = tmp_alloc(c, total * kw);
    double *sv = static_cast<double *>(tmp_alloc(c, total * 8));
    // the sort reads each item's bucket keys at a (chain keys computed as it reads them) and
    // its values at b
    sort_chain_values_batch(c, wide, whole, sk, reinterpret_cast<uint64_t *>(sv));
    // one lane per chain key of every item
    std::vector<BatchItem> chains(whole.size());
    for (size_t i = 0; i < whole.size(); ++i) {
      chains[i] = BatchItem{};
      chains[i].n = whole[i].n ? items[i].aux2 * 16 : 0;
      chains[i].aux = whole[i].aux;
      chains[i].aux2 = whole[i].n;
      chains[i].aux3 = items[i].aux3;
    }
    BatchPlan p = batch_plan(c, chains, kBlock);
    if (p.ntiles) {
      if (wide)
        k_chain_fold_keys_batch<uint64_t><<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item,
                                                                             static_cast<const uint64_t *>(sk), sv, stripes);
      else
        k_chain_fold_keys_batch<uint32_t><<<p.ntiles, kBlock, 0, c.stream>>>(p.items, p.tile_item,
                                                                             static_cast<const uint32_t *>(sk), sv, stripes);
    }
    check_hip(hipGetLastError(), "chain fold batch");
    batch_plan_free(c, p);
    tmp_free(c, sk);
    tmp_free(c, sv);
  }
  BatchPlan q = batch_plan(c, items, ~uint64_t(0) >> 1);  // just the item table
  k_combine_batch<<<uint32_t(nitems), kBlock, 0, c.stream>>>(q.items, stripes);
  check_hip(hipGetLastError(), "combine batch");
  batch_plan_free(c, q);
  tmp_free(c, stripes);
}

uint64_t reduce_into_add_f64(Ctx &c, const int64_t *keys, const double *vals, uint64_t n,
                             uint64_t buckets, double *out) {
  if (buckets == 0) return n == 0 ? ~uint64_t(0) : 0;  // every key is outside an empty range
  uint64_t bad = ~uint64_t(0);
  const uint64_t b = n - n % 16;
  if (n > b) {
    // the tail rows (after the last whole block of 16) are not sorted: checked on their own
    uint64_t *d_bad = c.scratch.slots + 25;
    check_keys(c, keys + b, n - b, buckets, d_bad);
    read_back(c, d_bad, 8, &bad);
    if (bad != ~uint64_t(0)) bad += b;
  }
  if (bad != ~uint64_t(0) && b == 0) return bad;
  // the stripe buffer (16 values per bucket) serves fewer than kBucketFoldMin buckets
  const bool stripe_form = buckets < kBucketFoldMin;
  double *stripes = nullptr;
  if (stripe_form) {
    stripes = static_cast<double *>(tmp_alloc(c, buckets * 16 * 8));
    k_fill<double><<<stream_grid(buckets * 16, kBlock, 4096), kBlock, 0, c.stream>>>(stripes, buckets * 16, 0.0);
  }
  bool combined = false;
  if (b > 0) {
    uint64_t bad_sorted = ~uint64_t(0);
    if (buckets * 16 <= 0xFFFFFFFFull)
      combined = fold_chains<uint32_t>(c, keys, vals, b, buckets, stripes, out, &bad_sorted);
    else
      combined = fold_chains<uint64_t>(c, keys, vals, b, buckets, stripes, out, &bad_sorted);
    if (bad_sorted != ~uint64_t(0)) bad = bad_sorted;  // rows below b come first
  }
  if (bad != ~uint64_t(0)) {
    if (stripes) tmp_free(c, stripes);
    return bad;
  }
  if (stripe_form)
    k_combine_stripes<<<stream_grid(buckets, kBlock, 8192), kBlock, 0, c.stream>>>(stripes, buckets, out);
  else if (!combined)
    k_fill<double><<<stream_grid(buckets, kBlock, 8192), kBlock, 0, c.stream>>>(out, buckets, 0.0);
  if (n > b) k_fold_tail<<<1, 64, 0, c.stream>>>(keys, vals, b, n, out);
  check_hip(hipGetLastError(), "f64 reduce_into");
  if (stripes) tmp_free(c, stripes);
  return bad;
}


// ------------------------------------------- reduce_into_add_f64 inside batch plans
// Keys are the group ids of a group_id statement of the same plan, and the bucket count of item
// i is that statement's group count, on the device. The (chain key, value) sort of each item's
// whole blocks runs as a plan sort (no host read); then one lane per chain key of every item
// folds its chain in row order from +0.0, the 16 lanes of a bucket combine pairwise
// (finish_lanes, reduce.rs:75-91), and the bucket's first lane adds the item's tail rows of that
// bucket in row order: the order of reduce_into_add_f64 (and of the stripes in
// reduce_into_add_f64_batch). Tiles of 16 buckets are planned on the device from the group
// counts.

namespace {

// Fold items: n = the bucket count (written by the device planner), a keys (i64), b values,
// out the output, aux the item's first row in the sorted buffers, aux2 its whole rows, aux3 its
// rows. A block takes tiles tile, tile + gridDim.x, ... (the grid is capped: a launch may not
// exceed 2^32 work-items).
template <typename K>
__device__ __forceinline__ void fold_combine_tile(const BatchItem &I, uint32_t tile, const K *sk, const double *sv) {
  const uint64_t ck = (tile - I.tile0) * kBlock + threadIdx.x;  // the lane's chain key
  const uint64_t bucket = ck / 16;
  const int j = static_cast<int>(ck & 15);
  const bool valid = bucket < I.n;
  double a = 0.0;
  if (valid) {
    const K *k_ = sk + I.aux;
    const double *v_ = sv + I.aux;
    const uint64_t rows = I.aux2;
    auto lower = [&](uint64_t key) {
      uint64_t lo = 0, hi = rows;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (static_cast<uint64_t>(k_[mid]) < key) lo = mid + 1;
        else hi = mid;
      }
      return lo;
    };
    const uint64_t s0 = lower(ck), e = lower(ck + 1);
    uint64_t q = s0;
    for (; q + 8 <= e; q += 8) {
      double u[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) u[m] = v_[q + m];
#pragma unroll
      for (int m = 0; m < 8; ++m) a = a + u[m];
    }
    for (; q < e; ++q) a = a + v_[q];
  }
  // finish_lanes over the bucket's 16 lanes: acc[j] += acc[j + step] for j < step
#pragma unroll
  for (int step = 8; step > 0; step /= 2) {
    const double o = __shfl_down(a, step, 16);
    if (j < step) a = a + o;
  }
  if (valid && j == 0) {
    const int64_t *keys = static_cast<const int64_t *>(I.a);
    const double *vals = static_cast<const double *>(I.b);
    for (uint64_t r = I.aux2; r < I.aux3; ++r)
      if (static_cast<uint64_t>(keys[r]) == bucket) a = a + vals[r];
    static_cast<double *>(I.out)[bucket] = a;
  }
}

constexpr uint32_t kFoldGrid = 65536;

template <typename K>
__global__ __launch_bounds__(kBlock) void k_fold_combine_plan(const BatchItem *items, const uint32_t *tile_item,
                                                              const uint32_t *ntiles, const K *sk,
                                                              const double *sv) {
  const uint32_t nt = *ntiles;
  for (uint32_t tile = blockIdx.x; tile < nt; tile += gridDim.x)  // uniform over the block
    fold_combine_tile(items[tile_item[tile]], tile, sk, sv);
}

}  // namespace

PlanFold *plan_fold_new(Ctx &c, std::vector<BatchItem> &items, const uint64_t *groups) {
  auto *f = new PlanFold{};
  const uint64_t ni = items.size();
  f->nitems = ni;
  f->groups = groups;
  uint64_t whole_total = 0, maxn = 0, max_tiles = 0;
  std::vector<BatchItem> sort_items(ni), fold(ni);
  for (size_t i = 0; i < ni; ++i) {
    const BatchItem &it = items[i];
    maxn = it.n > maxn ? it.n : maxn;
    sort_items[i] = BatchItem{};
    sort_items[i].a = it.a;
    sort_items[i].b = it.b;
    sort_items[i].n = it.n - it.n % 16;
    fold[i] = BatchItem{};
    fold[i].a = it.a;
    fold[i].b = it.b;
    fold[i].out = it.out;
    fold[i].n = it.n;  // the bucket count bound; the device planner writes the count
    fold[i].aux = whole_total;
    fold[i].aux2 = it.n - it.n % 16;
    fold[i].aux3 = it.n;
    whole_total += it.n - it.n % 16;
    max_tiles += (it.n + 15) / 16;
  }
  if (max_tiles > 0x7FFFFFFFull) usage_error("plan statement too large");
  f->wide = maxn * 16 > 0xFFFFFFFFull;
  f->sk = tmp_alloc(c, (whole_total ? whole_total : 1) * (f->wide ? 8 : 4));
  f->sv = static_cast<double *>(tmp_alloc(c, (whole_total ? whole_total : 1) * 8));
  f->sort = plan_sort_new(c, chain_key_type(), false, f->wide, sort_items, f->sk, reinterpret_cast<uint64_t *>(f->sv));
  DevPlan &dp = f->dp;
  dp.items = static_cast<BatchItem *>(tmp_alloc(c, (ni ? ni : 1) * sizeof(BatchItem)));
  if (ni)
    check_hip(hipMemcpyAsync(dp.items, fold.data(), ni * sizeof(BatchItem), hipMemcpyHostToDevice, c.stream),
              "fold items");
  check_hip(hipStreamSynchronize(c.stream), "fold items");
  dp.max_tiles = static_cast<uint32_t>(max_tiles);
  dp.tile_item = static_cast<uint32_t *>(tmp_alloc(c, (max_tiles ? max_tiles : 1) * 4));
  dp.ntiles = static_cast<uint32_t *>(tmp_alloc(c, 8));
  return f;
}

void plan_fold_run(Ctx &c, PlanFold &f, const uint32_t *epoch) {
  plan_sort_run(c, *f.sort, epoch);
  plan_tiles_dev(c, f.dp, static_cast<uint32_t>(f.nitems), f.groups, 1, kBlock / 16);
  if (f.dp.max_tiles) {
    const uint32_t grid = f.dp.max_tiles < kFoldGrid ? f.dp.max_tiles : kFoldGrid;
    if (f.wide)
      k_fold_combine_plan<uint64_t><<<grid, kBlock, 0, c.stream>>>(f.dp.items, f.dp.tile_item, f.dp.ntiles,
                                                                  static_cast<const uint64_t *>(f.sk), f.sv);
    else
      k_fold_combine_plan<uint32_t><<<grid, kBlock, 0, c.stream>>>(f.dp.items, f.dp.tile_item, f.dp.ntiles,
                                                                  static_cast<const uint32_t *>(f.sk), f.sv);
  }
  check_hip(hipGetLastError(), "plan reduce_into_add_f64");
}

void plan_fold_free(Ctx &c, PlanFold *f) {
  if (!f) return;
  plan_sort_free(c, f->sort);
  tmp_free(c, f->sk);
  tmp_free(c, f->sv);
  tmp_free(c, f->dp.items);
  tmp_free(c, f->dp.tile_item);
  tmp_free(c, f->dp.ntiles);
  delete f;
}

}  // namespace vk
