// Streaming reductions whose result does not depend on the order rows combine in:
// counts (count_<cmp>_<t>, fused.rs:142-179; count, reduce.rs:527-529), wrapping integer
// sums that widen (IntSum, reduce.rs:325-377), IEEE minimum/maximum (Extreme,
// reduce.rs:379-412), first-index arg extremes (ArgExtreme, reduce.rs:414-468) and bit
// folds (reduce.rs:531-575).
//
// One launch reads its rows once with 128-bit loads, one batch of loads per block, reduces
// per thread, per wave (shuffles) and per block (LDS) into one partial per block; a second,
// one-block launch folds the partials into the reducer's device state.
// HBM-bound: algorithmic bytes = n x width.
#include "chain.h"
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

// ---------------------------------------------------------------- accumulators

struct SumAcc {
  uint64_t s;
};
struct ExtAcc {
  uint64_t bits;  // value bits
  uint64_t has;
};
struct ArgAcc {
  uint64_t bits;  // best value bits
  int64_t idx;    // best index, INT64_MAX when none
  int64_t nan;    // first NaN index, INT64_MAX when none
};

constexpr int64_t kNone = INT64_MAX;

// Rows widen to u64 for sums: i32 sign-extends, u8 zero-extends (reduce.rs:375-377).
template <typename T>
__device__ __forceinline__ uint64_t widen(T v) {
  return static_cast<uint64_t>(static_cast<int64_t>(v));
}
template <>
__device__ __forceinline__ uint64_t widen<uint8_t>(uint8_t v) { return v; }
template <>
__device__ __forceinline__ uint64_t widen<uint64_t>(uint64_t v) { return v; }

// Ordered::less (reduce.rs:183-199): -0.0 ranks below +0.0; ints compare directly.
template <typename T>
__device__ __forceinline__ bool ord_less(T a, T b) { return a < b; }
template <>
__device__ __forceinline__ bool ord_less<double>(double a, double b) {
  return a < b || (a == b && __signbit(a) && !__signbit(b));
}
template <>
__device__ __forceinline__ bool ord_less<float>(float a, float b) {
  return a < b || (a == b && __signbit(a) && !__signbit(b));
}
// An integer image that orders like Ordered::less (reduce.rs:183-199) for non-NaN values:
// floats map negative values to their complemented magnitude (so -0.0 < +0.0), integers to
// themselves (u64 with the sign bit flipped).
template <typename T>
__device__ __forceinline__ int64_t ord_key(T v) {
  if constexpr (std::is_same<T, double>::value) {
    const int64_t b = static_cast<int64_t>(to_bits(v));
    return b < 0 ? (b ^ INT64_MAX) : b;
  } else if constexpr (std::is_same<T, float>::value) {
    const int32_t b = static_cast<int32_t>(to_bits(v));
    return b < 0 ? int64_t(b ^ INT32_MAX) : int64_t(b);
  } else if constexpr (std::is_same<T, uint64_t>::value) {
    return static_cast<int64_t>(v ^ 0x8000000000000000ull);
  } else {
    return static_cast<int64_t>(v);
  }
}

template <typename T>
__device__ __forceinline__ bool is_nan(T) { return false; }
template <>
__device__ __forceinline__ bool is_nan<double>(double v) { return v != v; }
template <>
__device__ __forceinline__ bool is_nan<float>(float v) { return v != v; }

// -------------------------------------------------------------- the operators

template <typename T>
struct OpSum {
  using Acc = SumAcc;
  __device__ Acc identity() const { return {0}; }
  __device__ void add(Acc &a, T v, uint64_t) const { a.s += widen(v); }
  __device__ void merge(Acc &a, const Acc &b) const { a.s += b.s; }
};

// IntProduct (reduce.rs:358-366): the wrapping product in the element type. The low bytes of
// a product depend only on the low bytes of its factors, so the fold multiplies widened rows
// modulo 2^64 and the caller keeps the element type's bytes.
template <typename T>
struct OpProd {
  using Acc = SumAcc;
  __device__ Acc identity() const { return {1}; }
  __device__ void add(Acc &a, T v, uint64_t) const { a.s *= widen(v); }
  __device__ void merge(Acc &a, const Acc &b) const { a.s *= b.s; }
};

// Num::mul over integers (elementwise.rs:380-397): the wrapping product in the element type.
template <typename T>
__device__ __forceinline__ T wrap_mul(T a, T b) {
  using U = typename std::make_unsigned<T>::type;
  return static_cast<T>(static_cast<U>(static_cast<U>(a) * static_cast<U>(b)));
}

// sumsq and condsum over integers (SumSq, CondSum fused.rs:108-140, 181-210, with IntSum): the
// row value in the element type, widened, added with wrapping.
template <typename T>
struct OpSumSq {
  using Acc = SumAcc;
  __device__ Acc identity() const { return {0}; }
  __device__ void add(Acc &a, T v, uint64_t) const { a.s += widen(wrap_mul(v, v)); }
  __device__ void merge(Acc &a, const Acc &b) const { a.s += b.s; }
};

template <typename T>
struct OpCondSum {
  using Acc = SumAcc;
  uint32_t keepset;  // the comparison outcomes (outcome()) that keep the row
  T c, e;
  __device__ Acc identity() const { return {0}; }
  __device__ void add(Acc &a, T v, uint64_t) const { a.s += widen((keepset & outcome(v, c)) != 0 ? v : e); }
  __device__ void merge(Acc &a, const Acc &b) const { a.s += b.s; }
};

template <typename T>
struct OpCount {
  using Acc = SumAcc;
  uint32_t keepset;  // the comparison outcomes counted (outcome())
  T c;
  __device__ Acc identity() const { return {0}; }
  __device__ void add(Acc &a, T v, uint64_t) const { a.s += (keepset & outcome(v, c)) != 0 ? 1 : 0; }
  __device__ void merge(Acc &a, const Acc &b) const { a.s += b.s; }
};

struct OpCountBool {
  using Acc = SumAcc;
  __device__ Acc identity() const { return {0}; }
  __device__ void add(Acc &a, uint8_t v, uint64_t) const { a.s += v != 0; }
  __device__ void merge(Acc &a, const Acc &b) const { a.s += b.s; }
};

template <typename T>
struct OpBits {
  using Acc = SumAcc;
  int op;  // 0 and, 1 or, 2 xor
  __device__ Acc identity() const {
    return {op == 0 ? (sizeof(T) == 8 ? ~uint64_t(0) : (uint64_t(1) << (8 * sizeof(T))) - 1) : 0};
  }
  __device__ void add(Acc &a, T v, uint64_t) const {
    uint64_t b = to_bits(v);
    a.s = op == 0 ? (a.s & b) : op == 1 ? (a.s | b) : (a.s ^ b);
  }
  __device__ void merge(Acc &a, const Acc &b) const {
    a.s = op == 0 ? (a.s & b.s) : op == 1 ? (a.s | b.s) : (a.s ^ b.s);
  }
};

template <typename T>
struct OpExtreme {
  using Acc = ExtAcc;
  bool want_max;
  __device__ Acc identity() const { return {0, 0}; }
  __device__ void add(Acc &a, T v, uint64_t) const {
    if (!a.has) {
      a.bits = to_bits(v);
      a.has = 1;
    } else {
      T cur = from_bits<T>(a.bits);
      a.bits = to_bits(want_max ? max2(cur, v) : min2(cur, v));
    }
  }
  __device__ void merge(Acc &a, const Acc &b) const {
    if (!b.has) return;
    add(a, from_bits<T>(b.bits), 0);
  }
};

template <typename T>
struct OpArg {
  using Acc = ArgAcc;
  bool want_max;
  uint64_t base;
  __device__ Acc identity() const { return {0, kNone, kNone}; }
  // Whether (v, i) beats (w, j): strictly more extreme, or equally ranked and earlier.
  __device__ bool beats(T v, int64_t i, T w, int64_t j) const {
    bool more = want_max ? ord_less(w, v) : ord_less(v, w);
    bool less = want_max ? ord_less(v, w) : ord_less(w, v);
    return more || (!less && i < j);
  }
  __device__ void add(Acc &a, T v, uint64_t row) const {
    int64_t i = static_cast<int64_t>(base + row);
    if (is_nan(v)) {
      if (i < a.nan) a.nan = i;
      return;
    }
    if (a.idx == kNone || beats(v, i, from_bits<T>(a.bits), a.idx)) {
      a.bits = to_bits(v);
      a.idx = i;
    }
  }
  // The same fold for rows arriving in increasing order: an equal value never wins, so a row
  // wins when its ordered image (ord_key) is strictly more extreme, one integer compare; the
  // first NaN is the smallest NaN row.
  __device__ void add_inorder(Acc &a, T v, uint64_t row) const {
    const int64_t i = static_cast<int64_t>(base + row);
    const bool nan = is_nan(v);
    a.nan = nan && i < a.nan ? i : a.nan;
    const int64_t kv = ord_key(v), kb = ord_key(from_bits<T>(a.bits));
    const bool win = !nan && (a.idx == kNone || (want_max ? kv > kb : kv < kb));
    a.bits = win ? to_bits(v) : a.bits;
    a.idx = win ? i : a.idx;
  }
  __device__ void merge(Acc &a, const Acc &b) const {
    if (b.nan < a.nan) a.nan = b.nan;
    if (b.idx == kNone) return;
    if (a.idx == kNone || beats(from_bits<T>(b.bits), b.idx, from_bits<T>(a.bits), a.idx)) {
      a.bits = b.bits;
      a.idx = b.idx;
    }
  }
};

// Rows a thread folds in increasing row order: an operator may offer add_inorder, a cheaper
// fold that relies on it (see OpArg); the others fold as usual.
template <class Op, class Acc, typename T>
__device__ __forceinline__ auto add_inorder(const Op &op, Acc &a, T v, uint64_t row, int)
    -> decltype(op.add_inorder(a, v, row), void()) {
  op.add_inorder(a, v, row);
}
template <class Op, class Acc, typename T>
__device__ __forceinline__ void add_inorder(const Op &op, Acc &a, T v, uint64_t row, long) {
  op.add(a, v, row);
}

// A fused region's root: the inner operator over the chain's values of the rows (ChainSpec).
// The chain runs in registers on each loaded row, so no intermediate array is written.
template <class Inner, typename T>
struct OpChain {
  using Acc = typename Inner::Acc;
  Inner in;
  ChainSpec ch;
  __device__ Acc identity() const { return in.identity(); }
  __device__ void add(Acc &a, T v, uint64_t row) const { in.add(a, chain_eval<T>(ch, v, v), row); }
  __device__ void add_inorder(Acc &a, T v, uint64_t row) const {
    vk::add_inorder(in, a, chain_eval<T>(ch, v, v), row, 0);
  }
  __device__ void merge(Acc &a, const Acc &b) const { in.merge(a, b); }
};

// ------------------------------------------------------------------- the kernel

template <class Op>
__device__ __forceinline__ typename Op::Acc block_merge(const Op &op, typename Op::Acc a) {
  using Acc = typename Op::Acc;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) op.merge(a, shfl_xor_pod(a, m));
  __shared__ Acc s_w[kBlock / kWave];
  const int w = threadIdx.x / kWave, l = lane_id();
  if (l == 0) s_w[w] = a;
  __syncthreads();
  // the four wave results meet in lanes 0..3 of wave 0 and merge with two more butterflies
  if (w == 0) {
    Acc b = l < kBlock / kWave ? s_w[l] : op.identity();
    op.merge(b, shfl_xor_pod(b, 1));
    op.merge(b, shfl_xor_pod(b, 2));
    a = b;
  }
  return a;  // valid in wave 0
}

// One batch of loads per block: a block of 1024 threads reads 8 x 16 bytes per thread (128 KB,
// all loads issued before the first is used) and folds them into one partial; the hardware
// starts a new block as each one ends, which keeps more loads in flight than a loop inside
// long-lived blocks (measured 6.4 against 5.5 TB/s for a count). Block 0 also folds the rows
// before the first 16-byte boundary and after the last whole 16 bytes. A one-block kernel merges
// the partials into the reducer state; the kernel boundary orders the partials before the merge.
constexpr int kRedThreads = 1024;
// 16-byte loads per thread: 8 for 8-byte rows, fewer for narrower rows so that a thread folds
// at most 32 rows (wider batches of narrow rows spill their unpacked rows)
template <typename T>
constexpr int red_loads() { return sizeof(T) >= 8 ? 8 : sizeof(T) == 4 ? 4 : 2; }
template <typename T>
constexpr uint64_t red_vecs() { return uint64_t(kRedThreads) * red_loads<T>(); }  // per block

template <class Op>
__device__ __forceinline__ typename Op::Acc block_merge_big(const Op &op, typename Op::Acc a) {
  using Acc = typename Op::Acc;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) op.merge(a, shfl_xor_pod(a, m));
  __shared__ Acc s_w[kRedThreads / kWave];
  const int w = threadIdx.x / kWave, l = lane_id();
  if (l == 0) s_w[w] = a;
  __syncthreads();
  // the 16 wave results meet in lanes 0..15 of wave 0
  if (w == 0) {
    Acc b = l < kRedThreads / kWave ? s_w[l] : op.identity();
#pragma unroll
    for (int m = 1; m < kRedThreads / kWave; m <<= 1) op.merge(b, shfl_xor_pod(b, m));
    a = b;
  }
  return a;  // valid in wave 0
}

template <typename T, class Op>
__global__ __launch_bounds__(kRedThreads) void k_reduce(Op op, const T *__restrict__ x, uint64_t n,
                                                        uint64_t head, typename Op::Acc *partials) {
  using Acc = typename Op::Acc;
  constexpr int V = 16 / sizeof(T);
  const int t = threadIdx.x;
  constexpr int kRedLoads = red_loads<T>();
  constexpr uint64_t kRedVecs = red_vecs<T>();
  const uint64_t nv = (n - head) / V;
  const Vec16<T> *xv = reinterpret_cast<const Vec16<T> *>(x + head);
  const uint64_t base = blockIdx.x * kRedVecs;
  Acc a = op.identity();
  Vec16<T> v[kRedLoads];
  if (base + kRedVecs <= nv) {
#pragma unroll
    for (int u = 0; u < kRedLoads; ++u) v[u] = xv[base + uint64_t(u) * kRedThreads + t];
#pragma unroll
    for (int u = 0; u < kRedLoads; ++u) {
      const uint64_t i = base + uint64_t(u) * kRedThreads + t;
#pragma unroll
      for (int k = 0; k < V; ++k) add_inorder(op, a, v[u].e[k], head + i * V + k, 0);
    }
  } else {
#pragma unroll
    for (int u = 0; u < kRedLoads; ++u) {
      const uint64_t i = base + uint64_t(u) * kRedThreads + t;
      if (i < nv) v[u] = xv[i];
    }
#pragma unroll
    for (int u = 0; u < kRedLoads; ++u) {
      const uint64_t i = base + uint64_t(u) * kRedThreads + t;
      if (i < nv) {
#pragma unroll
        for (int k = 0; k < V; ++k) add_inorder(op, a, v[u].e[k], head + i * V + k, 0);
      }
    }
  }
  if (blockIdx.x == 0) {
    // rows before the first 16-byte boundary and after the last: out of this thread's order,
    // so folded apart and merged
    Acc b = op.identity();
    if (uint64_t(t) < head) op.add(b, x[t], t);
    for (uint64_t r = head + nv * V + t; r < n; r += kRedThreads) op.add(b, x[r], r);
    op.merge(a, b);
  }
  a = block_merge_big(op, a);
  if (t == 0) partials[blockIdx.x] = a;
}

template <class Op>
__global__ __launch_bounds__(kRedThreads) void k_reduce_merge(Op op, const typename Op::Acc *partials,
                                                              uint32_t nparts, typename Op::Acc *state) {
  using Acc = typename Op::Acc;
  // thread t merges partials t, t + 1024, t + 2048, ... in that order, eight loads in flight
  constexpr int kAhead = 8;
  Acc b = op.identity();
  for (uint32_t i0 = threadIdx.x; i0 < nparts; i0 += kRedThreads * kAhead) {
    Acc p[kAhead];
#pragma unroll
    for (int k = 0; k < kAhead; ++k) {
      const uint32_t i = i0 + k * kRedThreads;
      p[k] = i < nparts ? partials[i] : op.identity();
    }
#pragma unroll
    for (int k = 0; k < kAhead; ++k)
      if (i0 + k * kRedThreads < nparts) op.merge(b, p[k]);
  }
  b = block_merge_big(op, b);
  if (threadIdx.x == 0) {
    Acc s = *state;
    op.merge(s, b);
    *state = s;
  }
}

template <typename T, class Op>
void launch(Ctx &c, const Op &op, const void *x, uint64_t n, void *state) {
  using Acc = typename Op::Acc;
  uint64_t head = ((16 - (reinterpret_cast<uintptr_t>(x) & 15)) & 15) / sizeof(T);
  if (head > n) head = n;
  const uint64_t nv = (n - head) / (16 / sizeof(T));
  const uint64_t blocks = nv == 0 ? 1 : (nv + red_vecs<T>() - 1) / red_vecs<T>();
  if (blocks > 0xFFFFFFFFull) usage_error("reduction too large");
  const uint32_t grid = static_cast<uint32_t>(blocks);
  Acc *parts = static_cast<Acc *>(partials(c, sizeof(Acc) * grid));
  k_reduce<T, Op><<<grid, kRedThreads, 0, c.stream>>>(op, static_cast<const T *>(x), n, head, parts);
  k_reduce_merge<Op><<<1, kRedThreads, 0, c.stream>>>(op, parts, grid, static_cast<Acc *>(state));
  check_hip(hipGetLastError(), "reduce launch");
}

// Dot over integers (Dot fused.rs:72-106 with IntSum): two inputs, read like k_reduce (16-byte
// loads, one batch per block, block 0 folding the rows outside whole 16-byte vectors) when both
// start at the same offset within 16 bytes, one row per load otherwise.
template <typename T, bool VEC>
__global__ __launch_bounds__(kRedThreads) void k_dot_int(const T *__restrict__ x, const T *__restrict__ y,
                                                         uint64_t n, uint64_t head, SumAcc *partials) {
  constexpr int V = VEC ? 16 / sizeof(T) : 1;
  constexpr int kL = red_loads<T>() / 2 > 0 ? red_loads<T>() / 2 : 1;
  const int t = threadIdx.x;
  const uint64_t h = VEC ? head : 0;
  const uint64_t nv = (n - h) / V;
  const uint64_t base = blockIdx.x * uint64_t(kRedThreads) * kL;
  uint64_t s = 0;
  if constexpr (VEC) {
    const Vec16<T> *xv = reinterpret_cast<const Vec16<T> *>(x + h);
    const Vec16<T> *yv = reinterpret_cast<const Vec16<T> *>(y + h);
    Vec16<T> a[kL], b[kL];
#pragma unroll
    for (int u = 0; u < kL; ++u) {
      const uint64_t i = base + uint64_t(u) * kRedThreads + t;
      if (i < nv) {
        a[u] = xv[i];
        b[u] = yv[i];
      }
    }
#pragma unroll
    for (int u = 0; u < kL; ++u) {
      const uint64_t i = base + uint64_t(u) * kRedThreads + t;
      if (i < nv) {
#pragma unroll
        for (int k = 0; k < V; ++k) s += widen(wrap_mul(a[u].e[k], b[u].e[k]));
      }
    }
  } else {
    T a[kL], b[kL];
#pragma unroll
    for (int u = 0; u < kL; ++u) {
      const uint64_t i = base + uint64_t(u) * kRedThreads + t;
      a[u] = i < nv ? x[i] : T(0);
      b[u] = i < nv ? y[i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < kL; ++u) s += widen(wrap_mul(a[u], b[u]));
  }
  if (VEC && blockIdx.x == 0) {
    if (uint64_t(t) < h) s += widen(wrap_mul(x[t], y[t]));
    for (uint64_t r = h + nv * V + t; r < n; r += kRedThreads) s += widen(wrap_mul(x[r], y[r]));
  }
  SumAcc acc{s};
  acc = block_merge_big(OpSum<T>{}, acc);
  if (t == 0) partials[blockIdx.x] = acc;
}

template <typename T>
void launch_dot(Ctx &c, const T *x, const T *y, uint64_t n, void *state) {
  const uintptr_t ax = reinterpret_cast<uintptr_t>(x) & 15, ay = reinterpret_cast<uintptr_t>(y) & 15;
  const bool vec = ax == ay;
  uint64_t head = ((16 - ax) & 15) / sizeof(T);
  if (head > n) head = n;
  const uint64_t per = uint64_t(kRedThreads) * (red_loads<T>() / 2 > 0 ? red_loads<T>() / 2 : 1);
  const uint64_t units = vec ? (n - head) / (16 / sizeof(T)) : n;
  const uint64_t blocks = units == 0 ? 1 : (units + per - 1) / per;
  if (blocks > 0xFFFFFFFFull) usage_error("reduction too large");
  const uint32_t grid = static_cast<uint32_t>(blocks);
  SumAcc *parts = static_cast<SumAcc *>(partials(c, sizeof(SumAcc) * grid));
  if (vec) k_dot_int<T, true><<<grid, kRedThreads, 0, c.stream>>>(x, y, n, head, parts);
  else k_dot_int<T, false><<<grid, kRedThreads, 0, c.stream>>>(x, y, n, 0, parts);
  k_reduce_merge<OpSum<T>><<<1, kRedThreads, 0, c.stream>>>(OpSum<T>{}, parts, grid, static_cast<SumAcc *>(state));
  check_hip(hipGetLastError(), "dot launch");
}

template <typename T>
OpCount<T> make_count(int cmp, uint64_t cbits) {
  T c;
  __builtin_memcpy(&c, &cbits, sizeof(T));
  return OpCount<T>{keepset_of(cmp), c};
}

__global__ void k_init_state(uint64_t *s, int words, uint64_t w0, uint64_t w1, uint64_t w2) {
  if (threadIdx.x == 0) {
    if (words > 0) s[0] = w0;
    if (words > 1) s[1] = w1;
    if (words > 2) s[2] = w2;
  }
}

}  // namespace

void reduce_state_init(Ctx &c, const ReduceSpec &s, void *state) {
  uint64_t *p = static_cast<uint64_t *>(state);
  switch (s.kind) {
    case ReduceSpec::BITS: {
      const uint32_t w = type_width(s.type);
      uint64_t all = w == 8 ? ~uint64_t(0) : (uint64_t(1) << (8 * w)) - 1;
      if (s.type == VK_BOOL) all = 1;  // the identity of bool `and` is true
      k_init_state<<<1, 64, 0, c.stream>>>(p, 1, s.bitop == 0 ? all : 0, 0, 0);
      break;
    }
    case ReduceSpec::PROD_INT: k_init_state<<<1, 64, 0, c.stream>>>(p, 1, 1, 0, 0); break;
    case ReduceSpec::EXTREME: k_init_state<<<1, 64, 0, c.stream>>>(p, 2, 0, 0, 0); break;
    case ReduceSpec::ARG_EXTREME:
      k_init_state<<<1, 64, 0, c.stream>>>(p, 3, 0, uint64_t(kNone), uint64_t(kNone));
      break;
    default: k_init_state<<<1, 64, 0, c.stream>>>(p, 1, 0, 0, 0); break;
  }
  check_hip(hipGetLastError(), "reduce init");
}

namespace {

// Calls f(op, (T *)nullptr) with the operator and element type of the spec.
template <class F>
void with_op(const ReduceSpec &s, F &&f) {
  switch (s.kind) {
    case ReduceSpec::SUM_INT:
      switch (s.type) {
        case VK_I64: f(OpSum<int64_t>{}, static_cast<int64_t *>(nullptr)); break;
        case VK_U64: f(OpSum<uint64_t>{}, static_cast<uint64_t *>(nullptr)); break;
        case VK_I32: f(OpSum<int32_t>{}, static_cast<int32_t *>(nullptr)); break;
        case VK_U8: f(OpSum<uint8_t>{}, static_cast<uint8_t *>(nullptr)); break;
        default: usage_error("unsupported integer sum type");
      }
      break;
    case ReduceSpec::COUNT_CMP:
      switch (s.type) {
        case VK_F64: f(make_count<double>(s.cmp, s.cbits), static_cast<double *>(nullptr)); break;
        case VK_F32: f(make_count<float>(s.cmp, s.cbits), static_cast<float *>(nullptr)); break;
        case VK_I64: f(make_count<int64_t>(s.cmp, s.cbits), static_cast<int64_t *>(nullptr)); break;
        case VK_U64: f(make_count<uint64_t>(s.cmp, s.cbits), static_cast<uint64_t *>(nullptr)); break;
        case VK_I32: f(make_count<int32_t>(s.cmp, s.cbits), static_cast<int32_t *>(nullptr)); break;
        case VK_U8: f(make_count<uint8_t>(s.cmp, s.cbits), static_cast<uint8_t *>(nullptr)); break;
        default: usage_error("unsupported count type");
      }
      break;
    case ReduceSpec::PROD_INT:
      switch (s.type) {
        case VK_I64: f(OpProd<int64_t>{}, static_cast<int64_t *>(nullptr)); break;
        case VK_U64: f(OpProd<uint64_t>{}, static_cast<uint64_t *>(nullptr)); break;
        case VK_I32: f(OpProd<int32_t>{}, static_cast<int32_t *>(nullptr)); break;
        case VK_U8: f(OpProd<uint8_t>{}, static_cast<uint8_t *>(nullptr)); break;
        default: usage_error("unsupported integer product type");
      }
      break;
    case ReduceSpec::FSUM_INT: {
      auto go = [&](auto *tag) {
        using T = typename std::remove_pointer<decltype(tag)>::type;
        if (s.fkind == 2) {
          f(OpSumSq<T>{}, tag);
        } else if (s.fkind == 3) {
          T cv, ev;
          __builtin_memcpy(&cv, &s.cbits, sizeof(T));
          __builtin_memcpy(&ev, &s.ebits, sizeof(T));
          f(OpCondSum<T>{keepset_of(s.cmp), cv, ev}, tag);
        } else {
          usage_error("an integer dot reads two inputs");
        }
      };
      switch (s.type) {
        case VK_I64: go(static_cast<int64_t *>(nullptr)); break;
        case VK_U64: go(static_cast<uint64_t *>(nullptr)); break;
        case VK_I32: go(static_cast<int32_t *>(nullptr)); break;
        default: usage_error("unsupported fused integer sum type");
      }
      break;
    }
    case ReduceSpec::COUNT_BOOL: f(OpCountBool{}, static_cast<uint8_t *>(nullptr)); break;
    case ReduceSpec::EXTREME:
    case ReduceSpec::ARG_EXTREME: {
      const bool arg = s.kind == ReduceSpec::ARG_EXTREME;
      auto go = [&](auto *tag) {
        using T = typename std::remove_pointer<decltype(tag)>::type;
        if (arg) f(OpArg<T>{s.want_max, s.base_row}, tag);
        else f(OpExtreme<T>{s.want_max}, tag);
      };
      switch (s.type) {
        case VK_F64: go(static_cast<double *>(nullptr)); break;
        case VK_F32: go(static_cast<float *>(nullptr)); break;
        case VK_I64: go(static_cast<int64_t *>(nullptr)); break;
        case VK_U64: go(static_cast<uint64_t *>(nullptr)); break;
        case VK_I32: go(static_cast<int32_t *>(nullptr)); break;
        case VK_U8: case VK_BOOL: go(static_cast<uint8_t *>(nullptr)); break;
        default: usage_error("unsupported element type");
      }
      break;
    }
    case ReduceSpec::BITS:
      switch (s.type) {
        case VK_I64: f(OpBits<int64_t>{s.bitop}, static_cast<int64_t *>(nullptr)); break;
        case VK_U64: f(OpBits<uint64_t>{s.bitop}, static_cast<uint64_t *>(nullptr)); break;
        case VK_I32: f(OpBits<int32_t>{s.bitop}, static_cast<int32_t *>(nullptr)); break;
        case VK_U8: case VK_BOOL: f(OpBits<uint8_t>{s.bitop}, static_cast<uint8_t *>(nullptr)); break;
        default: usage_error("unsupported bit fold type");
      }
      break;
  }
}

// ------------------------------------------------------------------ batched form
// One launch folds every tile of every item into a per-tile partial, a second merges each
// item's partials. Every operator here is commutative and associative (the arg operators
// break ties by index), so the merge order does not matter.
constexpr uint64_t kBatchTile = 8192;

// TILE rows per block (TILE / 256 rows per thread; batched plans pick it by autotuning). With
// tiles planned on the device, blocks past dev_ntiles exit.
template <typename T, class Op, uint64_t TILE = kBatchTile>
__global__ __launch_bounds__(kBlock) void k_reduce_tiles(Op op, const BatchItem *items,
                                                         const uint32_t *tile_item,
                                                         typename Op::Acc *partials,
                                                         const uint32_t *dev_ntiles = nullptr) {
  using Acc = typename Op::Acc;
  if (dev_ntiles && blockIdx.x >= *dev_ntiles) return;
  const BatchItem &I = items[tile_item[blockIdx.x]];
  const T *x = static_cast<const T *>(I.a);
  const uint64_t t0 = (blockIdx.x - I.tile0) * TILE;
  const uint64_t t1 = t0 + TILE < I.n ? t0 + TILE : I.n;
  Acc a = op.identity();
  for (uint64_t r = t0 + threadIdx.x; r < t1; r += kBlock) op.add(a, x[r], r);
  a = block_merge(op, a);
  if (threadIdx.x == 0) partials[blockIdx.x] = a;
}

template <class Op, uint64_t TILE = kBatchTile>
__global__ void k_reduce_items(Op op, const BatchItem *items, uint64_t nitems,
                               const typename Op::Acc *partials, typename Op::Acc *states) {
  const uint64_t it = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  if (it >= nitems) return;
  const BatchItem &I = items[it];
  const uint64_t tiles = (I.n + TILE - 1) / TILE;
  typename Op::Acc s = op.identity();
  for (uint64_t k = 0; k < tiles; ++k) op.merge(s, partials[I.tile0 + k]);
  states[it] = s;
}

}  // namespace

void reduce_rows(Ctx &c, const ReduceSpec &s, const void *x, uint64_t n, void *state) {
  if (n == 0) return;
  with_op(s, [&](auto op, auto *tag) {
    using T = typename std::remove_pointer<decltype(tag)>::type;
    launch<T>(c, op, x, n, state);
  });
}

void reduce_rows2(Ctx &c, const ReduceSpec &s, const void *x, const void *y, uint64_t n, void *state) {
  if (n == 0) return;
  switch (s.type) {
    case VK_I64: launch_dot(c, static_cast<const int64_t *>(x), static_cast<const int64_t *>(y), n, state); break;
    case VK_U64: launch_dot(c, static_cast<const uint64_t *>(x), static_cast<const uint64_t *>(y), n, state); break;
    case VK_I32: launch_dot(c, static_cast<const int32_t *>(x), static_cast<const int32_t *>(y), n, state); break;
    default: usage_error("unsupported integer dot type");
  }
}

void reduce_rows_chain(Ctx &c, const ReduceSpec &s, const ChainSpec &ch, const void *x, uint64_t n, void *state) {
  if (n == 0) return;
  with_op(s, [&](auto op, auto *tag) {
    using T = typename std::remove_pointer<decltype(tag)>::type;
    launch<T>(c, OpChain<decltype(op), T>{op, ch}, x, n, state);
  });
}

// Batched reductions over device-planned tiles of tile_rows rows (plan_tiles_dev with unit 1):
// tile_rows is 4096, 8192 or 16384; partials holds p.max_tiles accumulators (24 bytes each).
void reduce_batch_dev(Ctx &c, const ReduceSpec &s, uint64_t tile_rows, const DevPlan &p, uint32_t nitems,
                      void *partials, uint64_t *states) {
  if (nitems == 0) return;
  with_op(s, [&](auto op, auto *tag) {
    using T = typename std::remove_pointer<decltype(tag)>::type;
    using O = decltype(op);
    using Acc = typename O::Acc;
    auto *parts = static_cast<Acc *>(partials);
    auto go = [&](auto tile) {
      constexpr uint64_t TL = decltype(tile)::value;
      if (p.max_tiles)
        k_reduce_tiles<T, O, TL><<<p.max_tiles, kBlock, 0, c.stream>>>(op, p.items, p.tile_item, parts, p.ntiles);
      k_reduce_items<O, TL><<<stream_grid(nitems, kBlock, 1u << 20), kBlock, 0, c.stream>>>(
          op, p.items, nitems, parts, reinterpret_cast<Acc *>(states));
    };
    if (tile_rows == 4096) go(std::integral_constant<uint64_t, 4096>());
    else if (tile_rows == 16384) go(std::integral_constant<uint64_t, 16384>());
    else go(std::integral_constant<uint64_t, 8192>());
  });
  check_hip(hipGetLastError(), "reduce batch (device plan)");
}

uint32_t reduce_state_words(const ReduceSpec &s) {
  return s.kind == ReduceSpec::EXTREME ? 2 : s.kind == ReduceSpec::ARG_EXTREME ? 3 : 1;
}

void reduce_batch(Ctx &c, const ReduceSpec &s, std::vector<BatchItem> &items, uint64_t *states) {
  const uint64_t nitems = items.size();
  if (nitems == 0) return;
  BatchPlan p = batch_plan(c, items, kBatchTile);
  with_op(s, [&](auto op, auto *tag) {
    using T = typename std::remove_pointer<decltype(tag)>::type;
    using Acc = typename decltype(op)::Acc;
    Acc *parts = static_cast<Acc *>(tmp_alloc(c, (p.ntiles ? p.ntiles : 1) * sizeof(Acc)));
    if (p.ntiles)
      k_reduce_tiles<T, decltype(op)><<<p.ntiles, kBlock, 0, c.stream>>>(op, p.items, p.tile_item, parts);
    k_reduce_items<decltype(op)><<<stream_grid(nitems, kBlock, 1u << 20), kBlock, 0, c.stream>>>(
        op, p.items, nitems, parts, reinterpret_cast<Acc *>(states));
    check_hip(hipGetLastError(), "reduce batch");
    tmp_free(c, parts);
  });
  batch_plan_free(c, p);
}

}  // namespace vk
