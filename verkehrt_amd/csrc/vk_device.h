// Device-side helpers shared by the gfx950 kernels: wave64 primitives, agent-scope atomics
// for look-back, IEEE 754-2019 minimum/maximum, and the row iteration of streaming kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace vk {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of lanes below this one whose bit is set in `mask`.
__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Agent-scope atomics: visible across the eight XCDs, whose L2s are not coherent.
__device__ __forceinline__ uint64_t ld_acq(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_rlx(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rlx(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Look-back publication without cache maintenance. Release/acquire at agent scope write back
// or invalidate the whole per-XCD L2 on every publish and poll; look-back status is instead
// only ever touched by agent-scope atomic loads and stores (above), which the hardware keeps
// coherent across XCDs on their own. A publisher stores its payload words, then
// publish_fence() waits until every earlier store of the wave is acknowledged, then it stores
// the flag word: a wave that has seen the flag issues its payload loads after that and reads
// the payload. observe_fence() keeps the compiler from hoisting payload loads above the flag
// test. A publisher whose payload was stored by other waves of its block runs publish_fence()
// in each of those waves before a __syncthreads() and then stores the flag.
__device__ __forceinline__ void publish_fence() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ void observe_fence() { asm volatile("" ::: "memory"); }

// A dynamic tile id: tiles are numbered in the order their blocks start, so a tile only
// ever waits on tiles that are already running. The block that draws the last ticket
// resets the counter for the next launch of the context.
__device__ __forceinline__ uint32_t take_ticket(uint32_t *ticket, uint32_t ntiles) {
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) {
    uint32_t t = atomicAdd(ticket, 1u);
    if (t == ntiles - 1) atomicExch(ticket, 0u);
    s_tile = t;
  }
  __syncthreads();
  return s_tile;
}

__device__ __forceinline__ void backoff() { __builtin_amdgcn_s_sleep(1); }

// IEEE 754-2019 minimum/maximum with explicit compares: NaN propagates, -0.0 < +0.0
// (src/kernels/scalar/elementwise.rs:403-441). Hardware min/max are not used because their
// NaN and signed-zero rules differ.
__device__ __forceinline__ double qnan() { return __longlong_as_double(0x7ff8000000000000ll); }
__device__ __forceinline__ double min2(double a, double b) {
  if (a != a || b != b) return qnan();
  if (a == b) return __signbit(a) ? a : b;
  return a < b ? a : b;
}
__device__ __forceinline__ double max2(double a, double b) {
  if (a != a || b != b) return qnan();
  if (a == b) return __signbit(a) ? b : a;
  return a > b ? a : b;
}
__device__ __forceinline__ float min2(float a, float b) {
  if (a != a || b != b) return __int_as_float(0x7fc00000);
  if (a == b) return __signbit(a) ? a : b;
  return a < b ? a : b;
}
__device__ __forceinline__ float max2(float a, float b) {
  if (a != a || b != b) return __int_as_float(0x7fc00000);
  if (a == b) return __signbit(a) ? b : a;
  return a > b ? a : b;
}
template <typename T>
__device__ __forceinline__ T min2(T a, T b) { return b < a ? b : a; }
template <typename T>
__device__ __forceinline__ T max2(T a, T b) { return a < b ? b : a; }

// Cmp::test (src/kernels/scalar/fused.rs:42-55).
template <typename T>
__device__ __forceinline__ bool cmp_test(int cmp, T a, T b) {
  switch (cmp) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    default: return a >= b;
  }
}

// The comparison outcome of v against c as one bit: 1 less, 2 equal, 4 greater, 8 unordered.
// A comparison code keeps a set of outcomes (keepset_of), so a test is one AND, no branch.
template <typename T>
__device__ __forceinline__ uint32_t outcome(T v, T c) {
  return v < c ? 1u : v == c ? 2u : v > c ? 4u : 8u;
}
// Comparison codes 0 eq, 1 ne, 2 lt, 3 le, 4 gt, 5 ge as the outcome sets they accept.
inline uint32_t keepset_of(int cmp) {
  constexpr uint32_t sets[6] = {2, 1 | 4 | 8, 1, 1 | 2, 4, 4 | 2};
  return sets[cmp >= 0 && cmp < 6 ? cmp : 5];
}

template <typename T>
__device__ __forceinline__ T from_bits(uint64_t b) {
  T v;
  __builtin_memcpy(&v, &b, sizeof(T));
  return v;
}
template <typename T>
__device__ __forceinline__ uint64_t to_bits(T v) {
  uint64_t b = 0;
  __builtin_memcpy(&b, &v, sizeof(T));
  return b;
}

// 16 bytes of rows, the unit of a 128-bit load.
template <typename T>
struct alignas(16) Vec16 {
  T e[16 / sizeof(T)];
};

// Calls f(value, row) for every row of x[0..n), rows spread over the whole grid with
// 128-bit loads on the aligned body and scalar loads on the unaligned head and the tail.
template <typename T, typename F>
__device__ __forceinline__ void grid_rows(const T *__restrict__ x, uint64_t n, F &&f) {
  constexpr int V = 16 / sizeof(T);
  const uint64_t tid = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  const uint64_t nth = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  uint64_t head = ((16 - (reinterpret_cast<uintptr_t>(x) & 15)) & 15) / sizeof(T);
  if (head > n) head = n;
  if (tid < head) f(x[tid], tid);
  const uint64_t nv = (n - head) / V;
  const Vec16<T> *xv = reinterpret_cast<const Vec16<T> *>(x + head);
  // each block streams one contiguous chunk (a whole number of block-wide steps), eight
  // 16-byte loads in flight per thread (128 bytes: with eight blocks of 256 threads per
  // compute unit, 256 KB in flight); measured faster than a grid-wide stride
  const uint64_t bd = blockDim.x;
  const uint64_t per = ((nv + gridDim.x - 1) / gridDim.x + bd - 1) / bd * bd;
  const uint64_t b0 = blockIdx.x * per < nv ? blockIdx.x * per : nv;
  const uint64_t b1 = b0 + per < nv ? b0 + per : nv;
  uint64_t i = b0 + threadIdx.x;
  constexpr int U = 8;
  for (; i + (U - 1) * bd < b1; i += U * bd) {
    Vec16<T> v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = xv[i + u * bd];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int k = 0; k < V; ++k) f(v[u].e[k], head + (i + u * bd) * V + k);
  }
  for (; i < b1; i += bd) {
    Vec16<T> v = xv[i];
#pragma unroll
    for (int k = 0; k < V; ++k) f(v.e[k], head + i * V + k);
  }
  for (uint64_t r = head + nv * V + tid; r < n; r += nth) f(x[r], r);
}

// Shuffles a trivially copyable value word by word. The words are copied in and out with
// memcpy, never read through a cast pointer, so no access breaks type-based aliasing.
template <typename T>
struct PodWords {
  static_assert(sizeof(T) % 4 == 0, "pod shuffle needs whole words");
  static constexpr int N = static_cast<int>(sizeof(T) / 4);
  int w[N];
};
template <typename T>
__device__ __forceinline__ T shfl_xor_pod(const T &v, int mask) {
  PodWords<T> in, out;
  __builtin_memcpy(in.w, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < PodWords<T>::N; ++k) out.w[k] = __shfl_xor(in.w[k], mask);
  T r;
  __builtin_memcpy(&r, out.w, sizeof(T));
  return r;
}
template <typename T>
__device__ __forceinline__ T shfl_up_pod(const T &v, int delta) {
  PodWords<T> in, out;
  __builtin_memcpy(in.w, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < PodWords<T>::N; ++k) out.w[k] = __shfl_up(in.w[k], delta);
  T r;
  __builtin_memcpy(&r, out.w, sizeof(T));
  return r;
}
// v from lane `lane`, the same lane for the whole wave (readlane: no LDS round trip).
template <typename T>
__device__ __forceinline__ T bcast_pod(const T &v, int lane) {
  PodWords<T> in, out;
  __builtin_memcpy(in.w, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < PodWords<T>::N; ++k) out.w[k] = __builtin_amdgcn_readlane(in.w[k], lane);
  T r;
  __builtin_memcpy(&r, out.w, sizeof(T));
  return r;
}
template <typename T>
__device__ __forceinline__ T shfl_pod(const T &v, int src_lane) {
  PodWords<T> in, out;
  __builtin_memcpy(in.w, &v, sizeof(T));
#pragma unroll
  for (int k = 0; k < PodWords<T>::N; ++k) out.w[k] = __shfl(in.w[k], src_lane);
  T r;
  __builtin_memcpy(&r, out.w, sizeof(T));
  return r;
}

// Loads a trivially copyable value of 8-byte words through agent-coherent atomic loads.
template <typename T>
__device__ __forceinline__ T ld_rlx_pod(const T *p) {
  static_assert(sizeof(T) % 8 == 0, "pod load needs whole 8-byte words");
  uint64_t w[sizeof(T) / 8];
  const uint64_t *src = reinterpret_cast<const uint64_t *>(p);
#pragma unroll
  for (int k = 0; k < static_cast<int>(sizeof(T) / 8); ++k) w[k] = ld_rlx(src + k);
  T r;
  __builtin_memcpy(&r, w, sizeof(T));
  return r;
}

// Stores a trivially copyable value of 8-byte words through agent-coherent atomic stores.
template <typename T>
__device__ __forceinline__ void st_rlx_pod(T *p, const T &v) {
  static_assert(sizeof(T) % 8 == 0, "pod store needs whole 8-byte words");
  uint64_t w[sizeof(T) / 8];
  __builtin_memcpy(w, &v, sizeof(T));
  uint64_t *dst = reinterpret_cast<uint64_t *>(p);
#pragma unroll
  for (int k = 0; k < static_cast<int>(sizeof(T) / 8); ++k) st_rlx(dst + k, w[k]);
}

}  // namespace vk
