// Device evaluation of a fused region's elementwise chain (ChainSpec, vk_internal.h). The
// spec is a kernel argument, so every branch on an op or a source is uniform across the wave.
#pragma once

#include "vk_device.h"
#include "vk_internal.h"

namespace vk {

template <typename T>
__device__ __forceinline__ T chain_apply(int op, T a, T b) {
  if constexpr (std::is_integral<T>::value) {
    // wrapping arithmetic in the unsigned type of the width (elementwise.rs:380-397, 449-470)
    using W = typename std::make_unsigned<T>::type;
    const W x = static_cast<W>(a), y = static_cast<W>(b);
    switch (op) {
      case EW_ADD: return static_cast<T>(static_cast<W>(x + y));
      case EW_SUB: return static_cast<T>(static_cast<W>(x - y));
      case EW_MUL: return static_cast<T>(static_cast<W>(x * y));
      case EW_MIN: return min2(a, b);
      case EW_MAX: return max2(a, b);
      case EW_NEG: return static_cast<T>(static_cast<W>(W(0) - x));
      case EW_ABS: return a < T(0) ? static_cast<T>(static_cast<W>(W(0) - x)) : a;
      case EW_BAND: return static_cast<T>(x & y);
      case EW_BOR: return static_cast<T>(x | y);
      case EW_BXOR: return static_cast<T>(x ^ y);
      default: return static_cast<T>(static_cast<W>(~x));  // EW_BNOT
    }
  } else {
    switch (op) {
      case EW_ADD: return a + b;
      case EW_SUB: return a - b;
      case EW_MUL: return a * b;
      case EW_DIV: return a / b;
      case EW_MIN: return min2(a, b);
      case EW_MAX: return max2(a, b);
      case EW_NEG: return -a;
      case EW_ABS: return __builtin_fabs(a);
      default: return __builtin_sqrt(a);  // EW_SQRT
    }
  }
}

// The constant of an op from its low sizeof(T) bytes.
template <typename T>
__device__ __forceinline__ T chain_const(uint64_t bits) {
  if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, bits);
  else if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, static_cast<uint32_t>(bits));
  else return __builtin_bit_cast(T, static_cast<uint8_t>(bits));
}

// The chain over one row: a loop over the ops, kept rolled so that a kernel evaluating it for
// many rows stays small enough to keep its own row arrays unrolled in registers; ch.ops[i] is a
// uniform load from the kernel argument.
template <typename T>
__device__ __forceinline__ T chain_eval(const ChainSpec &ch, T x, T y) {
  T v = x;
  const int n = ch.n;
#pragma unroll 1
  for (int i = 0; i < n; ++i) {
    const int op = ch.ops[i].op, src = ch.ops[i].src;
    const bool swap = ch.ops[i].swap != 0;
    const T k = chain_const<T>(ch.ops[i].kbits);
    const T b = src == 0 ? k : src == 1 ? x : src == 2 ? y : v;
    v = swap ? chain_apply<T>(op, b, v) : chain_apply<T>(op, v, b);
  }
  return v;
}

}  // namespace vk
