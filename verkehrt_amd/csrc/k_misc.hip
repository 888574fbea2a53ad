// Construction kernels and the elementwise kernels the hot path's callers need.
//
// rand_f64 / rand_i64 (src/kernels/scalar/construct.rs:10-76, 233-251): element i is
// f(splitmix64(seed + i)), counter-based, so every row is independent. iota and fill
// (construct.rs:36-60). Elementwise binary ops (elementwise.rs:377-545): wrapping integer
// arithmetic, IEEE float arithmetic without FMA, IEEE comparisons to 0/1 bool bytes, bool
// logic on the 0/1 storage, IEEE minimum/maximum, and casts to i64.
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kBlock) void k_rand_f64(uint64_t n, uint64_t seed, double *out) {
  const double unit = 1.0 / 9007199254740992.0;  // 2^-53: the 53-bit value converts exactly
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = static_cast<double>(splitmix64(seed + i) >> 11) * unit;
}

__global__ __launch_bounds__(kBlock) void k_rand_i64(uint64_t n, uint64_t seed, int64_t *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = static_cast<int64_t>(splitmix64(seed + i));
}

__global__ __launch_bounds__(kBlock) void k_iota(uint64_t n, int64_t *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = static_cast<int64_t>(i);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_t(uint64_t n, T v, T *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = v;
}

// op codes (EwSpec.op)
enum {
  EW_ADD = 0, EW_SUB, EW_MUL, EW_EQ, EW_NE, EW_LT, EW_LE, EW_GT, EW_GE,
  EW_AND, EW_OR, EW_XOR, EW_NOT, EW_CAST_I64, EW_MIN, EW_MAX, EW_CAST_F64
};

template <typename T>
__device__ __forceinline__ T arith(int op, T a, T b) {
  if constexpr (std::is_integral<T>::value) {
    // integers wrap (elementwise.rs:380-395): compute in the unsigned type of the width
    using W = typename std::make_unsigned<T>::type;
    const W x = static_cast<W>(a), y = static_cast<W>(b);
    switch (op) {
      case EW_ADD: return static_cast<T>(static_cast<W>(x + y));
      case EW_SUB: return static_cast<T>(static_cast<W>(x - y));
      case EW_MUL: return static_cast<T>(static_cast<W>(x * y));
      case EW_MIN: return min2(a, b);
      default: return max2(a, b);
    }
  } else {
    switch (op) {
      case EW_ADD: return a + b;
      case EW_SUB: return a - b;
      case EW_MUL: return a * b;
      case EW_MIN: return min2(a, b);
      default: return max2(a, b);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ew(int op, const T *__restrict__ a, const T *__restrict__ b,
                                               bool as, bool bs, T av, T bv, uint64_t n, void *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const T x = as ? av : a[i];
    const T y = (op == EW_NOT || op == EW_CAST_I64 || op == EW_CAST_F64) ? T{} : (bs ? bv : b[i]);
    switch (op) {
      case EW_EQ: case EW_NE: case EW_LT: case EW_LE: case EW_GT: case EW_GE:
        static_cast<uint8_t *>(out)[i] = cmp_test(op - EW_EQ, x, y) ? 1 : 0;
        break;
      case EW_AND: static_cast<uint8_t *>(out)[i] = uint8_t(x) & uint8_t(y); break;
      case EW_OR: static_cast<uint8_t *>(out)[i] = uint8_t(x) | uint8_t(y); break;
      case EW_XOR: static_cast<uint8_t *>(out)[i] = uint8_t(x) ^ uint8_t(y); break;
      case EW_NOT: static_cast<uint8_t *>(out)[i] = uint8_t(x) ^ 1; break;
      case EW_CAST_I64: static_cast<int64_t *>(out)[i] = static_cast<int64_t>(x); break;
      case EW_CAST_F64: static_cast<double *>(out)[i] = static_cast<double>(x); break;
      default: static_cast<T *>(out)[i] = arith(op, x, y); break;
    }
  }
}

template <typename T>
void ew_launch(Ctx &c, const EwSpec &s, const void *a, const void *b, uint64_t n, void *out) {
  T av, bv;
  __builtin_memcpy(&av, &s.a_bits, sizeof(T));
  __builtin_memcpy(&bv, &s.b_bits, sizeof(T));
  k_ew<T><<<stream_grid(n, kBlock * 4, 8192), kBlock, 0, c.stream>>>(
      s.op, static_cast<const T *>(a), static_cast<const T *>(b), s.a_splat, s.b_splat, av, bv, n, out);
}

}  // namespace

void rand_f64(Ctx &c, uint64_t n, uint64_t seed, double *out) {
  if (n == 0) return;
  k_rand_f64<<<stream_grid(n, kBlock * 4, 8192), kBlock, 0, c.stream>>>(n, seed, out);
  check_hip(hipGetLastError(), "rand_f64");
}

void rand_i64(Ctx &c, uint64_t n, uint64_t seed, int64_t *out) {
  if (n == 0) return;
  k_rand_i64<<<stream_grid(n, kBlock * 4, 8192), kBlock, 0, c.stream>>>(n, seed, out);
  check_hip(hipGetLastError(), "rand_i64");
}

void iota(Ctx &c, uint64_t n, int64_t *out) {
  if (n == 0) return;
  k_iota<<<stream_grid(n, kBlock * 4, 8192), kBlock, 0, c.stream>>>(n, out);
  check_hip(hipGetLastError(), "iota");
}

void fill_bits(Ctx &c, uint32_t width, uint64_t n, uint64_t bits, void *out) {
  if (n == 0) return;
  const uint32_t g = stream_grid(n, kBlock * 4, 8192);
  switch (width) {
    case 8: k_fill_t<uint64_t><<<g, kBlock, 0, c.stream>>>(n, bits, static_cast<uint64_t *>(out)); break;
    case 4: k_fill_t<uint32_t><<<g, kBlock, 0, c.stream>>>(n, uint32_t(bits), static_cast<uint32_t *>(out)); break;
    default: k_fill_t<uint8_t><<<g, kBlock, 0, c.stream>>>(n, uint8_t(bits), static_cast<uint8_t *>(out)); break;
  }
  check_hip(hipGetLastError(), "fill");
}

void elementwise(Ctx &c, const EwSpec &s, const void *a, const void *b, uint64_t n, void *out) {
  if (n == 0) return;
  switch (s.in_type) {
    case VK_F64: ew_launch<double>(c, s, a, b, n, out); break;
    case VK_F32: ew_launch<float>(c, s, a, b, n, out); break;
    case VK_I64: ew_launch<int64_t>(c, s, a, b, n, out); break;
    case VK_U64: ew_launch<uint64_t>(c, s, a, b, n, out); break;
    case VK_I32: ew_launch<int32_t>(c, s, a, b, n, out); break;
    default: ew_launch<uint8_t>(c, s, a, b, n, out); break;
  }
  check_hip(hipGetLastError(), "elementwise");
}

}  // namespace vk
