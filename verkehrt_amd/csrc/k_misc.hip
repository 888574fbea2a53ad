// Construction kernels and the elementwise kernels of the kernel table.
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

// ---------------------------------------------------------------- elementwise
// One thread per row in a grid-stride loop; an operand is a row array or a scalar repeated
// for every row (Arg::Splat). The operation is uniform across the launch, so its switch is a
// scalar branch. Semantics of elementwise.rs:377-545 and :574-737.

template <typename T>
struct Operand {
  const T *p;
  T v;
  __device__ __forceinline__ T at(uint64_t i) const { return p ? p[i] : v; }
};

template <typename T>
Operand<T> operand(const void *p, bool splat, uint64_t bits) {
  Operand<T> o{splat ? nullptr : static_cast<const T *>(p), T{}};
  __builtin_memcpy(&o.v, &bits, sizeof(T));
  return o;
}

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
      case EW_MAX: return max2(a, b);
      case EW_BAND: return static_cast<T>(x & y);
      case EW_BOR: return static_cast<T>(x | y);
      default: return static_cast<T>(x ^ y);  // EW_BXOR
    }
  } else {
    switch (op) {
      case EW_ADD: return a + b;
      case EW_SUB: return a - b;
      case EW_MUL: return a * b;
      case EW_MIN: return min2(a, b);
      case EW_MAX: return max2(a, b);
      default: return a / b;  // EW_DIV, IEEE division (correctly rounded)
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ew_bin(int op, Operand<T> a, Operand<T> b, uint64_t n, T *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = arith(op, a.at(i), b.at(i));
}

// Comparisons write 0/1 bytes; floats compare as IEEE (NaN compares unequal, fused.rs:42-55).
// Bool operations work on the 0/1 storage (elementwise.rs:528-535).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ew_bool(int op, Operand<T> a, Operand<T> b, uint64_t n, uint8_t *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x) {
    const T x = a.at(i), y = b.at(i);
    uint8_t r;
    switch (op) {
      case EW_AND: r = uint8_t(x) & uint8_t(y); break;
      case EW_OR: r = uint8_t(x) | uint8_t(y); break;
      case EW_XOR: r = uint8_t(x) ^ uint8_t(y); break;
      default: r = cmp_test(op - EW_EQ, x, y) ? 1 : 0; break;
    }
    out[i] = r;
  }
}

// Rust's `as` from T to U (elementwise.rs:574-610): float to integer truncates toward zero,
// saturates at the target's bounds and maps NaN to 0; integer to integer keeps the low bytes
// (sign- or zero-extending first); conversions to a float round to nearest even.
template <typename T, typename U>
__device__ __forceinline__ U cast_as(T x) {
  if constexpr (std::is_floating_point<T>::value && std::is_integral<U>::value) {
    const double v = static_cast<double>(x);
    if (!(v == v)) return U(0);
    constexpr double lo = std::is_signed<U>::value ? -static_cast<double>(uint64_t(1) << (8 * sizeof(U) - 1)) : 0.0;
    constexpr double hi = std::is_signed<U>::value ? static_cast<double>(uint64_t(1) << (8 * sizeof(U) - 1))
                                                   : 2.0 * static_cast<double>(uint64_t(1) << (8 * sizeof(U) - 1));
    if (v <= lo) return std::is_signed<U>::value ? static_cast<U>(uint64_t(1) << (8 * sizeof(U) - 1)) : U(0);
    using UW = typename std::make_unsigned<U>::type;
    if (v >= hi) return static_cast<U>(static_cast<UW>(~UW(0) >> (std::is_signed<U>::value ? 1 : 0)));
    if constexpr (std::is_signed<U>::value) return static_cast<U>(static_cast<int64_t>(v));
    else return static_cast<U>(static_cast<uint64_t>(v));
  } else if constexpr (std::is_integral<T>::value && std::is_integral<U>::value) {
    using W = typename std::conditional<std::is_signed<T>::value, int64_t, uint64_t>::type;
    return static_cast<U>(static_cast<uint64_t>(static_cast<W>(x)));
  } else {
    return static_cast<U>(x);
  }
}

template <typename T, typename U>
__global__ __launch_bounds__(kBlock) void k_ew_cast(Operand<T> a, uint64_t n, U *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = cast_as<T, U>(a.at(i));
}

// Truncating integer division by a nonzero scalar (DivConstBody, elementwise.rs:612-737): the
// quotient of the magnitudes, negated when the signs differ, computed in the unsigned type of
// the width, so MIN / -1 wraps to MIN.
template <typename T>
__device__ __forceinline__ T div_trunc(T x, T d) {
  using W = typename std::make_unsigned<T>::type;
  if constexpr (std::is_signed<T>::value) {
    const W ux = x < 0 ? W(0) - W(x) : W(x), ud = d < 0 ? W(0) - W(d) : W(d);
    const W q = ux / ud;
    return static_cast<T>((x < 0) != (d < 0) ? W(0) - q : q);
  } else {
    return static_cast<T>(W(x) / W(d));
  }
}

template <typename T>
__device__ __forceinline__ T unary(int op, T x, T k) {
  if constexpr (std::is_integral<T>::value) {
    using W = typename std::make_unsigned<T>::type;
    switch (op) {
      case EW_NEG: return static_cast<T>(W(0) - W(x));                       // wrapping_neg
      case EW_ABS: return x < 0 ? static_cast<T>(W(0) - W(x)) : x;           // wrapping_abs
      case EW_BNOT: return static_cast<T>(~W(x));
      case EW_NOT: return static_cast<T>(uint8_t(x) ^ 1);                    // bool storage
      default: return div_trunc(x, k);                                       // EW_DIVC
    }
  } else {
    switch (op) {
      case EW_NEG: return -x;
      case EW_ABS: return __builtin_fabs(x);
      default: return __builtin_sqrt(x);  // EW_SQRT, correctly rounded
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_ew_un(int op, Operand<T> a, T k, uint64_t n, T *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = unary(op, a.at(i), k);
}

// select_<t> (elementwise.rs:105-120): row i is t[i] where cond[i] is nonzero, else e[i].
template <typename T>
__global__ __launch_bounds__(kBlock) void k_ew_select(Operand<uint8_t> c, Operand<T> t, Operand<T> e, uint64_t n,
                                                      T *out) {
  for (uint64_t i = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * blockDim.x)
    out[i] = c.at(i) ? t.at(i) : e.at(i);
}

inline uint32_t ew_grid(uint64_t n) { return stream_grid(n, kBlock * 4, 8192); }

template <typename T, typename U>
void cast_launch(Ctx &c, const EwSpec &s, const void *a, uint64_t n, void *out) {
  k_ew_cast<T, U><<<ew_grid(n), kBlock, 0, c.stream>>>(operand<T>(a, s.a_splat, s.a_bits), n, static_cast<U *>(out));
}

template <typename T>
void cast_from(Ctx &c, const EwSpec &s, const void *a, uint64_t n, void *out) {
  switch (s.out_type) {
    case VK_F64: cast_launch<T, double>(c, s, a, n, out); break;
    case VK_F32: cast_launch<T, float>(c, s, a, n, out); break;
    case VK_I64: cast_launch<T, int64_t>(c, s, a, n, out); break;
    case VK_U64: cast_launch<T, uint64_t>(c, s, a, n, out); break;
    case VK_I32: cast_launch<T, int32_t>(c, s, a, n, out); break;
    default: cast_launch<T, uint8_t>(c, s, a, n, out); break;
  }
}

template <typename T>
void ew_launch(Ctx &c, const EwSpec &s, const void *a, const void *b, const void *e, uint64_t n, void *out) {
  const Operand<T> A = operand<T>(a, s.a_splat, s.a_bits), B = operand<T>(b, s.b_splat, s.b_bits);
  switch (s.op) {
    case EW_EQ: case EW_NE: case EW_LT: case EW_LE: case EW_GT: case EW_GE:
    case EW_AND: case EW_OR: case EW_XOR:
      k_ew_bool<T><<<ew_grid(n), kBlock, 0, c.stream>>>(s.op, A, B, n, static_cast<uint8_t *>(out));
      break;
    case EW_NEG: case EW_ABS: case EW_SQRT: case EW_BNOT: case EW_NOT: case EW_DIVC:
      k_ew_un<T><<<ew_grid(n), kBlock, 0, c.stream>>>(s.op, A, B.v, n, static_cast<T *>(out));
      break;
    case EW_CAST: cast_from<T>(c, s, a, n, out); break;
    case EW_SELECT:
      // a is the condition (bool), b and e the two choices
      k_ew_select<T><<<ew_grid(n), kBlock, 0, c.stream>>>(operand<uint8_t>(a, s.a_splat, s.a_bits), B,
                                                          operand<T>(e, s.c_splat, s.c_bits), n, static_cast<T *>(out));
      break;
    default: k_ew_bin<T><<<ew_grid(n), kBlock, 0, c.stream>>>(s.op, A, B, n, static_cast<T *>(out)); break;
  }
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

void elementwise(Ctx &c, const EwSpec &s, const void *a, const void *b, const void *e, uint64_t n, void *out) {
  if (n == 0) return;
  switch (s.in_type) {
    case VK_F64: ew_launch<double>(c, s, a, b, e, n, out); break;
    case VK_F32: ew_launch<float>(c, s, a, b, e, n, out); break;
    case VK_I64: ew_launch<int64_t>(c, s, a, b, e, n, out); break;
    case VK_U64: ew_launch<uint64_t>(c, s, a, b, e, n, out); break;
    case VK_I32: ew_launch<int32_t>(c, s, a, b, e, n, out); break;
    default: ew_launch<uint8_t>(c, s, a, b, e, n, out); break;
  }
  check_hip(hipGetLastError(), "elementwise");
}

}  // namespace vk
