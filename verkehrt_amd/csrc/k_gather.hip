// gather_<t> (Gather<T, BOOL>, src/kernels/scalar/index.rs:27-73): out[i] = xs[idx[i]].
// An index outside 0..len is an error naming the first bad index in row order
// (first_out_of_range, index.rs:95-111). One pass: in-range rows copy their source row,
// out-of-range rows lower the device minimum of bad rows; the host names the value at that
// row. Algorithmic bytes: n x (8 + 2 x width). Each random source row costs a whole 128-byte
// line of HBM traffic once the table outgrows the caches, which bounds this kernel.
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

// Rows go in groups of R = 16 / width that fill one 16-byte store, aligned on the output
// (a fresh allocation; a misaligned start is taken row by row). Partial-line stores of single
// rows would make the memory write every output line twice. A bad row's output slot holds
// garbage: the call then fails and returns no array.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(const T *__restrict__ xs, uint64_t xlen,
                                                   const int64_t *__restrict__ idx, uint64_t n,
                                                   T *__restrict__ out, uint64_t *bad) {
  constexpr int R = 16 / sizeof(T);
  struct alignas(16) Vec {
    T v[R];
  };
  uint64_t my_bad = ~uint64_t(0);
  auto one = [&](uint64_t row) -> T {
    const int64_t i = idx[row];
    if (static_cast<uint64_t>(i) < xlen) return xs[i];
    if (row < my_bad) my_bad = row;
    return T(0);
  };
  const uint64_t mis = (reinterpret_cast<uintptr_t>(out) & 15) / sizeof(T);
  const uint64_t head = mis ? (R - mis < n ? R - mis : n) : 0;
  const uint64_t tid = blockIdx.x * uint64_t(blockDim.x) + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  if (tid < head) out[tid] = one(tid);
  const uint64_t groups = (n - head) / R;
  Vec *vout = reinterpret_cast<Vec *>(out + head);
  for (uint64_t g = tid; g < groups; g += stride) {
    const uint64_t r0 = head + g * R;
    int64_t ii[R];
#pragma unroll
    for (int k = 0; k < R; ++k) ii[k] = idx[r0 + k];
    Vec v;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (static_cast<uint64_t>(ii[k]) < xlen) {
        v.v[k] = xs[ii[k]];
      } else {
        v.v[k] = T(0);
        if (r0 + k < my_bad) my_bad = r0 + k;
      }
    }
    vout[g] = v;
  }
  const uint64_t tail0 = head + groups * R;
  if (tid < n - tail0) out[tail0 + tid] = one(tail0 + tid);
  // one atomic per wave that saw a bad row
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    uint64_t o = __shfl_xor(my_bad, m);
    my_bad = o < my_bad ? o : my_bad;
  }
  if (lane_id() == 0 && my_bad != ~uint64_t(0)) atomicMin(reinterpret_cast<unsigned long long *>(bad), my_bad);
}

template <typename T>
void launch(Ctx &c, const void *xs, uint64_t xlen, const int64_t *idx, uint64_t n, void *out,
            uint64_t *bad) {
  const uint32_t grid = stream_grid(n, kBlock * (16 / sizeof(T)) * 2, 8192);
  k_gather<T><<<grid, kBlock, 0, c.stream>>>(static_cast<const T *>(xs), xlen, idx, n,
                                             static_cast<T *>(out), bad);
  check_hip(hipGetLastError(), "gather launch");
}

}  // namespace

void gather_rows(Ctx &c, uint32_t width, const void *xs, uint64_t xlen, const int64_t *idx,
                 uint64_t n, void *out, uint64_t *d_bad) {
  fill_u64(c, d_bad, 1, ~uint64_t(0));
  if (n == 0) return;
  switch (width) {
    case 8: launch<uint64_t>(c, xs, xlen, idx, n, out, d_bad); break;
    case 4: launch<uint32_t>(c, xs, xlen, idx, n, out, d_bad); break;
    default: launch<uint8_t>(c, xs, xlen, idx, n, out, d_bad); break;
  }
}

}  // namespace vk
