// gather_<t> (Gather<T, BOOL>, src/kernels/scalar/index.rs:27-73): out[i] = xs[idx[i]].
// An index outside 0..len is an error naming the first bad index in row order
// (first_out_of_range, index.rs:95-111). One pass: rows read their index with 128-bit
// loads, in-range rows copy their source row, out-of-range rows lower the device minimum of
// bad rows; the host names the value at that row. Algorithmic bytes: n x (8 + 2 x width).
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(const T *__restrict__ xs, uint64_t xlen,
                                                   const int64_t *__restrict__ idx, uint64_t n,
                                                   T *__restrict__ out, uint64_t *bad) {
  uint64_t my_bad = ~uint64_t(0);
  grid_rows(idx, n, [&](int64_t i, uint64_t row) {
    if (static_cast<uint64_t>(i) < xlen) {
      out[row] = xs[i];
    } else if (row < my_bad) {
      my_bad = row;
    }
  });
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
  const uint32_t grid = stream_grid(n, kBlock * 2 * 4, 4096);
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
