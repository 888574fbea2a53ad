// Canonical-order float reductions: reduce_add_f64 / reduce_mul_f64 (CanonicalAcc,
// src/kernels/scalar/reduce.rs:249-323), dot_f64, sumsq_f64 and condsum_<cmp>_f64
// (fused.rs:72-210). Accumulator j folds rows j, j+16, ... in index order with plain IEEE
// adds (or multiplies); no FMA anywhere (built with -ffp-contract=off).
//
// This file holds the sequential form: one wave, lane j < 16 owning accumulator j. Its
// bound is the dependent-add latency, n/16 adds in a row, not HBM bandwidth.
#include "vk_device.h"
#include "vk_internal.h"

namespace vk {
namespace {

struct RowVal {
  int kind;  // 0 sum, 1 dot, 2 sumsq, 3 condsum, 4 mul
  int cmp;
  double c, e;
  __device__ __forceinline__ double operator()(const double *x, const double *y, uint64_t i) const {
    switch (kind) {
      case 1: return x[i] * y[i];
      case 2: return x[i] * x[i];
      case 3: { double v = x[i]; return cmp_test(cmp, v, c) ? v : e; }
      default: return x[i];
    }
  }
};

constexpr int kGroup = 8;  // 64-row chunks loaded ahead of the dependent chain

// acc[16] <- fold of rows [0, 16*blocks) in the canonical order.
__global__ __launch_bounds__(64) void k_canon_chain(RowVal rv, bool mul, const double *__restrict__ x,
                                                     const double *__restrict__ y, uint64_t blocks,
                                                     double *acc) {
  const int l = lane_id();
  const uint64_t rows = blocks * 16;
  double a = l < 16 ? acc[l] : 0.0;
  const uint64_t chunks = (rows + 63) / 64;
  double cur[kGroup], nxt[kGroup];
  auto load = [&](double *v, uint64_t g) {
#pragma unroll
    for (int k = 0; k < kGroup; ++k) {
      uint64_t r = (g * kGroup + k) * 64 + l;
      v[k] = r < rows ? rv(x, y, r) : 0.0;
    }
  };
  const uint64_t groups = (chunks + kGroup - 1) / kGroup;
  if (groups == 0) return;
  load(cur, 0);
  for (uint64_t g = 0; g < groups; ++g) {
    if (g + 1 < groups) load(nxt, g + 1);
#pragma unroll
    for (int k = 0; k < kGroup; ++k) {
      const uint64_t base = (g * kGroup + k) * 64;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double v = __shfl(cur[k], q * 16 + (l & 15));
        if (l < 16 && base + q * 16 < rows) a = mul ? a * v : a + v;
      }
    }
#pragma unroll
    for (int k = 0; k < kGroup; ++k) cur[k] = nxt[k];
  }
  if (l < 16) acc[l] = a;
}

__global__ void k_canon_rest(RowVal rv, const double *x, const double *y, uint64_t rest,
                             double *tail) {
  const uint64_t i = threadIdx.x;
  if (i < rest) tail[i] = rv(x, y, i);
}

// finish_lanes (reduce.rs:75-91): pairwise +8, +4, +2, +1, then the tail into acc 0.
__global__ void k_canon_finish(bool mul, const double *acc_in, const double *tail, int tail_len,
                               double *out) {
  if (threadIdx.x != 0) return;
  double acc[16];
  for (int j = 0; j < 16; ++j) acc[j] = acc_in[j];
  for (int step = 8; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) acc[j] = mul ? acc[j] * acc[j + step] : acc[j] + acc[j + step];
  double a = acc[0];
  for (int j = 0; j < tail_len; ++j) a = mul ? a * tail[j] : a + tail[j];
  *out = a;
}

__global__ void k_canon_init(double *acc, double id) {
  if (threadIdx.x < 16) acc[threadIdx.x] = id;
}

}  // namespace

void canon_init(Ctx &c, bool mul, double *acc) {
  k_canon_init<<<1, 64, 0, c.stream>>>(acc, mul ? 1.0 : 0.0);
  check_hip(hipGetLastError(), "canon init");
}

void canon_blocks(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t blocks,
                  double *acc) {
  if (blocks == 0) return;
  RowVal rv{s.kind, s.cmp, s.c, s.e};
  k_canon_chain<<<1, 64, 0, c.stream>>>(rv, s.kind == 4, x, y ? y : x, blocks, acc);
  check_hip(hipGetLastError(), "canon chain");
}

void canon_rest(Ctx &c, const CanonSpec &s, const double *x, const double *y, uint64_t rest,
                double *tail) {
  if (rest == 0) return;
  RowVal rv{s.kind, s.cmp, s.c, s.e};
  k_canon_rest<<<1, 64, 0, c.stream>>>(rv, x, y ? y : x, rest, tail);
  check_hip(hipGetLastError(), "canon rest");
}

void canon_finish(Ctx &c, bool mul, const double *acc, const double *tail, int tail_len,
                  double *out) {
  k_canon_finish<<<1, 64, 0, c.stream>>>(mul, acc, tail, tail_len, out);
  check_hip(hipGetLastError(), "canon finish");
}

}  // namespace vk
