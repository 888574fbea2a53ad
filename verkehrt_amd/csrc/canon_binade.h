// Exact parallel evaluation of one canonical accumulator chain, S <- fl(S + v) in index
// order (the chain of accumulator j of src/kernels/scalar/reduce.rs:40-62), without changing
// a single bit of the sequential result.
//
// The fact used: while the running value S and every exact sum S + v stay inside one binade
// [2^e, 2^(e+1)) of fixed sign, every representable value there is an integer multiple of
// u = 2^(e-52), so fl(S + v) = S + u * r(v, parity), where r rounds v/u to an integer, ties
// to the even result. r depends on S only through the parity of S/u. A run of elements is
// therefore a function of one bit: for each start parity p it adds D_p units of u, and its
// running offset stays within [lo_p, hi_p]. These functions compose associatively, so runs
// are summarised in parallel and joined in order; a start value S then checks the summary
// (S is in the binade, and S + lo_p, S + hi_p stay strictly inside it) and, when the check
// holds, the sequential result is exactly S + u * D_p. When it fails (a binade crossing, an
// element too large, NaN or infinity, a start of zero or a subnormal) the caller folds that
// run sequentially instead, so the result is always the sequential one.
//
// r(v, p) is computed with IEEE adds from reference points inside the binade: Z0 = +-1.5 * 2^e
// (an even multiple of u) and Z1 = Z0 + u (odd), r(v, p) = (fl(Zp + v) - Zp) / u, exact by
// Sterbenz whenever |v| < 2^(e-1). Offsets are integers held exactly in doubles.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vk {

// A binade hypothesis: exponent e of the running value and its sign. `ok` false means no
// hypothesis (the run is folded sequentially).
struct Binade {
  int e;
  int neg;
  int ok;
  int pad;
};

// The summary of a run under one hypothesis, in units of u, for start parities 0 and 1.
// An element too large for the binade, NaN or infinite makes the run unusable: its bounds
// become [-inf, +inf] (or NaN), which no start value passes, and joins keep them so.
struct RunFn {
  double d[2];   // offset after the run
  double lo[2];  // least offset after any element of the run (+inf for an empty run)
  double hi[2];  // greatest offset after any element (-inf for an empty run)
};

__host__ __device__ __forceinline__ double pow2(int k) {
  // 2^k for -1022 <= k <= 1023, built from its bits
  const uint64_t b = static_cast<uint64_t>(k + 1023) << 52;
  double r;
  __builtin_memcpy(&r, &b, 8);
  return r;
}

__host__ __device__ __forceinline__ uint64_t dbits_(double v) {
  uint64_t b;
  __builtin_memcpy(&b, &v, 8);
  return b;
}

// The hypothesis of a running value near `approx`; none for zero, subnormal, non-finite or
// extreme magnitudes (where u or 1/u would leave the normal range).
__host__ __device__ __forceinline__ Binade binade_of(double approx) {
  Binade h{0, 0, 0, 0};
  const uint64_t b = dbits_(approx);
  const int ex = static_cast<int>((b >> 52) & 0x7FF);
  if (ex == 0 || ex == 0x7FF) return h;
  const int e = ex - 1023;
  if (e < -960 || e > 960) return h;
  h.e = e;
  h.neg = static_cast<int>(b >> 63);
  h.ok = 1;
  return h;
}

__host__ __device__ __forceinline__ bool same_binade(const Binade &a, const Binade &b) {
  return a.ok && b.ok && a.e == b.e && a.neg == b.neg;
}

__host__ __device__ __forceinline__ RunFn run_identity() {
  const double inf = __builtin_inf();
  return RunFn{{0.0, 0.0}, {inf, inf}, {-inf, -inf}};
}

// Parity of an integer-valued double t: 1 when t is odd.
__host__ __device__ __forceinline__ int parity(double t) {
  const double h = t * 0.5;
  return h != __builtin_floor(h) ? 1 : 0;
}

// Per-run constants of a hypothesis.
struct BinadeRef {
  double z0, z1;   // reference points, even and odd multiples of u
  double inv_u;    // 1/u
  double lim;      // elements must satisfy |v| < lim = 2^(e-1)
  double half_u;   // u / 2
};

__host__ __device__ __forceinline__ BinadeRef binade_ref(const Binade &h) {
  const double two_e = pow2(h.e);
  const double u = pow2(h.e - 52);
  const double z0 = (h.neg ? -1.5 : 1.5) * two_e;
  return BinadeRef{z0, z0 + u, pow2(52 - h.e), 0.5 * two_e, pow2(h.e - 53)};
}

// A run summary being built: the summary so far, the current parity for each start parity
// (p xor parity(offset), tracked as bits from the increments' parities so that the chain from
// one element to the next is a single add), and whether an element the binade cannot take
// (too large, NaN, infinite) has been seen. The parity bits equal those of the offsets while
// the offsets stay exact, which every summary that passes a check does.
struct RunAcc {
  RunFn f;
  bool q[2];
  bool bad;
};

__host__ __device__ __forceinline__ RunAcc run_start() { return RunAcc{run_identity(), {false, true}, false}; }

__host__ __device__ __forceinline__ void run_push(RunAcc &a, const BinadeRef &r, double v) {
  const double a0 = ((r.z0 + v) - r.z0) * r.inv_u;  // rounding from an even value
  const double a1 = ((r.z1 + v) - r.z1) * r.inv_u;  // from an odd one
  const bool p0 = parity(a0) != 0, p1 = parity(a1) != 0;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const bool q = a.q[p];
    a.f.d[p] = a.f.d[p] + (q ? a1 : a0);
    a.q[p] = q ? !p1 : p0;
    // bounds are never NaN here (they start at +-inf and take minNum/maxNum with the offsets)
    a.f.lo[p] = __builtin_fmin(a.f.lo[p], a.f.d[p]);
    a.f.hi[p] = __builtin_fmax(a.f.hi[p], a.f.d[p]);
  }
  const double av = v < 0 ? -v : v;
  a.bad = a.bad || !(av < r.lim);
}

// The finished summary: a run with a bad element gets bounds [-inf, +inf], which no start
// value passes.
__host__ __device__ __forceinline__ RunFn run_end(const RunAcc &a) {
  RunFn f = a.f;
  if (a.bad) {
    const double inf = __builtin_inf();
    f.lo[0] = f.lo[1] = -inf;
    f.hi[0] = f.hi[1] = inf;
  }
  return f;
}

// The one-track summary of the first cnt (at most N) of the elements v, and their plain sum in
// index order (*asum). Without ties the rounding does not depend on the parity, so both start
// parities give one offset track. A tie (the rounding error of Z0 + v is exactly u / 2; Z0 + v
// rounds to a multiple of u and fl(Z0 + v) - Z0 - v is exact by Fast2Sum, |Z0| > |v|) makes
// the result false: the run needs the two-track form (run_of_two). NaN and infinite elements
// give no tie but make the run bad either way.
template <int N>
__host__ __device__ __forceinline__ bool run_of_one(const double (&v)[N], int cnt, const BinadeRef &r, RunFn *f,
                                                    double *asum) {
  const double inf = __builtin_inf();
  double d = 0.0, lo = inf, hi = -inf, a = 0.0;
  bool tie = false, bad = false;
  auto one = [&](double x) {
    const double w = (r.z0 + x) - r.z0;
    const double e = x - w;
    tie = tie || (e < 0 ? -e : e) == r.half_u;
    d = d + w * r.inv_u;
    lo = __builtin_fmin(lo, d);
    hi = __builtin_fmax(hi, d);
    const double ax = x < 0 ? -x : x;
    bad = bad || !(ax < r.lim);
    a = a + x;
  };
  if (cnt >= N) {  // a whole run: straight-line code
#pragma unroll
    for (int k = 0; k < N; ++k) one(v[k]);
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k)
      if (k < cnt) one(v[k]);
  }
  *asum = a;
  RunAcc acc{RunFn{{d, d}, {lo, lo}, {hi, hi}}, {false, true}, bad};
  *f = run_end(acc);
  return !tie;
}

// The two-track summary (start parities 0 and 1 followed separately) of the first cnt elements.
template <int N>
__host__ __device__ __forceinline__ RunFn run_of_two(const double (&v)[N], int cnt, const BinadeRef &r) {
  RunAcc a = run_start();
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (k < cnt) run_push(a, r, v[k]);
  return run_end(a);
}

// The summary of the first cnt (at most N) of the elements v.
template <int N>
__host__ __device__ __forceinline__ RunFn run_of(const double (&v)[N], int cnt, const BinadeRef &r) {
  RunFn f;
  double a;
  if (__builtin_expect(run_of_one(v, cnt, r, &f, &a), 1)) return f;
  return run_of_two(v, cnt, r);
}

// The summary of run a followed by run b (both under the same hypothesis).
__host__ __device__ __forceinline__ RunFn run_join(const RunFn &a, const RunFn &b) {
  RunFn o;
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int q = p ^ parity(a.d[p]);
    o.d[p] = a.d[p] + b.d[q];
    const double blo = a.d[p] + b.lo[q], bhi = a.d[p] + b.hi[q];
    o.lo[p] = a.lo[p] < blo ? a.lo[p] : blo;
    o.hi[p] = a.hi[p] > bhi ? a.hi[p] : bhi;
  }
  return o;
}

// Applies run f (under hypothesis h) to the exact start value s. Returns true and the exact
// end value in *out when the summary proves the sequential fold equals it.
__host__ __device__ __forceinline__ bool run_apply(const RunFn &f, const Binade &h, double s,
                                                   double *out) {
  if (!h.ok) return false;
  const Binade hs = binade_of(s);
  if (!same_binade(hs, h)) return false;
  const double inv_u = pow2(52 - h.e);
  const double k = s * inv_u;  // exact integer, 2^52 <= |k| < 2^53
  const int p = parity(k);
  const double two52 = 4503599627370496.0, two53 = 9007199254740992.0;
  const double lo = k + f.lo[p], hi = k + f.hi[p];
  bool ok;
  if (!h.neg) ok = lo >= two52 + 1.0 && hi <= two53 - 1.0;
  else ok = hi <= -(two52 + 1.0) && lo >= -(two53 - 1.0);
  // an empty run (lo = +inf, hi = -inf) leaves s unchanged
  if (f.lo[p] > f.hi[p]) {
    *out = s;
    return true;
  }
  if (!ok) return false;
  *out = (k + f.d[p]) * pow2(h.e - 52);
  return true;
}

}  // namespace vk
