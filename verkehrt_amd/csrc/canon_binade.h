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

// The precision of the chain being evaluated. Values are held in doubles either way; an f32
// chain's values are f32 values, and its adds round to f32 (rnd). For a sum of two f32 values
// the double sum rounded to f32 equals the f32 sum: a double has 53 >= 2 * 24 + 2 bits, which
// makes the double rounding innocuous. u = 2^(e - kMant) is the grid of binade e; binades with
// e outside [kEmin, kEmax] get no hypothesis (their chains are folded sequentially).
struct PrecF64 {
  static constexpr int kMant = 52, kEmin = -960, kEmax = 960;
  __host__ __device__ static __forceinline__ double rnd(double v) { return v; }
};
struct PrecF32 {
  static constexpr int kMant = 23, kEmin = -100, kEmax = 100;
  __host__ __device__ static __forceinline__ double rnd(double v) {
    return static_cast<double>(static_cast<float>(v));
  }
};

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
template <class P = PrecF64>
__host__ __device__ __forceinline__ Binade binade_of(double approx) {
  Binade h{0, 0, 0, 0};
  const uint64_t b = dbits_(approx);
  const int ex = static_cast<int>((b >> 52) & 0x7FF);
  if (ex == 0 || ex == 0x7FF) return h;
  const int e = ex - 1023;
  if (e < P::kEmin || e > P::kEmax) return h;
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

template <class P = PrecF64>
__host__ __device__ __forceinline__ BinadeRef binade_ref(const Binade &h) {
  const double two_e = pow2(h.e);
  const double u = pow2(h.e - P::kMant);
  const double z0 = (h.neg ? -1.5 : 1.5) * two_e;
  return BinadeRef{z0, z0 + u, pow2(P::kMant - h.e), 0.5 * two_e, pow2(h.e - P::kMant - 1)};
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

template <class P = PrecF64>
__host__ __device__ __forceinline__ void run_push(RunAcc &a, const BinadeRef &r, double v) {
  const double a0 = (P::rnd(r.z0 + v) - r.z0) * r.inv_u;  // rounding from an even value
  const double a1 = (P::rnd(r.z1 + v) - r.z1) * r.inv_u;  // from an odd one
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
template <int N, class P = PrecF64>
__host__ __device__ __forceinline__ bool run_of_one(const double (&v)[N], int cnt, const BinadeRef &r, RunFn *f,
                                                    double *asum) {
  const double inf = __builtin_inf();
  double d = 0.0, lo = inf, hi = -inf, a = 0.0;
  bool tie = false, bad = false;
  auto one = [&](double x) {
    const double w = P::rnd(r.z0 + x) - r.z0;
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
template <int N, class P = PrecF64>
__host__ __device__ __forceinline__ RunFn run_of_two(const double (&v)[N], int cnt, const BinadeRef &r) {
  RunAcc a = run_start();
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (k < cnt) run_push<P>(a, r, v[k]);
  return run_end(a);
}

// The summary of the first cnt (at most N) of the elements v.
template <int N, class P = PrecF64>
__host__ __device__ __forceinline__ RunFn run_of(const double (&v)[N], int cnt, const BinadeRef &r) {
  RunFn f;
  double a;
  if (__builtin_expect((run_of_one<N, P>(v, cnt, r, &f, &a)), 1)) return f;
  return run_of_two<N, P>(v, cnt, r);
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
template <class P = PrecF64>
__host__ __device__ __forceinline__ bool run_apply(const RunFn &f, const Binade &h, double s,
                                                   double *out) {
  if (!h.ok) return false;
  const Binade hs = binade_of<P>(s);
  if (!same_binade(hs, h)) return false;
  const double inv_u = pow2(P::kMant - h.e);
  const double k = s * inv_u;  // exact integer, 2^kMant <= |k| < 2^(kMant + 1)
  const int p = parity(k);
  const double two_m = pow2(P::kMant), two_m1 = pow2(P::kMant + 1);
  const double lo = k + f.lo[p], hi = k + f.hi[p];
  bool ok;
  if (!h.neg) ok = lo >= two_m + 1.0 && hi <= two_m1 - 1.0;
  else ok = hi <= -(two_m + 1.0) && lo >= -(two_m1 - 1.0);
  // an empty run (lo = +inf, hi = -inf) leaves s unchanged
  if (f.lo[p] > f.hi[p]) {
    *out = s;
    return true;
  }
  if (!ok) return false;
  *out = (k + f.d[p]) * pow2(h.e - P::kMant);
  return true;
}

// ------------------------------------------------------------------ wide summaries
// A stretch whose running value crosses binades has no run summary. It still has a small exact
// form. Let U be the largest ulp of any running value R_i of the fold. For delta a multiple of
// 2U, fl(x + delta) = fl(x) + delta whenever x and x + delta lie in one binade: both round on
// the same grid of that binade's ulp, and an even number of ulps keeps ties to even. So a fold
// started at S = Z + delta gives R_i + delta at every step, as long as each R_i + delta stays at
// least one ulp inside the binade of R_i (then the exact sum before the rounding stays inside it
// too). Its end value is then R_end + delta, exactly representable (a multiple of the last ulp,
// inside its binade), so one add gives it. R_end depends on the residue of S modulo P = 2U, so
// a table of P / u0 entries (u0 the ulp of the start binade), each from one exact simulation of
// the stretch from Z_r = B + r u0, covers every start near B. Each entry keeps R_end and the
// window [lo, hi] of the delta = S - Z_r that pass every step.
constexpr int kWideRes = 16;  // residues simulated: periods up to 16 u0 (three binades above the start)

struct WideHead {
  double b;      // B, a multiple of 16 u0 next to the approximate start (toward zero)
  int32_t e0;    // exponent of the start binade, u0 = 2^(e0 - 52)
  int32_t nres;  // table entries kept, 2^(M + 1) for M binades above the start; 0: no table
};

struct WideEnt {
  double end; // the simulation's end value R_end
  double lo;  // least admissible delta (+inf when the entry never applies)
  double hi;  // greatest admissible delta (-inf then)
};

// One residue's simulation: the running value, the segment of consecutive running values with
// one sign and exponent (key = the top 12 bits, 0xFFFFFFFF before the first step) with its least
// and greatest value, the delta window so far, and the greatest exponent seen.
struct WideAcc {
  double z, cur, mn, mx, lo, hi;
  uint32_t key;
  int emax;
  bool bad;
};

__host__ __device__ __forceinline__ uint32_t sign_exp_key(double v) { return static_cast<uint32_t>(dbits_(v) >> 52); }

__host__ __device__ __forceinline__ WideAcc wide_start(double z) {
  const double inf = __builtin_inf();
  return WideAcc{z, z, 0.0, 0.0, -inf, inf, 0xFFFFFFFFu, -2000, false};
}

// Folds the finished segment's constraint into the window: every running value v with this key
// must keep v + delta within [2^e + u, 2^(e+1) - u] (mirrored for negative values). The bounds
// are exact: both sides are multiples of u below 2^(e+1).
template <class P = PrecF64>
__host__ __device__ __forceinline__ void wide_flush(WideAcc &a) {
  if (a.key == 0xFFFFFFFFu) return;
  const int ex = static_cast<int>(a.key & 0x7FF);
  if (ex < 1023 + P::kEmin || ex > 1023 + P::kEmax) {  // zero, subnormal, extreme, infinite or NaN
    a.bad = true;
    return;
  }
  const int e = ex - 1023;
  a.emax = a.emax > e ? a.emax : e;
  const double u = pow2(e - P::kMant);
  const double bot = pow2(e) + u, top = pow2(e + 1) - u;
  double lo, hi;
  if (a.key & 0x800) {
    lo = -top - a.mn;
    hi = -bot - a.mx;
  } else {
    lo = bot - a.mn;
    hi = top - a.mx;
  }
  a.lo = a.lo > lo ? a.lo : lo;
  a.hi = a.hi < hi ? a.hi : hi;
}

// Records the new running value r (a.cur is set by the caller).
template <class P = PrecF64>
__host__ __device__ __forceinline__ void wide_visit(WideAcc &a, double r) {
  const uint32_t k = sign_exp_key(r);
  if (__builtin_expect(k != a.key, 0)) {
    wide_flush<P>(a);
    a.key = k;
    a.mn = r;
    a.mx = r;
  } else {
    a.mn = __builtin_fmin(a.mn, r);
    a.mx = __builtin_fmax(a.mx, r);
  }
}

template <class P = PrecF64>
__host__ __device__ __forceinline__ void wide_push(WideAcc &a, double v) {
  a.cur = P::rnd(a.cur + v);
  wide_visit<P>(a, a.cur);
}

// N elements at once, with the same result as N wide_push calls: the running values first (the
// only dependent chain), then, when all of them keep the current segment's key, one min and max
// over them; otherwise they are visited one by one.
template <int N, class P = PrecF64>
__host__ __device__ __forceinline__ void wide_push_n(WideAcc &a, const double (&v)[N]) {
  double r[N];
  r[0] = P::rnd(a.cur + v[0]);
#pragma unroll
  for (int m = 1; m < N; ++m) r[m] = P::rnd(r[m - 1] + v[m]);
  a.cur = r[N - 1];
  bool same = true;
  double mn = r[0], mx = r[0];
#pragma unroll
  for (int m = 0; m < N; ++m) {
    same = same && sign_exp_key(r[m]) == a.key;
    mn = __builtin_fmin(mn, r[m]);
    mx = __builtin_fmax(mx, r[m]);
  }
  if (__builtin_expect(same, 1)) {
    a.mn = __builtin_fmin(a.mn, mn);
    a.mx = __builtin_fmax(a.mx, mx);
  } else {
#pragma unroll
    for (int m = 0; m < N; ++m) wide_visit<P>(a, r[m]);
  }
}

// The finished entry; an entry whose simulation met a value with no binade never applies.
template <class P = PrecF64>
__host__ __device__ __forceinline__ WideEnt wide_end(WideAcc &a) {
  wide_flush<P>(a);
  const double inf = __builtin_inf();
  if (a.bad || !(a.lo <= a.hi)) return WideEnt{0.0, inf, -inf};
  return WideEnt{a.cur, a.lo, a.hi};
}

// B for the approximate start `approx` under hypothesis h (h.ok): approx less its residue
// modulo 16 u0, exact.
template <class P = PrecF64>
__host__ __device__ __forceinline__ double wide_base(double approx, const Binade &h) {
  return approx - __builtin_fmod(approx, 16.0 * pow2(h.e - P::kMant));
}

// The number of entries to keep from the 16 simulated ones, given M, the greatest number of
// binades any valid entry's path rose above the start (at least 0); 0 when M > 3, or when no
// entry is valid (m < -1000, wide_start's value).
__host__ __device__ __forceinline__ int wide_nres(int m) { return m < -1000 || m > 3 ? 0 : 2 << (m < 0 ? 0 : m); }

// Applies the table (head, entries) to the exact start s: returns true and the exact end value
// when s falls in the entry's window. s and B must have one sign and lie within a factor of two
// of each other (so s - B is exact), and s must be a multiple of u0.
template <class P = PrecF64>
__host__ __device__ __forceinline__ bool wide_apply(const WideHead &hd, const WideEnt *ents, double s,
                                                    double *out) {
  if (hd.nres <= 0) return false;
  const double b = hd.b;
  const bool near = b > 0 ? (s >= 0.5 * b && s <= 2.0 * b) : (s <= 0.5 * b && s >= 2.0 * b);
  if (!near) return false;
  const double u0 = pow2(hd.e0 - P::kMant), inv_u0 = pow2(P::kMant - hd.e0);
  // t = s - B is exact (Sterbenz) and below 2^(e0 + 1) in magnitude, so t / u0 is an integer
  // below 2^(kMant + 1) when s is a multiple of u0; its low bits give the residue (nres is a power of 2)
  const double t = s - b;
  const double k = t * inv_u0;
  if (k != __builtin_floor(k)) return false;
  const int r = static_cast<int>(static_cast<int64_t>(k) & (hd.nres - 1));
  const WideEnt e = ents[r];
  const double delta = t - static_cast<double>(r) * u0;
  if (!(delta >= e.lo && delta <= e.hi)) return false;
  *out = e.end + delta;
  return true;
}

}  // namespace vk
