/*
 * vk_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C restatement of verkehrt's scalar kernel entries (the reference every
 * variant must match bit for bit). Only the test suite, the smoke check and the CPU
 * baseline leg of the bench load this library; the HIP backend never links or calls it.
 *
 * Every function cites the reference file:line it restates. Build flags: -O2
 * -ffp-contract=off (no FMA anywhere, as the reference's canonical order forbids it).
 *
 * Parity of each function here is pinned by tests/test_oracle_golden.py against the
 * reference's golden vectors and known-answer tests, plus closed forms and brute force.
 * The three float primitives the reference does not define (f64 scan add, f64
 * reduce_into add, f64 argsort) have no reference vectors: their parity is UNPINNED by
 * the reference and is pinned only by the order this project states for them and by
 * brute-force restatements in the tests.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define LANES 16 /* REDUCE_BLOCK, src/kernels.rs:47 */

/* ------------------------------------------------------------------ RNG */

/* splitmix64 finalizer on x + golden gamma: src/kernels/scalar/construct.rs:10-17 */
uint64_t vko_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* rand_f64: element i = (splitmix64(seed + i) >> 11) * 2^-53, construct.rs:19-20,63-76,242-251 */
void vko_rand_f64(uint64_t n, int64_t seed, double *out) {
  const double unit = 1.0 / 9007199254740992.0;
  for (uint64_t i = 0; i < n; ++i)
    out[i] = (double)(vko_splitmix64((uint64_t)seed + i) >> 11) * unit;
}

/* rand_i64: element i = splitmix64(seed + i) as i64, construct.rs:233-241 */
void vko_rand_i64(uint64_t n, int64_t seed, int64_t *out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = (int64_t)vko_splitmix64((uint64_t)seed + i);
}

/* ------------------------------------------------ canonical float reductions */

/*
 * The canonical 16-accumulator order, reduce.rs:40-91 (update_lanes_with, finish_lanes)
 * and the state machine CanonicalAcc reduce.rs:249-323: accumulator j adds rows
 * j, j+16, ... below b = n - n%16 in index order, the accumulators combine pairwise
 * (+8, +4, +2, +1), then the tail rows fold into accumulator 0 in index order.
 * `kind` selects the row value: 0 x[i] (reduce_add), 1 x[i]*y[i] (dot, fused.rs:72-106),
 * 2 x[i]*x[i] (sumsq, fused.rs:108-140), 3 condsum pick (fused.rs:181-210).
 * `mul` selects the product (identity 1.0) over the sum (identity 0.0).
 */
typedef struct {
  double acc[LANES];
  double tail[LANES];
  int tail_len;
  int mul;
} vko_canon_f64;

static inline double canon_op(int mul, double a, double b) { return mul ? a * b : a + b; }

void vko_canon_init(vko_canon_f64 *s, int mul) {
  double id = mul ? 1.0 : 0.0;
  for (int j = 0; j < LANES; ++j) { s->acc[j] = id; s->tail[j] = id; }
  s->tail_len = 0;
  s->mul = mul;
}

/* feed_blocks, reduce.rs:284-308. Returns 0, or -1 for "canonical update after a partial
 * block". `vals` are the already-combined row values. */
int vko_canon_feed(vko_canon_f64 *s, const double *vals, uint64_t n) {
  if (n == 0) return 0;
  if (s->tail_len != 0) return -1;
  uint64_t blocks = n / LANES, b = blocks * LANES;
  for (uint64_t k = 0; k < blocks; ++k)
    for (int j = 0; j < LANES; ++j) s->acc[j] = canon_op(s->mul, s->acc[j], vals[k * LANES + j]);
  s->tail_len = (int)(n - b);
  for (int j = 0; j < s->tail_len; ++j) s->tail[j] = vals[b + j];
  return 0;
}

/* finish_lanes, reduce.rs:75-91 */
double vko_canon_finish(const vko_canon_f64 *s) {
  double acc[LANES];
  memcpy(acc, s->acc, sizeof acc);
  for (int step = LANES / 2; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) acc[j] = canon_op(s->mul, acc[j], acc[j + step]);
  double a = acc[0];
  for (int j = 0; j < s->tail_len; ++j) a = canon_op(s->mul, a, s->tail[j]);
  return a;
}

static double canon_whole(const double *vals, uint64_t n, int mul) {
  vko_canon_f64 s;
  vko_canon_init(&s, mul);
  vko_canon_feed(&s, vals, n);
  return vko_canon_finish(&s);
}

/* canonical_sum / canonical_product, reduce.rs:152-166 */
double vko_canonical_sum_f64(const double *x, uint64_t n) { return canon_whole(x, n, 0); }
double vko_canonical_product_f64(const double *x, uint64_t n) { return canon_whole(x, n, 1); }

/* f32 canonical sum, accumulation in f32: reduce.rs:40-91 instantiated at f32 */
float vko_canonical_sum_f32(const float *x, uint64_t n) {
  float acc[LANES];
  for (int j = 0; j < LANES; ++j) acc[j] = 0.0f;
  uint64_t b = n - n % LANES;
  for (uint64_t k = 0; k < b; k += LANES)
    for (int j = 0; j < LANES; ++j) acc[j] = acc[j] + x[k + j];
  for (int step = LANES / 2; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) acc[j] = acc[j] + acc[j + step];
  float a = acc[0];
  for (uint64_t i = b; i < n; ++i) a = a + x[i];
  return a;
}

/* The row value of `kind` (see above). Row products round once each: no FMA. */
static inline double row_value(int kind, const double *x, const double *y, uint64_t i,
                               int cmp, double c, double e);

/* Cmp::test, fused.rs:42-55: 0 eq, 1 ne, 2 lt, 3 le, 4 gt, 5 ge, IEEE comparisons (NaN
 * false except for ne). */
static inline int cmp_f64(int cmp, double a, double b) {
  switch (cmp) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    default: return a >= b;
  }
}
static inline int cmp_i64(int cmp, int64_t a, int64_t b) {
  switch (cmp) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    default: return a >= b;
  }
}

static inline double row_value(int kind, const double *x, const double *y, uint64_t i,
                               int cmp, double c, double e) {
  switch (kind) {
    case 1: return x[i] * y[i];
    case 2: return x[i] * x[i];
    case 3: return cmp_f64(cmp, x[i], c) ? x[i] : e;
    default: return x[i];
  }
}

/* The canonical sum of row values of `kind` fed in tiles cut at `cuts` (ascending row
 * ends, each a multiple of 16; ncuts may be 0 for one update). Returns 0 and the sum in
 * *out, or -1 when a cut leaves a partial block before a later update. */
int vko_canonical_fold_f64(int kind, const double *x, const double *y, uint64_t n, int cmp,
                           double c, double e, const uint64_t *cuts, int ncuts, int mul,
                           double *out) {
  vko_canon_f64 s;
  vko_canon_init(&s, mul);
  double *vals = (double *)malloc((n ? n : 1) * sizeof(double));
  for (uint64_t i = 0; i < n; ++i) vals[i] = row_value(kind, x, y, i, cmp, c, e);
  uint64_t start = 0;
  for (int t = 0; t <= ncuts; ++t) {
    uint64_t end = t < ncuts ? cuts[t] : n;
    if (end > n) end = n;
    if (end > start) {
      if (vko_canon_feed(&s, vals + start, end - start) != 0) { free(vals); return -1; }
    }
    start = end;
  }
  *out = vko_canon_finish(&s);
  free(vals);
  return 0;
}

double vko_dot_f64(const double *x, const double *y, uint64_t n) {
  double r;
  vko_canonical_fold_f64(1, x, y, n, 0, 0, 0, 0, 0, 0, &r);
  return r;
}
double vko_sumsq_f64(const double *x, uint64_t n) {
  double r;
  vko_canonical_fold_f64(2, x, x, n, 0, 0, 0, 0, 0, 0, &r);
  return r;
}
double vko_condsum_f64(const double *x, uint64_t n, int cmp, double c, double e) {
  double r;
  vko_canonical_fold_f64(3, x, x, n, cmp, c, e, 0, 0, 0, &r);
  return r;
}

/* The canonical fold at T = f32 (update_lanes_with / finish_lanes reduce.rs:40-91 and
 * CanonicalAcc reduce.rs:249-323 instantiated at f32, Reducible reduce.rs:567-575), with the
 * fused row values of fused.rs:72-210 at f32: kind 0 x[i] (reduce_add_f32), 1 x[i]*y[i]
 * (dot_f32), 2 x[i]*x[i] (sumsq_f32), 3 condsum pick (condsum_<cmp>_f32), 4 the product of
 * x[i] (reduce_mul_f32, identity 1). Every operation is an f32 operation (no FMA: built with
 * -ffp-contract=off). Returns the f32 result's bits. */
static inline int cmp_f32(int cmp, float a, float b) {
  switch (cmp) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    default: return a >= b;
  }
}
uint32_t vko_canonical_fold_f32(int kind, const float *x, const float *y, uint64_t n, int cmp, float c, float e) {
  const int mul = kind == 4;
  float acc[LANES];
  for (int j = 0; j < LANES; ++j) acc[j] = mul ? 1.0f : 0.0f;
  uint64_t b = n - n % LANES;
  for (uint64_t i = 0; i < n; ++i) {
    float v;
    switch (kind) {
      case 1: v = x[i] * y[i]; break;
      case 2: v = x[i] * x[i]; break;
      case 3: v = cmp_f32(cmp, x[i], c) ? x[i] : e; break;
      default: v = x[i]; break;
    }
    if (i < b) {
      acc[i % LANES] = mul ? acc[i % LANES] * v : acc[i % LANES] + v;
    } else {
      if (i == b)
        for (int step = LANES / 2; step > 0; step /= 2)
          for (int j = 0; j < step; ++j) acc[j] = mul ? acc[j] * acc[j + step] : acc[j] + acc[j + step];
      acc[0] = mul ? acc[0] * v : acc[0] + v;
    }
  }
  if (b == n)
    for (int step = LANES / 2; step > 0; step /= 2)
      for (int j = 0; j < step; ++j) acc[j] = mul ? acc[j] * acc[j + step] : acc[j] + acc[j + step];
  uint32_t bits;
  memcpy(&bits, &acc[0], 4);
  return bits;
}

/* ----------------------------------------------------- integer reductions */

/* IntSum wrapping in the widened type, reduce.rs:325-377 */
int64_t vko_sum_i64(const int64_t *x, uint64_t n) {
  uint64_t s = 0;
  for (uint64_t i = 0; i < n; ++i) s += (uint64_t)x[i];
  return (int64_t)s;
}
uint64_t vko_sum_u64(const uint64_t *x, uint64_t n) {
  uint64_t s = 0;
  for (uint64_t i = 0; i < n; ++i) s += x[i];
  return s;
}
int64_t vko_sum_i32(const int32_t *x, uint64_t n) {
  uint64_t s = 0;
  for (uint64_t i = 0; i < n; ++i) s += (uint64_t)(int64_t)x[i];
  return (int64_t)s;
}
uint64_t vko_sum_u8(const uint8_t *x, uint64_t n) {
  uint64_t s = 0;
  for (uint64_t i = 0; i < n; ++i) s += x[i];
  return s;
}
/* IntProduct<T> wrapping in the element type T (i64, u64, i32 or u8, `width` bytes),
 * reduce.rs:358-366, identity 1: returns the product's bits in T, zero-extended. */
uint64_t vko_product_int(const void *x, uint32_t width, uint64_t n) {
  const uint8_t *b = (const uint8_t *)x;
  uint64_t p = 1;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t v = 0;
    memcpy(&v, b + i * width, width);
    p *= v; /* the low `width` bytes of a product depend only on the low bytes of its factors */
  }
  return width == 8 ? p : p & ((1ull << (8 * width)) - 1);
}

/* dot, sumsq and condsum over integers (Dot, SumSq, CondSum fused.rs:72-210 with the IntSum
 * accumulator reduce.rs:325-377): the row value in the element type (Num::mul is a wrapping
 * product, elementwise.rs:380-397), widened (i32 sign-extends to i64) and added with wrapping.
 * kind 1 x*y, 2 x*x, 3 (x <cmp> c ? x : e). type: 2 i64, 3 u64, 4 i32 (the element type codes
 * below); c and e are the constants' bits. Returns the sum's 64 bits. */
uint64_t vko_fused_sum_int(int kind, int type, const void *x, const void *y, uint64_t n, int cmp, uint64_t c,
                           uint64_t e) {
  uint64_t s = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t w;
    if (type == 4) {
      const int32_t a = ((const int32_t *)x)[i], bb = kind == 1 ? ((const int32_t *)y)[i] : a;
      int32_t v;
      if (kind == 3) v = cmp_i64(cmp, a, (int32_t)(uint32_t)c) ? a : (int32_t)(uint32_t)e;
      else v = (int32_t)((uint32_t)a * (uint32_t)bb);
      w = (uint64_t)(int64_t)v;
    } else if (type == 3) {
      const uint64_t a = ((const uint64_t *)x)[i], bb = kind == 1 ? ((const uint64_t *)y)[i] : a;
      if (kind == 3) {
        int k;
        switch (cmp) {
          case 0: k = a == c; break;
          case 1: k = a != c; break;
          case 2: k = a < c; break;
          case 3: k = a <= c; break;
          case 4: k = a > c; break;
          default: k = a >= c; break;
        }
        w = k ? a : e;
      } else {
        w = a * bb;
      }
    } else {
      const int64_t a = ((const int64_t *)x)[i], bb = kind == 1 ? ((const int64_t *)y)[i] : a;
      if (kind == 3) w = cmp_i64(cmp, a, (int64_t)c) ? (uint64_t)a : e;
      else w = (uint64_t)a * (uint64_t)bb;
    }
    s += w;
  }
  return s;
}

/* count_cmp, fused.rs:142-147 (CountCmp fused.rs:149-179) */
int64_t vko_count_cmp_f64(const double *x, uint64_t n, int cmp, double c) {
  int64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) k += cmp_f64(cmp, x[i], c);
  return k;
}
int64_t vko_count_cmp_i64(const int64_t *x, uint64_t n, int cmp, int64_t c) {
  int64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) k += cmp_i64(cmp, x[i], c);
  return k;
}
static inline int cmp_u64(int cmp, uint64_t a, uint64_t b) {
  switch (cmp) {
    case 0: return a == b;
    case 1: return a != b;
    case 2: return a < b;
    case 3: return a <= b;
    case 4: return a > b;
    default: return a >= b;
  }
}
int64_t vko_count_cmp_u64(const uint64_t *x, uint64_t n, int cmp, uint64_t c) {
  int64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) k += cmp_u64(cmp, x[i], c);
  return k;
}
/* BoolCount, reduce.rs:527-529: bytes != 0 */
int64_t vko_count_bool(const uint8_t *m, uint64_t n) {
  int64_t k = 0;
  for (uint64_t i = 0; i < n; ++i) k += m[i] != 0;
  return k;
}

/* ------------------------------------------------------------- extremes */

/* IEEE 754-2019 minimum / maximum, elementwise.rs:412-438 (Num::min2 / max2 for floats) */
double vko_min2_f64(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == b) return signbit(a) ? a : b;
  return a < b ? a : b;
}
double vko_max2_f64(double a, double b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == b) return signbit(a) ? b : a;
  return a > b ? a : b;
}

/* Extreme<T, MAX>, reduce.rs:379-412: 0 ok, -1 empty ("reduce_min of an empty array") */
int vko_reduce_extreme_f64(const double *x, uint64_t n, int want_max, double *out) {
  if (n == 0) return -1;
  double cur = x[0];
  for (uint64_t i = 1; i < n; ++i) cur = want_max ? vko_max2_f64(cur, x[i]) : vko_min2_f64(cur, x[i]);
  *out = cur;
  return 0;
}
int vko_reduce_extreme_i64(const int64_t *x, uint64_t n, int want_max, int64_t *out) {
  if (n == 0) return -1;
  int64_t cur = x[0];
  for (uint64_t i = 1; i < n; ++i)
    cur = want_max ? (x[i] > cur ? x[i] : cur) : (x[i] < cur ? x[i] : cur);
  *out = cur;
  return 0;
}

int vko_reduce_extreme_u64(const uint64_t *x, uint64_t n, int want_max, uint64_t *out) {
  if (n == 0) return -1;
  uint64_t cur = x[0];
  for (uint64_t i = 1; i < n; ++i)
    cur = want_max ? (x[i] > cur ? x[i] : cur) : (x[i] < cur ? x[i] : cur);
  *out = cur;
  return 0;
}

/* Ordered::less / greater for floats, reduce.rs:183-199: -0.0 ranks below +0.0 */
static inline int less_f64(double a, double b) {
  return a < b || (a == b && signbit(a) && !signbit(b));
}
static inline int greater_f64(double a, double b) {
  return a > b || (a == b && !signbit(a) && signbit(b));
}

/* ArgExtreme<T, MAX>, reduce.rs:414-468: index of the first NaN, else the first index of
 * the extreme. 0 ok, -1 empty. */
int vko_arg_extreme_f64(const double *x, uint64_t n, int want_max, int64_t *out) {
  if (n == 0) return -1;
  int64_t best = -1;
  double bv = 0;
  for (uint64_t i = 0; i < n; ++i) {
    double v = x[i];
    if (isnan(v)) { *out = (int64_t)i; return 0; }
    int better = best < 0 || (want_max ? greater_f64(v, bv) : less_f64(v, bv));
    if (better) { best = (int64_t)i; bv = v; }
  }
  *out = best;
  return 0;
}
int vko_arg_extreme_i64(const int64_t *x, uint64_t n, int want_max, int64_t *out) {
  if (n == 0) return -1;
  int64_t best = 0;
  for (uint64_t i = 1; i < n; ++i)
    if (want_max ? x[i] > x[best] : x[i] < x[best]) best = (int64_t)i;
  *out = best;
  return 0;
}

int vko_arg_extreme_u64(const uint64_t *x, uint64_t n, int want_max, int64_t *out) {
  if (n == 0) return -1;
  int64_t best = 0;
  for (uint64_t i = 1; i < n; ++i)
    if (want_max ? x[i] > x[best] : x[i] < x[best]) best = (int64_t)i;
  *out = best;
  return 0;
}

/* ------------------------------------------------------------ selections */

/* Filter<T, C>, fused.rs:212-246: the rows with x <cmp> c in row order (append_chunks /
 * predicated, select.rs:24-77, restated as a plain ordered copy). Returns the count. */
uint64_t vko_filter_f64(const double *x, uint64_t n, int cmp, double c, double *out) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (cmp_f64(cmp, x[i], c)) out[k++] = x[i];
  return k;
}
uint64_t vko_filter_i64(const int64_t *x, uint64_t n, int cmp, int64_t c, int64_t *out) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (cmp_i64(cmp, x[i], c)) out[k++] = x[i];
  return k;
}

/* compress_append, select.rs:79-92: rows whose mask byte is nonzero, any row width. */
uint64_t vko_compress(const void *x, uint32_t width, const uint8_t *mask, uint64_t n, void *out) {
  uint64_t k = 0;
  const uint8_t *src = (const uint8_t *)x;
  uint8_t *dst = (uint8_t *)out;
  for (uint64_t i = 0; i < n; ++i)
    if (mask[i]) { memcpy(dst + k * width, src + i * width, width); ++k; }
  return k;
}

/* where_append, select.rs:94-104: base + i for nonzero mask rows */
uint64_t vko_where(const uint8_t *mask, uint64_t n, uint64_t base, int64_t *out) {
  uint64_t k = 0;
  for (uint64_t i = 0; i < n; ++i)
    if (mask[i]) out[k++] = (int64_t)(base + i);
  return k;
}

/* ------------------------------------------------------------------ scans */

/* Scan<i64, ScanSum<i64>, BY>, scan.rs:282-326 with ScanSum scan.rs:142-161: inclusive
 * wrapping prefix sum; with starts, a nonzero start byte restarts (row 0 always starts). */
void vko_scan_add_i64(const int64_t *x, uint64_t n, int64_t *out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { acc = i == 0 ? (uint64_t)x[0] : acc + (uint64_t)x[i]; out[i] = (int64_t)acc; }
}
void vko_scan_by_add_i64(const int64_t *x, const uint8_t *starts, uint64_t n, int64_t *out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    acc = (i == 0 || starts[i]) ? (uint64_t)x[i] : acc + (uint64_t)x[i];
    out[i] = (int64_t)acc;
  }
}
/* ScanExtreme over f64 (IEEE minimum/maximum), scan.rs:164-183 */
void vko_scan_extreme_f64(const double *x, const uint8_t *starts, uint64_t n, int want_max, double *out) {
  double acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (i == 0 || (starts && starts[i])) acc = x[i];
    else acc = want_max ? vko_max2_f64(acc, x[i]) : vko_min2_f64(acc, x[i]);
    out[i] = acc;
  }
}
void vko_scan_extreme_i64(const int64_t *x, const uint8_t *starts, uint64_t n, int want_max, int64_t *out) {
  int64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (i == 0 || (starts && starts[i])) acc = x[i];
    else acc = want_max ? (x[i] > acc ? x[i] : acc) : (x[i] < acc ? x[i] : acc);
    out[i] = acc;
  }
}

void vko_scan_extreme_u64(const uint64_t *x, const uint8_t *starts, uint64_t n, int want_max, uint64_t *out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    if (i == 0 || (starts && starts[i])) acc = x[i];
    else acc = want_max ? (x[i] > acc ? x[i] : acc) : (x[i] < acc ? x[i] : acc);
    out[i] = acc;
  }
}

/*
 * f64 scan add: NOT defined by the reference (src/check/testdata/builtins.dd:50-54). This
 * project defines it by the tier 2 candidate order (docs/superpowers/specs/
 * 2026-10-31-verkehrt-tier2-design.md:172): out[i] is the canonical reduction order applied
 * to the prefix x[0..=i], so out[n-1] == sum(x). Parity unpinned by the reference.
 */
void vko_scan_add_f64(const double *x, uint64_t n, double *out) {
  double acc[LANES];
  for (int j = 0; j < LANES; ++j) acc[j] = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t len = i + 1, b = len - len % LANES;
    if (len % LANES == 0) /* row i completes a block: its 16 rows join the accumulators */
      for (int j = 0; j < LANES; ++j) acc[j] = acc[j] + x[len - LANES + j];
    double t[LANES];
    memcpy(t, acc, sizeof t);
    for (int step = LANES / 2; step > 0; step /= 2)
      for (int j = 0; j < step; ++j) t[j] = t[j] + t[j + step];
    double a = t[0];
    for (uint64_t r = b; r < len; ++r) a = a + x[r];
    out[i] = a;
  }
}

/* ---------------------------------------------------------------- gather */

/* Gather<T, BOOL>, index.rs:27-73 with first_out_of_range index.rs:95-111: every index in
 * 0..xlen, else the error names the first bad index value in row order. Returns 0, or -1
 * with *bad set to that value. */
int vko_gather(const void *xs, uint32_t width, uint64_t xlen, const int64_t *idx, uint64_t n,
               void *out, int64_t *bad) {
  for (uint64_t i = 0; i < n; ++i)
    if ((uint64_t)idx[i] >= xlen) { *bad = idx[i]; return -1; }
  const uint8_t *src = (const uint8_t *)xs;
  uint8_t *dst = (uint8_t *)out;
  for (uint64_t i = 0; i < n; ++i) memcpy(dst + i * width, src + (uint64_t)idx[i] * width, width);
  return 0;
}

/* ----------------------------------------------------- buckets (histogram) */

/* checked_keys, index.rs:113-126: -2 negative bucket count, -1 key out of range (*bad). */
static int check_keys(const int64_t *keys, uint64_t n, int64_t buckets, int64_t *bad) {
  if (buckets < 0) return -2;
  for (uint64_t i = 0; i < n; ++i)
    if ((uint64_t)keys[i] >= (uint64_t)buckets) { *bad = keys[i]; return -1; }
  return 0;
}

/* Histogram, index.rs:137-155 */
int vko_histogram(const int64_t *keys, uint64_t n, int64_t buckets, int64_t *out, int64_t *bad) {
  int r = check_keys(keys, n, buckets, bad);
  if (r) return r;
  for (int64_t b = 0; b < buckets; ++b) out[b] = 0;
  for (uint64_t i = 0; i < n; ++i) out[keys[i]] += 1;
  return 0;
}

/* ReduceInto<i64, ScanSum>, index.rs:192-222: wrapping add per bucket, identity 0 */
int vko_reduce_into_add_i64(const int64_t *keys, const int64_t *vals, uint64_t n, int64_t buckets,
                            int64_t *out, int64_t *bad) {
  int r = check_keys(keys, n, buckets, bad);
  if (r) return r;
  for (int64_t b = 0; b < buckets; ++b) out[b] = 0;
  for (uint64_t i = 0; i < n; ++i) out[keys[i]] = (int64_t)((uint64_t)out[keys[i]] + (uint64_t)vals[i]);
  return 0;
}
/* ReduceInto with ScanExtreme on f64, identity +inf (min) / -inf (max), scan.rs:164-183 */
int vko_reduce_into_extreme_f64(const int64_t *keys, const double *vals, uint64_t n, int64_t buckets,
                                int want_max, double *out, int64_t *bad) {
  int r = check_keys(keys, n, buckets, bad);
  if (r) return r;
  for (int64_t b = 0; b < buckets; ++b) out[b] = want_max ? -INFINITY : INFINITY;
  for (uint64_t i = 0; i < n; ++i)
    out[keys[i]] = want_max ? vko_max2_f64(out[keys[i]], vals[i]) : vko_min2_f64(out[keys[i]], vals[i]);
  return 0;
}

/* ReduceInto with ScanExtreme on i64 (index.rs:192-222, scan.rs:185-202): identity
 * i64::MAX (min, Extremes::HIGHEST) / i64::MIN (max, LOWEST) in empty buckets. */
int vko_reduce_into_extreme_i64(const int64_t *keys, const int64_t *vals, uint64_t n, int64_t buckets,
                                int want_max, int64_t *out, int64_t *bad) {
  int r = check_keys(keys, n, buckets, bad);
  if (r) return r;
  for (int64_t b = 0; b < buckets; ++b) out[b] = want_max ? INT64_MIN : INT64_MAX;
  for (uint64_t i = 0; i < n; ++i) {
    int64_t *o = &out[keys[i]];
    if (want_max ? vals[i] > *o : vals[i] < *o) *o = vals[i];
  }
  return 0;
}

/*
 * f64 reduce_into add: NOT defined by the reference (builtins.dd:67-71). This project
 * defines it by the tier 2 candidate order (tier2-design.md:171): with b = n - n%16, row
 * i < b folds into stripe i%16 of its bucket in index order, the 16 stripes of a bucket
 * combine pairwise (+8,+4,+2,+1) from +0.0, then the rows i >= b add to the bucket in index
 * order. With every key equal the result is sum(vals). Parity unpinned by the reference.
 */
int vko_reduce_into_add_f64(const int64_t *keys, const double *vals, uint64_t n, int64_t buckets,
                            double *out, int64_t *bad) {
  int r = check_keys(keys, n, buckets, bad);
  if (r) return r;
  double *stripes = (double *)calloc((size_t)(buckets ? buckets : 1) * LANES, sizeof(double));
  uint64_t b = n - n % LANES;
  for (uint64_t i = 0; i < b; ++i) {
    double *s = &stripes[(uint64_t)keys[i] * LANES + (i % LANES)];
    *s = *s + vals[i];
  }
  for (int64_t k = 0; k < buckets; ++k) {
    double *a = &stripes[(uint64_t)k * LANES];
    for (int step = LANES / 2; step > 0; step /= 2)
      for (int j = 0; j < step; ++j) a[j] = a[j] + a[j + step];
    out[k] = a[0];
  }
  for (uint64_t i = b; i < n; ++i) out[keys[i]] = out[keys[i]] + vals[i];
  free(stripes);
  return 0;
}

/* --------------------------------------------------------------- argsort */

/* Key images, radix.rs:51-101: i64/i32 flip the sign bit, u64/u8 are themselves.
 * f64 (not defined by the reference, builtins.dd:73-77) uses this project's total order:
 * negative values (sign bit set) flip all bits, others flip the sign bit, so
 * -NaN < -inf < ... < -0.0 < +0.0 < ... < +inf < +NaN (NaNs further ordered by payload).
 * Parity of f64 argsort is unpinned by the reference. */
static inline uint64_t image_f64(uint64_t bits) {
  return (bits >> 63) ? ~bits : bits ^ 0x8000000000000000ull;
}

/* The LSD byte radix of radix.rs:187-259 on 64-bit images: one counting pass for every
 * digit, constant digits skipped, a stable scatter per remaining digit; inputs of at most
 * `small` rows use a comparison sort with the position as tiebreak (radix.rs:137-149).
 * Both give the unique stable ascending permutation. */
typedef struct { uint64_t k; uint64_t p; } kp_t;

static int kp_cmp(const void *a, const void *b) {
  const kp_t *x = (const kp_t *)a, *y = (const kp_t *)b;
  if (x->k != y->k) return x->k < y->k ? -1 : 1;
  return x->p < y->p ? -1 : (x->p > y->p);
}

static void argsort_images(const uint64_t *img, uint64_t n, int passes, uint64_t small, int64_t *out) {
  if (n == 0) return;
  if (n <= small) {
    kp_t *v = (kp_t *)malloc(n * sizeof(kp_t));
    for (uint64_t i = 0; i < n; ++i) { v[i].k = img[i]; v[i].p = i; }
    qsort(v, n, sizeof(kp_t), kp_cmp);
    for (uint64_t i = 0; i < n; ++i) out[i] = (int64_t)v[i].p;
    free(v);
    return;
  }
  uint64_t (*counts)[256] = calloc(8, sizeof *counts);
  for (uint64_t i = 0; i < n; ++i)
    for (int d = 0; d < passes; ++d) counts[d][(img[i] >> (8 * d)) & 255]++;
  kp_t *a = (kp_t *)malloc(n * sizeof(kp_t)), *b = (kp_t *)malloc(n * sizeof(kp_t));
  for (uint64_t i = 0; i < n; ++i) { a[i].k = img[i]; a[i].p = i; }
  for (int d = 0; d < passes; ++d) {
    uint64_t first_digit = (img[0] >> (8 * d)) & 255;
    if (counts[d][first_digit] == n) continue; /* constant digit, radix.rs:203-211 */
    uint64_t off[256], sum = 0;
    for (int v = 0; v < 256; ++v) { off[v] = sum; sum += counts[d][v]; }
    for (uint64_t i = 0; i < n; ++i) b[off[(a[i].k >> (8 * d)) & 255]++] = a[i];
    kp_t *t = a; a = b; b = t;
  }
  for (uint64_t i = 0; i < n; ++i) out[i] = (int64_t)a[i].p;
  free(a); free(b); free(counts);
}

/* ArgSort<T>, index.rs:157-173; SMALL per type radix.rs:64-101 */
void vko_argsort_u64(const uint64_t *x, uint64_t n, int64_t *out) {
  argsort_images(x, n, 8, 1024, out);
}
void vko_argsort_i64(const int64_t *x, uint64_t n, int64_t *out) {
  uint64_t *img = (uint64_t *)malloc((n ? n : 1) * 8);
  for (uint64_t i = 0; i < n; ++i) img[i] = (uint64_t)x[i] ^ 0x8000000000000000ull;
  argsort_images(img, n, 8, 1024, out);
  free(img);
}
void vko_argsort_i32(const int32_t *x, uint64_t n, int64_t *out) {
  uint64_t *img = (uint64_t *)malloc((n ? n : 1) * 8);
  for (uint64_t i = 0; i < n; ++i) img[i] = (uint64_t)((uint32_t)x[i] ^ 0x80000000u);
  argsort_images(img, n, 4, 256, out);
  free(img);
}
void vko_argsort_u8(const uint8_t *x, uint64_t n, int64_t *out) {
  uint64_t *img = (uint64_t *)malloc((n ? n : 1) * 8);
  for (uint64_t i = 0; i < n; ++i) img[i] = x[i];
  argsort_images(img, n, 1, 32, out);
  free(img);
}
void vko_argsort_f64(const double *x, uint64_t n, int64_t *out) {
  uint64_t *img = (uint64_t *)malloc((n ? n : 1) * 8);
  for (uint64_t i = 0; i < n; ++i) { uint64_t b; memcpy(&b, &x[i], 8); img[i] = image_f64(b); }
  argsort_images(img, n, 8, 1024, out);
  free(img);
}

/* -------------------------------------------------------------- group_id */

/*
 * group_id: the dense id of each row's key numbered by first occurrence (group.rs:1-5,
 * probe_into group.rs:199-257). The ids do not depend on the hash or the table, so this
 * restatement uses one linear-probe table sized at twice the next power of two of n
 * instead of the reference's growth policy (group.rs:164-197). Keys are compared as
 * 64-bit patterns (Column::bits, group.rs:84-116); a pair key compares both words.
 */
static uint64_t mixh(uint64_t x) { return (x ^ (x >> 32)) * 0x9E3779B97F4A7C15ull; }

typedef struct { uint64_t a, b; int64_t id; } vko_slot;

static void group_ids(const uint64_t *a, const uint64_t *b, uint64_t n, int64_t *out) {
  if (n == 0) return;
  uint64_t slots = 2;
  while (slots < 2 * n) slots <<= 1;
  vko_slot *tab = (vko_slot *)malloc(slots * sizeof *tab);
  for (uint64_t s = 0; s < slots; ++s) tab[s].id = -1;
  int64_t next = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t h = b ? mixh(mixh(a[i]) ^ b[i]) : mixh(a[i]);
    uint64_t s = h & (slots - 1);
    for (;;) {
      vko_slot *t = &tab[s];
      if (t->id < 0) { t->a = a[i]; t->b = b ? b[i] : 0; t->id = next++; out[i] = t->id; break; }
      if (t->a == a[i] && (!b || t->b == b[i])) { out[i] = t->id; break; }
      s = (s + 1) & (slots - 1);
    }
  }
  free(tab);
}

void vko_group_id_u64(const uint64_t *x, uint64_t n, int64_t *out) { group_ids(x, 0, n, out); }
void vko_group_id_pair_u64(const uint64_t *x, const uint64_t *y, uint64_t n, int64_t *out) {
  group_ids(x, y, n, out);
}

/* ------------------------------------------------------------ elementwise */

/* Element type codes, in the order of ScalarType's storage types used here (vk_type):
 * 0 f64, 1 f32, 2 i64, 3 u64, 4 i32, 5 u8, 6 bool (u8 storage). */
static int type_width(int t) { return t <= 1 ? (t == 0 ? 8 : 4) : t <= 3 ? 8 : t == 4 ? 4 : 1; }

/* Float ops, elementwise.rs:398-414 and :443-504 (Num for floats, Signed, Float): 0 add, 1 sub,
 * 2 mul, 3 div, 4 min (IEEE minimum), 5 max, 6 neg, 7 abs, 8 sqrt. b is unused by the unary ops. */
void vko_ew_f64(int op, const double *a, const double *b, uint64_t n, double *out) {
  for (uint64_t i = 0; i < n; ++i) {
    double x = a[i], y = b ? b[i] : 0.0, r;
    switch (op) {
      case 0: r = x + y; break;
      case 1: r = x - y; break;
      case 2: r = x * y; break;
      case 3: r = x / y; break;
      case 4: r = vko_min2_f64(x, y); break;
      case 5: r = vko_max2_f64(x, y); break;
      case 6: r = -x; break;
      case 7: r = fabs(x); break;
      default: r = sqrt(x); break;
    }
    out[i] = r;
  }
}
static float min2_f32(float a, float b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == b) return signbit(a) ? a : b;
  return a < b ? a : b;
}
static float max2_f32(float a, float b) {
  if (isnan(a) || isnan(b)) return NAN;
  if (a == b) return signbit(a) ? b : a;
  return a > b ? a : b;
}
void vko_ew_f32(int op, const float *a, const float *b, uint64_t n, float *out) {
  for (uint64_t i = 0; i < n; ++i) {
    float x = a[i], y = b ? b[i] : 0.0f, r;
    switch (op) {
      case 0: r = x + y; break;
      case 1: r = x - y; break;
      case 2: r = x * y; break;
      case 3: r = x / y; break;
      case 4: r = min2_f32(x, y); break;
      case 5: r = max2_f32(x, y); break;
      case 6: r = -x; break;
      case 7: r = fabsf(x); break;
      default: r = sqrtf(x); break;
    }
    out[i] = r;
  }
}

static int64_t load_int(const void *p, int t, uint64_t i) {
  switch (t) {
    case 2: return ((const int64_t *)p)[i];
    case 3: return (int64_t)((const uint64_t *)p)[i];
    case 4: return ((const int32_t *)p)[i];
    default: return ((const uint8_t *)p)[i];
  }
}
static void store_int(void *p, int t, uint64_t i, uint64_t v) {
  switch (t) {
    case 2: case 3: ((uint64_t *)p)[i] = v; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
    default: ((uint8_t *)p)[i] = (uint8_t)v; break;
  }
}

/* Integer ops on type t (2 i64, 3 u64, 4 i32, 5 u8), elementwise.rs:380-397 (wrapping Num),
 * :449-470 (wrapping neg/abs), :536-545 (bits), :612-737 (div_const, truncating, MIN / -1
 * wraps): 0 add, 1 sub, 2 mul, 3 div_const by b[i] (nonzero), 4 min, 5 max, 6 neg, 7 abs,
 * 8 and, 9 or, 10 xor, 11 not. Results keep the low bytes of the exact value. */
void vko_ew_int(int op, int t, const void *a, const void *b, uint64_t n, void *out) {
  const int sgn = t == 2 || t == 4, bitsz = 8 * type_width(t);
  for (uint64_t i = 0; i < n; ++i) {
    int64_t x = load_int(a, t, i), y = b ? load_int(b, t, i) : 0;
    uint64_t ux = (uint64_t)x, uy = (uint64_t)y, r;
    switch (op) {
      case 0: r = ux + uy; break;
      case 1: r = ux - uy; break;
      case 2: r = ux * uy; break;
      case 3: {
        uint64_t mx = sgn && x < 0 ? 0 - ux : ux, my = sgn && y < 0 ? 0 - uy : uy;
        if (!sgn && bitsz < 64) { mx &= (1ull << bitsz) - 1; my &= (1ull << bitsz) - 1; }
        uint64_t q = mx / my;
        r = sgn && ((x < 0) != (y < 0)) ? 0 - q : q;
        break;
      }
      case 4: r = sgn ? (uint64_t)(x < y ? x : y) : (ux < uy ? ux : uy); break;
      case 5: r = sgn ? (uint64_t)(x > y ? x : y) : (ux > uy ? ux : uy); break;
      case 6: r = 0 - ux; break;
      case 7: r = x < 0 ? 0 - ux : ux; break;
      case 8: r = ux & uy; break;
      case 9: r = ux | uy; break;
      case 10: r = ux ^ uy; break;
      default: r = ~ux; break;
    }
    store_int(out, t, i, r);
  }
}

/* Rust `as` casts (elementwise.rs:574-610, CastOp): float to integer truncates toward zero,
 * saturates at the bounds and maps NaN to 0; integer to integer keeps the low bytes of the
 * sign- or zero-extended value; to a float rounds to nearest even; bool reads its byte. */
void vko_cast(int from, int to, const void *x, uint64_t n, void *out) {
  for (uint64_t i = 0; i < n; ++i) {
    if (to <= 1) {
      double d;
      float f;
      if (from == 0) { d = ((const double *)x)[i]; f = (float)d; }
      else if (from == 1) { f = ((const float *)x)[i]; d = f; }
      else if (from == 3) { uint64_t u = ((const uint64_t *)x)[i]; d = (double)u; f = (float)u; }
      else { int64_t v = load_int(x, from == 6 ? 5 : from, i); d = (double)v; f = (float)v; }
      if (to == 0) ((double *)out)[i] = d; else ((float *)out)[i] = f;
      continue;
    }
    uint64_t r;
    if (from <= 1) {
      double v = from == 0 ? ((const double *)x)[i] : ((const float *)x)[i];
      int bitsz = 8 * type_width(to), sgn = to == 2 || to == 4;
      double lo = sgn ? -ldexp(1.0, bitsz - 1) : 0.0, hi = sgn ? ldexp(1.0, bitsz - 1) : ldexp(1.0, bitsz);
      if (isnan(v)) r = 0;
      else if (v <= lo) r = sgn ? (uint64_t)1 << (bitsz - 1) : 0;
      else if (v >= hi) r = sgn ? ((uint64_t)1 << (bitsz - 1)) - 1 : (bitsz == 64 ? ~0ull : (1ull << bitsz) - 1);
      else r = sgn ? (uint64_t)(int64_t)trunc(v) : (uint64_t)trunc(v);
    } else {
      r = (uint64_t)load_int(x, from == 6 ? 5 : from, i);
    }
    store_int(out, to, i, r);
  }
}
