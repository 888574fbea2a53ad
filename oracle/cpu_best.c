/*
 * cpu_best.c -- TEST / BASELINE INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A C restatement of verkehrt's best x86 CPU backend ("intrinsics, falling back to auto, on
 * x86-64-v4"), used only as the timed CPU baseline of bench.py and checked bit for bit
 * against the oracle (vk_oracle.c) by the tests before any timing. Built with
 * -O3 -march=x86-64-v4 -mavx512vbmi2 -ffp-contract=off (no FMA).
 *
 * - auto entries (src/kernels/auto.rs:11-121): the scalar loops compiled for x86-64-v4; the
 *   canonical 16-accumulator order is written as two 8-lane vectors (reduce.rs:40-62, the
 *   "2 zmm for f64" tiling of the design's canonical order);
 * - filter / compress / where with AVX-512 compress stores (src/kernels/x86/compress/
 *   avx512.rs:48-279): a count pass sizes the output, then vcompresspd per 8 rows;
 * - gather with a vectorised bounds fold and vpgatherqq (src/kernels/x86/gather.rs:44-102).
 */
#include <immintrin.h>
#include <stdint.h>
#include <string.h>

/* canonical sum / dot, reduce.rs:40-91 */
static double finish16(__m512d a0, __m512d a1, const double *tail, int tail_len) {
  double acc[16];
  _mm512_storeu_pd(acc, a0);
  _mm512_storeu_pd(acc + 8, a1);
  for (int step = 8; step > 0; step /= 2)
    for (int j = 0; j < step; ++j) acc[j] = acc[j] + acc[j + step];
  double a = acc[0];
  for (int j = 0; j < tail_len; ++j) a = a + tail[j];
  return a;
}

double vkb_sum_f64(const double *x, uint64_t n) {
  __m512d a0 = _mm512_setzero_pd(), a1 = _mm512_setzero_pd();
  uint64_t b = n - n % 16;
  for (uint64_t i = 0; i < b; i += 16) {
    a0 = _mm512_add_pd(a0, _mm512_loadu_pd(x + i));
    a1 = _mm512_add_pd(a1, _mm512_loadu_pd(x + i + 8));
  }
  return finish16(a0, a1, x + b, (int)(n - b));
}

double vkb_dot_f64(const double *x, const double *y, uint64_t n) {
  __m512d a0 = _mm512_setzero_pd(), a1 = _mm512_setzero_pd();
  uint64_t b = n - n % 16;
  for (uint64_t i = 0; i < b; i += 16) {
    a0 = _mm512_add_pd(a0, _mm512_mul_pd(_mm512_loadu_pd(x + i), _mm512_loadu_pd(y + i)));
    a1 = _mm512_add_pd(a1, _mm512_mul_pd(_mm512_loadu_pd(x + i + 8), _mm512_loadu_pd(y + i + 8)));
  }
  double tail[16];
  for (uint64_t i = b; i < n; ++i) tail[i - b] = x[i] * y[i];
  return finish16(a0, a1, tail, (int)(n - b));
}

/* count_gt_f64, fused.rs:142-179 */
int64_t vkb_count_gt_f64(const double *x, uint64_t n, double c) {
  const __m512d cv = _mm512_set1_pd(c);
  int64_t k = 0;
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8) k += __builtin_popcount(_mm512_cmp_pd_mask(_mm512_loadu_pd(x + i), cv, _CMP_GT_OQ));
  for (; i < n; ++i) k += x[i] > c;
  return k;
}

/* reduce_min_f64 (IEEE minimum) and argmax_f64 (first NaN, else first max), auto entries */
static inline double min2(double a, double b) {
  if (a != a || b != b) return __builtin_nan("");
  if (a == b) return __builtin_signbit(a) ? a : b;
  return a < b ? a : b;
}
double vkb_min_f64(const double *x, uint64_t n) {
  double cur = x[0];
  for (uint64_t i = 1; i < n; ++i) cur = min2(cur, x[i]);
  return cur;
}
int64_t vkb_argmax_f64(const double *x, uint64_t n) {
  int64_t best = 0;
  double bv = x[0];
  if (bv != bv) return 0;
  for (uint64_t i = 1; i < n; ++i) {
    double v = x[i];
    if (v != v) return (int64_t)i;
    if (v > bv || (v == bv && !__builtin_signbit(v) && __builtin_signbit(bv))) { bv = v; best = (int64_t)i; }
  }
  return best;
}

/* scan_add_i64, scan.rs:282-326 (the scalar loop is the reference's fastest) */
void vkb_scan_add_i64(const int64_t *x, uint64_t n, int64_t *out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) { acc += (uint64_t)x[i]; out[i] = (int64_t)acc; }
}

/* filter_gt_f64 with AVX-512 compress, select_loops.rs:105-204 + avx512.rs:144-165 */
uint64_t vkb_filter_gt_f64(const double *x, uint64_t n, double c, double *out) {
  const __m512d cv = _mm512_set1_pd(c);
  uint64_t k = 0, i = 0;
  for (; i + 8 <= n; i += 8) {
    __m512d v = _mm512_loadu_pd(x + i);
    __mmask8 m = _mm512_cmp_pd_mask(v, cv, _CMP_GT_OQ);
    _mm512_mask_compressstoreu_pd(out + k, m, v);
    k += __builtin_popcount(m);
  }
  for (; i < n; ++i) if (x[i] > c) out[k++] = x[i];
  return k;
}

/* compress_f64 with a bool mask, avx512.rs:75-88 */
uint64_t vkb_compress_f64(const double *x, const uint8_t *mask, uint64_t n, double *out) {
  uint64_t k = 0, i = 0;
  for (; i + 8 <= n; i += 8) {
    __m128i mb = _mm_loadl_epi64((const __m128i *)(mask + i));
    __mmask8 m = _mm_test_epi8_mask(mb, mb);
    _mm512_mask_compressstoreu_pd(out + k, m, _mm512_loadu_pd(x + i));
    k += __builtin_popcount(m);
  }
  for (; i < n; ++i) if (mask[i]) out[k++] = x[i];
  return k;
}

/* gather of 64-bit rows with a bounds fold first, x86/gather.rs:44-102. Returns 0, or -1
 * with *bad set to the first out-of-range index in row order. */
int vkb_gather_f64(const double *xs, uint64_t xlen, const int64_t *idx, uint64_t n, double *out,
                   int64_t *bad) {
  const __m512i lim = _mm512_set1_epi64((long long)xlen);
  __mmask8 any = 0;
  uint64_t i = 0;
  for (; i + 8 <= n; i += 8)
    any |= _mm512_cmp_epu64_mask(_mm512_loadu_si512(idx + i), lim, _MM_CMPINT_NLT);
  int tail_bad = 0;
  for (uint64_t r = i; r < n; ++r) tail_bad |= (uint64_t)idx[r] >= xlen;
  if (any || tail_bad) {
    for (uint64_t r = 0; r < n; ++r)
      if ((uint64_t)idx[r] >= xlen) { *bad = idx[r]; return -1; }
  }
  for (i = 0; i + 8 <= n; i += 8)
    _mm512_storeu_pd(out + i, _mm512_i64gather_pd(_mm512_loadu_si512(idx + i), xs, 8));
  for (; i < n; ++i) out[i] = xs[idx[i]];
  return 0;
}
