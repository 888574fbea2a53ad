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
 * - gather with a vectorised bounds fold and vpgatherqq (src/kernels/x86/gather.rs:44-102);
 * - the scalar entries the reference has no intrinsics for: segmented scan_by add
 *   (scan.rs:282-326), the LSD byte radix argsort (index/radix.rs:123-259) and the hash
 *   group_id (group.rs:164-257), plus reduce_into_add_f64 in this backend's defined order.
 */
#include <immintrin.h>
#include <stdint.h>
#include <stdlib.h>
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

/* filter_gt_f64, select_loops.rs:105-204 + avx512.rs:144-165: a count pass sizes the output
 * (append_chunks reserves exactly, select.rs:24-57), then AVX-512 compress stores. */
uint64_t vkb_filter_gt_f64(const double *x, uint64_t n, double c, double *out) {
  const __m512d cv = _mm512_set1_pd(c);
  uint64_t k = 0, i = 0;
  volatile int64_t sized = vkb_count_gt_f64(x, n, c);
  (void)sized;
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

/* scan_by_add_i64, scan.rs:282-326: a nonzero start byte restarts the sum (row 0 always). */
void vkb_scan_by_add_i64(const int64_t *x, const uint8_t *starts, uint64_t n, int64_t *out) {
  uint64_t acc = 0;
  for (uint64_t i = 0; i < n; ++i) {
    acc = (i == 0 || starts[i]) ? (uint64_t)x[i] : acc + (uint64_t)x[i];
    out[i] = (int64_t)acc;
  }
}

/* ---------------------------------------------------------------- argsort */

/* radix_into, radix.rs:187-259, with u32 positions (n < 2^32): one counting pass for every
 * byte digit of the key images, constant digits skipped, (image, position) pairs moved through
 * two buffers by stable scatters, the last pass writing i64 positions. `kind` picks the image:
 * 0 u64 (itself), 1 i64 (sign bit flipped, radix.rs:51-101), 2 f64 (this backend's total
 * order: negative values flip every bit, others the sign bit). */
typedef struct { uint64_t k; uint32_t p; } vkb_pair;

static inline uint64_t vkb_image(int kind, uint64_t b) {
  if (kind == 1) return b ^ 0x8000000000000000ull;
  if (kind == 2) return (b >> 63) ? ~b : b ^ 0x8000000000000000ull;
  return b;
}

int vkb_argsort_64(int kind, const uint64_t *x, uint64_t n, int64_t *out) {
  if (n == 0) return 0;
  uint64_t (*counts)[256] = calloc(8, sizeof *counts);
  if (!counts) return -1;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t k = vkb_image(kind, x[i]);
    for (int d = 0; d < 8; ++d) counts[d][(k >> (8 * d)) & 255]++;
  }
  int active[8], passes = 0;
  uint64_t first = vkb_image(kind, x[0]);
  for (int d = 0; d < 8; ++d)
    if (counts[d][(first >> (8 * d)) & 255] != n) active[passes++] = d;
  if (passes == 0) {
    for (uint64_t i = 0; i < n; ++i) out[i] = (int64_t)i;
    free(counts);
    return 0;
  }
  vkb_pair *a = malloc(n * sizeof *a), *b = passes > 2 ? malloc(n * sizeof *b) : NULL;
  if (!a || (passes > 2 && !b)) { free(a); free(b); free(counts); return -1; }
  uint64_t off[256];
  for (int q = 0; q < passes; ++q) {
    const int d = active[q];
    uint64_t sum = 0;
    for (int v = 0; v < 256; ++v) { off[v] = sum; sum += counts[d][v]; }
    if (q == 0 && passes == 1) {
      for (uint64_t i = 0; i < n; ++i) out[off[(vkb_image(kind, x[i]) >> (8 * d)) & 255]++] = (int64_t)i;
    } else if (q == 0) {
      for (uint64_t i = 0; i < n; ++i) {
        uint64_t k = vkb_image(kind, x[i]);
        vkb_pair *dst = &a[off[(k >> (8 * d)) & 255]++];
        dst->k = k;
        dst->p = (uint32_t)i;
      }
    } else if (q == passes - 1) {
      for (uint64_t i = 0; i < n; ++i) out[off[(a[i].k >> (8 * d)) & 255]++] = a[i].p;
    } else {
      for (uint64_t i = 0; i < n; ++i) b[off[(a[i].k >> (8 * d)) & 255]++] = a[i];
      vkb_pair *t = a; a = b; b = t;
    }
  }
  free(a); free(b); free(counts);
  return 0;
}

/* --------------------------------------------------------------- group_id */

/* probe_into, group.rs:199-257 with mix (group.rs:39-46) and grow (group.rs:164-197): slots
 * of (key, tag = id + 1, u32), linear probing on the high bits of the hash, a first table of
 * at most 256 slots that grows right after an insert leaves it more than half full, 2^3-fold
 * when nearly every row so far opened a group, never past nextpow2(2n). The multiplier is a
 * fixed odd constant (the reference draws one per call; the ids do not depend on it). */
typedef struct { uint64_t k; uint32_t tag; } vkb_slot;

static inline uint64_t vkb_mix(uint64_t x) { return (x ^ (x >> 32)) * 0x9E3779B97F4A7C15ull; }

static uint64_t vkb_pow2_at_least(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

int vkb_group_id_64(const uint64_t *x, uint64_t n, int64_t *out) {
  if (n == 0) return 0;
  uint64_t max_slots = vkb_pow2_at_least(2 * n);
  uint64_t slots = max_slots < 256 ? (max_slots < 2 ? 2 : max_slots) : 256;
  int shift = 64 - __builtin_ctzll(slots);
  vkb_slot *tab = calloc(slots, sizeof *tab);
  if (!tab) return -1;
  uint64_t groups = 0;
  for (uint64_t row = 0; row < n; ++row) {
    const uint64_t k = x[row], mask = slots - 1;
    uint64_t i = vkb_mix(k) >> shift;
    for (;;) {
      if (tab[i].tag == 0) {
        tab[i].k = k;
        tab[i].tag = (uint32_t)(groups + 1);
        out[row] = (int64_t)groups++;
        break;
      }
      if (tab[i].k == k) { out[row] = (int64_t)tab[i].tag - 1; break; }
      i = (i + 1) & mask;
    }
    if (2 * groups > slots) {
      const int distinct = 8 * groups >= 7 * (row + 1);
      uint64_t grown = slots << (distinct ? 3 : 1);
      if (grown > max_slots) grown = max_slots;
      const int new_shift = shift - __builtin_ctzll(grown / slots);
      vkb_slot *nt = calloc(grown, sizeof *nt);
      if (!nt) { free(tab); return -1; }
      for (uint64_t s = 0; s < slots; ++s) {
        if (tab[s].tag == 0) continue;
        uint64_t j = vkb_mix(tab[s].k) >> new_shift;
        while (nt[j].tag != 0) j = (j + 1) & (grown - 1);
        nt[j] = tab[s];
      }
      free(tab);
      tab = nt;
      slots = grown;
      shift = new_shift;
    }
  }
  free(tab);
  return 0;
}

/* reduce_into_add_f64 in this backend's defined order (the tier 2 candidate): row i below
 * b = n - n % 16 folds into stripe i % 16 of its bucket in row order, the 16 stripes of a
 * bucket combine pairwise, then rows from b add in row order. Keys are in 0..buckets. */
int vkb_reduce_into_add_f64(const int64_t *keys, const double *vals, uint64_t n, uint64_t buckets, double *out) {
  double *st = calloc((buckets ? buckets : 1) * 16, sizeof(double));
  if (!st) return -1;
  const uint64_t b = n - n % 16;
  for (uint64_t i = 0; i < b; ++i) {
    double *s = &st[(uint64_t)keys[i] * 16 + (i & 15)];
    *s = *s + vals[i];
  }
  for (uint64_t k = 0; k < buckets; ++k) {
    double *a = &st[k * 16];
    for (int step = 8; step > 0; step /= 2)
      for (int j = 0; j < step; ++j) a[j] = a[j] + a[j + step];
    out[k] = a[0];
  }
  for (uint64_t i = b; i < n; ++i) out[keys[i]] = out[keys[i]] + vals[i];
  free(st);
  return 0;
}
