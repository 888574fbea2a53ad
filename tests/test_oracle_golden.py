"""Pins the CPU oracle (oracle/vk_oracle.c) before anything is checked against it.

Every check here comes from outside the oracle's own formulas: the reference's golden
vectors and known-answer tests (cited file:line in the reference), closed forms, or brute
force written differently from the oracle (Python loops, sorted(), dicts, numpy), chosen
so that a dropped term, a wrong index or a swapped operand fails one of them.
"""
import math
import struct

import numpy as np
import pytest

from oracle import oracle as O


def bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def fbits(x) -> int:
    return struct.unpack("<I", struct.pack("<f", float(x)))[0]


# ------------------------------------------------------------------------ RNG

def test_splitmix_and_rand():
    # src/kernels/tests.rs:1167-1181
    assert O.splitmix64(0) == 0xE220A8397B1DCDAF
    r = O.rand_i64(3, 7)
    assert int(r[1]) & (2**64 - 1) == O.splitmix64(8)
    x = O.rand_f64(1000, 1)
    assert np.all((x >= 0) & (x < 1))
    want = (O.splitmix64(1) >> 11) * 2.0**-53
    assert bits(x[0]) == bits(want)
    # tests/programs/random.dd:4-9
    assert list(O.rand_i64(2, 0)) == [-2152535657050944081, -7995527694508729151]


def test_splitmix_closed_form():
    # The published SplitMix64 finalizer, written independently in Python.
    def sm(x):
        m = 2**64 - 1
        z = (x + 0x9E3779B97F4A7C15) & m
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return z ^ (z >> 31)
    for v in [1, 2, 12345, 2**63, 2**64 - 1]:
        assert O.splitmix64(v) == sm(v)
    got = O.rand_i64(50, -3)
    assert [int(g) & (2**64 - 1) for g in got] == [sm((-3 + i) & (2**64 - 1)) for i in range(50)]


# ------------------------------------------------------ canonical float order

def varied(n):
    # src/kernels/tests.rs:667-675
    out = []
    for i in range(n):
        unit = (O.splitmix64(i) >> 11) * 2.0**-53
        out.append((unit - 0.5) * 10.0 ** ((i % 9) * 2))
    return np.array(out)


def reference_sum(xs):
    # The spec's order bullet by bullet, src/kernels/tests.rs:678-697, in Python.
    n = len(xs)
    b = n - n % 16
    acc = [0.0] * 16
    for j in range(16):
        i = j
        while i < b:
            acc[j] += float(xs[i])
            i += 16
    for step in [8, 4, 2, 1]:
        for j in range(step):
            acc[j] += acc[j + step]
    for x in xs[b:]:
        acc[0] += float(x)
    return acc[0]


def test_canonical_sum_golden_values():
    # src/kernels/tests.rs:761-791
    xs = [1e16, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.1, 1.2, 1.3, 1.4, 1.5, 1.6,
          -1e16, 1.7, 1.8, 1.9, 2.1, 2.2, 2.3, 2.4, 2.5, 2.6, 2.7, 2.8, 2.9, 3.1, 3.2, 3.3,
          0.01, 0.02, 0.03]
    assert bits(O.canonical_sum_f64(xs)) == 0x4049147AE147AE14
    naive = 0.0
    for x in xs:
        naive += x
    assert naive == 49.56
    ys = [16777216.0, 0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5, 11.5, 12.5, 13.5, 14.5,
          1.0, 0.25, 0.75, 1.25, 1.75, 2.25, 2.75, 3.25, 3.75, 4.25, 4.75, 5.25, 5.75, 6.25, 6.75, 7.25,
          3.0, 3.0, 3.0]
    assert fbits(O.canonical_sum_f32(ys)) == 0x4B80005A
    assert fbits(np.float32(O.canonical_sum_f64(ys))) == 0x4B800059


def test_canonical_sum_follows_the_spec_order():
    # src/kernels/tests.rs:699-718
    xs = [1e16 if i == 0 else (-1e16 if i == 16 else 1.0) for i in range(35)]
    assert bits(O.canonical_sum_f64(xs)) == bits(reference_sum(xs))
    naive = 0.0
    for x in xs:
        naive += x
    assert bits(naive) != bits(reference_sum(xs))
    for n in [0, 1, 15, 16, 17, 35, 100, 1000]:
        v = varied(n)
        assert bits(O.canonical_sum_f64(v)) == bits(reference_sum(v)), n


def test_canonical_identities_and_product():
    # src/kernels/tests.rs:720-725, 816-825
    assert bits(O.canonical_sum_f64([])) == bits(0.0)
    assert bits(O.canonical_product_f64([])) == bits(1.0)
    assert O.canonical_product_f64([2.0, 3.0, 4.0]) == 24.0


def test_canonical_state_update_finish_matches_one_shot():
    # src/kernels/tests.rs:794-814: tiles of multiples of 16 give the one-shot bits
    xs = varied(1000)
    whole = O.canonical_fold_f64("sum", xs)
    assert bits(O.canonical_fold_f64("sum", xs, cuts=[64, 384, 992])) == bits(whole)
    assert bits(whole) == bits(reference_sum(xs))
    with pytest.raises(O.OracleError, match="partial block"):
        O.canonical_fold_f64("sum", xs, cuts=[17, 64])


def test_dot_sumsq_condsum_are_sums_of_row_values():
    # fused.rs:72-210: the fused kernels equal reduce_add over the materialized rows
    x, y = varied(300), varied(300)[::-1].copy()
    assert bits(O.dot_f64(x, y)) == bits(reference_sum([a * b for a, b in zip(x, y)]))
    assert bits(O.sumsq_f64(x)) == bits(reference_sum([a * a for a in x]))
    picked = [a if a > 0.25 else -7.0 for a in x]
    assert bits(O.condsum_f64(x, "gt", 0.25, -7.0)) == bits(reference_sum(picked))


def reference_sum_f32(vals, mul=False):
    """The canonical order at f32 written bullet by bullet with numpy float32 scalars:
    accumulator j takes rows j, j + 16, ... of the whole blocks, the accumulators combine
    pairwise (+8, +4, +2, +1), the tail rows fold into accumulator 0 (reduce.rs:40-91)."""
    f = np.float32
    op = (lambda a, b: f(a * b)) if mul else (lambda a, b: f(a + b))
    v = [f(x) for x in vals]
    b = len(v) - len(v) % 16
    acc = [f(1.0) if mul else f(0.0)] * 16
    for i in range(b):
        acc[i % 16] = op(acc[i % 16], v[i])
    step = 8
    while step:
        for j in range(step):
            acc[j] = op(acc[j], acc[j + step])
        step //= 2
    a = acc[0]
    for x in v[b:]:
        a = op(a, x)
    return a


def test_canonical_fold_f32_kinds():
    # reduce.rs:40-91 and fused.rs:72-210 instantiated at f32 (Reducible, reduce.rs:567-575)
    ys = [16777216.0, 0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5, 11.5, 12.5, 13.5, 14.5,
          1.0, 0.25, 0.75, 1.25, 1.75, 2.25, 2.75, 3.25, 3.75, 4.25, 4.75, 5.25, 5.75, 6.25, 6.75, 7.25,
          3.0, 3.0, 3.0]
    assert fbits(O.canonical_fold_f32("sum", ys)) == 0x4B80005A  # tests.rs:761-791
    rng = np.random.default_rng(5)
    with np.errstate(all="ignore"):
        for n in [0, 1, 15, 16, 17, 35, 100, 333]:
            x = (rng.standard_normal(n) * 10.0 ** rng.integers(-3, 4, n)).astype(np.float32)
            y = rng.random(n).astype(np.float32)
            if n > 20:
                x[3], x[7], x[19] = np.float32(1e30), np.float32(-0.0), np.float32(1e-40)
            f = np.float32
            assert fbits(O.canonical_fold_f32("sum", x)) == fbits(reference_sum_f32(x)), n
            assert fbits(O.canonical_fold_f32("dot", x, y)) == fbits(
                reference_sum_f32([f(a * b) for a, b in zip(x, y)])), n
            assert fbits(O.canonical_fold_f32("sumsq", x)) == fbits(reference_sum_f32([f(a * a) for a in x])), n
            picked = [a if a <= f(0.25) else f(-7.0) for a in x]
            assert fbits(O.canonical_fold_f32("condsum", x, cmp="le", c=0.25, e=-7.0)) == fbits(
                reference_sum_f32(picked)), n
            z = (1.0 + rng.random(n) * 0.01).astype(np.float32)
            assert fbits(O.canonical_fold_f32("mul", z)) == fbits(reference_sum_f32(z, mul=True)), n
    # identities and a NaN
    assert fbits(O.canonical_fold_f32("sum", [])) == 0
    assert fbits(O.canonical_fold_f32("mul", [])) == fbits(1.0)
    assert math.isnan(O.canonical_fold_f32("sum", [1.0, float("nan"), 2.0]))


@pytest.mark.parametrize("dt", [np.int64, np.uint64, np.int32])
def test_fused_integer_sums_wrap(dt):
    # Dot, SumSq, CondSum (fused.rs:72-210) with IntSum (reduce.rs:325-377): wrapping products
    # in the element type (elementwise.rs:380-397), widened, summed with wrapping; Python ints
    w = np.dtype(dt).itemsize * 8
    signed = np.dtype(dt).kind == "i"
    rng = np.random.default_rng(9)
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, 500, dtype=dt, endpoint=True)
    y = rng.integers(info.min, info.max, 500, dtype=dt, endpoint=True)
    x[:4] = [info.min, info.max, 0, 1 if not signed else -1]
    sum_bits = 64

    def elem(v):  # wrap to the element type
        v &= (1 << w) - 1
        return v - (1 << w) if signed and v >> (w - 1) else v

    def total(vals):
        s = sum(vals) & ((1 << sum_bits) - 1)
        return s - (1 << 64) if signed and s >> 63 else s

    xs, ys = [int(v) for v in x], [int(v) for v in y]
    assert O.fused_sum_int("dot", x, y) == total(elem(a * b) for a, b in zip(xs, ys))
    assert O.fused_sum_int("sumsq", x) == total(elem(a * a) for a in xs)
    c = int(x[10])
    e = int(info.min) if signed else int(info.max)
    for cmp, fn in [("eq", lambda a: a == c), ("ne", lambda a: a != c), ("lt", lambda a: a < c),
                    ("le", lambda a: a <= c), ("gt", lambda a: a > c), ("ge", lambda a: a >= c)]:
        assert O.fused_sum_int("condsum", x, cmp=cmp, c=c, e=e) == total(a if fn(a) else e for a in xs), cmp
    assert O.fused_sum_int("dot", x[:0], y[:0]) == 0


# --------------------------------------------------- 1e6 / 1e7 goldens

def test_readme_example_goldens():
    # README.md:55-69: plan outputs of the example program
    x = O.rand_f64(1_000_000, 1)
    y = O.rand_f64(1_000_000, 2)
    assert repr(O.dot_f64(x, y)) == "249724.39002046228"
    assert O.count_cmp(x, "gt", 0.5) == 499521
    big = O.filter_cmp(x, "gt", 0.99)
    assert big.size == 9851 and big.nbytes == 78808


def test_ten_million_goldens():
    # docs/measurements/2026-11-03-m5/bench-corgi-tier1.txt:5,17; bench-tier1.txt:110,149,186,227
    x = O.rand_f64(10_000_000, 1)
    y = O.rand_f64(10_000_000, 2)
    assert repr(O.canonical_sum_f64(x)) == "4998770.953740556"
    assert repr(O.dot_f64(x, y)) == "2499047.034215136"
    assert O.arg_extreme(x, False) == 5618431
    assert repr(float(O.reduce_extreme(x, False))) == "2.5056049013372217e-08"
    assert [O.count_cmp(x, "gt", c) for c in (0.99, 0.5, 0.01)] == [99695, 4998931, 9899130]
    assert O.filter_cmp(x, "gt", 0.01).size == 9899130


# --------------------------------------------------------- integer kernels

def test_integer_reductions_widen_and_wrap():
    # src/kernels/tests.rs:834-858
    assert O.sum_int(np.array([2**31 - 1, 1], np.int32)) == 2147483648
    assert O.sum_int(np.array([200, 100], np.uint8)) == 300
    assert O.sum_int(np.array([2**63 - 1, 1], np.int64)) == -(2**63)
    r = np.random.default_rng(5).integers(-2**62, 2**62, 1000, dtype=np.int64)
    assert O.sum_int(r) == int(np.sum(r, dtype=np.int64))  # numpy wraps too


def test_count_cmp_brute_force():
    x = np.array([0.0, -0.0, 1.0, np.nan, -np.inf, np.inf, 0.5, 0.5])
    for cmp, f in [("eq", lambda a, b: a == b), ("ne", lambda a, b: a != b), ("lt", lambda a, b: a < b),
                   ("le", lambda a, b: a <= b), ("gt", lambda a, b: a > b), ("ge", lambda a, b: a >= b)]:
        for c in [0.5, 0.0, np.nan]:
            assert O.count_cmp(x, cmp, c) == sum(1 for v in x if f(float(v), float(c))), (cmp, c)


def test_count_cmp_i64_brute_force():
    # count_<cmp>_i64 (fused.rs:142-179) for all six comparisons, with the extremes of i64 as
    # rows and as constants: Python's own integer comparisons are the independent check.
    lo, hi = -2**63, 2**63 - 1
    x = np.array([lo, hi, 0, -1, 1, lo + 1, hi - 1, 5, 5, -5], np.int64)
    rng = np.random.default_rng(7)
    x = np.concatenate([x, rng.integers(-6, 7, 200).astype(np.int64)])
    ops = {"eq": lambda a, b: a == b, "ne": lambda a, b: a != b, "lt": lambda a, b: a < b,
           "le": lambda a, b: a <= b, "gt": lambda a, b: a > b, "ge": lambda a, b: a >= b}
    for cmp, f in ops.items():
        for c in [lo, hi, 0, -1, 5, lo + 1]:
            want = sum(1 for v in x.tolist() if f(v, c))
            assert O.count_cmp(x, cmp, c) == want, (cmp, c)
    assert O.count_cmp(np.array([], np.int64), "eq", 0) == 0


def test_u64_extremes_and_counts_compare_unsigned():
    # u64 rows order as unsigned integers (src/kernels/tests.rs:1322-1338: max 1 << 63)
    x = np.array([3, 1 << 63, 5, 2**64 - 1, 3], np.uint64)
    assert O.reduce_extreme(x, True) == 2**64 - 1 and O.reduce_extreme(x, False) == 3
    assert O.arg_extreme(x, True) == 3 and O.arg_extreme(x, False) == 0
    rng = np.random.default_rng(13)
    y = np.concatenate([x, rng.integers(0, 2**64, 300, dtype=np.uint64)])
    Y = [int(v) for v in y]
    assert O.reduce_extreme(y, True) == max(Y) and O.reduce_extreme(y, False) == min(Y)
    assert O.arg_extreme(y, True) == Y.index(max(Y)) and O.arg_extreme(y, False) == Y.index(min(Y))
    ops = {"eq": lambda a, b: a == b, "ne": lambda a, b: a != b, "lt": lambda a, b: a < b,
           "le": lambda a, b: a <= b, "gt": lambda a, b: a > b, "ge": lambda a, b: a >= b}
    for cmp, f in ops.items():
        for c in [0, 3, 1 << 63, 2**64 - 1]:
            assert O.count_cmp(y, cmp, c) == sum(1 for v in Y if f(v, c)), (cmp, c)


def test_integer_products_wrap():
    # src/kernels/tests.rs:847-850,1337: reduce_mul wraps in the element type, identity 1
    assert O.product_int(np.array([16, 17], np.uint8)) == 16
    assert O.product_int(np.array([], np.int32)) == 1
    assert O.product_int(np.array([2**63, 2], np.uint64)) == 0
    assert O.product_int(np.array([2.0, 3.0, 4.0]).astype(np.int64)) == 24
    # brute force: Python integers reduced modulo 2^width, read back signed for i64/i32
    rng = np.random.default_rng(8)
    for dt, w in [(np.int64, 64), (np.uint64, 64), (np.int32, 32), (np.uint8, 8)]:
        info = np.iinfo(dt)
        for n in [1, 2, 7, 300]:
            x = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
            p = 1
            for v in x.tolist():
                p = (p * v) % (1 << w)
            if info.min < 0 and p >> (w - 1):
                p -= 1 << w
            assert O.product_int(x) == p, (dt, n)


# --------------------------------------------------------------- extremes

def test_min_max_follow_ieee_minimum():
    # src/kernels/tests.rs:861-872
    assert math.isnan(O.reduce_extreme(np.array([1.0, np.nan, 0.0]), False))
    assert bits(O.reduce_extreme(np.array([0.0, -0.0]), False)) == bits(-0.0)
    assert bits(O.reduce_extreme(np.array([-0.0, 0.0]), True)) == bits(0.0)
    with pytest.raises(O.OracleError, match="reduce_min of an empty array"):
        O.reduce_extreme(np.array([], np.float64), False)
    assert bits(O.min2(-0.0, 0.0)) == bits(-0.0) and bits(O.min2(0.0, -0.0)) == bits(-0.0)
    assert bits(O.max2(-0.0, 0.0)) == bits(0.0) and bits(O.max2(0.0, -0.0)) == bits(0.0)


def test_argmin_argmax_tie_and_nan_rules():
    # src/kernels/tests.rs:874-891
    a = lambda v, mx: O.arg_extreme(np.array(v, np.float64), mx)
    assert a([3.0, 1.0, 1.0], False) == 1
    assert a([3.0, np.nan, np.nan], False) == 1
    assert a([0.0, -0.0], False) == 1
    assert a([-0.0, 0.0], True) == 1
    assert a([1.0, 5.0, 5.0], True) == 1
    assert a([1.0, np.nan, 5.0], True) == 1
    with pytest.raises(O.OracleError, match="argmax of an empty array"):
        a([], True)
    assert O.arg_extreme(np.array([4, 9, 9, 2], np.int64), True) == 1


def test_extremes_brute_force():
    rng = np.random.default_rng(1)
    x = rng.normal(size=1000)
    assert O.reduce_extreme(x, False) == x.min() and O.reduce_extreme(x, True) == x.max()
    assert O.arg_extreme(x, False) == int(np.argmin(x)) and O.arg_extreme(x, True) == int(np.argmax(x))
    i = rng.integers(-5, 5, 1000)
    assert O.arg_extreme(i, True) == int(np.argmax(i))  # numpy returns the first max too


# -------------------------------------------------------------- selections

def test_compress_and_where():
    # src/kernels/tests.rs:911-948
    assert list(O.compress(np.array([1.0, 2.0, 3.0]), [0, 1, 1])) == [2.0, 3.0]
    assert O.compress(np.array([1.0, 2.0]), [0, 0]).size == 0
    with pytest.raises(O.OracleError, match="length mismatch: 2 values vs 1 mask rows"):
        O.compress(np.array([1.0, 2.0]), [1])
    assert list(O.compress(np.array([1, 0, 1], np.uint8), [1, 1, 0])) == [1, 0]
    assert list(O.compress(np.array([7, 8, 9], np.uint8), [1, 0, 1])) == [7, 9]
    assert list(O.where([0, 1, 1])) == [1, 2]
    assert list(O.where([0, 1, 1], base=100)) == [101, 102]


def test_selection_brute_force():
    rng = np.random.default_rng(2)
    x = rng.random(5000)
    m = (rng.random(5000) < 0.3).astype(np.uint8)
    assert np.array_equal(O.compress(x, m), x[m != 0])
    assert np.array_equal(O.where(m), np.nonzero(m)[0])
    for c in (0.99, 0.5, 0.01):
        assert np.array_equal(O.filter_cmp(x, "gt", c), x[x > c])
    xi = rng.integers(-100, 100, 5000)
    assert np.array_equal(O.filter_cmp(xi, "le", 3), xi[xi <= 3])


# ---------------------------------------------------------------------- scans

def test_scans():
    # src/kernels/tests.rs:1416-1490
    assert list(O.scan_add_i64([1, 2, 3, 4])) == [1, 3, 6, 10]
    assert list(O.scan_add_i64([2**63 - 1, 1])) == [2**63 - 1, -(2**63)]
    assert list(O.scan_by_add_i64([1, 2, 3, 4, 5], [0, 0, 1, 0, 1])) == [1, 3, 3, 7, 5]
    assert list(O.scan_by_add_i64([1, 2], [1, 0])) == [1, 3]
    nan = O.scan_extreme(np.array([1.0, np.nan, 3.0]), True)
    assert nan[0] == 1.0 and np.isnan(nan[1]) and np.isnan(nan[2])
    z = O.scan_extreme(np.array([-0.0, 0.0]), True)
    assert [bits(v) for v in z] == [bits(-0.0), bits(0.0)]
    assert list(O.scan_extreme(np.array([3, 1, 2], np.int64), False)) == [3, 1, 1]


def test_scan_brute_force_and_last_equals_sum():
    rng = np.random.default_rng(3)
    x = rng.integers(-2**62, 2**62, 3000, dtype=np.int64)
    s = O.scan_add_i64(x)
    assert np.array_equal(s, np.cumsum(x, dtype=np.int64))
    assert int(s[-1]) == O.sum_int(x)  # tests.rs:1492-1510
    starts = (rng.random(3000) < 1 / 16).astype(np.uint8)
    want, acc = [], 0
    for i in range(3000):
        acc = int(x[i]) if (i == 0 or starts[i]) else acc + int(x[i])
        acc = (acc + 2**63) % 2**64 - 2**63
        want.append(acc)
    assert list(O.scan_by_add_i64(x, starts)) == want


def test_scan_extremes_segmented_brute_force():
    # ScanExtreme with starts (scan.rs:164-202): a start restarts at the row itself; u64 rows
    # compare unsigned, i64 signed
    rng = np.random.default_rng(12)
    n = 500
    st = (rng.random(n) < 0.1).astype(np.uint8)
    for x in [rng.integers(0, 2**64, n, dtype=np.uint64), rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)]:
        for mx in (False, True):
            want, acc = [], None
            for i, v in enumerate(x.tolist()):
                acc = v if (i == 0 or st[i]) else (max(acc, v) if mx else min(acc, v))
                want.append(acc)
            assert O.scan_extreme(x, mx, st).tolist() == want
            assert O.scan_extreme(x, mx).tolist() == (np.maximum.accumulate(x) if mx else np.minimum.accumulate(x)).tolist()


def test_scan_add_f64_defined_order():
    # Build-defined order (tier 2 candidate): out[i] = canonical sum of x[0..=i].
    v = varied(100)
    out = O.scan_add_f64(v)
    for i in range(100):
        assert bits(out[i]) == bits(reference_sum(v[: i + 1])), i


# --------------------------------------------------------------------- gather

def test_gather():
    # src/kernels/tests.rs:1512-1541
    assert list(O.gather(np.array([10, 20, 30]), [2, 0, 2])) == [30, 10, 30]
    assert O.gather(np.array([10, 20, 30]), []).size == 0
    assert list(O.gather(np.array([1, 0, 0], np.uint8), [1, 0, 1, 0])) == [0, 1, 0, 1]
    assert list(O.gather(np.array([1.5, 2.5]), [1, 1, 0])) == [2.5, 2.5, 1.5]
    with pytest.raises(O.OracleError, match="gather index 3 out of range for length 3"):
        O.gather(np.array([10, 20, 30]), [0, 3])
    with pytest.raises(O.OracleError, match="gather index -1 out of range for length 3"):
        O.gather(np.array([10, 20, 30]), [-1, 0])
    with pytest.raises(O.OracleError, match="gather index 0 out of range for length 0"):
        O.gather(np.array([], np.int64), [0])
    # the FIRST bad value in row order is named
    with pytest.raises(O.OracleError, match="gather index 7 out of range"):
        O.gather(np.array([10, 20, 30]), [0, 7, -1, 9])


# -------------------------------------------------------------------- buckets

def test_histogram_and_reduce_into():
    # src/kernels/tests.rs:1573-1600, 1625-1640; tests/programs/grouping.dd (buckets)
    assert list(O.histogram([0, 2, 2, 3], 5)) == [1, 0, 2, 1, 0]
    assert O.histogram([], 0).size == 0
    assert list(O.histogram([], 3)) == [0, 0, 0]
    for key, n in [(3, 3), (-1, 3), (0, 0)]:
        first = 0 if n == 0 else key
        with pytest.raises(O.OracleError, match=f"^bucket key {first} out of range for {n} buckets$"):
            O.histogram([0, key], n)
    with pytest.raises(O.OracleError, match="^negative bucket count -1$"):
        O.histogram([], -1)
    assert list(O.reduce_into([1, 0, 1], np.array([5, 7, 9], np.int64), 3)) == [7, 14, 0]
    keys = [1, 0, 1, 3, 1]
    assert list(O.histogram(keys, 5)) == [1, 3, 0, 1, 0]
    assert list(O.reduce_into(keys, np.array([5, 7, 9, 2, 1], np.int64), 4)) == [7, 15, 0, 2]
    mn = O.reduce_into(keys, np.array([1.5, 2.5, 0.5, 9.0, 4.0]), 4, "min")
    assert list(mn) == [2.5, 0.5, math.inf, 9.0]
    with pytest.raises(O.OracleError, match="^length mismatch: 2 keys for 1 values$"):
        O.reduce_into([0, 0], np.array([1], np.int64), 1)


def ieee_max(a, b):
    # IEEE 754-2019 maximum written from the standard: NaN wins, +0.0 above -0.0
    if math.isnan(a) or math.isnan(b):
        return math.nan
    if a == b:
        return b if math.copysign(1.0, a) < 0 else a
    return a if a > b else b


def ieee_min(a, b):
    if math.isnan(a) or math.isnan(b):
        return math.nan
    if a == b:
        return a if math.copysign(1.0, a) < 0 else b
    return a if a < b else b


def test_reduce_into_extremes_identities():
    # src/kernels/tests.rs:1653-1672: an empty bucket holds the operator's identity
    assert list(O.reduce_into([1], np.array([2.5]), 3, "min")) == [math.inf, 2.5, math.inf]
    mx = O.reduce_into([1], np.array([2.5]), 3, "max")
    assert list(mx) == [-math.inf, 2.5, -math.inf]
    assert list(O.reduce_into([1], np.array([4], np.int64), 2, "max")) == [-2**63, 4]
    assert list(O.reduce_into([1], np.array([4], np.int64), 2, "min")) == [2**63 - 1, 4]


@pytest.mark.parametrize("op", ["min", "max"])
def test_reduce_into_extremes_brute_force(op):
    # f64: NaN, +-0 and +-inf among the values, folded by ieee_min/ieee_max per bucket in row
    # order from the identity; i64: the extremes of i64 among the values.
    rng = np.random.default_rng(9 + (op == "max"))
    n, nb = 2000, 37
    keys = rng.integers(0, nb, n)
    v = rng.normal(size=n)
    special = rng.integers(0, 10, n)
    v = np.where(special == 0, np.nan, np.where(special == 1, 0.0, np.where(special == 2, -0.0,
                 np.where(special == 3, math.inf, np.where(special == 4, -math.inf, v)))))
    keys[:3] = [5, 5, 6]
    v[:3] = [0.0, -0.0, -0.0]  # bucket 5: {+0, -0}; bucket 6 starts with -0
    f = ieee_max if op == "max" else ieee_min
    want = [(-math.inf if op == "max" else math.inf)] * nb
    for k, x in zip(keys.tolist(), v.tolist()):
        want[k] = f(want[k], x)
    got = O.reduce_into(keys, v, nb + 2, op)
    assert len(got) == nb + 2
    for k in range(nb):
        w = want[k]
        assert (math.isnan(w) and math.isnan(got[k])) or bits(w) == bits(float(got[k])), (op, k)
    assert list(got[nb:]) == [-math.inf if op == "max" else math.inf] * 2  # empty buckets
    iv = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64, endpoint=True)
    iv[:4] = [-2**63, 2**63 - 1, -2**63, 2**63 - 1]
    wi = [(-2**63 if op == "max" else 2**63 - 1)] * nb
    for k, x in zip(keys.tolist(), iv.tolist()):
        wi[k] = max(wi[k], x) if op == "max" else min(wi[k], x)
    assert list(O.reduce_into(keys, iv, nb, op)) == wi


def test_reduce_into_f64_add_defined_order():
    # Build-defined order (tier 2 candidate): every key equal gives sum's bits.
    v = varied(1000)
    assert bits(O.reduce_into(np.zeros(1000, np.int64), v, 1)[0]) == bits(reference_sum(v))
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 5, 1000)
    got = O.reduce_into(keys, v, 5)
    b = 1000 - 1000 % 16
    for k in range(5):
        st = [0.0] * 16
        for i in range(b):
            if keys[i] == k:
                st[i % 16] += float(v[i])
        for step in [8, 4, 2, 1]:
            for j in range(step):
                st[j] += st[j + step]
        a = st[0]
        for i in range(b, 1000):
            if keys[i] == k:
                a += float(v[i])
        assert bits(got[k]) == bits(a), k


# -------------------------------------------------------------------- argsort

def test_argsort_is_stable():
    # src/kernels/tests.rs:1839-1852
    assert list(O.argsort(np.array([3, 1, 3, 1], np.int64))) == [1, 3, 0, 2]
    assert O.argsort(np.array([], np.int64)).size == 0
    assert list(O.argsort(np.array([0, -1, -2**31, 2**31 - 1], np.int32))) == [2, 1, 0, 3]
    assert list(O.argsort(np.array([2**63, 1, 2**63, 0], np.uint64))) == [3, 1, 0, 2]
    assert list(O.argsort(np.array([200, 5, 200, 5, 0], np.uint8))) == [4, 1, 3, 0, 2]


@pytest.mark.parametrize("n", [0, 1, 17, 1024, 1025, 5000])
def test_argsort_brute_force(n):
    # The stable ascending permutation is unique: sorted() on (key, position).
    rng = np.random.default_rng(n)
    for x in [rng.integers(0, 2**64, n, dtype=np.uint64),
              rng.integers(-3, 3, n).astype(np.int64),
              (rng.integers(0, 2**64, n, dtype=np.uint64) & np.uint64(0xFF00)),  # constant digits
              rng.integers(-2**31, 2**31, n).astype(np.int32),
              rng.integers(0, 256, n).astype(np.uint8)]:
        want = sorted(range(n), key=lambda i: (int(x[i]), i))
        assert list(O.argsort(x)) == want


def test_argsort_f64_defined_order():
    # Build-defined total order: -NaN < -inf < ... < -0.0 < +0.0 < ... < +inf < +NaN.
    x = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, -np.nan, 2.5, -0.0, 0.0])
    order = O.argsort(x)
    def rank(v):
        b = bits(v)
        neg = b >> 63
        return (0, -(b & (2**63 - 1))) if neg else (1, b)
    want = sorted(range(x.size), key=lambda i: (rank(float(x[i])), i))
    assert list(order) == want
    assert list(x[order][1:8].astype(str)) == ["-inf", "-1.0", "-0.0", "-0.0", "0.0", "0.0", "1.0"]


# ------------------------------------------------------------------- group_id

def test_group_id_goldens():
    # tests/programs/grouping.dd (group_id)
    assert list(O.group_id(np.array([7, -1, 7, 3, -1, 7], np.int64))) == [0, 1, 0, 2, 1, 0]
    assert list(O.group_id(np.array([1, 2, 1, 2], np.int64), np.array([2, 1, 2, 2], np.int64))) == [0, 1, 0, 2]
    assert list(O.group_id(np.array([5, 5, 9], np.uint8), np.array([-1, -1, -1], np.int32))) == [0, 0, 1]
    assert O.group_id(np.array([], np.int64)).size == 0


def test_group_id_brute_force():
    rng = np.random.default_rng(6)
    for d in [1, 3, 100, 5000]:
        x = rng.integers(-2**63, 2**63 - 1, d, dtype=np.int64)[rng.integers(0, d, 5000)]
        first = {}
        want = [first.setdefault(int(v), len(first)) for v in x]
        assert list(O.group_id(x)) == want


def test_group_sorted_program():
    # tests/programs/grouping.dd (group_sorted): scan_by + compress through the oracle
    k = np.array([1, 1, 2, 2, 2, 5], np.int64)
    v = np.array([1, 2, 3, 4, 5, 6], np.int64)
    n = 6
    prev = np.concatenate([[0], k[:-1]])
    s = ((k != prev) | (np.arange(n) == 0)).astype(np.uint8)
    nxt = O.gather(k, np.minimum(np.arange(n) + 1, n - 1))
    e = ((k != nxt) | (np.arange(n) == n - 1)).astype(np.uint8)
    assert list(O.compress(O.scan_by_add_i64(v, s), e)) == [3, 12, 6]
    ids = O.scan_add_i64(s.astype(np.int64)) - 1
    assert list(O.reduce_into(ids, v, O.count_bool(s))) == [3, 12, 6]
    assert list(O.compress(k, s)) == [1, 2, 5]


# ---------------------------------------------------------------- elementwise

def _f32_round(v: int) -> float:
    # an integer rounded to the nearest float32, ties to even, by integer arithmetic
    m, sgn = abs(v), -1 if v < 0 else 1
    e = m.bit_length() - 24
    if e > 0:
        q, r = divmod(m, 1 << e)
        half = 1 << (e - 1)
        if r > half or (r == half and q & 1):
            q += 1
        m = q << e
    return float(np.float32(sgn * float(m)))


def test_elementwise_floats_follow_ieee():
    rng = np.random.default_rng(21)
    x = np.concatenate([rng.normal(size=300) * 1e3, [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-310, -1e-310]])
    y = np.concatenate([rng.normal(size=300), [-0.0, 0.0, np.inf, 1.0, 2.0, np.nan, 3.0]])
    with np.errstate(all="ignore"):
        for op, f in [("add", np.add), ("sub", np.subtract), ("mul", np.multiply), ("div", np.divide)]:
            got = O.elementwise(op, "f64", x, y)
            want = f(x, y)
            assert np.array_equal(np.isnan(got), np.isnan(want))
            assert np.array_equal(got[~np.isnan(got)].view(np.uint64), want[~np.isnan(want)].view(np.uint64)), op
        got = O.elementwise("sqrt", "f64", np.abs(x))
        assert np.array_equal(got.view(np.uint64), np.sqrt(np.abs(x)).view(np.uint64))
        assert bits(O.elementwise("neg", "f64", [0.0])[0]) == bits(-0.0)
        assert bits(O.elementwise("abs", "f64", [-0.0])[0]) == bits(0.0)
        for op, f in [("min", ieee_min), ("max", ieee_max)]:
            got = O.elementwise(op, "f64", x, y)
            for g, a, b in zip(got.tolist(), x.tolist(), y.tolist()):
                w = f(a, b)
                assert (math.isnan(w) and math.isnan(g)) or bits(w) == bits(g), (op, a, b)
        x32, y32 = x.astype(np.float32), y.astype(np.float32)
        got = O.elementwise("div", "f32", x32, y32)
        want = x32 / y32
        assert np.array_equal(got[~np.isnan(want)].view(np.uint32), want[~np.isnan(want)].view(np.uint32))
        assert np.array_equal(O.elementwise("lt", "f64", x, y), (x < y).astype(np.uint8))
        assert O.elementwise("eq", "f64", [np.nan], [np.nan])[0] == 0
        assert O.elementwise("ne", "f64", [np.nan], [np.nan])[0] == 1


@pytest.mark.parametrize("t,w,signed", [("i64", 64, True), ("u64", 64, False), ("i32", 32, True), ("u8", 8, False)])
def test_elementwise_integers_wrap(t, w, signed):
    # Python integers reduced modulo 2^w (elementwise.rs:380-397, 449-470, 536-545, 612-737)
    rng = np.random.default_rng(22 + w)
    dt = O.NP[t]
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, 400, dtype=dt, endpoint=True)
    y = rng.integers(info.min, info.max, 400, dtype=dt, endpoint=True)
    x[:4] = [info.min, info.max, 0, info.min]
    y[:4] = [info.max, info.min, 1, -1 if signed else 1]
    y[y == 0] = 3

    def norm(v):
        v %= 1 << w
        return v - (1 << w) if signed and v >> (w - 1) else v

    def trunc_div(a, b):
        q = abs(a) // abs(b)
        return norm(-q if (a < 0) != (b < 0) else q)

    X, Y = x.tolist(), y.tolist()
    for op, f in [("add", lambda a, b: norm(a + b)), ("sub", lambda a, b: norm(a - b)),
                  ("mul", lambda a, b: norm(a * b)), ("div_const", trunc_div), ("min", min), ("max", max),
                  ("and", lambda a, b: norm(a & b)), ("or", lambda a, b: norm(a | b)),
                  ("xor", lambda a, b: norm(a ^ b))]:
        assert O.elementwise(op, t, x, y).tolist() == [f(a, b) for a, b in zip(X, Y)], op
    assert O.elementwise("not", t, x).tolist() == [norm(~a) for a in X]
    if signed:
        assert O.elementwise("neg", t, x).tolist() == [norm(-a) for a in X]
        assert O.elementwise("abs", t, x).tolist() == [norm(abs(a)) for a in X]
        assert O.elementwise("div_const", t, [info.min], [-1]).tolist() == [info.min]  # MIN / -1 wraps
    with pytest.raises(O.OracleError, match="^div_const: division by zero$"):
        O.elementwise("div_const", t, x, 0)
    # tests/programs/arith.dd (div_sort): [5, 3, 9, 3, 7] / 2, negated dividend, negated divisor
    if signed:
        c = np.array([5, 3, 9, 3, 7], dt)
        assert O.elementwise("div_const", t, c, 2).tolist() == [2, 1, 4, 1, 3]
        assert O.elementwise("div_const", t, -c, 2).tolist() == [-2, -1, -4, -1, -3]
        assert O.elementwise("div_const", t, c, -2).tolist() == [-2, -1, -4, -1, -3]


def test_select():
    c = np.array([1, 0, 1, 0], np.uint8)
    assert O.select(c, "i64", [1, 2, 3, 4], [10, 20, 30, 40]).tolist() == [1, 20, 3, 40]
    assert O.select(c, "f64", 1.5, [10.0, 20.0, 30.0, 40.0]).tolist() == [1.5, 20.0, 1.5, 40.0]


def test_casts_follow_rust_as():
    # float to int saturates and maps NaN to 0; int to int keeps the low bytes; to float rounds
    # to nearest even (elementwise.rs:574-610)
    f = np.array([np.nan, np.inf, -np.inf, 1e300, -1e300, 2.9, -2.9, -0.5, 255.5, 256.0, -1.0,
                  2.0**63, -2.0**63, 2.0**64, 2.0**31 - 0.5, -2.0**31 - 1])
    bounds = {"i64": (-2**63, 2**63 - 1), "u64": (0, 2**64 - 1), "i32": (-2**31, 2**31 - 1), "u8": (0, 255)}
    for t, (lo, hi) in bounds.items():
        want = [0 if math.isnan(v) else max(lo, min(hi, math.trunc(v))) if math.isfinite(v) else (hi if v > 0 else lo)
                for v in f.tolist()]
        assert O.cast(f, "f64", t).tolist() == want, t
        assert O.cast(f.astype(np.float32), "f32", t).tolist() == [
            0 if math.isnan(v) else max(lo, min(hi, math.trunc(v))) if math.isfinite(v) else (hi if v > 0 else lo)
            for v in f.astype(np.float32).astype(np.float64).tolist()], t
    i = np.array([-1, 2**40 + 7, -2**63, 2**63 - 1, 255, 256, -129], np.int64)
    assert O.cast(i, "i64", "u8").tolist() == [v % 256 for v in i.tolist()]
    assert O.cast(i, "i64", "i32").tolist() == [((v % 2**32) ^ 2**31) - 2**31 for v in i.tolist()]
    assert O.cast(i, "i64", "u64").tolist() == [v % 2**64 for v in i.tolist()]
    assert O.cast(np.array([-1, 7], np.int32), "i32", "u64").tolist() == [2**64 - 1, 7]  # sign-extends
    assert O.cast(np.array([200], np.uint8), "u8", "i32").tolist() == [200]             # zero-extends
    assert O.cast(np.array([1, 0], np.uint8), "bool", "f64").tolist() == [1.0, 0.0]
    big = np.array([2**53 + 1, -(2**53 + 1), 2**63 - 1, 2**24 + 1, 2**62 + 2**38 + 1], np.int64)
    assert [bits(v) for v in O.cast(big, "i64", "f64").tolist()] == [bits(float(v)) for v in big.tolist()]
    assert [fbits(v) for v in O.cast(big, "i64", "f32").tolist()] == [fbits(_f32_round(v)) for v in big.tolist()]
    u = np.array([2**64 - 1, 2**63 + 2**11 + 1], np.uint64)
    assert [bits(v) for v in O.cast(u, "u64", "f64").tolist()] == [bits(float(v)) for v in u.tolist()]
    assert [fbits(v) for v in O.cast(np.array([1e40, 1.0000001], np.float64), "f64", "f32").tolist()] == \
        [fbits(np.inf), fbits(np.float32(1.0000001))]
