"""Parity of the hip variant with the oracle (the CPU restatement of the reference's scalar
entries), through the C ABI, on seeded inputs.

Modelled on the reference's tests/isa_variants.rs: row counts around vector and tile widths
(0, 1, 3, 4, 7, 8, 15, 16, 17, 31, 32, 63, 64, 65, 300, 1000, 4500), views at offsets 1..7,
float inputs mixing finite values, subnormals, +-0, +-inf and NaN, tiles cut at 64/384/992
for reductions and 37/100 for selections and scans, one forced out-of-range index or key in
some cases, and repeated keys for group_id. The bar is bit-exact results (a NaN result only
needs to be NaN) and identical error strings.
"""
import math
import struct

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

LENGTHS = [0, 1, 3, 4, 7, 8, 15, 16, 17, 31, 32, 63, 64, 65, 300, 1000, 4500]
CMPS = ["eq", "ne", "lt", "le", "gt", "ge"]


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def S(t, v):
    from verkehrt_amd import Scalar
    return Scalar(t, v)


def fbits(v):
    return struct.unpack("<Q", struct.pack("<d", float(v)))[0]


def same_float(want, got):
    if math.isnan(want):
        return math.isnan(got)
    return fbits(want) == fbits(got)


def same_array(want, got):
    want, got = np.asarray(want), np.asarray(got)
    if want.shape != got.shape:
        return False
    if want.dtype.kind == "f":
        wn, gn = np.isnan(want), np.isnan(got)
        return bool(np.array_equal(wn, gn) and np.array_equal(want[~wn].view(np.uint64) if want.dtype == np.float64 else want[~wn],
                                                               got[~gn].view(np.uint64) if got.dtype == np.float64 else got[~gn]))
    return bool(np.array_equal(want, got))


def floats(rng, n, special):
    """isa_variants.rs:55-81: finite values in -1e3..1e3, wide normals/subnormals, specials."""
    finite = rng.uniform(-1e3, 1e3, n)
    wide = np.ldexp(rng.uniform(-1, 1, n), rng.integers(-1074, 1000, n))
    if special:
        specials = np.array([0.0, -0.0, np.inf, -np.inf, np.nan])
        pick = rng.choice(3, n, p=[12 / 21, 4 / 21, 5 / 21])
        sp = specials[rng.integers(0, 5, n)]
    else:
        specials = np.array([0.0, -0.0])
        pick = rng.choice(3, n, p=[30 / 35, 4 / 35, 1 / 35])
        sp = specials[rng.integers(0, 2, n)]
    return np.where(pick == 0, finite, np.where(pick == 1, wide, sp))


def view(D, host, offset, type=None):
    """A device view of `host` starting `offset` rows into a larger buffer."""
    host = np.asarray(host)
    pad = np.concatenate([np.zeros(offset, host.dtype), host]) if offset else host
    a = D.array(pad, type)
    return a.slice(offset, host.size) if offset else a


def err_of(f):
    from verkehrt_amd import KernelError
    try:
        return ("ok", f())
    except (KernelError, O.OracleError) as e:
        return ("err", str(e))


# ------------------------------------------------------------------ example

def test_readme_example_plan(D):
    # README.md:55-69: rand_f64 x2, dot_f64, count_gt_f64, length(filter_gt_f64)
    x = D.call("rand_f64", [S("i64", 1_000_000), S("i64", 1)])
    y = D.call("rand_f64", [S("i64", 1_000_000), S("i64", 2)])
    assert np.array_equal(x.to_numpy().view(np.uint64), O.rand_f64(1_000_000, 1).view(np.uint64))
    dot = D.call("dot_f64", [x, y])
    assert repr(dot.value) == "249724.39002046228"
    assert D.call("count_gt_f64", [x, S("f64", 0.5)]).value == 499521
    big = D.call("filter_gt_f64", [x, S("f64", 0.99)])
    assert len(big) == 9851 and big.nbytes == 78808
    assert D.call("length_f64", [big]).value == 9851


def test_ten_million_goldens(D):
    n = 10_000_000
    x = D.call("rand_f64", [S("i64", n), S("i64", 1)])
    y = D.call("rand_f64", [S("i64", n), S("i64", 2)])
    assert repr(D.call("reduce_add_f64", [x]).value) == "4998770.953740556"
    assert repr(D.call("dot_f64", [x, y]).value) == "2499047.034215136"
    assert D.call("argmin_f64", [x]).value == 5618431
    assert repr(D.call("reduce_min_f64", [x]).value) == "2.5056049013372217e-08"
    for c, want in [(0.99, 99695), (0.5, 4998931), (0.01, 9899130)]:
        assert D.call("count_gt_f64", [x, S("f64", c)]).value == want
        assert len(D.call("filter_gt_f64", [x, S("f64", c)])) == want


def test_rand_i64(D):
    r = D.call("rand_i64", [S("i64", 2), S("i64", 0)]).to_numpy()
    assert list(r) == [-2152535657050944081, -7995527694508729151]
    r = D.call("rand_i64", [S("i64", 5000), S("i64", -3)]).to_numpy()
    assert np.array_equal(r, O.rand_i64(5000, -3))
    e = err_of(lambda: D.call("rand_f64", [S("i64", -1), S("i64", 0)]))
    assert e == ("err", "negative length -1")


# ------------------------------------------------------- canonical reductions

@pytest.mark.parametrize("special", [False, True])
def test_canonical_float_reductions(D, special):
    rng = np.random.default_rng(11 + special)
    for n in LENGTHS:
        for off in (0, 3):
            x = floats(rng, n, special)
            y = floats(rng, n, special)
            dx, dy = view(D, x, off), view(D, y, (off + 2) % 8)
            assert same_float(O.canonical_sum_f64(x), D.call("reduce_add_f64", [dx]).value), n
            assert same_float(O.dot_f64(x, y), D.call("dot_f64", [dx, dy]).value), n
            assert same_float(O.sumsq_f64(x), D.call("sumsq_f64", [dx]).value), n
            c = float(x[n // 2]) if n else 0.5
            for cmp in ("gt", "le", "ne"):
                want = O.condsum_f64(x, cmp, c, -7.5)
                got = D.call(f"condsum_{cmp}_f64", [dx, S("f64", c), S("f64", -7.5)]).value
                assert same_float(want, got), (n, cmp)
            z = 1.0 + rng.uniform(-1e-3, 1e-3, n)
            assert same_float(O.canonical_product_f64(z), D.call("reduce_mul_f64", [view(D, z, off)]).value)


def test_canonical_reducer_tiles(D):
    # reduce.rs contract: tiles of multiples of 16 reproduce the one-shot bits
    rng = np.random.default_rng(5)
    x = floats(rng, 1000, False)
    y = floats(rng, 1000, False)
    dx, dy = D.array(x), D.array(y)
    for name, args, want in [("reduce_add_f64", lambda a, b: [dx.slice(a, b - a)], O.canonical_sum_f64(x)),
                             ("dot_f64", lambda a, b: [dx.slice(a, b - a), dy.slice(a, b - a)], O.dot_f64(x, y))]:
        r = D.reducer(name)
        for a, b in [(0, 64), (64, 384), (384, 992), (992, 1000)]:
            r.update(args(a, b))
        assert same_float(want, r.finish().value)
    r = D.reducer("reduce_add_f64")
    r.update([dx.slice(0, 17)])
    e = err_of(lambda: r.update([dx.slice(17, 47)]))
    assert e == ("err", "canonical update after a partial block")


def test_golden_canonical_vector(D):
    xs = [1e16, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9, 1.1, 1.2, 1.3, 1.4, 1.5, 1.6,
          -1e16, 1.7, 1.8, 1.9, 2.1, 2.2, 2.3, 2.4, 2.5, 2.6, 2.7, 2.8, 2.9, 3.1, 3.2, 3.3,
          0.01, 0.02, 0.03]
    assert fbits(D.call("reduce_add_f64", [D.array(np.array(xs))]).value) == 0x4049147AE147AE14
    assert fbits(D.call("reduce_add_f64", [D.array(np.zeros(0))]).value) == fbits(0.0)
    assert D.call("reduce_mul_f64", [D.array(np.array([2.0, 3.0, 4.0]))]).value == 24.0


# ------------------------------------------------------ streaming reductions

def test_integer_sums_widen_and_wrap(D):
    assert D.call("reduce_add_i32", [D.array(np.array([2**31 - 1, 1], np.int32))]) == S("i64", 2147483648)
    assert D.call("reduce_add_u8", [D.array(np.array([200, 100], np.uint8))]) == S("u64", 300)
    assert D.call("reduce_add_i64", [D.array(np.array([2**63 - 1, 1], np.int64))]) == S("i64", -(2**63))
    rng = np.random.default_rng(3)
    for n in LENGTHS + [1 << 20]:
        for dt, t in [(np.int64, "i64"), (np.uint64, "u64"), (np.int32, "i32"), (np.uint8, "u8")]:
            info = np.iinfo(dt)
            x = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
            got = D.call(f"reduce_add_{t}", [view(D, x, n % 7)]).value
            want = O.sum_int(x)
            assert got == want, (t, n)


def test_counts(D):
    rng = np.random.default_rng(4)
    for n in LENGTHS + [123457]:
        x = floats(rng, n, True)
        dx = view(D, x, n % 5)
        for cmp in CMPS:
            for c in [0.5, 0.0, float("nan"), float(x[0]) if n else 1.0]:
                assert D.call(f"count_{cmp}_f64", [dx, S("f64", c)]).value == O.count_cmp(x, cmp, c)
        xi = rng.integers(-5, 5, n).astype(np.int64)
        for cmp in CMPS:
            assert D.call(f"count_{cmp}_i64", [view(D, xi, 1), S("i64", 1)]).value == O.count_cmp(xi, cmp, 1)
        m = (rng.random(n) < 0.3).astype(np.uint8)
        assert D.call("count", [view(D, m, 2, "bool")]).value == O.count_bool(m)


def test_extremes(D):
    rng = np.random.default_rng(6)
    for n in LENGTHS:
        for special in (False, True):
            x = floats(rng, n, special)
            dx = view(D, x, n % 7)
            for mx in (False, True):
                name = "max" if mx else "min"
                want = err_of(lambda: float(O.reduce_extreme(x, mx)))
                got = err_of(lambda: D.call(f"reduce_{name}_f64", [dx]).value)
                assert want[0] == got[0] and (want[0] == "err" and want == got or same_float(want[1], got[1])), (n, name)
                want = err_of(lambda: O.arg_extreme(x, mx))
                got = err_of(lambda: D.call(f"arg{name}_f64", [dx]).value)
                assert want == got, (n, name, special)
        xi = rng.integers(-3, 3, n).astype(np.int64)
        for mx in (False, True):
            name = "max" if mx else "min"
            assert err_of(lambda: O.arg_extreme(xi, mx)) == err_of(lambda: D.call(f"arg{name}_i64", [view(D, xi, 1)]).value)
    # tie and NaN rules, src/kernels/tests.rs:874-891
    a = lambda name, v: D.call(name, [D.array(np.array(v, np.float64))]).value
    assert a("argmin_f64", [3.0, 1.0, 1.0]) == 1
    assert a("argmin_f64", [3.0, np.nan, np.nan]) == 1
    assert a("argmin_f64", [0.0, -0.0]) == 1
    assert a("argmax_f64", [-0.0, 0.0]) == 1
    assert a("argmax_f64", [1.0, np.nan, 5.0]) == 1
    assert fbits(a("reduce_min_f64", [0.0, -0.0])) == fbits(-0.0)
    assert fbits(a("reduce_max_f64", [-0.0, 0.0])) == fbits(0.0)


def test_arg_extreme_tiles_count_rows(D):
    rng = np.random.default_rng(7)
    x = floats(rng, 1000, False)
    dx = D.array(x)
    for mx in (False, True):
        r = D.reducer("argmax_f64" if mx else "argmin_f64")
        for a, b in [(0, 64), (64, 384), (384, 992), (992, 1000)]:
            r.update([dx.slice(a, b - a)])
        assert r.finish().value == O.arg_extreme(x, mx)


def test_bool_and_bit_reductions(D):
    b = lambda name, v: D.call(name, [D.array(np.array(v, np.uint8), "bool")]).value
    assert b("any", []) is False and b("all", []) is True
    assert b("any", [0, 1]) is True and b("all", [1, 0]) is False
    assert b("reduce_xor_bool", [1, 1]) is False and b("reduce_xor_bool", [1, 1, 1]) is True
    rng = np.random.default_rng(8)
    x = rng.integers(0, 2**64, 3000, dtype=np.uint64)
    assert D.call("reduce_xor_u64", [D.array(x)]).value == int(np.bitwise_xor.reduce(x))
    assert D.call("reduce_and_u64", [D.array(np.zeros(0, np.uint64))]).value == 2**64 - 1


# ----------------------------------------------------------------- selections

@pytest.mark.parametrize("t,dt", [("f64", np.float64), ("i64", np.int64), ("u64", np.uint64),
                                  ("i32", np.int32), ("f32", np.float32), ("u8", np.uint8)])
def test_compress(D, t, dt):
    rng = np.random.default_rng(9)
    for n in LENGTHS + [100_000]:
        x = rng.integers(0, 200, n).astype(dt)
        for density in (0.0, 0.3, 1.0):
            m = (rng.random(n) < density).astype(np.uint8)
            got = D.call(f"compress_{t}", [view(D, x, n % 7), view(D, m, (n + 3) % 7, "bool")]).to_numpy()
            assert np.array_equal(got, O.compress(x, m)), (n, density)
    e = err_of(lambda: D.call(f"compress_{t}", [D.array(np.zeros(2, dt)), D.array(np.ones(1, np.uint8), "bool")]))
    assert e == ("err", "length mismatch: 2 vs 1")


def test_compress_bool_and_where(D):
    got = D.call("compress_bool", [D.array(np.array([1, 0, 1], np.uint8), "bool"),
                                   D.array(np.array([1, 1, 0], np.uint8), "bool")])
    assert got.type == "bool" and list(got.to_numpy()) == [1, 0]
    assert list(D.call("where", [D.array(np.array([0, 1, 1], np.uint8), "bool")]).to_numpy()) == [1, 2]
    rng = np.random.default_rng(10)
    for n in LENGTHS + [70_000]:
        m = (rng.random(n) < 0.4).astype(np.uint8)
        assert np.array_equal(D.call("where", [view(D, m, 5, "bool")]).to_numpy(), O.where(m))


def test_filters(D):
    rng = np.random.default_rng(12)
    for n in LENGTHS + [200_003]:
        x = floats(rng, n, True)
        dx = view(D, x, n % 7)
        for cmp in CMPS:
            for c in [0.5, float("nan"), float(x[n // 3]) if n else 0.0]:
                got = D.call(f"filter_{cmp}_f64", [dx, S("f64", c)]).to_numpy()
                assert same_array(O.filter_cmp(x, cmp, c), got), (n, cmp, c)
        xi = rng.integers(-10, 10, n).astype(np.int64)
        got = D.call("filter_le_i64", [view(D, xi, 2), S("i64", 3)]).to_numpy()
        assert np.array_equal(got, O.filter_cmp(xi, "le", 3))


def test_selection_tiles(D):
    # the output of a selection does not depend on the tile split, and `where` keeps global rows
    rng = np.random.default_rng(13)
    n = 4500
    x = rng.random(n)
    m = (rng.random(n) < 0.5).astype(np.uint8)
    dx, dm = D.array(x), D.array(m, "bool")
    for name, args, want in [("filter_gt_f64", lambda a, b: [dx.slice(a, b - a), S("f64", 0.3)], O.filter_cmp(x, "gt", 0.3)),
                             ("compress_f64", lambda a, b: [dx.slice(a, b - a), dm.slice(a, b - a)], O.compress(x, m)),
                             ("where", lambda a, b: [dm.slice(a, b - a)], O.where(m))]:
        out = D.builder("i64" if name == "where" else "f64")
        for a, b in [(0, 37), (37, 100), (100, n)]:
            D.select_tile(name, args(a, b), a, out)
        assert np.array_equal(out.finish().to_numpy(), want), name


# ----------------------------------------------------------------------- scans

@pytest.mark.parametrize("t,dt", [("i64", np.int64), ("u64", np.uint64), ("i32", np.int32), ("u8", np.uint8)])
def test_scan_add(D, t, dt):
    rng = np.random.default_rng(14)
    info = np.iinfo(dt)
    for n in LENGTHS + [1_000_003]:
        x = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
        got = D.call(f"scan_add_{t}", [view(D, x, n % 7)]).to_numpy()
        want = np.cumsum(x.astype(np.int64 if dt in (np.int64, np.int32) else np.uint64))
        assert np.array_equal(got, want), (t, n)
        if t == "i64":
            assert np.array_equal(got, O.scan_add_i64(x))
            s = (rng.random(n) < 1 / 16).astype(np.uint8)
            got = D.call("scan_by_add_i64", [view(D, x, 1), view(D, s, 2, "bool")]).to_numpy()
            assert np.array_equal(got, O.scan_by_add_i64(x, s)), n


def test_scan_goldens_and_tiles(D):
    # src/kernels/tests.rs:1416-1490
    sc = lambda name, *a: D.call(name, list(a)).to_numpy()
    assert list(sc("scan_add_i64", D.array(np.array([1, 2, 3, 4], np.int64)))) == [1, 3, 6, 10]
    assert list(sc("scan_by_add_i64", D.array(np.array([1, 2, 3, 4, 5], np.int64)),
                   D.array(np.array([0, 0, 1, 0, 1], np.uint8), "bool"))) == [1, 3, 3, 7, 5]
    assert list(sc("scan_by_add_i64", D.array(np.array([1, 2], np.int64)),
                   D.array(np.array([1, 0], np.uint8), "bool"))) == [1, 3]
    nan = sc("scan_max_f64", D.array(np.array([1.0, np.nan, 3.0])))
    assert nan[0] == 1.0 and np.isnan(nan[1]) and np.isnan(nan[2])
    z = sc("scan_max_f64", D.array(np.array([-0.0, 0.0])))
    assert [fbits(v) for v in z] == [fbits(-0.0), fbits(0.0)]
    rng = np.random.default_rng(15)
    n = 20_000
    x = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    s = (rng.random(n) < 0.05).astype(np.uint8)
    dx, ds = D.array(x), D.array(s, "bool")
    from verkehrt_amd import Array
    for name, want, by in [("scan_add_i64", O.scan_add_i64(x), False), ("scan_by_add_i64", O.scan_by_add_i64(x, s), True)]:
        out = Array.empty(D.ctx, "i64", n)
        sc_ = D.scanner(name)
        for a, b in [(0, 37), (37, 100), (100, 4500), (4500, n)]:
            args = [dx.slice(a, b - a)] + ([ds.slice(a, b - a)] if by else [])
            sc_.update(args, out.slice(a, b - a))
        assert np.array_equal(out.to_numpy(), want), name
    f = floats(rng, 50_000, True)
    for mx in (False, True):
        got = D.call("scan_max_f64" if mx else "scan_min_f64", [D.array(f)]).to_numpy()
        assert same_array(O.scan_extreme(f, mx), got)


def test_scan_add_f64_defined_order(D):
    rng = np.random.default_rng(16)
    for n in LENGTHS:
        x = floats(rng, n, False)
        assert same_array(O.scan_add_f64(x), D.call("scan_add_f64", [view(D, x, 1)]).to_numpy()), n


@pytest.mark.parametrize("n", [16384, 16384 * 3 + 5, 1_000_003, 4_000_037])
def test_scan_add_f64_tiles(D, n):
    # from 1024 whole blocks the exact tile starts come from the canonical fold's walk and the
    # tiles run in parallel: uniform, zero-mean, specials (NaN, inf, subnormals, +-0), parity
    # ties, and a sum drifting through zero
    rng = np.random.default_rng(n)
    for name, x in fast_path_sets(rng, n):
        got = D.call("scan_add_f64", [view(D, x, 1)]).to_numpy()
        assert same_array(O.scan_add_f64(x), got), (name, n)


def test_scan_add_f64_tiles_scanner(D):
    # a scanner fed in updates of whole blocks (the last one partial): each update continues
    # from the accumulators the previous left, above and below the tiled form's threshold
    from verkehrt_amd import Array
    rng = np.random.default_rng(160)
    n = 3_000_007
    x = rng.normal(size=n) * 100.0 + 0.125
    dx = D.array(x)
    out = Array.empty(D.ctx, "f64", n)
    sc = D.scanner("scan_add_f64")
    cuts = [0, 16 * 50, 16 * 70_000, 16 * 71_000, 16 * 150_000, n]
    for a, b in zip(cuts, cuts[1:]):
        sc.update([dx.slice(a, b - a)], out.slice(a, b - a))
    assert same_array(O.scan_add_f64(x), out.to_numpy())


# ---------------------------------------------------------------------- gather

@pytest.mark.parametrize("t,dt", [("f64", np.float64), ("i64", np.int64), ("i32", np.int32), ("u8", np.uint8)])
def test_gather(D, t, dt):
    rng = np.random.default_rng(17)
    for n in LENGTHS + [300_000]:
        xs = rng.integers(0, 100, max(n, 1)).astype(dt)
        idx = rng.integers(0, xs.size, n).astype(np.int64)
        if n and n % 3 == 0:  # one forced bad index, isa_variants.rs:173-190
            idx[rng.integers(0, n)] = xs.size if n % 2 else -1
        want = err_of(lambda: O.gather(xs, idx))
        got = err_of(lambda: D.call(f"gather_{t}", [view(D, xs, 3), view(D, idx, 1)]).to_numpy())
        assert want[0] == got[0], n
        if want[0] == "err":
            assert want == got
        else:
            assert np.array_equal(want[1], got[1])
    e = err_of(lambda: D.call(f"gather_{t}", [D.array(np.array([10, 20, 30], dt)), D.array(np.array([0, 7, -1, 9], np.int64))]))
    assert e == ("err", "gather index 7 out of range for length 3")
    e = err_of(lambda: D.call(f"gather_{t}", [D.array(np.zeros(0, dt)), D.array(np.array([0], np.int64))]))
    assert e == ("err", "gather index 0 out of range for length 0")


# --------------------------------------------------------------------- buckets

def test_histogram_and_reduce_into(D):
    I = lambda v: D.array(np.array(v, np.int64))
    assert list(D.call("histogram", [I([0, 2, 2, 3]), S("i64", 5)]).to_numpy()) == [1, 0, 2, 1, 0]
    assert len(D.call("histogram", [I([]), S("i64", 0)])) == 0
    for key, nb in [(3, 3), (-1, 3), (0, 0)]:
        first = 0 if nb == 0 else key
        want = f"bucket key {first} out of range for {nb} buckets"
        assert err_of(lambda: D.call("histogram", [I([0, key]), S("i64", nb)])) == ("err", want)
        assert err_of(lambda: D.call("reduce_into_add_i64", [I([0, key]), I([1, 1]), S("i64", nb)])) == ("err", want)
    assert err_of(lambda: D.call("histogram", [I([]), S("i64", -1)])) == ("err", "negative bucket count -1")
    assert err_of(lambda: D.call("reduce_into_add_i64", [I([0, 0]), I([1]), S("i64", 1)])) == \
        ("err", "length mismatch: 2 keys for 1 values")
    keys = [1, 0, 1, 3, 1]
    assert list(D.call("reduce_into_add_i64", [I(keys), I([5, 7, 9, 2, 1]), S("i64", 4)]).to_numpy()) == [7, 15, 0, 2]
    mn = D.call("reduce_into_min_f64", [I(keys), D.array(np.array([1.5, 2.5, 0.5, 9.0, 4.0])), S("i64", 4)]).to_numpy()
    assert list(mn) == [2.5, 0.5, math.inf, 9.0]
    rng = np.random.default_rng(18)
    for n in LENGTHS + [500_000]:
        for nb in (1, 7, 64, 5000):
            k = rng.integers(0, nb, n).astype(np.int64)
            v = rng.integers(-2**40, 2**40, n).astype(np.int64)
            assert np.array_equal(D.call("histogram", [view(D, k, 1), S("i64", nb)]).to_numpy(), O.histogram(k, nb))
            assert np.array_equal(D.call("reduce_into_add_i64", [view(D, k, 2), view(D, v, 3), S("i64", nb)]).to_numpy(),
                                  O.reduce_into(k, v, nb))
            f = floats(rng, n, True)
            for op in ("min", "max"):
                got = D.call(f"reduce_into_{op}_f64", [view(D, k, 1), D.array(f), S("i64", nb)]).to_numpy()
                assert same_array(O.reduce_into(k, f, nb, op), got)


def test_reduce_into_add_f64_bad_keys(D):
    # the f64 add checks its keys inside the sort's counting pass (rows of whole 16-row blocks)
    # and on its own for the tail rows: the first bad row is named wherever it falls, for
    # u32 chain keys (few buckets) and u64 chain keys (more than 2^28 buckets)
    rng = np.random.default_rng(31)
    cases = [(4, 3, {1: 7}), (40, 3, {5: 3}), (40, 3, {35: -1}), (40, 3, {35: 9, 7: -4}),
             (100_000, 50, {99_999: 50}), (100_000, 50, {70_001: -1, 123: 1 << 40}),
             (100_000, 70_000, {16: 70_000, 99_990: 80_000}), (33, 0, {}), (37, 1 << 29, {20: -2, 34: 1 << 29})]
    for n, nb, bad in cases:
        k = rng.integers(0, max(nb, 1), n).astype(np.int64)
        for r, key in bad.items():
            k[r] = key
        v = floats(rng, n, False)
        want = err_of(lambda: O.reduce_into(k, v, nb))
        assert want[0] == "err"
        got = err_of(lambda: D.call("reduce_into_add_f64", [D.array(k), D.array(v), S("i64", nb)]))
        assert got == want, (n, nb, got, want)
    # the context stays usable, with the same keys made good
    n, nb = 100_000, 50
    k = rng.integers(0, nb, n).astype(np.int64)
    v = floats(rng, n, False)
    got = D.call("reduce_into_add_f64", [D.array(k), D.array(v), S("i64", nb)]).to_numpy()
    assert same_array(O.reduce_into(k, v, nb), got)


def test_reduce_into_add_f64_defined_order(D):
    rng = np.random.default_rng(19)
    for n in LENGTHS + [100_000]:
        for nb in (1, 5, 300):
            k = rng.integers(0, nb, n).astype(np.int64)
            v = floats(rng, n, False)
            got = D.call("reduce_into_add_f64", [view(D, k, 1), view(D, v, 2), S("i64", nb)]).to_numpy()
            assert same_array(O.reduce_into(k, v, nb), got), (n, nb)
    # many long chains (one lane per chain): 300 buckets x 16 stripes over 2M rows, ragged
    n = 2_000_003
    k = rng.integers(0, 300, n).astype(np.int64)
    v = floats(rng, n, False)
    got = D.call("reduce_into_add_f64", [view(D, k, 1), view(D, v, 2), S("i64", 300)]).to_numpy()
    assert same_array(O.reduce_into(k, v, 300), got)
    # many buckets (one lane per bucket, no stripe buffer), some of them empty
    n = 300_007
    k = rng.integers(0, 100_000, n).astype(np.int64)
    v = floats(rng, n, False)
    got = D.call("reduce_into_add_f64", [view(D, k, 3), view(D, v, 1), S("i64", 100_000)]).to_numpy()
    assert same_array(O.reduce_into(k, v, 100_000), got)
    # many buckets, nearly all empty (whole waves of empty buckets), rows in the first and last
    # buckets and a few between; and fewer rows than one 16-row block
    for n, nb, keys in [(1_003, 200_000, None), (4_099, 131_072, [0, 131_071, 65_536, 7]), (9, 70_000, None)]:
        k = (rng.integers(0, nb, n) if keys is None else rng.choice(np.array(keys), n)).astype(np.int64)
        v = floats(rng, n, False)
        got = D.call("reduce_into_add_f64", [D.array(k), D.array(v), S("i64", nb)]).to_numpy()
        assert same_array(O.reduce_into(k, v, nb), got), (n, nb)
    # every key equal gives sum's bits
    v = floats(rng, 10_007, False)
    got = D.call("reduce_into_add_f64", [D.array(np.zeros(v.size, np.int64)), D.array(v), S("i64", 1)]).to_numpy()
    assert fbits(got[0]) == fbits(O.canonical_sum_f64(v))


# --------------------------------------------------------------------- argsort

@pytest.mark.parametrize("t,dt", [("u64", np.uint64), ("i64", np.int64), ("i32", np.int32), ("u8", np.uint8)])
def test_argsort(D, t, dt):
    rng = np.random.default_rng(20)
    info = np.iinfo(dt)
    for n in LENGTHS + [5000, 123_457]:
        for x in [rng.integers(info.min, info.max, n, dtype=dt, endpoint=True),
                  rng.integers(0, 3, n).astype(dt)]:
            got = D.call(f"argsort_{t}", [view(D, x, n % 7)]).to_numpy()
            assert np.array_equal(got, O.argsort(x)), (t, n)
    assert list(D.call("argsort_i64", [D.array(np.array([3, 1, 3, 1], np.int64))]).to_numpy()) == [1, 3, 0, 2]
    assert list(D.call("argsort_u8", [D.array(np.array([200, 5, 200, 5, 0], np.uint8))]).to_numpy()) == [4, 1, 3, 0, 2]


def test_argsort_f64_defined_order(D):
    rng = np.random.default_rng(21)
    for n in LENGTHS + [50_000]:
        x = floats(rng, n, True)
        got = D.call("argsort_f64", [view(D, x, 1)]).to_numpy()
        assert np.array_equal(got, O.argsort(x)), n


# -------------------------------------------------------------------- group_id

def test_group_id(D):
    assert list(D.call("group_id_i64", [D.array(np.array([7, -1, 7, 3, -1, 7], np.int64))]).to_numpy()) == [0, 1, 0, 2, 1, 0]
    assert list(D.call("group_id2_i64_i64", [D.array(np.array([1, 2, 1, 2], np.int64)),
                                             D.array(np.array([2, 1, 2, 2], np.int64))]).to_numpy()) == [0, 1, 0, 2]
    assert list(D.call("group_id2_u8_i32", [D.array(np.array([5, 5, 9], np.uint8)),
                                            D.array(np.array([-1, -1, -1], np.int32))]).to_numpy()) == [0, 0, 1]
    rng = np.random.default_rng(22)
    for n in LENGTHS + [200_000]:
        for d in (1, 3, 64, max(n, 1)):
            base = rng.integers(-2**63, 2**63 - 1, d, dtype=np.int64)
            x = base[(np.arange(n) * 0x9E3779B9) % d]
            assert np.array_equal(D.call("group_id_i64", [view(D, x, n % 7)]).to_numpy(), O.group_id(x)), (n, d)
            y = rng.integers(0, 2, n).astype(np.int32)
            assert np.array_equal(D.call("group_id2_i64_i32", [view(D, x, 1), view(D, y, 2)]).to_numpy(),
                                  O.group_id(x, y)), (n, d)
        xu = rng.integers(0, 256, n).astype(np.uint8)
        assert np.array_equal(D.call("group_id_u8", [view(D, xu, 3)]).to_numpy(), O.group_id(xu))


def test_group_id_large(D):
    """Columns past 2^20 rows: a few distinct keys (the small global table, every key held by
    the blocks' LDS tables), more keys than an LDS table takes, more keys than the small table
    takes (its overflow, then the full table), and the LDS empty marker's bit pattern as a key."""
    rng = np.random.default_rng(40)
    n = 3_000_000
    marker = np.array([0xFFF8DEADBEEF1357], np.uint64).view(np.int64)[0]
    for d in (1, 7, 5000, 100_000, n):
        base = rng.integers(-2**63, 2**63 - 1, d, dtype=np.int64)
        base[d // 2] = marker
        x = base[rng.integers(0, d, n)] if d < n else base
        got = D.call("group_id_i64", [view(D, x, 1)]).to_numpy()
        assert np.array_equal(got, O.group_id(x)), d
    xi = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    assert np.array_equal(D.call("group_id_i32", [D.array(xi)]).to_numpy(), O.group_id(xi))
    xu = rng.integers(0, 256, n).astype(np.uint8)
    assert np.array_equal(D.call("group_id_u8", [D.array(xu)]).to_numpy(), O.group_id(xu))


def test_group_id_sketch_sized_table(D):
    """More keys than the small table takes but far fewer than rows: the full table is sized from
    the sketched count of distinct keys (here 2^23 slots for about 3M keys, against 2^24 for 2n),
    with the empty marker's key among them, and i32 keys of the same shape."""
    rng = np.random.default_rng(44)
    n = 8_000_003
    marker = np.array([0xFFF8DEADBEEF1357], np.uint64).view(np.int64)[0]
    base = rng.integers(-2**63, 2**63 - 1, 3_000_000, dtype=np.int64)
    base[1234] = marker
    x = base[rng.integers(0, base.size, n)]
    assert np.array_equal(D.call("group_id_i64", [view(D, x, 3)]).to_numpy(), O.group_id(x))
    xi = rng.integers(-2**31, 2**31 - 1, 2_500_000, dtype=np.int64).astype(np.int32)[rng.integers(0, 2_500_000, n)]
    assert np.array_equal(D.call("group_id_i32", [D.array(xi)]).to_numpy(), O.group_id(xi))


# ----------------------------------------------------------------- elementwise

def test_elementwise(D):
    rng = np.random.default_rng(23)
    x, y = floats(rng, 1000, True), floats(rng, 1000, True)
    dx, dy = D.array(x), D.array(y)
    with np.errstate(invalid="ignore"):
        for cmp, f in [("gt", np.greater), ("lt", np.less), ("eq", np.equal), ("ne", np.not_equal),
                       ("ge", np.greater_equal), ("le", np.less_equal)]:
            assert np.array_equal(D.call(f"{cmp}_f64", [dx, dy]).to_numpy(), f(x, y).astype(np.uint8))
            assert np.array_equal(D.call(f"{cmp}_f64", [dx, S("f64", 0.5)]).to_numpy(), f(x, 0.5).astype(np.uint8))
        assert same_array(x * y, D.call("mul_f64", [dx, dy]).to_numpy())
        assert same_array(x + 2.0, D.call("add_f64", [dx, S("f64", 2.0)]).to_numpy())
    xi = rng.integers(-2**63, 2**63 - 1, 1000, dtype=np.int64)
    assert np.array_equal(D.call("add_i64", [D.array(xi), D.array(xi)]).to_numpy(), xi + xi)
    b = (rng.random(1000) < 0.5).astype(np.uint8)
    assert np.array_equal(D.call("cast_bool_i64", [D.array(b, "bool")]).to_numpy(), b.astype(np.int64))
    assert np.array_equal(D.call("not_bool", [D.array(b, "bool")]).to_numpy(), 1 - b)


def test_errors_as_strings(D):
    e = err_of(lambda: D.call("dot_f64", [D.array(np.zeros(3)), D.array(np.zeros(2))]))
    assert e == ("err", "length mismatch: 3 vs 2")
    e = err_of(lambda: D.call("reduce_add_f64", [D.array(np.zeros(3, np.int64))]))
    assert e == ("err", "expected f64[], got i64[]")
    e = err_of(lambda: D.call("count_gt_f64", [D.array(np.zeros(3)), S("i64", 1)]))
    assert e == ("err", "expected f64, got i64")
    e = err_of(lambda: D.call("argmin_f64", [D.array(np.zeros(0))]))
    assert e == ("err", "argmin of an empty array")
    e = err_of(lambda: D.call("reduce_max_i64", [D.array(np.zeros(0, np.int64))]))
    assert e == ("err", "reduce_max of an empty array")


# ------------------------------------------------------------------ plan replay

def test_readme_plan_replay_with_trace(D):
    # README.md:55-69: the four plan lines, their outputs, and the trace line format
    from verkehrt_amd.plan import Plan
    P = Plan(D)
    D.ctx.set_trace(True)
    D.ctx.take_trace()
    P.run("(rand_f64 1000000 1)", bind="x")
    dot = P.run("(dot_f64 x (rand_f64 1000000 2))")
    cnt = P.run("(count_gt_f64 x 0.5)")
    ln = P.run("(length_f64 (filter_gt_f64 x 0.99))")
    trace = D.ctx.take_trace().strip().splitlines()
    D.ctx.set_trace(False)
    assert repr(dot.value) == "249724.39002046228" and cnt.value == 499521 and ln.value == 9851
    names = [t.split()[1] for t in trace]
    assert names == ["rand_f64", "rand_f64", "dot_f64", "count_gt_f64", "filter_gt_f64", "length_f64"]
    import re
    pat = re.compile(r"^trace \w+ isa=hip n=\d+ read=\d+ write=\d+ time=\d+\.\d gbps=\d+\.\d\d$")
    assert all(pat.match(t) for t in trace), trace
    f = [t for t in trace if t.split()[1] == "filter_gt_f64"][0]
    assert "n=1000000 read=8000000 write=78808" in f


# --------------------------------------------- exact fast canonical fold (large n)

def fast_path_sets(rng, n):
    yield "uniform", rng.random(n)
    yield "normal", rng.normal(size=n)
    yield "varied", (rng.random(n) - 0.5) * 10.0 ** ((np.arange(n) % 9) * 2)
    yield "specials-free", floats(rng, n, False)
    yield "with-specials", floats(rng, n, True)
    yield "halves", rng.integers(-4, 5, n) * 0.5 + (np.arange(n) < 16) * 2.0**52  # parity ties
    drift = rng.random(n) - 0.5001
    drift[:16] = 40.0
    yield "drift-through-zero", drift


@pytest.mark.parametrize("n", [8192, 16384, 16384 + 13, 65536 + 17, 1_000_003, 4_000_037])
def test_canonical_fast_path(D, n):
    rng = np.random.default_rng(n)
    for name, x in fast_path_sets(rng, n):
        dx = view(D, x, 1)
        assert same_float(O.canonical_sum_f64(x), D.call("reduce_add_f64", [dx]).value), (name, n)
        y = rng.random(n)
        assert same_float(O.dot_f64(x, y), D.call("dot_f64", [dx, D.array(y)]).value), (name, n)
        want = O.condsum_f64(x, "gt", 0.25, -1.0)
        got = D.call("condsum_gt_f64", [dx, S("f64", 0.25), S("f64", -1.0)]).value
        assert same_float(want, got), (name, n)


def test_many_tiles(D):
    # every look-back kernel over thousands of tiles
    rng = np.random.default_rng(5)
    n = 6_000_011
    x = rng.random(n)
    dx = D.array(x)
    for c in (0.01, 0.5, 0.99):
        got = D.call("filter_gt_f64", [dx, S("f64", c)]).to_numpy()
        assert np.array_equal(got, x[x > c]), c
    m = (x < 0.3).astype(np.uint8)
    assert np.array_equal(D.call("where", [D.array(m, "bool")]).to_numpy(), np.flatnonzero(m))
    xi = rng.integers(-2**40, 2**40, n).astype(np.int64)
    dxi = D.array(xi)
    assert np.array_equal(D.call("scan_add_i64", [dxi]).to_numpy(), np.cumsum(xi))
    st = (rng.random(n) < 1e-4).astype(np.uint8)
    assert np.array_equal(D.call("scan_by_add_i64", [dxi, D.array(st, "bool")]).to_numpy(),
                          O.scan_by_add_i64(xi, st))
    assert np.array_equal(D.call("scan_max_i64", [dxi]).to_numpy(), np.maximum.accumulate(xi))
    k = rng.integers(0, 1 << 20, n).astype(np.int64)
    assert np.array_equal(D.call("argsort_i64", [D.array(k)]).to_numpy(), np.argsort(k, kind="stable"))
    assert D.call("count_gt_f64", [dx, S("f64", 0.5)]).value == int((x > 0.5).sum())
    assert D.call("argmax_f64", [dx]).value == int(np.argmax(x))


def test_mask_selections_long(D):
    # compress and where over at least 512 tiles take a count pass over the mask bytes and a scan
    # of the tile counts instead of the look-back: aligned and unaligned masks, a partial last
    # tile, nonzero bytes other than 1, densities from none to all
    rng = np.random.default_rng(15)
    n = 9_000_013
    x = rng.random(n)
    for density in (0.0, 0.01, 0.5, 1.0):
        m = (rng.random(n) < density).astype(np.uint8)
        if density == 0.5:
            m[m != 0] = rng.integers(1, 256, int(m.sum())).astype(np.uint8)
        for off in (0, 3):
            dx, dm = view(D, x, off), view(D, m, (off * 5) % 7, "bool")
            got = D.call("compress_f64", [dx, dm]).to_numpy()
            assert np.array_equal(got.view(np.uint64), x[m != 0].view(np.uint64)), (density, off)
            w = D.call("where", [dm]).to_numpy()
            assert np.array_equal(w, np.flatnonzero(m)), (density, off)
    xi = rng.integers(-5, 5, n).astype(np.int32)
    m = (rng.random(n) < 0.3).astype(np.uint8)
    assert np.array_equal(D.call("compress_i32", [D.array(xi), D.array(m, "bool")]).to_numpy(), xi[m != 0])


def test_filters_staged_long(D):
    # filters over 2^24 rows or more that keep few rows (a sample of every 1024th row decides)
    # write each tile's survivors to a staging slot, scan the tile counts and move them: none
    # kept, a few in the last partial tile only, 1% and 5% kept, and 30% (the look-back form)
    rng = np.random.default_rng(16)
    n = (1 << 24) + 4099
    x = rng.random(n)
    dx = D.array(x)
    for c in (2.0, 0.99, 0.95, 0.7):
        got = D.call("filter_gt_f64", [dx, S("f64", c)]).to_numpy()
        assert np.array_equal(got.view(np.uint64), x[x > c].view(np.uint64)), c
    y = np.zeros(n)
    y[-3:] = [5.0, -1.0, 7.0]
    got = D.call("filter_gt_f64", [view(D, y, 1), S("f64", 1.0)]).to_numpy()
    assert list(got) == [5.0, 7.0]
    xi = rng.integers(0, 1000, n).astype(np.int32)
    got = D.call("filter_lt_i32", [D.array(xi), S("i32", 7)]).to_numpy()
    assert np.array_equal(got, xi[xi < 7])


def test_canonical_streamed_long(D):
    # above 64 tiles of 16K rows the streamed form runs; 80M rows make 4883 tiles, so the
    # walk over groups of 64 tiles goes past one wave's 64 groups
    rng = np.random.default_rng(80)
    n = 80_000_021
    y = rng.random(n)
    dy = D.array(y)
    for name, x in (("uniform", rng.random(n)), ("normal", rng.normal(size=n)),
                    ("drift-through-zero", np.where(np.arange(n) < 16, 4000.0, rng.random(n) - 0.5001))):
        dx = D.array(x)
        assert same_float(O.canonical_sum_f64(x), D.call("reduce_add_f64", [dx]).value), name
        assert same_float(O.dot_f64(x, y), D.call("dot_f64", [dx, dy]).value), name
        del dx


def test_canonical_zero_mean_large(D):
    # zero-mean rows: every chain's running sum returns near zero again and again, so thousands
    # of tiles per chain cross binades; the walk takes them through their wide tables
    rng = np.random.default_rng(26)
    n = (1 << 26) + 5
    x = rng.normal(size=n)
    y = rng.normal(size=n)
    dx, dy = D.array(x), D.array(y)
    assert same_float(O.canonical_sum_f64(x), D.call("reduce_add_f64", [dx]).value)
    assert same_float(O.dot_f64(x, y), D.call("dot_f64", [dx, dy]).value)
    want = O.condsum_f64(x, "lt", 0.5, -0.25)
    assert same_float(want, D.call("condsum_lt_f64", [dx, S("f64", 0.5), S("f64", -0.25)]).value)
    # the same fold over x * 3 - y as one fused region
    z = x * 3.0 - y
    fused, one_pass = D.region_reduce([("mul_f64", [("in", 0), ("in", 2)]), ("sub_f64", [("k", 0), ("in", 1)])],
                                      "reduce_add_f64", [("k", 1)], [dx, dy, S("f64", 3.0)])
    assert one_pass and same_float(O.canonical_sum_f64(z), fused.value)


def test_canonical_streamed_slices(D):
    # the streamed form continuing a running state, between slices of the single-pass form
    rng = np.random.default_rng(78)
    n = 5_000_003
    x = rng.normal(size=n) * 1e3 + 0.25
    dx = D.array(x)
    r = D.reducer("reduce_add_f64")
    cuts = [0, 16 * 40000, 16 * 40000 + 16 * 70000, 16 * 300000, n]
    for a, b in zip(cuts, cuts[1:]):
        r.update([dx.slice(a, b - a)])
    assert same_float(O.canonical_sum_f64(x), r.finish().value)


def test_canonical_fast_path_tiles(D):
    rng = np.random.default_rng(77)
    n = 300_000 + 5
    x = rng.normal(size=n) * 1e3
    dx = D.array(x)
    r = D.reducer("reduce_add_f64")
    cuts = [0, 16 * 1000, 16 * 9000, 16 * 9001, 16 * 18000, n]
    for a, b in zip(cuts, cuts[1:]):
        r.update([dx.slice(a, b - a)])
    assert same_float(O.canonical_sum_f64(x), r.finish().value)


def test_grouping_programs_replayed(D):
    # tests/programs/grouping.dd (group_sorted, buckets) as the kernel sequences a plan runs
    I = lambda v: D.array(np.array(v, np.int64))
    k, v = I([1, 1, 2, 2, 2, 5]), I([1, 2, 3, 4, 5, 6])
    n = D.call("length_i64", [k]).value
    it = D.call("iota", [S("i64", n)])
    s = D.call("or_bool", [D.call("ne_i64", [k, D.call("prev_i64", [k, S("i64", 0)])]),
                           D.call("eq_i64", [it, S("i64", 0)])])
    nxt = D.call("gather_i64", [k, D.call("min_i64", [D.call("add_i64", [it, S("i64", 1)]), S("i64", n - 1)])])
    e = D.call("or_bool", [D.call("ne_i64", [k, nxt]), D.call("eq_i64", [it, S("i64", n - 1)])])
    assert list(D.call("compress_i64", [D.call("scan_by_add_i64", [v, s]), e]).to_numpy()) == [3, 12, 6]
    ids = D.call("sub_i64", [D.call("scan_add_i64", [D.call("cast_bool_i64", [s])]), S("i64", 1)])
    ng = D.call("count", [s])
    assert list(D.call("reduce_into_add_i64", [ids, v, ng]).to_numpy()) == [3, 12, 6]
    assert list(D.call("compress_i64", [k, s]).to_numpy()) == [1, 2, 5]
    g = D.call("group_id_i64", [I([7, -1, 7, 3, -1, 7])])
    assert list(g.to_numpy()) == [0, 1, 0, 2, 1, 0]
    assert D.call("reduce_max_i64", [g]).value + 1 == 3


def test_prev_and_tiles(D):
    rng = np.random.default_rng(31)
    x = rng.integers(-100, 100, 5000).astype(np.int64)
    got = D.call("prev_i64", [D.array(x), S("i64", -9)]).to_numpy()
    assert got[0] == -9 and np.array_equal(got[1:], x[:-1])
    from verkehrt_amd import Array
    out = Array.empty(D.ctx, "i64", x.size)
    sc = D.scanner("prev_i64")
    dx = D.array(x)
    for a, b in [(0, 37), (37, 100), (100, 5000)]:
        sc.update([dx.slice(a, b - a), S("i64", -9)], out.slice(a, b - a))
    assert np.array_equal(out.to_numpy(), got)
