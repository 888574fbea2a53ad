"""Parity of the f32 canonical reductions and the integer fused sums with the oracle, through
the C ABI.

- reduce_add_f32, reduce_mul_f32, dot_f32, sumsq_f32, condsum_<cmp>_f32: the canonical order at
  f32 (reduce.rs:40-91, 249-323 and fused.rs:72-210 instantiated at f32), bit for bit, on
  uniform, zero-mean, wide-magnitude and special rows, at view offsets, through the sequential
  and the streamed form, across reducer updates, against the reference's f32 golden vector and
  against hip's own single-op patterns (tests/fused_oracle.rs:178-222).
- dot, sumsq and condsum over i64, u64 and i32 (fused.rs:72-210 with IntSum): wrapping
  products, widened sums, both alignments of two-input rows, and the patterns they replace.
"""
import struct

import numpy as np
import pytest

from oracle import oracle as O

from test_parity_gpu import CMPS, S, err_of, same_float, view

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def f32bits(v):
    return struct.unpack("<I", struct.pack("<f", float(v)))[0]


def same_f32(want, got):
    if np.isnan(want):
        return bool(np.isnan(got))
    return f32bits(want) == f32bits(got)


LENGTHS = [0, 1, 15, 16, 17, 1000, 32_767, 32_768 + 5, 200_003, 1_048_593]


def data(rng, kind, n):
    f = np.float32
    if kind == "uniform":
        return rng.random(n).astype(f)
    if kind == "normal":
        return rng.standard_normal(n).astype(f)
    if kind == "wide":
        with np.errstate(over="ignore"):
            v = np.ldexp(rng.uniform(-1, 1, n), rng.integers(-149, 60, n)).astype(f)
        return np.where(rng.random(n) < 0.7, rng.uniform(-1e3, 1e3, n).astype(f), v)
    # specials: NaN, infinities, signed zeros among finite rows
    v = rng.uniform(-1e3, 1e3, n).astype(f)
    if n:
        sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan], f)
        k = rng.random(n) < 0.001
        v[k] = sp[rng.integers(0, 5, int(k.sum()))]
    return v


def test_f32_golden_vector(D):
    # src/kernels/tests.rs:761-791: f32 accumulation gives 0x4B80005A (f64 would give ...59)
    ys = [16777216.0, 0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5, 9.5, 10.5, 11.5, 12.5, 13.5, 14.5,
          1.0, 0.25, 0.75, 1.25, 1.75, 2.25, 2.75, 3.25, 3.75, 4.25, 4.75, 5.25, 5.75, 6.25, 6.75, 7.25,
          3.0, 3.0, 3.0]
    got = D.call("reduce_add_f32", [D.array(np.array(ys, np.float32))])
    assert got.type == "f32" and f32bits(got.value) == 0x4B80005A
    assert f32bits(D.call("reduce_add_f32", [D.array(np.zeros(0, np.float32))]).value) == 0
    assert f32bits(D.call("reduce_mul_f32", [D.array(np.zeros(0, np.float32))]).value) == f32bits(1.0)


@pytest.mark.parametrize("kind", ["uniform", "normal", "wide", "special"])
def test_f32_canonical_kinds(D, kind):
    rng = np.random.default_rng({"uniform": 1, "normal": 2, "wide": 3, "special": 4}[kind])
    for n in LENGTHS:
        x, y = data(rng, kind, n), data(rng, kind, n)
        off = int(rng.integers(0, 4))
        dx, dy = view(D, x, off), view(D, y, 3 - off)
        with np.errstate(all="ignore"):
            assert same_f32(O.canonical_fold_f32("sum", x), D.call("reduce_add_f32", [dx]).value), (kind, n)
            assert same_f32(O.canonical_fold_f32("dot", x, y), D.call("dot_f32", [dx, dy]).value), (kind, n)
            assert same_f32(O.canonical_fold_f32("sumsq", x), D.call("sumsq_f32", [dx]).value), (kind, n)
            c, e = (float(x[n // 2]) if n else 0.5), -2.5
            for cmp in CMPS:
                want = O.canonical_fold_f32("condsum", x, cmp=cmp, c=c, e=e)
                got = D.call(f"condsum_{cmp}_f32", [dx, S("f32", c), S("f32", e)]).value
                assert same_f32(want, got), (kind, cmp, n)
            if n <= 200_003:
                z = (1.0 + (x if kind != "special" else np.nan_to_num(x)) * np.float32(1e-6)).astype(np.float32)
                assert same_f32(O.canonical_fold_f32("mul", z), D.call("reduce_mul_f32", [D.array(z)]).value), (kind, n)


@pytest.mark.parametrize("kind,start", [("uniform", 0.0), ("normal", 0.0), ("uniform", 6.0e6)])
def test_f32_large(D, kind, start):
    # 2^24 + 9 rows through the streamed form; a start of 6e6 puts every chain's running value
    # where the f32 ulp is 0.5 to 1, so most rows round
    rng = np.random.default_rng(11)
    n = (1 << 24) + 9
    x = data(rng, kind, n)
    if start:
        x[:16] = np.float32(start)
    assert same_f32(O.canonical_fold_f32("sum", x), D.call("reduce_add_f32", [D.array(x)]).value)


def test_f32_normal_2p28(D):
    # 2^28 normal rows: tens of thousands of tile chains cross binades, more than the 32768 slots
    # whose rows the table kernel keeps, so evaluated tile chains read both kept rows and x
    rng = np.random.default_rng(13)
    x = rng.standard_normal(1 << 28, dtype=np.float32)
    dx = D.array(x)
    assert same_f32(O.canonical_fold_f32("sum", x), D.call("reduce_add_f32", [dx]).value)
    y = x[::-1].copy()
    assert same_f32(O.canonical_fold_f32("dot", x, y), D.call("dot_f32", [dx, D.array(y)]).value)


def test_f32_reducer_tiles(D):
    # CanonicalAcc::feed_blocks (reduce.rs:284-308): updates in multiples of 16 give the
    # one-shot bits, across the sequential (short) and streamed (long) forms
    rng = np.random.default_rng(6)
    x = data(rng, "normal", 600_000)
    y = data(rng, "uniform", 600_000)
    dx, dy = D.array(x), D.array(y)
    cuts = [(0, 64), (64, 40_064), (40_064, 500_000), (500_000, 500_016), (500_016, 600_000)]
    for name, args, want in [
            ("reduce_add_f32", lambda a, b: [dx.slice(a, b - a)], O.canonical_fold_f32("sum", x)),
            ("dot_f32", lambda a, b: [dx.slice(a, b - a), dy.slice(a, b - a)], O.canonical_fold_f32("dot", x, y))]:
        r = D.reducer(name)
        for a, b in cuts:
            r.update(args(a, b))
        assert same_f32(want, r.finish().value), name
    r = D.reducer("reduce_add_f32")
    r.update([dx.slice(0, 17)])
    assert err_of(lambda: r.update([dx.slice(17, 47)])) == ("err", "canonical update after a partial block")


def test_f32_sums_equal_their_patterns(D):
    # dot_f32 = (reduce_add_f32 (mul_f32 ?x ?y)), sumsq and condsum likewise (fused.rs:341-379)
    rng = np.random.default_rng(7)
    for n in [0, 17, 5000, 70_001, 1_048_593]:
        x, y = data(rng, "wide", n), data(rng, "special", n)
        dx, dy = D.array(x), D.array(y)
        want = D.call("reduce_add_f32", [D.call("mul_f32", [dx, dy])]).value
        assert same_f32(want, D.call("dot_f32", [dx, dy]).value), n
        want = D.call("reduce_add_f32", [D.call("mul_f32", [dx, dx])]).value
        assert same_f32(want, D.call("sumsq_f32", [dx]).value), n
        for cmp in CMPS:
            mask = D.call(f"{cmp}_f32", [dx, S("f32", 0.5)])
            want = D.call("reduce_add_f32", [D.call("select_f32", [mask, dx, S("f32", -1.0)])]).value
            assert same_f32(want, D.call(f"condsum_{cmp}_f32", [dx, S("f32", 0.5), S("f32", -1.0)]).value), (cmp, n)


def test_f32_errors(D):
    a, b = D.array(np.zeros(5, np.float32)), D.array(np.zeros(6, np.float32))
    assert err_of(lambda: D.call("dot_f32", [a, b])) == ("err", "length mismatch: 5 vs 6")


INT_TYPES = {"i64": np.int64, "u64": np.uint64, "i32": np.int32}


def int_rows(rng, t, n):
    dt = INT_TYPES[t]
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
    if n >= 4:
        x[:4] = [info.min, info.max, 0, 1]
    small = rng.integers(0, 5, n).astype(dt)
    return np.where(rng.random(n) < 0.3, small, x).astype(dt)


@pytest.mark.parametrize("t", list(INT_TYPES))
def test_integer_fused_sums(D, t):
    rng = np.random.default_rng(21)
    for n in [0, 1, 3, 15, 16, 17, 1000, 70_001, 1_048_579]:
        x, y = int_rows(rng, t, n), int_rows(rng, t, n)
        for ox, oy in [(0, 0), (1, 0), (2, 2)]:  # unequal offsets take one row per load
            dx, dy = view(D, x, ox), view(D, y, oy)
            got = D.call(f"dot_{t}", [dx, dy])
            assert got.value == O.fused_sum_int("dot", x, y), (t, n, ox, oy)
            assert got.type == ("u64" if t == "u64" else "i64")
            assert D.call(f"sumsq_{t}", [dx]).value == O.fused_sum_int("sumsq", x), (t, n, ox)
        c = int(x[n // 2]) if n else 1
        e = int(np.iinfo(INT_TYPES[t]).min)
        for cmp in CMPS:
            got = D.call(f"condsum_{cmp}_{t}", [dx, S(t, c), S(t, e)]).value
            assert got == O.fused_sum_int("condsum", x, cmp=cmp, c=c, e=e), (t, cmp, n)


@pytest.mark.parametrize("t", list(INT_TYPES))
def test_integer_fused_sums_equal_their_patterns(D, t):
    rng = np.random.default_rng(22)
    x, y = int_rows(rng, t, 100_003), int_rows(rng, t, 100_003)
    dx, dy = D.array(x), D.array(y)
    sum_name = f"reduce_add_{t}"
    assert D.call(f"dot_{t}", [dx, dy]).value == D.call(sum_name, [D.call(f"mul_{t}", [dx, dy])]).value
    assert D.call(f"sumsq_{t}", [dx]).value == D.call(sum_name, [D.call(f"mul_{t}", [dx, dx])]).value
    c = S(t, int(x[7]))
    for cmp in CMPS:
        mask = D.call(f"{cmp}_{t}", [dx, c])
        want = D.call(sum_name, [D.call(f"select_{t}", [mask, dx, S(t, 3)])]).value
        assert D.call(f"condsum_{cmp}_{t}", [dx, c, S(t, 3)]).value == want, cmp
    r = D.reducer(f"dot_{t}")
    for a, b in [(0, 5), (5, 70_000), (70_000, 100_003)]:
        r.update([dx.slice(a, b - a), dy.slice(a, b - a)])
    assert r.finish().value == O.fused_sum_int("dot", x, y)
    assert err_of(lambda: D.call(f"dot_{t}", [dx, dy.slice(0, 5)])) == ("err", "length mismatch: 100003 vs 5")
