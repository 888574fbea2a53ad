"""CPU check of the exact parallel canonical fold (verkehrt_amd/csrc/canon_binade.h): the
host-side model of the device algorithm must return the sequential canonical order's bits
(the oracle) on data built to break it: binade crossings, cancellation, ties (values whose
rounding depends on the parity of the running sum), subnormals, signed zeros, infinities,
NaN, huge elements and sums hovering around zero. Fast and slow path counts show that both
paths are exercised."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "verkehrt_amd", "lib", "libvk_model.so")


@pytest.fixture(scope="module")
def model():
    L = C.CDLL(LIB)
    L.vkm_canonical_fold.restype = C.c_int
    L.vkm_canonical_fold.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, C.c_void_p]
    L.vkm_canonical_fold_streamed.restype = C.c_int
    L.vkm_canonical_fold_streamed.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int, C.c_int,
                                              C.c_void_p]
    L.vkm_c32_fold.restype = C.c_int
    L.vkm_c32_fold.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    L.vkm_wide_table_apply.restype = C.c_int
    L.vkm_wide_table_apply.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_void_p, C.c_uint64, C.c_void_p,
                                       C.c_void_p, C.POINTER(C.c_int32)]
    return L


def bits(v):
    return struct.unpack("<Q", struct.pack("<d", float(v)))[0]


def chains_oracle(vals, acc):
    """Sequential chain folds of blocks of 16 (reduce.rs:40-62); numpy's 1-D cumulative sum
    adds strictly left to right."""
    vals = np.asarray(vals, np.float64)
    return [float(np.cumsum(np.concatenate([[acc[j]], vals[j::16]]))[-1]) for j in range(16)]


def run(model, vals, acc, lag, streamed=False, records=1):
    vals = np.ascontiguousarray(vals, np.float64)
    a = np.array(acc, np.float64)
    stats = np.zeros(8, np.uint64)
    if streamed:
        model.vkm_canonical_fold_streamed(vals.ctypes.data, vals.size // 16, a.ctypes.data, lag, int(records),
                                          stats.ctypes.data)
    else:
        model.vkm_canonical_fold(vals.ctypes.data, vals.size // 16, a.ctypes.data, lag, stats.ctypes.data)
    return a, stats


def datasets():
    rng = np.random.default_rng(1)
    n = 16 * 1024 * 40 + 16 * 300  # 40 tiles and a partial one
    yield "uniform", rng.random(n), np.zeros(16)
    yield "uniform-from-large-start", rng.random(n), np.full(16, 1e6)
    yield "negative", -rng.random(n), np.zeros(16)
    yield "mixed-sign", rng.normal(size=n), np.zeros(16)
    yield "drift-through-zero", rng.random(n) - 0.5001, np.full(16, 30.0)
    # ties: halves and quarters land exactly between representable values of large sums
    big = np.full(16, 2.0**53 - 2**10)
    yield "ties", rng.integers(-4, 5, n) * 0.5, np.full(16, 2.0**52 + 3.0)
    yield "ties-near-top", np.full(n, 0.5), big / 4
    yield "crossing-up", np.full(n, 1.0), np.full(16, 2.0**40 - 2**12)
    yield "varied-magnitudes", (rng.random(n) - 0.5) * 10.0 ** ((np.arange(n) % 9) * 2), np.zeros(16)
    sub = rng.random(n) * 5e-324 * 1000
    yield "subnormals", sub, np.zeros(16)
    sp = rng.random(n)
    sp[rng.integers(0, n, 5)] = np.inf
    sp[rng.integers(0, n, 5)] = -0.0
    yield "with-inf", sp, np.zeros(16)
    sp2 = rng.random(n)
    sp2[n // 2] = np.nan
    yield "with-nan", sp2, np.zeros(16)
    huge = rng.random(n)
    huge[rng.integers(0, n, 20)] = 1e300
    yield "huge-elements", huge, np.full(16, 5.0)
    yield "products", rng.random(n) * rng.random(n), np.zeros(16)


@pytest.mark.parametrize("lag", [1, 3, 64])
def test_model_matches_sequential_chains(model, lag):
    """`lag`: the look-back finds exact values published from `lag` tiles back."""
    total_fast = total_slow = 0
    for name, vals, acc in datasets():
        got, stats = run(model, vals, acc, lag)
        want = chains_oracle(vals, acc)
        for j in range(16):
            if np.isnan(want[j]):
                assert np.isnan(got[j]), (name, j)
            else:
                assert bits(got[j]) == bits(want[j]), (name, j, got[j], want[j])
        total_fast += int(stats[0])
        total_slow += int(stats[1])
    assert total_fast > 0 and total_slow > 0


def test_model_gives_canonical_sum(model):
    x = O.rand_f64(16 * 1024 * 100 + 7, 1)
    b = x.size - x.size % 16
    acc, stats = run(model, x[:b], np.zeros(16), 8)
    # finish_lanes: pairwise +8,+4,+2,+1, then the tail
    a = list(acc)
    for step in (8, 4, 2, 1):
        for j in range(step):
            a[j] = a[j] + a[j + step]
    s = a[0]
    for v in x[b:]:
        s = s + float(v)
    assert bits(s) == bits(O.canonical_sum_f64(x))
    assert stats[0] > 10 * stats[1]  # uniform data takes the fast path almost everywhere


def check_bits(name, got, want):
    for j in range(16):
        if np.isnan(want[j]):
            assert np.isnan(got[j]), (name, j)
        else:
            assert bits(got[j]) == bits(want[j]), (name, j, got[j], want[j])


@pytest.mark.parametrize("records", [1, 0, 3, 2])
@pytest.mark.parametrize("group_tiles", [1, 4, 64])
def test_streamed_model_matches_sequential_chains(model, group_tiles, records):
    """The streamed form (prefix hypotheses, clean groups, the walk) on the same data; small
    groups make joined spans, failed spans and segments all occur on 41 tiles. Tiles whose
    summary fails go through their wide table (records & 2), then run records (records & 1) or
    the fallback wave."""
    st = np.zeros(8, np.uint64)
    for name, vals, acc in datasets():
        got, stats = run(model, vals, acc, group_tiles, streamed=True, records=records)
        check_bits(name, got, chains_oracle(vals, acc))
        st += stats
    assert st[0] > 0 and st[3] > 0  # joined spans applied, tiles evaluated exactly
    if 1 < group_tiles < 64:
        assert st[2] > 0  # segments of non-clean groups applied at once
    if records == 1:
        assert st[4] > 0 and st[5] > 0  # run segments applied, runs folded
    if records & 2:
        assert st[6] > 0 and st[7] > 0  # wide tables applied, and some that could not apply


def test_streamed_model_long_uniform(model):
    """Uniform data over 300 tiles: everything in joined clean spans except the tiles where a
    chain's running sum crosses into the next binade (about 512 per tile, so from 2^9 to 2^17:
    9 crossings) and the first tile, which starts at zero."""
    x = O.rand_f64(16 * 1024 * 300, 5)
    got, stats = run(model, x, np.zeros(16), 64, streamed=True)
    check_bits("uniform-300", got, chains_oracle(x, np.zeros(16)))
    assert 0 < stats[3] <= 16 * 11
    assert stats[0] > 0


def wide(model, chain, approx, starts):
    chain = np.ascontiguousarray(chain, np.float64)
    starts = np.ascontiguousarray(starts, np.float64)
    out = np.zeros(starts.size)
    applied = np.zeros(starts.size, np.uint8)
    nres = C.c_int32(0)
    model.vkm_wide_table_apply(chain.ctypes.data, chain.size, approx, starts.ctypes.data, starts.size,
                               out.ctypes.data, applied.ctypes.data, C.byref(nres))
    return out, applied.astype(bool), nres.value


def seq_fold(start, chain):
    return float(np.cumsum(np.concatenate([[start], chain]))[-1])


def wide_cases():
    rng = np.random.default_rng(21)
    # from 40 down through zero (tiny binades, a change of sign) and back up
    drift = np.concatenate([np.full(512, -40.0 / 500), np.full(512, 40.0 / 500)])
    yield "through-zero", rng.normal(size=1024) * 0.3 + drift, 40.0, True
    # zero-mean steps from 0.37: the running value rises about six binades, beyond the table
    yield "normal-from-small", rng.normal(size=1024), 0.37, False
    # a start just under 2^40 with steps of 1/2 + multiples of 2^-14: ties on the start binade's
    # grid (u0 = 2^-13) and a rise into the next binade, so the end depends on the start's residue
    yield "tie-crossing-up", 0.5 + rng.integers(0, 4, 1024) * 2.0 ** -14, 2.0 ** 40 - 300.0, True
    # a walk that falls two binades and climbs back
    yield "down-and-up", np.concatenate([np.full(512, -3.0), np.full(512, 3.0)]), 2000.3, True
    # the same from 2000: the running value lands on 512 = 2^9 exactly, where the margin rule
    # refuses every start below (conservative there, still exact where it applies)
    yield "lands-on-power-of-two", np.concatenate([np.full(512, -3.0), np.full(512, 3.0)]), 2000.0, None
    # late zero-mean steps with a large running value
    yield "normal-large", rng.normal(size=1024), 4100.3, True
    # a rise of six binades (beyond the 16-entry table)
    yield "steep-rise", np.full(1024, 1.0), 20.0, False
    # a NaN step, an infinite step, and a fold that lands on zero exactly
    c = rng.random(1024)
    c[500] = np.nan
    yield "nan", c, 3.0, False
    c = rng.random(1024)
    c[10] = np.inf
    yield "inf", c, 3.0, False
    yield "hits-zero", np.concatenate([[-1.0] * 4, rng.random(1020)]), 4.0, False


@pytest.mark.parametrize("case", [c[0] for c in wide_cases()])
def test_wide_table_matches_sequential_folds(model, case):
    """The wide table of a chain (built from an approximate start, as the walk's dirty tiles are)
    applied to exact starts around it: wherever it applies, the bits of the sequential fold.
    The starts include every residue modulo 16 ulps and offsets of up to 400 ulps, small beside
    every case's steps (as an approximate start's error is beside real data's steps), which the
    table must nearly all cover, and offsets of up to 10^6 ulps, which can move a crossing to
    another step and must then be refused."""
    name, chain, approx, expect_apply = next(c for c in wide_cases() if c[0] == case)
    e = int(np.frexp(approx)[1]) - 1
    u0 = 2.0 ** (e - 52)
    rng = np.random.default_rng(5)
    near = np.concatenate([np.arange(-40, 40), rng.integers(-400, 400, 200)])
    offs = np.concatenate([near, rng.integers(-10 ** 6, 10 ** 6, 300)])
    starts = approx + offs * u0
    got, applied, nres = wide(model, chain, approx, starts)
    for s, g, a in zip(starts, got, applied):
        if a:
            assert bits(g) == bits(seq_fold(s, chain)), (name, s)
    if expect_apply:
        assert nres > 0 and applied[:near.size].mean() > 0.9, (name, nres, applied[:near.size].mean())
    elif expect_apply is None:
        assert applied.any() and not applied.all(), name
    else:
        assert not applied.any(), name
    if name == "tie-crossing-up":
        # the residue matters: the folds from these starts do not all move by one offset
        ends = np.array([seq_fold(s, chain) for s in starts[:64]]) - starts[:64]
        assert len(set(ends.tolist())) > 1 and nres >= 4


# ------------------------------------------------------------------ f32 chains

def chains_oracle_f32(vals, acc):
    """Sequential f32 chain folds: numpy's float32 cumulative sum adds left to right in f32."""
    vals = np.asarray(vals, np.float32)
    return [np.cumsum(np.concatenate([[np.float32(acc[j])], vals[j::16]]).astype(np.float32),
                      dtype=np.float32)[-1] for j in range(16)]


def f32_datasets():
    rng = np.random.default_rng(3)
    n = 16 * 1024 * 12 + 16 * 300  # 12 tiles and a partial one
    f = np.float32
    yield "uniform", rng.random(n).astype(f), np.zeros(16)
    yield "uniform-from-large-start", rng.random(n).astype(f), np.full(16, 3.0e5)
    yield "negative", (-rng.random(n)).astype(f), np.zeros(16)
    yield "zero-mean", rng.standard_normal(n).astype(f), np.zeros(16)
    v = rng.standard_normal(n).astype(f) * f(1e-3)
    v[:: 997] = f(3.0e4)
    yield "spikes", v, np.zeros(16)
    # ties: halves and quarters against a running sum whose ulp is 1 or 2
    t = (rng.integers(0, 8, n) * 0.25).astype(f)
    yield "ties", t, np.full(16, 8388608.0)
    w = rng.random(n).astype(f)
    w[5000] = np.float32(np.inf)
    w[90001] = np.float32(np.nan)
    yield "inf-nan", w, np.zeros(16)
    s = (rng.random(n) * 1e-38).astype(f)  # subnormal f32 values
    yield "subnormal", s, np.zeros(16)
    z = np.zeros(n, f)
    z[1::2] = f(-0.0)
    yield "signed-zeros", z, np.zeros(16)
    big = rng.random(n).astype(f) * f(3e38)
    yield "overflow", big, np.zeros(16)


@pytest.mark.parametrize("name,vals,acc", list(f32_datasets()), ids=lambda p: p if isinstance(p, str) else "")
def test_f32_streamed_model_is_exact(model, name, vals, acc):
    a = np.array(acc, np.float64)
    st = np.zeros(5, np.uint64)
    v = np.ascontiguousarray(vals.astype(np.float64))
    with np.errstate(all="ignore"):
        want = chains_oracle_f32(vals, acc)
    model.vkm_c32_fold(v.ctypes.data, v.size // 16, a.ctypes.data, st.ctypes.data)
    got = a.astype(np.float32)
    for j in range(16):
        assert np.float32(got[j]).tobytes() == np.float32(want[j]).tobytes(), (name, j, got[j], want[j])
        assert float(np.float32(a[j])) == a[j] or np.isnan(a[j])  # the state holds f32 values
    if name in ("uniform", "uniform-from-large-start", "negative"):
        assert st[0] > st[1], st  # most tiles apply their summaries
    if name == "zero-mean":
        assert st[4] > st[1] > 0, st  # most crossing tiles take their wide table, the rest the rows
