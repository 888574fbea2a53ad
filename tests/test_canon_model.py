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
    return L


def bits(v):
    return struct.unpack("<Q", struct.pack("<d", float(v)))[0]


def chains_oracle(vals, acc):
    """Sequential chain folds of blocks of 16 (reduce.rs:40-62); numpy's 1-D cumulative sum
    adds strictly left to right."""
    vals = np.asarray(vals, np.float64)
    return [float(np.cumsum(np.concatenate([[acc[j]], vals[j::16]]))[-1]) for j in range(16)]


def run(model, vals, acc, lag, streamed=False, records=True):
    vals = np.ascontiguousarray(vals, np.float64)
    a = np.array(acc, np.float64)
    stats = np.zeros(6, np.uint64)
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


@pytest.mark.parametrize("records", [True, False])
@pytest.mark.parametrize("group_tiles", [1, 4, 64])
def test_streamed_model_matches_sequential_chains(model, group_tiles, records):
    """The streamed form (prefix hypotheses, clean groups, the walk) on the same data; small
    groups make joined spans, failed spans and segments all occur on 41 tiles. Tiles whose
    summary fails go through run records or, without them, the fallback wave."""
    st = np.zeros(6, np.uint64)
    for name, vals, acc in datasets():
        got, stats = run(model, vals, acc, group_tiles, streamed=True, records=records)
        check_bits(name, got, chains_oracle(vals, acc))
        st += stats
    assert st[0] > 0 and st[3] > 0  # joined spans applied, tiles evaluated exactly
    if 1 < group_tiles < 64:
        assert st[2] > 0  # segments of non-clean groups applied at once
    if records:
        assert st[4] > 0 and st[5] > 0  # run segments applied, runs folded


def test_streamed_model_long_uniform(model):
    """Uniform data over 300 tiles: everything in joined clean spans except the tiles where a
    chain's running sum crosses into the next binade (about 512 per tile, so from 2^9 to 2^17:
    9 crossings) and the first tile, which starts at zero."""
    x = O.rand_f64(16 * 1024 * 300, 5)
    got, stats = run(model, x, np.zeros(16), 64, streamed=True)
    check_bits("uniform-300", got, chains_oracle(x, np.zeros(16)))
    assert 0 < stats[3] <= 16 * 11
    assert stats[0] > 0
