"""The CPU baseline (oracle/cpu_best.c, the restated x86-64-v4 backend) returns the oracle's
bits on the bench's inputs before it is ever timed (as benches/tier1.rs:403-433 asserts)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O


def p(a):
    return a.ctypes.data_as(C.c_void_p)


@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 1000, 10_007])
def test_cpu_best_matches_oracle(n):
    B = O.best()
    x, y = O.rand_f64(n, 1), O.rand_f64(n, 2)
    assert B.vkb_sum_f64(p(x), n) == O.canonical_sum_f64(x) or n == 0
    assert np.float64(B.vkb_dot_f64(p(x), p(y), n)).tobytes() == np.float64(O.dot_f64(x, y)).tobytes()
    assert B.vkb_count_gt_f64(p(x), n, 0.5) == O.count_cmp(x, "gt", 0.5)
    if n:
        assert B.vkb_min_f64(p(x), n) == O.reduce_extreme(x, False)
        assert B.vkb_argmax_f64(p(x), n) == O.arg_extreme(x, True)
    i = O.rand_i64(n, 4)
    out = np.empty(n, np.int64)
    B.vkb_scan_add_i64(p(i), n, p(out))
    assert np.array_equal(out, O.scan_add_i64(i))
    for c in (0.99, 0.5, 0.01):
        o = np.empty(n + 8, np.float64)
        k = B.vkb_filter_gt_f64(p(x), n, c, p(o))
        assert np.array_equal(o[:k], O.filter_cmp(x, "gt", c))
        m = (y > c).astype(np.uint8)
        k = B.vkb_compress_f64(p(x), p(m), n, p(o))
        assert np.array_equal(o[:k], O.compress(x, m))
    t = O.rand_f64(64, 3)
    idx = (O.rand_i64(n, 5).view(np.uint64) % np.uint64(64)).astype(np.int64)
    o = np.empty(n, np.float64)
    bad = np.zeros(1, np.int64)
    assert B.vkb_gather_f64(p(t), 64, p(idx), n, p(o), p(bad)) == 0
    assert np.array_equal(o, O.gather(t, idx))
    if n > 3:
        idx[3] = 64
        assert B.vkb_gather_f64(p(t), 64, p(idx), n, p(o), p(bad)) == -1 and bad[0] == 64


@pytest.mark.parametrize("n", [0, 1, 17, 300, 5000, 70_000])
def test_cpu_best_sort_and_group_match_oracle(n):
    B = O.best()
    rng = np.random.default_rng(n)
    for kind, x in [(0, rng.integers(0, 2**64, n, dtype=np.uint64)),
                    (1, rng.integers(-3, 3, n).astype(np.int64)),
                    (1, (rng.zipf(1.1, n).clip(max=1 << 20)).astype(np.int64)),
                    (2, np.where(rng.random(n) < 0.1, -0.0, rng.normal(size=n)))]:
        out = np.empty(n, np.int64)
        assert B.vkb_argsort_64(kind, p(np.ascontiguousarray(x)), n, p(out)) == 0
        assert np.array_equal(out, O.argsort(x)), kind
    for d in (1, 7, 1000, max(n, 1)):
        k = rng.integers(-2**63, 2**63 - 1, d, dtype=np.int64)[rng.integers(0, d, n)]
        ids = np.empty(n, np.int64)
        assert B.vkb_group_id_64(p(k), n, p(ids)) == 0
        assert np.array_equal(ids, O.group_id(k)), d
        g = int(ids.max()) + 1 if n else 0
        v = O.rand_f64(n, 3)
        out = np.empty(max(g, 1), np.float64)
        assert B.vkb_reduce_into_add_f64(p(ids), p(v), n, g, p(out)) == 0
        assert np.array_equal(out[:g].view(np.uint64), O.reduce_into(ids, v, g).view(np.uint64))
    x = O.rand_i64(n, 4)
    st = (rng.random(n) < 1 / 16).astype(np.uint8)
    out = np.empty(n, np.int64)
    B.vkb_scan_by_add_i64(p(x), p(st), n, p(out))
    assert np.array_equal(out, O.scan_by_add_i64(x, st))
