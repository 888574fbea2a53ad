"""The exact checks bench.py applies to full-size argsort and group_id results accept the
oracle's answers and reject each plausible mistake (a swapped tie, an unsorted pair, a
repeated position, relabelled, merged or split groups). The Zipf input generator agrees
with a direct inverse-CDF restatement."""
import numpy as np

import bench
from oracle import oracle as O


def test_check_argsort():
    rng = np.random.default_rng(1)
    for kind, keys in [(0, rng.integers(0, 50, 3000).astype(np.uint64)),
                       (1, rng.integers(-20, 20, 3000).astype(np.int64)),
                       (2, np.where(rng.random(3000) < 0.2, -0.0, rng.integers(-3, 3, 3000).astype(np.float64)))]:
        perm = O.argsort(keys)
        assert bench.check_argsort(keys, kind, perm)
        img = bench.key_image(keys, kind)[perm]
        tie = int(np.nonzero(img[1:] == img[:-1])[0][0])
        bad = perm.copy()
        bad[[tie, tie + 1]] = bad[[tie + 1, tie]]  # equal keys out of position order
        assert not bench.check_argsort(keys, kind, bad)
        step = int(np.nonzero(img[1:] != img[:-1])[0][0])
        bad = perm.copy()
        bad[[step, step + 1]] = bad[[step + 1, step]]  # a larger key before a smaller one
        assert not bench.check_argsort(keys, kind, bad)
        bad = perm.copy()
        bad[5] = bad[6]  # not a permutation
        assert not bench.check_argsort(keys, kind, bad)
        assert not bench.check_argsort(keys, kind, perm[:-1])
    # the f64 order puts -0.0 before +0.0
    z = np.array([0.0, -0.0])
    assert bench.check_argsort(z, 2, np.array([1, 0])) and not bench.check_argsort(z, 2, np.array([0, 1]))


def test_check_group_ids():
    rng = np.random.default_rng(2)
    keys = rng.integers(-2**63, 2**63 - 1, 50, dtype=np.int64)[rng.integers(0, 50, 4000)]
    ids = O.group_id(keys)
    assert bench.check_group_ids(keys, ids)
    g = int(ids.max()) + 1
    swap = np.arange(g)
    swap[[3, 4]] = [4, 3]  # two groups relabelled: not first-occurrence order
    assert not bench.check_group_ids(keys, swap[ids])
    merged = ids.copy()
    merged[merged == g - 1] = g - 2  # two keys share one id
    assert not bench.check_group_ids(keys, merged)
    split = ids.copy()
    last = int(np.nonzero(ids == 0)[0][-1])
    split[last] = g  # one key gets two ids
    assert not bench.check_group_ids(keys, split)
    assert bench.check_group_ids(np.zeros(0, np.int64), np.zeros(0, np.int64))


def test_zipf_inverse_cdf():
    cdf = bench.zipf_cdf(1000, 1.1)
    u = O.rand_f64(5000, 9)
    got = bench.zipf_from_uniform(u, cdf)
    w = np.arange(1, 1001, dtype=np.float64) ** -1.1
    c = np.cumsum(w) / w.sum()
    want = [next(k + 1 for k in range(1000) if v < c[k]) for v in u[:500]]
    assert got[:500].tolist() == want
    assert got.min() >= 1 and got.max() <= 1000


def test_argsort_alg_bytes():
    # 192n for 8 varying byte digits, 72n for 3 (the sizes quoted for the sort workload)
    n = 4096
    assert bench.argsort_alg_bytes(bench.splitmix64_np(7, 0, n), 0) == 192 * n
    z = np.random.default_rng(3).integers(1, 1 << 20, n).astype(np.int64)
    z[0], z[1] = 1, (1 << 20)
    assert bench.argsort_alg_bytes(z, 1) == 72 * n
