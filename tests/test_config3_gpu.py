"""The sort and grouping workload at full size against the oracle, bit for bit.

- argsort of 2^27 keys: u64 uniform (splitmix64(7 + i)), i64 Zipf(1.1) over 2^20 values (the
  bench's inverse-CDF keys), f64 normal(0,1) with +0.0 and -0.0 injected; the device result
  against the oracle's LSD radix restatement (radix.rs:187-259).
- group-by sum over i64 keys pool[splitmix64(12 + i) mod G] with G = 1K, 1M and 64M at 2^28
  rows (the bench's size): group_id against the oracle's first-occurrence table,
  reduce_max + 1, then reduce_into_add_f64 of rand_f64 values against the oracle's defined
  order.

Each case is one test, so a slow host shows which case it is.
"""
import numpy as np
import pytest

import bench
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def S(t, v):
    from verkehrt_amd import Scalar
    return Scalar(t, v)


@pytest.mark.parametrize("case", ["u64", "i64_zipf", "f64_normal"])
def test_argsort_full_size(D, case):
    from verkehrt_amd import Array
    n = 1 << 27
    if case == "u64":
        keys = D.call("cast_i64_u64", [D.call("rand_i64", [S("i64", n), S("i64", 7)])])
        name = "argsort_u64"
    elif case == "i64_zipf":
        keys = bench.zipf_device(D, n, 9, bench.zipf_cdf())
        name = "argsort_i64"
    else:
        rng = np.random.default_rng(3)
        f = rng.normal(size=n)
        f[rng.integers(0, n, 1000)] = 0.0
        f[rng.integers(0, n, 1000)] = -0.0
        keys = Array.from_numpy(D.ctx, f)
        del f
        name = "argsort_f64"
    got = D.call(name, [keys]).to_numpy()
    host = keys.to_numpy()
    del keys
    if case == "i64_zipf":
        assert np.array_equal(host[:1 << 16], bench.zipf_from_uniform(O.rand_f64(1 << 16, 9), bench.zipf_cdf()))
        assert host.min() >= 1 and host.max() <= 1 << 20
    want = O.argsort(host)
    assert np.array_equal(got, want), case


@pytest.mark.parametrize("groups", [1 << 10, 1 << 20, 1 << 26])
def test_group_sum_full_size(D, groups):
    n = 1 << 28
    pool = D.call("rand_i64", [S("i64", groups), S("i64", 11)])
    sel = D.call("and_i64", [D.call("rand_i64", [S("i64", n), S("i64", 12)]), S("i64", groups - 1)])
    keys = D.call("gather_i64", [pool, sel])
    del pool, sel
    ids = D.call("group_id_i64", [keys])
    ng = D.call("reduce_max_i64", [ids]).value + 1
    vals = D.call("rand_f64", [S("i64", n), S("i64", 1)])
    sums = D.call("reduce_into_add_f64", [ids, vals, S("i64", ng)]).to_numpy()
    kh = keys.to_numpy()
    hpool = bench.splitmix64_np(11, 0, groups).view(np.int64)
    assert np.array_equal(kh[:1 << 16], hpool[(bench.splitmix64_np(12, 0, 1 << 16) % np.uint64(groups)).astype(np.int64)])
    want_ids = O.group_id(kh)
    del kh
    assert np.array_equal(ids.to_numpy(), want_ids), groups
    assert ng == int(want_ids.max()) + 1
    want = O.reduce_into(want_ids, O.rand_f64(n, 1), ng)
    assert np.array_equal(sums.view(np.uint64), want.view(np.uint64)), groups
