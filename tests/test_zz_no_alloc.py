"""vk_select_into, the selection entry that allocates nothing, against vk_call's selection on
the same inputs, bit for bit: two calls in a row (the second on a smaller input), every
selection kind, an empty input, the size limit and destination errors, enough calls to wrap
the look-back epoch of its fixed scratch array, and no growth of the context's memory pool."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def S(t, v):
    from verkehrt_amd import Scalar
    return Scalar(t, v)


def into(D, name, args, out_type, rows):
    from verkehrt_amd import Array
    out = Array.empty(D.ctx, out_type, rows)
    k = D.select_into(name, args, out)
    return out.slice(0, k).to_numpy() if k else np.zeros(0, out.to_numpy().dtype), k


def same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_select_into_equals_vk_call(D):
    rng = np.random.default_rng(60)
    x = D.array(rng.random(3_000_017))
    # two calls in a row, the second on a smaller input
    for n, c in [(3_000_017, 0.5), (100_003, 0.9)]:
        xs = x.slice(0, n)
        got, k = into(D, "filter_gt_f64", [xs, S("f64", c)], "f64", n)
        want = D.call("filter_gt_f64", [xs, S("f64", c)]).to_numpy()
        assert k == want.size and same(got, want), n
    xi = D.array(rng.integers(-2**62, 2**62, 1_000_003).astype(np.int64))
    m = D.array((rng.random(1_000_003) < 0.3).astype(np.uint8), "bool")
    for n in [1_000_003, 40_000, 17, 0]:
        args = [xi.slice(0, n), m.slice(0, n)]
        got, k = into(D, "compress_i64", args, "i64", n)
        assert same(got, D.call("compress_i64", args).to_numpy()), n
        got, k = into(D, "where", [m.slice(0, n)], "i64", n)
        assert same(got, D.call("where", [m.slice(0, n)]).to_numpy()), n
    # a destination longer than the input is fine: only the kept rows are written
    from verkehrt_amd import Array
    out = Array.empty(D.ctx, "f64", 5000)
    k = D.select_into("filter_lt_f64", [x.slice(0, 1000), S("f64", 0.25)], out)
    assert same(out.slice(0, k).to_numpy(), D.call("filter_lt_f64", [x.slice(0, 1000), S("f64", 0.25)]).to_numpy())


def test_select_into_errors(D):
    from verkehrt_amd import Array, KernelError
    x = D.array(np.arange(100, dtype=np.float64))
    with pytest.raises(KernelError, match="^destination of 99 rows is shorter than the input's 100 rows$"):
        D.select_into("filter_gt_f64", [x, S("f64", 1.0)], Array.empty(D.ctx, "f64", 99))
    with pytest.raises(KernelError, match="^filter_gt_f64 produces f64 rows, got a i64 destination$"):
        D.select_into("filter_gt_f64", [x, S("f64", 1.0)], Array.empty(D.ctx, "i64", 100))
    with pytest.raises(KernelError, match="^dot_f64 is not a selection$"):
        D.select_into("dot_f64", [x, x], Array.empty(D.ctx, "f64", 100))
    with pytest.raises(KernelError, match="^expected f64, got i64$"):
        D.select_into("filter_gt_f64", [x, S("i64", 1)], Array.empty(D.ctx, "f64", 100))
    # beyond the limit: a real (zero) mask of 2^31 + 1 rows and a short destination, so that
    # even the destination check alone would stop the call before any launch
    big = D.call("fill_bool", [S("i64", (1 << 31) + 1), S("bool", False)])
    with pytest.raises(KernelError, match="^vk_select_into takes at most 2147483648 rows, got 2147483649$"):
        D.select_into("where", [big], Array.empty(D.ctx, "i64", 10))
    del big


def test_select_into_allocates_nothing_and_wraps_its_epoch():
    from verkehrt_amd import Array, Registry
    D = Registry().resolve(0)  # a fresh context: its pool holds only what this test allocates
    rng = np.random.default_rng(61)
    n = 20_000
    xh = rng.random(n)
    x = D.array(xh)
    out = Array.empty(D.ctx, "f64", n)
    cs = [0.1, 0.5, 0.9, 0.99]
    want = {c: xh[xh > c] for c in cs}
    D.ctx.sync()
    held0, used0 = D.ctx.pool_bytes()
    # 70,000 calls take 70,000 epochs of the device's 16-bit epoch: at least one wrap
    for k in range(70_000):
        c = cs[k % 4]
        kept = D.select_into("filter_gt_f64", [x, S("f64", c)], out)
        assert kept == want[c].size, k
        if k % 4096 == 0 or 65_520 <= k <= 65_560 or k >= 69_996:
            assert np.array_equal(out.slice(0, kept).to_numpy(), want[c]), k
    D.ctx.sync()
    assert D.ctx.pool_bytes() == (held0, used0)
