"""Batched calls (vk_batch_call): every item's result equals the single call's result and the
oracle's, on batches of uneven arrays (empty, shorter than a tile, several tiles).

The batched form has no counterpart in the reference; its parity bar is the single entries':
bit-exact results and identical error strings.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 15, 16, 17, 300, 4095, 4096, 4097, 16383, 16384, 16385, 70000, 200003]


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def S(t, v):
    from verkehrt_amd import Scalar
    return Scalar(t, v)


def bits(v):
    return np.float64(v).view(np.uint64)


def uneven(rng, count, lo=0, hi=70000):
    sizes = list(SIZES) + list(rng.integers(lo, hi, count))
    rng.shuffle(sizes)
    return sizes


@pytest.mark.parametrize("c", [0.1, 0.5, 0.9])
def test_filter_batch(D, c):
    rng = np.random.default_rng(11)
    sizes = uneven(rng, 40)
    xs = [rng.random(n) for n in sizes]
    arrs = [D.array(x) for x in xs]
    got = D.batch_call("filter_gt_f64", [[a, S("f64", c)] for a in arrs])
    for x, a, g in zip(xs, arrs, got):
        want = O.filter_cmp(x, "gt", c)
        assert np.array_equal(g.to_numpy().view(np.uint64), want.view(np.uint64))
        one = D.call("filter_gt_f64", [a, S("f64", c)])
        assert np.array_equal(one.to_numpy(), g.to_numpy())


@pytest.mark.parametrize("t", ["f64", "i64", "i32", "u8"])
def test_compress_batch(D, t):
    rng = np.random.default_rng(12)
    sizes = uneven(rng, 30)
    dt = {"f64": np.float64, "i64": np.int64, "i32": np.int32, "u8": np.uint8}[t]
    xs = [rng.integers(0, 200, n).astype(dt) for n in sizes]
    ms = [(rng.random(n) < rng.random()).astype(np.uint8) for n in sizes]
    items = [[D.array(x), D.array(m, "bool")] for x, m in zip(xs, ms)]
    got = D.batch_call(f"compress_{t}", items)
    for x, m, g in zip(xs, ms, got):
        assert np.array_equal(g.to_numpy(), O.compress(x, m))


def test_where_batch(D):
    rng = np.random.default_rng(13)
    sizes = uneven(rng, 30)
    ms = [(rng.random(n) < 0.3).astype(np.uint8) for n in sizes]
    got = D.batch_call("where", [[D.array(m, "bool")] for m in ms])
    for m, g in zip(ms, got):
        assert np.array_equal(g.to_numpy(), np.flatnonzero(m).astype(np.int64))


@pytest.mark.parametrize("name", ["reduce_add_f64", "dot_f64", "sumsq_f64", "condsum_gt_f64"])
def test_canonical_batch(D, name):
    rng = np.random.default_rng(14)
    sizes = uneven(rng, 40, hi=300000)
    xs = [rng.normal(3.0, 5.0, n) for n in sizes]
    ys = [rng.random(n) for n in sizes]
    items, want = [], []
    for x, y in zip(xs, ys):
        a = D.array(x)
        if name == "reduce_add_f64":
            items.append([a])
            want.append(O.canonical_sum_f64(x))
        elif name == "dot_f64":
            items.append([a, D.array(y)])
            want.append(O.dot_f64(x, y))
        elif name == "sumsq_f64":
            items.append([a])
            want.append(O.sumsq_f64(x))
        else:
            items.append([a, S("f64", 1.0), S("f64", 0.25)])
            want.append(O.condsum_f64(x, "gt", 1.0, 0.25))
    got = D.batch_call(name, items)
    for it, w, g in zip(items, want, got):
        assert bits(g.value) == bits(w)
        assert bits(D.call(name, it).value) == bits(w)


def test_batch_errors_and_fallbacks(D):
    # mixed filter constants run item by item and still match
    xs = [np.arange(n, dtype=np.float64) for n in (5, 9000)]
    got = D.batch_call("filter_gt_f64", [[D.array(xs[0]), S("f64", 2.0)], [D.array(xs[1]), S("f64", 100.0)]])
    assert list(got[0].to_numpy()) == [3.0, 4.0]
    assert np.array_equal(got[1].to_numpy(), xs[1][xs[1] > 100.0])
    # a kind error carries the single call's text
    from verkehrt_amd import KernelError
    with pytest.raises(KernelError) as e1:
        D.batch_call("reduce_add_f64", [[D.array(xs[0])], [S("f64", 1.0)]])
    with pytest.raises(KernelError) as e2:
        D.call("reduce_add_f64", [S("f64", 1.0)])
    assert str(e1.value) == str(e2.value)
    # arguments of different lengths within an item
    with pytest.raises(KernelError) as e3:
        D.batch_call("dot_f64", [[D.array(xs[0]), D.array(xs[1])]])
    assert "length mismatch: 5 vs 9000" in str(e3.value)
    # an ordinary kernel runs item by item
    got = D.batch_call("count_gt_f64", [[D.array(x), S("f64", 3.0)] for x in xs])
    assert [g.value for g in got] == [1, 8996]
