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


@pytest.mark.parametrize("t", ["u64", "i64", "i32", "u8", "f64"])
def test_argsort_batch(D, t):
    rng = np.random.default_rng(15)
    sizes = uneven(rng, 30, hi=100000)
    xs = []
    for n in sizes:
        if t == "f64":
            x = rng.normal(size=n)
            if n > 10:
                x[rng.integers(0, n, 3)] = 0.0
                x[rng.integers(0, n, 3)] = -0.0
        elif t == "u8":
            x = rng.integers(0, 256, n).astype(np.uint8)
        elif t == "i32":
            x = rng.integers(-1000, 1000, n).astype(np.int32)
        elif t == "u64":
            x = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
        else:
            x = rng.integers(-2**40, 2**40, n).astype(np.int64)
        xs.append(x)
    got = D.batch_call(f"argsort_{t}", [[D.array(x, t)] for x in xs])
    for x, g in zip(xs, got):
        assert np.array_equal(g.to_numpy(), O.argsort(x)), (t, x.size)


def test_argsort_batch_constant_digits(D):
    # every item has the same high bytes: those passes are skipped for the whole batch; one
    # item varies only in a byte that is constant in the others
    xs = [np.arange(n, dtype=np.int64) % 200 for n in (5, 300, 20000)]
    xs.append((np.arange(5000, dtype=np.int64) % 3) << 16)
    xs.append(np.full(17, 7, np.int64))
    got = D.batch_call("argsort_i64", [[D.array(x)] for x in xs])
    for x, g in zip(xs, got):
        assert np.array_equal(g.to_numpy(), np.argsort(x, kind="stable"))


REDUCTIONS = ["reduce_add_i64", "reduce_add_u8", "reduce_min_f64", "reduce_max_f64", "argmin_f64",
              "argmax_f64", "reduce_max_i64", "argmin_i32", "reduce_and_u8", "reduce_xor_i64",
              "reduce_or_bool", "count", "any", "all", "count_gt_f64", "count_le_i64"]


def reduction_input(rng, name, n):
    if name.endswith("_f64"):
        x = rng.normal(size=n)
        if n > 100:
            x[rng.integers(0, n)] = np.nan if name.startswith("arg") else x[0]
        return x, "f64"
    if name in ("count", "any", "all") or name.endswith("_bool"):
        return (rng.random(n) < (0.999 if name == "all" else 0.001)).astype(np.uint8), "bool"
    if name.endswith("_u8"):
        return rng.integers(0, 256, n).astype(np.uint8), "u8"
    if name.endswith("_i32"):
        return rng.integers(-50, 50, n).astype(np.int32), "i32"
    return rng.integers(-2**40, 2**40, n).astype(np.int64), "i64"


@pytest.mark.parametrize("name", REDUCTIONS)
def test_reduction_batch(D, name):
    rng = np.random.default_rng(16)
    sizes = [n for n in uneven(rng, 20, lo=1, hi=50000) if n > 0]
    args = []
    for n in sizes:
        x, t = reduction_input(rng, name, n)
        a = [D.array(x, t)]
        if name.startswith("count_"):
            a.append(S(t, 0.25 if t == "f64" else 7))
        args.append(a)
    got = D.batch_call(name, args)
    for a, g in zip(args, got):
        want = D.call(name, a)
        if isinstance(want.value, float) and np.isnan(want.value):
            assert np.isnan(g.value)
        else:
            assert g.value == want.value and g.type == want.type, (name, len(a[0]))


def test_reduction_batch_empty_error(D):
    from verkehrt_amd import KernelError
    xs = [D.array(np.arange(5, dtype=np.float64)), D.array(np.zeros(0)), D.array(np.ones(3))]
    with pytest.raises(KernelError) as e1:
        D.batch_call("reduce_min_f64", [[x] for x in xs])
    with pytest.raises(KernelError) as e2:
        D.call("reduce_min_f64", [xs[1]])
    assert str(e1.value) == str(e2.value)
    assert [g.value for g in D.batch_call("reduce_add_i64", [[D.array(np.zeros(0, np.int64))]])] == [0]


@pytest.mark.parametrize("t", ["i64", "u64", "i32", "u8"])
def test_group_id_batch(D, t):
    rng = np.random.default_rng(17)
    sizes = uneven(rng, 25, hi=60000)
    xs = []
    for i, n in enumerate(sizes):
        g = [3, 256, 5000][i % 3]
        pool = {"i64": rng.integers(-2**62, 2**62, g), "u64": rng.integers(0, 2**63, g, dtype=np.uint64),
                "i32": rng.integers(-2**30, 2**30, g).astype(np.int32),
                "u8": rng.integers(0, 256, g).astype(np.uint8)}[t]
        xs.append(pool[rng.integers(0, len(pool), n)])
    got = D.batch_call(f"group_id_{t}", [[D.array(x, t)] for x in xs])
    for x, g in zip(xs, got):
        assert np.array_equal(g.to_numpy(), O.group_id(x)), (t, x.size)


def test_group_id_batch_small_tables_overflow(D):
    # items whose distinct keys overflow the first small tables (over 8192 slots' worth) beside
    # items that fit them, and items with every key distinct
    rng = np.random.default_rng(23)
    xs = []
    for i, n in enumerate([200000, 3000, 150000, 70000, 1, 100000, 40000]):
        g = [20000, 7, 100000, 256, 1, None, 9000][i]
        if g is None:
            xs.append(rng.permutation(n).astype(np.int64) * 7919 - 3)
        else:
            pool = rng.integers(-2**62, 2**62, g)
            xs.append(pool[rng.integers(0, g, n)])
    got = D.batch_call("group_id_i64", [[D.array(x)] for x in xs])
    for x, g in zip(xs, got):
        assert np.array_equal(g.to_numpy(), O.group_id(x)), x.size


def test_reduce_into_add_f64_batch(D):
    rng = np.random.default_rng(18)
    sizes = uneven(rng, 25, hi=80000)
    items, want = [], []
    for i, n in enumerate(sizes):
        nb = [1, 7, 256, 3000][i % 4]
        k = rng.integers(0, nb, n).astype(np.int64)
        v = rng.normal(2.0, 3.0, n)
        items.append([D.array(k), D.array(v), S("i64", nb)])
        want.append(O.reduce_into(k, v, nb, "add"))
    got = D.batch_call("reduce_into_add_f64", items)
    for it, w, g in zip(items, want, got):
        assert np.array_equal(g.to_numpy().view(np.uint64), np.asarray(w, np.float64).view(np.uint64))
        assert np.array_equal(D.call("reduce_into_add_f64", it).to_numpy().view(np.uint64),
                              np.asarray(w, np.float64).view(np.uint64))


def test_reduce_into_batch_errors(D):
    from verkehrt_amd import KernelError
    ok = [D.array(np.array([0, 1, 1], np.int64)), D.array(np.ones(3)), S("i64", 2)]
    badk = [D.array(np.array([0, 5, 1, 9], np.int64)), D.array(np.ones(4)), S("i64", 3)]
    with pytest.raises(KernelError) as e1:
        D.batch_call("reduce_into_add_f64", [ok, badk])
    with pytest.raises(KernelError) as e2:
        D.call("reduce_into_add_f64", badk)
    assert str(e1.value) == str(e2.value) and "bucket key 5" in str(e1.value)
    neg = [D.array(np.array([0], np.int64)), D.array(np.ones(1)), S("i64", -1)]
    with pytest.raises(KernelError) as e3:
        D.batch_call("reduce_into_add_f64", [ok, neg])
    with pytest.raises(KernelError) as e4:
        D.call("reduce_into_add_f64", neg)
    assert str(e3.value) == str(e4.value)


def test_batched_workload_matches_plain(D):
    """The batched form of the many-arrays program returns the plain dispatch's results."""
    from verkehrt_amd import batch as B
    wl = B.make_workload(D, count=48, lo=1 << 10, hi=1 << 16, seed=7)
    p = B.plain(D, wl)
    b = B.batched(D, wl)
    assert B.same(p, b)
    # spot-check against the oracle
    for i in (0, 17, 47):
        x, k = (a.to_numpy() for a in wl.item(i))
        assert b.sums[i].bits() == np.float64(O.canonical_sum_f64(x[x > 0.9])).view(np.uint64)
        assert np.array_equal(b.perms[i].to_numpy(), O.argsort(k))
        g = O.group_id(k)
        want = O.reduce_into(g, x, int(g.max()) + 1, "add")
        assert np.array_equal(b.groups[i].to_numpy().view(np.uint64), want.view(np.uint64))


def test_batch_mixes_long_items(D):
    # items of 2^22 rows or more run as single calls inside a batched argsort, group_id and
    # reduce_into_add_f64; results, and the first failing item's error, stay the single calls'
    from verkehrt_amd import KernelError
    rng = np.random.default_rng(90)
    sizes = [1000, (1 << 22) + 5, 37, 1 << 22, 0, 5000]
    keys = [rng.integers(-150, 150, n).astype(np.int64) for n in sizes]
    vals = [rng.normal(size=n) for n in sizes]
    dk, dv = [D.array(k) for k in keys], [D.array(v) for v in vals]
    for p, k in zip(D.batch_call("argsort_i64", [[k] for k in dk]), keys):
        assert np.array_equal(p.to_numpy(), O.argsort(k))
    gs = D.batch_call("group_id_i64", [[k] for k in dk])
    ids = [g.to_numpy() for g in gs]
    for g, k in zip(ids, keys):
        assert np.array_equal(g, O.group_id(k))
    ng = [int(g.max()) + 1 if g.size else 0 for g in ids]
    rs = D.batch_call("reduce_into_add_f64", [[g, v, S("i64", n)] for g, v, n in zip(gs, dv, ng)])
    for r, g, v, n in zip(rs, ids, vals, ng):
        assert np.array_equal(r.to_numpy().view(np.uint64), O.reduce_into(g, v, n).view(np.uint64))
    # a bad key in the long item 1 and in the short item 2: item 1's error, as in order
    bad = [k.copy() for k in ids]
    bad[1][123] = 10**6
    bad[2][3] = -1
    items = [[D.array(b), v, S("i64", n)] for b, v, n in zip(bad, dv, ng)]
    with pytest.raises(KernelError) as e1:
        D.batch_call("reduce_into_add_f64", items)
    with pytest.raises(KernelError) as e2:
        D.call("reduce_into_add_f64", items[1])
    assert str(e1.value) == str(e2.value) and "bucket key 1000000" in str(e1.value)
