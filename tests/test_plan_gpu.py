"""Batch plans (vk_bplan_*): statements over many uneven resident arrays, tiles planned on the
device, the launch sequence captured into one hipGraph and replayed. Every replay's result for
every item must equal the single call on that item, bit for bit (and the oracle's), with the
program's own eager work and synchronisations between replays; errors keep the single call's
text.

The plans have no counterpart in the reference; their parity bar is the single entries'.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 15, 16, 17, 300, 4095, 4096, 4097, 8191, 8192, 8193, 16383, 16384, 16385, 70000,
         200003, 1048577]


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def S(t, v):
    from verkehrt_amd import Scalar
    return Scalar(t, v)


def same_value(a, b):
    from verkehrt_amd import Scalar
    if isinstance(a, Scalar):
        return isinstance(b, Scalar) and a.type == b.type and a.bits() == b.bits()
    x, y = a.to_numpy(), b.to_numpy()
    return x.shape == y.shape and np.array_equal(x.view(np.uint8), y.view(np.uint8))


def single(D, name, args):
    from verkehrt_amd import KernelError
    try:
        return D.call(name, args)
    except KernelError as e:
        return e


@pytest.mark.parametrize("tune", [False, True])
def test_plan_filter_sum_and_reductions(D, tune):
    from verkehrt_amd import Input, KernelError, Stmt
    rng = np.random.default_rng(21)
    sizes = list(SIZES) + list(rng.integers(0, 90000, 30))
    rng.shuffle(sizes)
    xs = [rng.random(n) for n in sizes]
    arrs = [D.array(x) for x in xs]
    plan = D.batch_plan([[a] for a in arrs])
    stmts = [
        ("filter_gt_f64", [Input(0), S("f64", 0.9)]),
        ("reduce_add_f64", [Stmt(0)]),
        ("count_gt_f64", [Input(0), S("f64", 0.5)]),
        ("reduce_max_f64", [Stmt(0)]),
        ("argmax_f64", [Input(0)]),
        ("reduce_add_f64", [Input(0)]),
        ("filter_lt_f64", [Stmt(0), S("f64", 0.95)]),
        ("sumsq_f64", [Stmt(6)]),
    ]
    ids = [plan.add(n, a) for n, a in stmts]
    assert ids == list(range(len(stmts)))
    plan.compile(tune=tune)
    if tune:
        lines = plan.tuning().strip().splitlines()
        assert len(lines) == 5 and all("chose" in l for l in lines), lines
    # single calls, argument by argument as the plan refers to them
    want = []
    for i, a in enumerate(arrs):
        f = D.call("filter_gt_f64", [a, S("f64", 0.9)])
        g = D.call("filter_lt_f64", [f, S("f64", 0.95)])
        want.append([f, single(D, "reduce_add_f64", [f]), single(D, "count_gt_f64", [a, S("f64", 0.5)]),
                     single(D, "reduce_max_f64", [f]), single(D, "argmax_f64", [a]),
                     single(D, "reduce_add_f64", [a]), g, single(D, "sumsq_f64", [g])])
    for rep in range(3):
        plan.run()
        for i, x in enumerate(xs):
            for k in range(len(stmts)):
                w = want[i][k]
                if isinstance(w, KernelError):
                    with pytest.raises(KernelError) as e:
                        plan.result(k, i)
                    assert e.value.message == w.message and e.value.kind == w.kind
                else:
                    assert same_value(plan.result(k, i), w), (rep, i, sizes[i], stmts[k][0])
            # the oracle for the filter and its canonical sum
            kept = O.filter_cmp(x, "gt", 0.9)
            assert plan.result(1, i).bits() == np.float64(O.canonical_sum_f64(kept)).view(np.uint64)
        # every item's results at once: the kept counts and the sums' bits
        assert [int(v) for v in plan.bits(0)] == [len(want[i][0]) for i in range(len(xs))]
        assert [int(v) for v in plan.bits(1)] == [want[i][1].bits() for i in range(len(xs))]
        # eager work between replays: other calls on the same context, then a sync
        junk = D.call("filter_gt_f64", [arrs[-1], S("f64", 0.3)])
        D.call("reduce_add_f64", [junk])
        D.sync()
    plan.close()


def test_plan_compress_and_dot(D):
    from verkehrt_amd import Input, Stmt
    rng = np.random.default_rng(22)
    sizes = [0, 5, 16, 1000, 16385, 300000]
    xs = [rng.random(n) for n in sizes]
    ys = [rng.random(n) for n in sizes]
    ms = [(rng.random(n) < 0.3).astype(np.uint8) for n in sizes]
    items = [[D.array(x), D.array(y), D.array(m, "bool")] for x, y, m in zip(xs, ys, ms)]
    plan = D.batch_plan(items)
    plan.add("compress_f64", [Input(0), Input(2)])
    plan.add("dot_f64", [Input(0), Input(1)])
    plan.add("reduce_add_f64", [Stmt(0)])
    plan.add("count", [Input(2)])
    plan.compile(tune=True)
    for _ in range(2):
        plan.run()
        for i, (x, y, m) in enumerate(zip(xs, ys, ms)):
            c = plan.result(0, i).to_numpy()
            assert np.array_equal(c.view(np.uint64), x[m != 0].view(np.uint64))
            assert plan.result(1, i).bits() == np.float64(O.dot_f64(x, y)).view(np.uint64)
            assert plan.result(2, i).bits() == np.float64(O.canonical_sum_f64(x[m != 0])).view(np.uint64)
            assert plan.result(3, i).value == int((m != 0).sum())
    plan.close()


def test_plan_errors(D):
    from verkehrt_amd import Input, KernelError, Stmt
    a, b = D.array(np.arange(10.0)), D.array(np.arange(12.0))
    plan = D.batch_plan([[a, b]])
    with pytest.raises(KernelError, match="expected i64\\[\\], got f64\\[\\]"):
        plan.add("reduce_add_i64", [Input(0)])
    with pytest.raises(KernelError, match="length mismatch: 10 vs 12"):
        plan.add("dot_f64", [Input(0), Input(1)])
    s = plan.add("reduce_add_f64", [Input(0)])
    with pytest.raises(KernelError, match="has no array result"):
        plan.add("reduce_add_f64", [Stmt(s)])
    with pytest.raises(KernelError, match="cannot be planned"):
        plan.add("scan_add_f64", [Input(0)])
    with pytest.raises(KernelError, match="group count of a group_id statement"):
        plan.add("reduce_into_add_f64", [Input(0), Input(0), S("i64", 4)])
    with pytest.raises(KernelError, match="not compiled"):
        plan.run()
    plan.compile()
    plan.run()
    assert plan.result(s, 0).value == 45.0
    with pytest.raises(KernelError, match="already compiled"):
        plan.add("reduce_add_f64", [Input(1)])
    plan.close()


def test_plan_sort_and_group(D):
    """argsort, group_id and reduce_into_add_f64 over the group ids and group counts inside one
    captured plan, beside filter+sum: every replay against the single calls and the oracle."""
    from verkehrt_amd import Input, Stmt
    rng = np.random.default_rng(23)
    # item 12 has 1.2M distinct keys: more bucket tiles (of 16 buckets) than the fold's grid
    sizes = [0, 1, 15, 16, 17, 300, 4095, 4096, 4097, 8193, 16385, 70000, 1_200_000, 1048577]
    sizes += [int(v) for v in rng.integers(0, 120000, 12)]
    items, host = [], []
    pool = rng.integers(-(1 << 40), 1 << 40, 256)
    for j, n in enumerate(sizes):
        if j % 3 == 0:  # distinct keys: the small group tables overflow past 4096 rows
            k = rng.integers(-(1 << 62), 1 << 62, n)
        elif j % 3 == 1:
            k = pool[rng.integers(0, 256, n)]
        else:
            k = rng.integers(0, 3, n) * (1 << 33)  # few keys, constant low digits
        x = rng.random(n)
        f = rng.standard_normal(n)
        if n > 4:
            f[:4] = [0.0, -0.0, -0.0, 0.0]
        i32 = rng.integers(-(1 << 31), (1 << 31) - 1, n).astype(np.int32)
        u8 = rng.integers(0, 256, n).astype(np.uint8)
        host.append((k, x, f, i32, u8))
        items.append([D.array(k), D.array(x), D.array(f), D.array(i32), D.array(u8, "u8")])
    plan = D.batch_plan(items)
    st_sort = plan.add("argsort_i64", [Input(0)])
    st_gid = plan.add("group_id_i64", [Input(0)])
    st_fold = plan.add("reduce_into_add_f64", [Stmt(st_gid), Input(1), Stmt(st_gid)])
    st_filt = plan.add("filter_gt_f64", [Input(1), S("f64", 0.9)])
    st_sum = plan.add("reduce_add_f64", [Stmt(st_filt)])
    st_fsum = plan.add("reduce_add_f64", [Stmt(st_fold)])
    st_f64 = plan.add("argsort_f64", [Input(2)])
    st_i32 = plan.add("argsort_i32", [Input(3)])
    st_u8 = plan.add("argsort_u8", [Input(4)])
    plan.compile(tune=True)
    want = []
    for it, (k, x, f, i32, u8) in zip(items, host):
        ids = D.call("group_id_i64", [it[0]])
        g = int(ids.to_numpy().max()) + 1 if len(k) else 0
        fold = D.call("reduce_into_add_f64", [ids, it[1], S("i64", g)])
        kept = D.call("filter_gt_f64", [it[1], S("f64", 0.9)])
        want.append({st_sort: D.call("argsort_i64", [it[0]]), st_gid: ids, st_fold: fold, st_filt: kept,
                     st_sum: D.call("reduce_add_f64", [kept]), st_fsum: D.call("reduce_add_f64", [fold]),
                     st_f64: D.call("argsort_f64", [it[2]]), st_i32: D.call("argsort_i32", [it[3]]),
                     st_u8: D.call("argsort_u8", [it[4]]), "groups": g})
    for rep in range(3):
        plan.run()
        for i, (k, x, f, i32, u8) in enumerate(host):
            for st in [st_sort, st_gid, st_fold, st_filt, st_sum, st_fsum, st_f64, st_i32, st_u8]:
                assert same_value(plan.result(st, i), want[i][st]), (rep, i, sizes[i], st)
            if i % 5 == 0:  # the oracle on some items
                assert np.array_equal(plan.result(st_sort, i).to_numpy(), O.argsort(k)), (i, sizes[i])
                ids = O.group_id(k)
                assert np.array_equal(plan.result(st_gid, i).to_numpy(), ids)
                g = want[i]["groups"]
                assert np.array_equal(plan.result(st_fold, i).to_numpy().view(np.uint64),
                                      O.reduce_into(ids, x, g, "add").view(np.uint64))
        assert [int(v) for v in plan.bits(st_gid)] == [w["groups"] for w in want]
        assert [int(v) for v in plan.bits(st_fold)] == [w["groups"] for w in want]
        junk = D.call("argsort_i64", [items[-1][0]])
        D.call("group_id_i64", [junk])
        D.sync()
    plan.close()
