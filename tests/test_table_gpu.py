"""Parity of the rest of the hip kernel table with the oracle, through the C ABI.

- Every elementwise kernel of the table (elementwise.rs:740-912 over this backend's element
  types): rows at view offsets, scalar operands in either position, float specials, the
  integer extremes, and the `div_const` zero-divisor error text.
- Integer products (`reduce_mul_<t>`, reduce.rs:358-366), `reduce_into_{min,max}_i64`
  (index.rs:192-222 with ScanExtreme identities), min/max scans with and without starts.
- The fused kernels against hip's own single-op kernels: each fused kernel must equal the
  pattern it replaces evaluated kernel by kernel on this backend (the check
  tests/fused_oracle.rs:178-222 makes for the scalar registry), on finite and special inputs.
"""
import numpy as np
import pytest

from oracle import oracle as O

from test_parity_gpu import CMPS, LENGTHS, S, err_of, floats, same_array, same_float, view

pytestmark = pytest.mark.gpu

TYPES = ["f64", "f32", "i64", "u64", "i32", "u8", "bool"]


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def rows(rng, t, n, special=True):
    if t == "f64":
        return floats(rng, n, special)
    if t == "f32":
        with np.errstate(over="ignore"):
            return floats(rng, n, special).astype(np.float32)
    if t == "bool":
        return (rng.random(n) < 0.5).astype(np.uint8)
    dt = O.NP[t]
    info = np.iinfo(dt)
    x = rng.integers(info.min, info.max, n, dtype=dt, endpoint=True)
    if n >= 4:
        x[:4] = [info.min, info.max, 0, 1]
    small = rng.integers(0, 5, n).astype(dt)  # repeated small values so comparisons tie
    return np.where(rng.random(n) < 0.3, small, x).astype(dt)


def scalar_of(t, v):
    v = v.item() if hasattr(v, "item") else v
    return S(t, bool(v) if t == "bool" else v)


def parse(name):
    """(op, type, target) of an elementwise kernel name."""
    if name.startswith("cast_"):
        _, src, dst = name.split("_")
        return "cast", src, dst
    if name.startswith("div_const_"):
        return "div_const", name[len("div_const_"):], None
    op, t = name.rsplit("_", 1)
    return op, t, None


def oracle_of(op, t, dst, args):
    if op == "cast":
        return O.cast(args[0], t, dst)
    if op == "select":
        return O.select(args[0], t, args[1], args[2])
    if t == "bool":
        a, b = (np.asarray(args[0], np.uint8), np.asarray(args[1], np.uint8)) if len(args) > 1 else (np.asarray(args[0], np.uint8), None)
        if op == "not":
            return (a ^ 1).astype(np.uint8)
        if op in ("eq", "ne"):
            return O.elementwise(op, "u8", a, b)
        return {"and": a & b, "or": a | b, "xor": a ^ b}[op].astype(np.uint8)
    return O.elementwise(op, t, *args)


def elementwise_names(D):
    return [n for n in D.registry.names() if D.registry.kernel_class(n) == "elementwise"]


def test_every_elementwise_kernel(D):
    rng = np.random.default_rng(40)
    names = elementwise_names(D)
    assert len(names) > 150, len(names)
    for name in names:
        op, t, dst = parse(name)
        arity = 3 if op == "select" else 1 if op in ("neg", "abs", "sqrt", "not", "cast") else 2
        for n in [0, 1, 17, 1000, 4500]:
            if op == "select":
                host = [rows(rng, "bool", n), rows(rng, t, n), rows(rng, t, n)]
                types = ["bool", t, t]
            else:
                host = [rows(rng, t, n) for _ in range(arity)]
                types = [t] * arity
            if op == "div_const":
                d = host[1][0] if n else 7
                d = d if d != 0 else 3
                host[1] = np.full(n, d, host[1].dtype)
            if op == "sqrt":
                host[0] = np.abs(host[0])
            # every operand as rows at an offset; then each operand (but div_const's divisor,
            # always a scalar) once as a scalar
            variants = [None] + list(range(arity)) if n else [None]
            for sc in variants:
                if op == "div_const" and sc is None:
                    sc = 1
                args, ohost = [], []
                for k in range(arity):
                    if k == sc or (op == "div_const" and k == 1):
                        v = host[k][0] if n else host[k].dtype.type(1)
                        args.append(scalar_of(types[k], v))
                        ohost.append(np.full(n, v, host[k].dtype))
                    else:
                        args.append(view(D, host[k], (n + k) % 5, types[k]))
                        ohost.append(host[k])
                if all(not hasattr(a, "handle") for a in args):
                    continue
                want = oracle_of(op, t, dst, ohost)
                got = D.call(name, args).to_numpy()
                assert same_array(want, got), (name, n, sc)
    # div_const by zero is the kernel's runtime error, with the reference's text
    for t in ["i64", "u64", "i32", "u8"]:
        e = err_of(lambda: D.call(f"div_const_{t}", [D.array(np.ones(3, O.NP[t])), S(t, 0)]))
        assert e == ("err", "div_const: division by zero"), t
    # tests/programs/arith.dd (div_sort, bitwise)
    c = D.array(np.array([5, 3, 9, 3, 7], np.int64))
    assert D.call("div_const_i64", [c, S("i64", 2)]).to_numpy().tolist() == [2, 1, 4, 1, 3]
    assert D.call("div_const_i64", [c, S("i64", -2)]).to_numpy().tolist() == [-2, -1, -4, -1, -3]
    x = D.array(np.array([12, 10, 6], np.int64))
    y = D.array(np.array([10, 3, 5], np.int64))
    assert D.call("and_i64", [x, S("i64", 10)]).to_numpy().tolist() == [8, 10, 2]
    xo = D.call("xor_i64", [D.call("or_i64", [x, y]), D.call("and_i64", [x, y])])
    assert xo.to_numpy().tolist() == [6, 9, 3]
    assert D.call("not_i64", [x]).to_numpy().tolist() == [-13, -11, -7]


# --------------------------------------------------------- table reductions

def test_integer_products(D):
    # src/kernels/tests.rs:847-850,1337
    assert D.call("reduce_mul_u8", [D.array(np.array([16, 17], np.uint8))]) == S("u8", 16)
    assert D.call("reduce_mul_i32", [D.array(np.zeros(0, np.int32))]) == S("i32", 1)
    assert D.call("reduce_mul_u64", [D.array(np.array([1 << 63, 2], np.uint64))]) == S("u64", 0)
    rng = np.random.default_rng(41)
    for n in LENGTHS + [1 << 20]:
        for t in ["i64", "u64", "i32", "u8"]:
            dt = O.NP[t]
            x = rng.integers(1, 4, n).astype(dt) if n > 64 else rng.integers(np.iinfo(dt).min, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
            if n > 64:  # mostly odd factors so the product does not vanish
                x |= dt(1)
                x[rng.integers(0, n, 5)] = dt(np.iinfo(dt).max)
            got = D.call(f"reduce_mul_{t}", [view(D, x, n % 7)])
            assert got.type == t and got.value == O.product_int(x), (t, n)


def test_reduce_into_extremes_i64(D):
    I = lambda v: D.array(np.array(v, np.int64))
    # src/kernels/tests.rs:1668: an empty bucket holds the identity
    assert D.call("reduce_into_max_i64", [I([1]), I([4]), S("i64", 2)]).to_numpy().tolist() == [-2**63, 4]
    assert D.call("reduce_into_min_i64", [I([1]), I([4]), S("i64", 2)]).to_numpy().tolist() == [2**63 - 1, 4]
    rng = np.random.default_rng(42)
    for n in LENGTHS + [300_000]:
        for nb in (1, 7, 64, 5000, 100_000):
            k = rng.integers(0, nb, n).astype(np.int64)
            v = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64, endpoint=True)
            for op in ("min", "max"):
                got = D.call(f"reduce_into_{op}_i64", [view(D, k, 1), view(D, v, 2), S("i64", nb)]).to_numpy()
                assert np.array_equal(got, O.reduce_into(k, v, nb, op)), (n, nb, op)
        if n and n % 3 == 0:
            k = rng.integers(0, 5, n).astype(np.int64)
            k[rng.integers(0, n)] = 5
            want = err_of(lambda: O.reduce_into(k, np.zeros(n, np.int64), 5, "max"))
            got = err_of(lambda: D.call("reduce_into_max_i64", [D.array(k), D.array(np.zeros(n, np.int64)), S("i64", 5)]))
            assert want == got


@pytest.mark.parametrize("t", ["f64", "i64", "u64"])
def test_min_max_scans(D, t):
    rng = np.random.default_rng(43)
    for n in LENGTHS + [1_000_003]:
        x = rows(rng, t, n, special=True)
        st = (rng.random(n) < 1 / 16).astype(np.uint8)
        for mx in (False, True):
            nm = "max" if mx else "min"
            got = D.call(f"scan_{nm}_{t}", [view(D, x, n % 7)]).to_numpy()
            assert same_array(O.scan_extreme(x, mx), got), (t, n, nm)
            got = D.call(f"scan_by_{nm}_{t}", [view(D, x, 1), view(D, st, 2, "bool")]).to_numpy()
            assert same_array(O.scan_extreme(x, mx, st), got), (t, n, nm, "by")
    # scan_by over tiles: the state carries across updates and the split never shows
    from verkehrt_amd import Array
    n = 20_000
    x = rows(rng, t, n)
    st = (rng.random(n) < 0.05).astype(np.uint8)
    dx, ds = D.array(x), D.array(st, "bool")
    for mx in (False, True):
        nm = "max" if mx else "min"
        out = Array.empty(D.ctx, t, n)
        sc = D.scanner(f"scan_by_{nm}_{t}")
        for a, b in [(0, 37), (37, 100), (100, 4500), (4500, n)]:
            sc.update([dx.slice(a, b - a), ds.slice(a, b - a)], out.slice(a, b - a))
        assert same_array(O.scan_extreme(x, mx, st), out.to_numpy()), nm


# ---------------------------------------------- fused kernels = their patterns

def fused_lengths(rng):
    return sorted(set([0, 1, 15, 16, 17, 1000] + list(rng.integers(0, 5000, 6)) + [70_001, 1_048_593, 4_194_321]))


def finite_mix(rng, n):
    """Finite values whose products and sums stay finite: -1e3..1e3, and magnitudes from
    subnormal to 2^200 (Mix::Finite of tests/fused_oracle.rs keeps float sums finite)."""
    wide = np.ldexp(rng.uniform(-1, 1, n), rng.integers(-1074, 200, n))
    pick = rng.random(n)
    return np.where(pick < 0.8, rng.uniform(-1e3, 1e3, n), np.where(pick < 0.95, wide, 0.0))


@pytest.mark.parametrize("special", [False, True])
def test_fused_float_sums_equal_their_patterns(D, special):
    rng = np.random.default_rng(44 + special)
    for n in fused_lengths(rng):
        x, y = (floats(rng, n, True), floats(rng, n, True)) if special else (finite_mix(rng, n), finite_mix(rng, n))
        dx, dy = D.array(x), D.array(y)
        # dot_f64 = (reduce_add_f64 (mul_f64 ?x ?y)), fused.rs:341-352
        want = D.call("reduce_add_f64", [D.call("mul_f64", [dx, dy])]).value
        assert same_float(want, D.call("dot_f64", [dx, dy]).value), ("dot", n)
        # sumsq_f64 = (reduce_add_f64 (mul_f64 ?x ?x))
        want = D.call("reduce_add_f64", [D.call("mul_f64", [dx, dx])]).value
        assert same_float(want, D.call("sumsq_f64", [dx]).value), ("sumsq", n)
        if not special:
            assert np.isfinite(want) or n == 0
        # condsum_<cmp>_f64 = (reduce_add_f64 (select_f64 (<cmp>_f64 ?x (Bcast ?c)) ?x (Bcast ?e)))
        for cmp in CMPS:
            for c, e in [(0.5, 0.0), (float(x[0]) if n else 0.0, -2.5)]:
                mask = D.call(f"{cmp}_f64", [dx, S("f64", c)])
                want = D.call("reduce_add_f64", [D.call("select_f64", [mask, dx, S("f64", e)])]).value
                got = D.call(f"condsum_{cmp}_f64", [dx, S("f64", c), S("f64", e)]).value
                assert same_float(want, got), ("condsum", cmp, n)


@pytest.mark.parametrize("t", ["f64", "f32", "i64", "u64", "i32"])
def test_fused_counts_and_filters_equal_their_patterns(D, t):
    rng = np.random.default_rng(46)
    for n in fused_lengths(rng)[:-1]:
        x = rows(rng, t, n)
        dx = D.array(x)
        for cmp in CMPS:
            c = scalar_of(t, x[n // 2] if n else 1)
            mask = D.call(f"{cmp}_{t}", [dx, c])
            # count_<cmp>_<t> = (count (<cmp>_<t> ?x (Bcast ?c))), fused.rs:293-307
            assert D.call(f"count_{cmp}_{t}", [dx, c]).value == D.call("count", [mask]).value, (cmp, n)
            # filter_<cmp>_<t> = (compress_<t> ?x (<cmp>_<t> ?x (Bcast ?c)))
            want = D.call(f"compress_{t}", [dx, mask]).to_numpy()
            got = D.call(f"filter_{cmp}_{t}", [dx, c]).to_numpy()
            assert same_array(want, got), (cmp, n)
