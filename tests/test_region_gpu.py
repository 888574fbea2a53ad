"""Fused regions (vk_region_reduce): elementwise chains feeding one reduction, run as one device
pass, against the oracle evaluating the same kernels one by one (and against this backend's
own kernel-by-kernel run), bit for bit, with the same errors. Includes the reference's
tests/programs/sum_rewrites.dd (sums of an array combined with a broadcast scalar, wrapping in
every width) and shapes that are not one chain, which must run kernel by kernel."""
import math

import numpy as np
import pytest

from oracle import oracle as O

from test_parity_gpu import floats, same_float

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


def S(t, v):
    from verkehrt_amd import Scalar
    return Scalar(t, v)


def host_value(v):
    from verkehrt_amd import Scalar
    if isinstance(v, Scalar):
        return v.value
    return v.to_numpy()


def oracle_region(kernels, root, root_args, inputs, types):
    """The region evaluated on the host with the oracle: each kernel in order, then the root."""
    vals = []

    def get(s):
        kind, i = s
        return vals[i] if kind == "k" else host_value(inputs[i])

    for name, srcs in kernels:
        op, t = name.rsplit("_", 1)
        args = [get(s_) for s_ in srcs]
        n = next(np.size(a) for a in args if np.ndim(a) > 0)  # the rows of the array operands
        args = [np.full(n, a, O.NP[t]) if np.ndim(a) == 0 else a for a in args]
        vals.append(O.elementwise(op, t, *args))
    r = [get(s_) for s_ in root_args]
    x = r[0]
    if root == "reduce_add_f64":
        return O.canonical_sum_f64(x)
    if root.startswith("reduce_add_"):
        return O.sum_int(np.ascontiguousarray(x, O.NP[root.rsplit("_", 1)[1]]))
    if root.startswith("reduce_min_") or root.startswith("reduce_max_"):
        return O.reduce_extreme(x, root.startswith("reduce_max_"))
    if root.startswith("argmin_") or root.startswith("argmax_"):
        return O.arg_extreme(x, root.startswith("argmax_"))
    if root == "count":
        return O.count_bool(x)
    cmp = root.split("_")[1]
    return O.count_cmp(x, cmp, r[1])


def run(D, kernels, root, root_args, inputs):
    """(fused result, fused flag, kernel-by-kernel result) on the device, errors as strings."""
    from verkehrt_amd import KernelError
    try:
        got, fused = D.region_reduce(kernels, root, root_args, inputs)
    except KernelError as e:
        got, fused = ("err", str(e)), None
    vals = []

    def get(s_):
        kind, i = s_
        return vals[i] if kind == "k" else inputs[i]

    try:
        for name, srcs in kernels:
            vals.append(D.call(name, [get(s_) for s_ in srcs]))
        want = D.call(root, [get(s_) for s_ in root_args])
    except KernelError as e:
        want = ("err", str(e))
    return got, fused, want


def same(a, b):
    # bit for bit, except that a NaN result only needs to be NaN (tests/isa_variants.rs): NaN
    # payloads depend on the order two NaNs meet in an add, which the fused pass may change
    from verkehrt_amd import Scalar
    if isinstance(a, tuple) or isinstance(b, tuple):
        return a == b
    if not (isinstance(a, Scalar) and isinstance(b, Scalar) and a.type == b.type):
        return False
    if a.type in ("f64", "f32") and math.isnan(a.value):
        return math.isnan(b.value)
    return a.bits() == b.bits()


def check_value(got, want_host):
    if isinstance(got, tuple):
        return False
    if isinstance(want_host, float) or isinstance(want_host, np.floating):
        return same_float(float(want_host), float(got.value))
    return int(got.value) == int(want_host)


F64_SHAPES = [
    # sum((x * 2.0) + 1.0)
    ([("mul_f64", [("in", 0), ("in", 2)]), ("add_f64", [("k", 0), ("in", 3)])], "reduce_add_f64", [("k", 1)], True),
    # sum(1.0 - x * y), a second input under the canonical sum
    ([("mul_f64", [("in", 0), ("in", 1)]), ("sub_f64", [("in", 3), ("k", 0)])], "reduce_add_f64", [("k", 1)], True),
    # sum((x - y) * (x - y)): the chain value twice
    ([("sub_f64", [("in", 0), ("in", 1)]), ("mul_f64", [("k", 0), ("k", 0)])], "reduce_add_f64", [("k", 1)], True),
    # sum(sqrt(abs(x)) / 3.0), unary ops
    ([("abs_f64", [("in", 0)]), ("sqrt_f64", [("k", 0)]), ("div_f64", [("k", 1), ("in", 4)])], "reduce_add_f64",
     [("k", 2)], True),
    # min(max(x, 0.5) * -1.0), max(neg(x) + x)
    ([("max_f64", [("in", 0), ("in", 5)]), ("mul_f64", [("k", 0), ("in", 6)])], "reduce_min_f64", [("k", 1)], True),
    ([("neg_f64", [("in", 0)]), ("add_f64", [("k", 0), ("in", 0)])], "reduce_max_f64", [("k", 1)], True),
    # argmax(x * 2.0), argmin(2.0 - x)
    ([("mul_f64", [("in", 0), ("in", 2)])], "argmax_f64", [("k", 0)], True),
    ([("sub_f64", [("in", 2), ("in", 0)])], "argmin_f64", [("k", 0)], True),
    # count((x * 2.0) > 1.0), count(0.5 < x + 0.25), count_le_f64(x * x, 0.25)
    ([("mul_f64", [("in", 0), ("in", 2)]), ("gt_f64", [("k", 0), ("in", 3)])], "count", [("k", 1)], True),
    ([("add_f64", [("in", 0), ("in", 7)]), ("lt_f64", [("in", 5), ("k", 0)])], "count", [("k", 1)], True),
    ([("mul_f64", [("in", 0), ("in", 0)])], "count_le_f64", [("k", 0), ("in", 7)], True),
    # not one chain: (x + 1.0) * (y + 2.0); min over two inputs; a comparison with an array
    ([("add_f64", [("in", 0), ("in", 3)]), ("add_f64", [("in", 1), ("in", 2)]), ("mul_f64", [("k", 0), ("k", 1)])],
     "reduce_add_f64", [("k", 2)], False),
    ([("mul_f64", [("in", 0), ("in", 1)])], "reduce_min_f64", [("k", 0)], False),
    ([("gt_f64", [("in", 0), ("in", 1)])], "count", [("k", 0)], False),
]


@pytest.mark.parametrize("special", [False, True])
def test_f64_regions(D, special):
    rng = np.random.default_rng(70 + special)
    for n in [0, 1, 15, 16, 17, 1000, 70_001, (1 << 20) + 5, (4 << 20) + 3]:
        xh = floats(rng, n, special) if special else rng.normal(size=n) * 10
        yh = floats(rng, n, special) if special else rng.normal(size=n)
        inputs = [D.array(xh), D.array(yh), S("f64", 2.0), S("f64", 1.0), S("f64", 3.0), S("f64", 0.5),
                  S("f64", -1.0), S("f64", 0.25)]
        for kernels, root, rargs, chain in F64_SHAPES:
            got, fused, want = run(D, kernels, root, rargs, inputs)
            assert same(got, want), (n, root, kernels, got, want)
            if isinstance(got, tuple):
                assert n == 0 and root.startswith(("reduce_m", "arg")), got
                continue
            assert fused == chain, (root, kernels)
            assert check_value(got, oracle_region(kernels, root, rargs, inputs, None)), (n, root, kernels)


@pytest.mark.parametrize("t", ["i64", "u64", "i32", "u8"])
def test_integer_regions(D, t):
    rng = np.random.default_rng(72)
    info = np.iinfo(O.NP[t])
    wide = {"i64": "i64", "u64": "u64", "i32": "i64", "u8": "u64"}[t]
    for n in [0, 1, 17, 4500, 300_001]:
        xh = rng.integers(info.min, info.max, n, dtype=O.NP[t], endpoint=True)
        yh = rng.integers(info.min, info.max, n, dtype=O.NP[t], endpoint=True)
        k = int(info.max)
        inputs = [D.array(xh), D.array(yh), S(t, k), S(t, 3), S(t, 5)]
        shapes = [
            ([("add_" + t, [("in", 0), ("in", 2)])], "reduce_add_" + t, [("k", 0)], True),
            ([("sub_" + t, [("in", 2), ("in", 0)]), ("mul_" + t, [("k", 0), ("in", 3)])], "reduce_add_" + t, [("k", 1)], True),
            ([("xor_" + t, [("in", 0), ("in", 4)]), ("and_" + t, [("k", 0), ("in", 2)])], "reduce_max_" + t, [("k", 1)], True),
            ([("mul_" + t, [("in", 0), ("in", 0)])], "argmin_" + t, [("k", 0)], True),
            ([("add_" + t, [("in", 0), ("in", 3)]), ("gt_" + t, [("k", 0), ("in", 4)])], "count", [("k", 1)], True),
            ([("add_" + t, [("in", 0), ("in", 1)])], "reduce_add_" + t, [("k", 0)], False),
        ]
        if t in ("i64", "i32"):
            shapes.append(([("neg_" + t, [("in", 0)]), ("abs_" + t, [("k", 0)])], "reduce_min_" + t, [("k", 1)], True))
        for kernels, root, rargs, chain in shapes:
            got, fused, want = run(D, kernels, root, rargs, inputs)
            assert same(got, want), (t, n, root, kernels)
            if isinstance(got, tuple):
                assert n == 0
                continue
            assert fused == chain, (t, root, kernels)
            assert got.type == (wide if root.startswith("reduce_add") else "i64" if root.startswith(("arg", "count")) else t)
            assert check_value(got, oracle_region(kernels, root, rargs, inputs, None)), (t, n, root, kernels)


def test_sum_rewrites_program(D):
    # tests/programs/sum_rewrites.dd: sum(a + k), sum(k + a), sum(a - k), sum(k - a), sum(a * k),
    # sum(k * a) for each width, as one region each, with the file's expected outputs
    cases = {
        "i64": ([2**63 - 1, 2**63 - 1, 5, -2**63], 2**63 - 1,
                [2**63 - 1, 2**63 - 1, -9223372036854775801, 9223372036854775801, -3, -3]),
        "u64": ([2**64 - 1, 2**64 - 1, 5], 2**64 - 1, [0, 0, 6, 18446744073709551610, 18446744073709551613,
                                                    18446744073709551613]),
        "i32": ([2**31 - 1, 2**31 - 1, 5, -2**31], 2**31 - 1, [-2147483649, -2147483649, -2147483641, 2147483641, -3, -3]),
        "u8": ([255, 255, 5], 255, [512, 512, 6, 250, 253, 253]),
    }
    for t, (a, k, want) in cases.items():
        inputs = [D.array(np.array(a, O.NP[t])), S(t, k)]
        got = []
        for op in ("add", "sub", "mul"):
            for order in ([("in", 0), ("in", 1)], [("in", 1), ("in", 0)]):
                r, fused = D.region_reduce([(f"{op}_{t}", order)], f"reduce_add_{t}", [("k", 0)], inputs)
                assert fused
                got.append(r.value)
        # the file prints sum(a+k), sum(k+a), sum(a-k), sum(k-a), sum(a*k), sum(k*a)
        assert got == want, (t, got)
    # sum(filter(a, |v| v == 1) + k) and sum(k - filter(...)): the empty-array result
    a = D.array(np.array([2**63 - 1, 2**63 - 1, 5, -2**63], np.int64))
    none = D.call("filter_eq_i64", [a, S("i64", 1)])
    for order in ([("in", 0), ("in", 1)], [("in", 1), ("in", 0)]):
        r, _ = D.region_reduce([("add_i64" if order[0][1] == 0 else "sub_i64", order)], "reduce_add_i64",
                               [("k", 0)], [none, S("i64", 2**63 - 1)])
        assert r.value == 0


def test_region_errors_match_kernel_by_kernel(D):
    x = D.array(np.arange(10, dtype=np.float64))
    y = D.array(np.arange(9, dtype=np.float64))
    e = np.zeros(0)
    for kernels, root, rargs, inputs in [
        ([("mul_f64", [("in", 0), ("in", 1)])], "reduce_add_f64", [("k", 0)], [x, y]),           # lengths
        ([("mul_f64", [("in", 0), ("in", 1)])], "reduce_add_f64", [("k", 0)], [x, S("i64", 2)]),  # kinds
        ([("mul_f64", [("in", 0), ("in", 1)])], "reduce_min_f64", [("k", 0)], [D.array(e), S("f64", 2.0)]),
        ([("mul_f64", [("in", 0), ("in", 1)])], "argmax_f64", [("k", 0)], [D.array(e), S("f64", 2.0)]),
    ]:
        got, fused, want = run(D, kernels, root, rargs, inputs)
        assert isinstance(want, tuple) and got == want, (got, want)
