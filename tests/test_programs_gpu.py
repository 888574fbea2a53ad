"""Plan replay of the reference's program tests (tests/programs/*.dd) under the hip variant.

Each block's statements are written as the kernel plans a planner produces for them (the
README's plan format, `(dot_f64 x y)`), evaluated through verkehrt_amd.plan.Plan on the device,
printed in the reference's `print` format and compared with the block's expected output lines
(tests/golden/programs_expected.json, written by tests/golden/make_programs.py). Runtime errors
compare the kernel's message with the text after the reference's `Runtime error at L:C: `.

The plans are derived by hand from each program. Statements whose value is a scalar-only
expression (`div_const(100, 7)`, `and(12, 10)`) are evaluated by the reference's scalar
entries, which this backend does not provide; those lines are skipped (marked SKIP below).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXPECTED = json.load(open(os.path.join(HERE, "golden", "programs_expected.json")))
SKIP = ("skip",)

F = np.float64
I = np.int64

PROGRAMS = {
    "arrays.dd:dot": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0, 5.0, 6.0], F),
        ("print", "(dot_f64 x y)")],
    "arrays.dd:lambda": [
        ("bind", "x", [1.0, -2.0, 3.0], F), ("bind", "y", [10.0, 20.0, 30.0], F),
        ("print", "(select_f64 (gt_f64 x 0.0) x (sub_f64 0.0 x))"),
        ("print", "(add_f64 x y)"), ("print", "(mul_f64 x 2.0)"), ("print", "(mul_f64 x 0.5)")],
    "arrays.dd:filter": [
        ("bind", "x", [1.0, -2.0, 3.0, -4.0, 5.0], F),
        ("let", "big", "(filter_gt_f64 x 0.0)"), ("print", "big"),
        ("print", "(count_gt_f64 x 0.0)"), ("print", "(where (gt_f64 x 0.0))"),
        ("let", "none", "(filter_gt_f64 x 100.0)"), ("print", "(reduce_add_f64 none)"), ("print", "none")],
    "arrays.dd:gather": [
        ("let", "x", "(iota 10)"), ("let", "s", "(slice x 2 8)"), ("print", "s"), ("print", "(slice s 1 3)"),
        ("bind", "idx", [9, 0, 4, 4], I), ("print", "(gather_i64 (mul_i64 x 10) idx)"),
        ("print", "(reduce_add_i64 (gather_i64 x (slice x 0 5)))")],
    "arrays.dd:prev": [
        ("let", "x", "(iota 5)"), ("bind", "a", [1, 2, 3], I), ("print", "(prev_i64 a 0)"),
        ("print", "(sub_i64 x (prev_i64 x 0))")],
    "arrays.dd:string_len": [
        ("bind", "ends", [3, 3, 7, 7, 8], I), ("let", "len", "(sub_i64 ends (prev_i64 ends 0))"),
        ("print", "len"), ("print", "(ne_i64 len 0)")],
    "arrays.dd:reduce": [
        ("bind", "x", [3.0, 1.0, 4.0, 1.0, 5.0], F), ("print", "(reduce_max_f64 x)"),
        ("print", "(reduce_min_f64 x)"), ("print", "(reduce_mul_f64 x)"),
        ("bind", "b", [200, 100], np.uint8), ("print", "(reduce_add_u8 b)"), ("print", "(argmin_f64 x)")],
    "arith.dd:bitwise": [
        ("bind", "x", [12, 10, 6], I), ("bind", "y", [10, 3, 5], I), ("print", "(and_i64 x 10)"),
        ("print", "(xor_i64 (or_i64 x y) (and_i64 x y))"), ("print", "(not_i64 x)")],
    "arith.dd:div_lambda": [
        ("bind", "x", [60, 125, 7, 3600], I), ("print", "(div_const_i64 x 60)"),
        ("print", "(div_const_i64 x 60)"), ("print", "(div_const_i64 x 3)"),
        ("print", "(div_const_i64 (sub_i64 x 1) 10)"), SKIP],
    "arith.dd:div_sort": [
        ("bind", "c", [5, 3, 9, 3, 7], I), ("print", "(div_const_i64 c 2)"),
        ("print", "(div_const_i64 (sub_i64 0 c) 2)"), ("print", "(div_const_i64 c -2)"),
        ("print", "(div_const_i64 c 3)"), SKIP, ("print", "(argsort_i64 c)"),
        ("print", "(argsort_i64 (sub_i64 0 c))"), ("print", "(gather_i64 c (argsort_i64 c))"),
        ("print", "(argsort_u8 (cast_i64_u8 c))")],
    "arith.dd:bitwise_on_integers": [
        SKIP, SKIP, SKIP, SKIP, SKIP,
        ("bind", "a", [12, 5], I), ("bind", "b", [10, 3], I), ("print", "(and_i64 a b)"),
        ("print", "(or_i64 a b)"), ("print", "(xor_i64 a b)"), ("print", "(not_i64 a)"),
        ("bind", "z", [0, 255], np.uint8), ("print", "(not_u8 z)"),
        ("bind", "u", [12, 5], np.uint64), ("print", "(and_u64 u 10u64)"),
        ("bind", "p", [12], np.int32), ("bind", "q", [10], np.int32), ("print", "(xor_i32 p q)")],
    "arith.dd:xor_on_bool": [
        ("bind", "a", [1, 1, 0], "bool"), ("bind", "b", [1, 0, 0], "bool"), ("print", "(xor_bool a b)")],
    "arith.dd:reduce_bitwise_on_integers": [
        ("bind", "x", [12, 10, 6], I), ("print", "(reduce_and_i64 x)"), ("print", "(reduce_or_i64 x)"),
        ("print", "(reduce_xor_i64 x)"), ("let", "e", "(filter_gt_i64 x 100)"), ("print", "(reduce_and_i64 e)"),
        ("print", "(reduce_or_i64 e)"), ("print", "(reduce_xor_i64 e)"),
        ("bind", "one", [1], np.uint8), ("let", "u", "(compress_u8 one (gt_u8 one 100u8))"),
        ("print", "(reduce_and_u8 u)")],
    "grouping.dd:scan": [
        ("bind", "x", [1, 2, 3, 4, 5], I), ("print", "(scan_add_i64 x)"),
        ("bind", "s", [0, 0, 1, 0, 1], "bool"), ("print", "(scan_by_add_i64 x s)")],
    "grouping.dd:group_id": [
        ("bind", "k", [7, -1, 7, 3, -1, 7], I), ("let", "g", "(group_id_i64 k)"), ("print", "g"),
        ("print", "(reduce_max_i64 (add_i64 g 1))"),
        ("bind", "a", [1, 2, 1, 2], I), ("bind", "b", [2, 1, 2, 2], I), ("print", "(group_id2_i64_i64 a b)"),
        ("bind", "c", [5, 5, 9], np.uint8), ("bind", "d", [-1, -1, -1], np.int32), ("print", "(group_id2_u8_i32 c d)"),
        ("print", "(length_i64 (group_id_i64 (compress_i64 k (gt_i64 k 100))))")],
    "zip_map.dd:zip_map_first_operand": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0, 5.0, 6.0], F), ("print", "(select_f64 true x y)")],
    "zip_map.dd:zip_map_first_operand_length_mismatch": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0], F), ("error", "(select_f64 true x y)")],
    "zip_map.dd:zip_map_second_operand": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0, 5.0, 6.0], F), ("print", "(select_f64 true y x)")],
    "zip_map.dd:zip_map_second_operand_length_mismatch": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0], F), ("error", "(select_f64 true y x)")],
    "zip_map.dd:zip_map_constant_body": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0, 5.0, 6.0], F),
        ("print", "(select_f64 true 1.0 (select_f64 true x y))")],
    "zip_map.dd:zip_map_constant_body_length_mismatch": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0], F),
        ("error", "(select_f64 true 1.0 (select_f64 true x y))")],
    "zip_map.dd:zip_map_sum": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0, 5.0, 6.0], F), ("print", "(add_f64 x y)")],
    "zip_map.dd:zip_map_sum_length_mismatch": [
        ("bind", "x", [1.0, 2.0, 3.0], F), ("bind", "y", [4.0], F), ("error", "(add_f64 x y)")],
    "zip_map.dd:zip_map_fused_with_a_product": [
        ("bind", "x", [1.0, 2.0], F), ("bind", "y", [3.0, 4.0], F), ("print", "(dot_f64 x y)")],
}

# sum_rewrites.dd: sum(a + k), sum(k + a), sum(a - k), sum(k - a), sum(a * k), sum(k * a), kernel
# by kernel (the fused form is tests/test_region_gpu.py)
for _t, _a, _k in [("i64", [2**63 - 1, 2**63 - 1, 5, -2**63], "9223372036854775807i64"),
                   ("u64", [2**64 - 1, 2**64 - 1, 5], "18446744073709551615u64"),
                   ("i32", [2**31 - 1, 2**31 - 1, 5, -2**31], "2147483647i32"),
                   ("u8", [255, 255, 5], "255u8")]:
    _steps = [("bind", "a", _a, {"i64": I, "u64": np.uint64, "i32": np.int32, "u8": np.uint8}[_t])]
    for _op in ("add", "sub", "mul"):
        _steps.append(("print", f"(reduce_add_{_t} ({_op}_{_t} a {_k}))"))
        _steps.append(("print", f"(reduce_add_{_t} ({_op}_{_t} {_k} a))"))
    if _t == "i64":
        _steps += [("let", "none", "(filter_eq_i64 a 1)"), ("print", f"(reduce_add_i64 (add_i64 none {_k}))"),
                   ("print", f"(reduce_add_i64 (sub_i64 {_k} none))")]
    PROGRAMS[f"sum_rewrites.dd:{_t}"] = _steps


@pytest.fixture(scope="module")
def D():
    from verkehrt_amd import Registry
    return Registry().resolve(0)


@pytest.mark.parametrize("block", sorted(PROGRAMS))
def test_program_block(D, block):
    from verkehrt_amd import KernelError
    from verkehrt_amd.plan import Plan, format_value
    want = EXPECTED[block]
    P = Plan(D)
    out = []
    for step in PROGRAMS[block]:
        if step[0] == "bind":
            _, name, rows, t = step
            P.bind(name, np.array(rows, np.uint8 if t == "bool" else t), "bool" if t == "bool" else None)
        elif step[0] == "let":
            P.run(step[2], bind=step[1])
        elif step[0] == "print":
            out.append(format_value(P.run(step[1])))
        elif step[0] == "error":
            with pytest.raises(KernelError) as e:
                P.run(step[1])
            out.append(str(e.value))
        else:
            out.append(None)
    assert want["kind"] == ("error" if any(s[0] == "error" for s in PROGRAMS[block]) else "run")
    assert len(want["expected"]) == len(out), (block, out)
    for i in range(len(out)):
        got, exp = out[i], want["expected"][i]
        if got is None:
            continue
        if want["kind"] == "error":
            exp = exp.split(": ", 1)[1]
        assert got == exp, (block, i, got, exp)
