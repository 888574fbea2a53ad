"""Batched statements over many uneven arrays.

Every array i of a batch runs the same program, as a verkehrt plan would run it once per
array (the reference's interpreter executes statements one at a time, single threaded,
src/exec.rs:341-365):

    f = filter_gt_f64(x, 0.9);  s = reduce_add_f64(f)            # sum of the kept rows
    p = argsort_i64(k)
    g = group_id_i64(k);  ng = reduce_max_i64(g) + 1;  r = reduce_into_add_f64(g, x, ng)

`plain` dispatches one call per array and statement (Dispatch::call for each); `batched`
issues one vk_batch_call per statement for all arrays, so each statement is one launch per
kernel phase over every array's tiles. Both return the same results, bit for bit: the sum
keeps the canonical order over each array's kept rows, and every other result is exact.

The workload: `count` arrays with lengths log-uniform in [lo, hi] rows (seeded), f64 values
rand_f64 (uniform [0, 1)), and i64 keys drawn from a pool of `keys` distinct splitmix64 words.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List

import numpy as np

from .kernels import Array, Dispatch, Scalar

MASK64 = (1 << 64) - 1


def splitmix64(x: int) -> int:
    """splitmix64 of one word (src/kernels/scalar/construct.rs:10-17)."""
    z = (x + 0x9E3779B97F4A7C15) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


def splitmix64_np(seed: int, start: int, count: int) -> np.ndarray:
    """splitmix64(seed + i) for i in [start, start + count), vectorised."""
    z = (np.arange(start, start + count, dtype=np.uint64) + np.uint64(seed & MASK64)
         + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def lengths(count: int, seed: int, lo: int, hi: int) -> List[int]:
    """Row counts log-uniform in [lo, hi]: 2^(log2 lo + u * log2(hi / lo)) rounded, u from the
    top 53 bits of splitmix64(seed + i)."""
    a, b = math.log2(lo), math.log2(hi)
    out = []
    for i in range(count):
        u = (splitmix64((seed + i) & MASK64) >> 11) * 2.0 ** -53
        out.append(int(round(2.0 ** (a + u * (b - a)))))
    return out


@dataclass
class Workload:
    x: Array          # all arrays' values, concatenated
    k: Array          # all arrays' keys, concatenated
    offsets: List[int]
    lengths: List[int]
    views: List[tuple] = None  # array i's (x, k) views, made once: the resident inputs

    def __post_init__(self):
        if self.views is None:
            self.views = [(self.x.slice(o, n), self.k.slice(o, n)) for o, n in zip(self.offsets, self.lengths)]

    @property
    def rows(self) -> int:
        return sum(self.lengths)

    def item(self, i: int):
        return self.views[i]


def make_workload(D: Dispatch, count: int = 4096, lo: int = 1 << 10, hi: int = 1 << 20,
                  keys: int = 256, seed: int = 41, first: int = 0, take: int | None = None) -> Workload:
    """Arrays first .. first + take - 1 of the `count`-array workload (all by default). Row r of
    the arrays' concatenation holds x = rand_f64 element splitmix64(seed + 1 + r) and
    k = pool[splitmix64(seed + 3 + r) mod keys] with pool[j] = splitmix64(seed + 2 + j), so any
    run of consecutive arrays is generated alone with the values it has in the whole workload.
    Generated on the device (keys: a gather from the pool by masked rand_i64 words)."""
    all_ns = lengths(count, seed, lo, hi)
    take = count - first if take is None else take
    ns = all_ns[first:first + take]
    base = int(sum(all_ns[:first]))
    offs = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64).tolist()
    total = int(sum(ns))
    S = Scalar
    x = D.call("rand_f64", [S("i64", total), S("i64", seed + 1 + base)])
    pool = D.call("rand_i64", [S("i64", keys), S("i64", seed + 2)])
    sel = D.call("and_i64", [D.call("rand_i64", [S("i64", total), S("i64", seed + 3 + base)]), S("i64", keys - 1)])
    k = D.call("gather_i64", [pool, sel])
    del sel
    D.sync()
    return Workload(x, k, offs, ns)


def chunks(count: int, lo: int, hi: int, seed: int = 41, max_rows: int = 1 << 29) -> List[tuple]:
    """(first, take) runs of consecutive arrays of at most max_rows rows each (an array longer
    than max_rows is a run of its own)."""
    ns = lengths(count, seed, lo, hi)
    out, first, rows = [], 0, 0
    for i, n in enumerate(ns):
        if i > first and rows + n > max_rows:
            out.append((first, i - first))
            first, rows = i, 0
        rows += n
    out.append((first, count - first))
    return out


@dataclass
class Results:
    sums: List[Scalar]
    perms: List[Array]
    groups: List[Array]


def plain(D: Dispatch, wl: Workload) -> Results:
    """One call per array and statement."""
    S = Scalar
    sums, perms, groups = [], [], []
    for i in range(len(wl.lengths)):
        x, k = wl.item(i)
        f = D.call("filter_gt_f64", [x, S("f64", 0.9)])
        sums.append(D.call("reduce_add_f64", [f]))
        perms.append(D.call("argsort_i64", [k]))
        g = D.call("group_id_i64", [k])
        ng = D.call("reduce_max_i64", [g]).value + 1
        groups.append(D.call("reduce_into_add_f64", [g, x, S("i64", ng)]))
    D.sync()
    return Results(sums, perms, groups)


def batched(D: Dispatch, wl: Workload) -> Results:
    """One batched call per statement."""
    S = Scalar
    items = [wl.item(i) for i in range(len(wl.lengths))]
    fs = D.batch_call("filter_gt_f64", [[x, S("f64", 0.9)] for x, _ in items])
    sums = D.batch_call("reduce_add_f64", [[f] for f in fs])
    perms = D.batch_call("argsort_i64", [[k] for _, k in items])
    gs = D.batch_call("group_id_i64", [[k] for _, k in items])
    ms = D.batch_call("reduce_max_i64", [[g] for g in gs])
    groups = D.batch_call("reduce_into_add_f64",
                          [[g, x, S("i64", m.value + 1)] for g, (x, _), m in zip(gs, items, ms)])
    D.sync()
    return Results(sums, perms, groups)


def filter_sum_plain(D: Dispatch, wl: Workload) -> List[Scalar]:
    """The statement pair s = reduce_add_f64(filter_gt_f64(x, 0.9)), one call per array."""
    out = []
    for i in range(len(wl.lengths)):
        x, _ = wl.item(i)
        out.append(D.call("reduce_add_f64", [D.call("filter_gt_f64", [x, Scalar("f64", 0.9)])]))
    D.sync()
    return out


def filter_sum_batched(D: Dispatch, wl: Workload) -> List[Scalar]:
    """The same pair as two batched calls."""
    xs = [wl.item(i)[0] for i in range(len(wl.lengths))]
    fs = D.batch_call("filter_gt_f64", [[x, Scalar("f64", 0.9)] for x in xs])
    out = D.batch_call("reduce_add_f64", [[f] for f in fs])
    D.sync()
    return out


def filter_sum_plan(D: Dispatch, wl: Workload, tune: bool = True):
    """The same pair as a batch plan: statement 0 filters, statement 1 sums its rows."""
    from .kernels import Input, Stmt
    plan = D.batch_plan([[wl.item(i)[0]] for i in range(len(wl.lengths))])
    plan.add("filter_gt_f64", [Input(0), Scalar("f64", 0.9)])
    plan.add("reduce_add_f64", [Stmt(0)])
    plan.compile(tune=tune)
    return plan


def program_plan(D: Dispatch, wl: Workload, tune: bool = True):
    """The whole program as one batch plan, replayed as one hipGraph: filter and sum, argsort,
    group_id, and reduce_into_add_f64 whose bucket count is the group_id statement's group count
    on the device (reduce_max_i64(g) + 1 by definition of first-occurrence ids)."""
    from .kernels import Input, Stmt
    plan = D.batch_plan([[x, k] for x, k in (wl.item(i) for i in range(len(wl.lengths)))])
    plan.add("filter_gt_f64", [Input(0), Scalar("f64", 0.9)])
    plan.add("reduce_add_f64", [Stmt(0)])
    plan.add("argsort_i64", [Input(1)])
    plan.add("group_id_i64", [Input(1)])
    plan.add("reduce_into_add_f64", [Stmt(3), Input(0), Stmt(3)])
    plan.compile(tune=tune)
    return plan


def program_plan_run(D: Dispatch, plan) -> np.ndarray:
    """One replay of the program plan; waits for it by reading every item's sum bits."""
    plan.run()
    return plan.bits(1)


def program_plan_results(plan, every: int = 1) -> Results:
    """The program plan's results after a replay: all sums, every `every`-th item's
    permutation and group sums."""
    bits = plan.bits(1)
    n = len(bits)
    sums = [Scalar.from_bits("f64", int(b)) for b in bits]
    perms = [plan.result(2, i) for i in range(0, n, every)]
    groups = [plan.result(4, i) for i in range(0, n, every)]
    return Results(sums, perms, groups)


def plan_sums(D: Dispatch, plan) -> np.ndarray:
    """One replay of a filter-then-sum plan; its sums as f64 bits (one read-back for all)."""
    plan.run()
    return plan.bits(1)


def same(a: Results, b: Results) -> bool:
    """Bit-identical results."""
    if len(a.sums) != len(b.sums):
        return False
    for s, t in zip(a.sums, b.sums):
        if s.bits() != t.bits():
            return False
    for u, v in zip(a.perms + a.groups, b.perms + b.groups):
        p, q = u.to_numpy(), v.to_numpy()
        if p.shape != q.shape or not np.array_equal(p.view(np.uint64), q.view(np.uint64)):
            return False
    return True


def timed(f, *args, repeat: int = 1):
    """(best seconds, last result) of `repeat` runs."""
    best, out = None, None
    for _ in range(repeat):
        t0 = time.perf_counter()
        out = f(*args)
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    return best, out
