"""Bench of the hip backend on the streaming-primitives workload at 2^28 rows.

Workload (one step): over inputs resident in HBM -- x1..x2 = rand_f64(2^28, seeds 1, 2),
i1 = i64 uniform in [-2^40, 2^40) = (splitmix64(4 + i) >> 23) - 2^40, masks m_c = x2 > c for
c in {0.99, 0.5, 0.01}, a gather table T = rand_f64(2^26, 3) and indices idx =
splitmix64(5 + i) mod 2^26 -- run dot_f64(x1, x2), reduce_add_f64(x1), count_gt_f64(x1, 0.5),
reduce_min_f64(x1), argmax_f64(x1), scan_add_i64(i1), filter_gt_f64(x1, c) and
compress_f64(x1, m_c) for each c, gather_f64(T, idx) and gather_i64(Ti, idx).

value = the step's algorithmic bytes (inputs read + outputs written, per primitive as listed
in ALG_BYTES) / the step's wall time, in GB/s. Per-primitive times come from HIP events on
the library's stream around each entry. One JSON line is printed.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
SELECTIVITIES = (0.99, 0.5, 0.01)  # filter_gt constants: ~1%, ~50%, ~99% kept


def alg_bytes(name: str, n: int, kept: int = 0) -> int:
    """Algorithmic bytes of one call at n rows (reads of every input row + writes of every
    output row); `kept` is the survivor count of a selection."""
    if name == "dot_f64":
        return 16 * n
    if name in ("reduce_add_f64", "count_gt_f64", "reduce_min_f64", "argmax_f64", "reduce_add_i64"):
        return 8 * n
    if name == "scan_add_i64":
        return 16 * n
    if name.startswith("filter_gt_f64"):
        return 8 * n + 8 * kept
    if name.startswith("compress_f64"):
        return 9 * n + 8 * kept
    if name.startswith("gather"):
        return 24 * n
    if name == "copy_f64":
        return 16 * n
    raise KeyError(name)


def splitmix64_np(seed: int, start: int, count: int) -> np.ndarray:
    """splitmix64(seed + i) for i in [start, start + count), vectorised (construct.rs:10-17)."""
    z = (np.arange(start, start + count, dtype=np.uint64) + np.uint64(seed & (2**64 - 1))
         + np.uint64(0x9E3779B97F4A7C15))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _chunked(n: int, f) -> np.ndarray:
    out = np.empty(n, np.int64)
    step = 1 << 24
    for s0 in range(0, n, step):
        out[s0:s0 + step] = f(s0, min(step, n - s0))
    return out


def i64_uniform_40(n: int, seed: int) -> np.ndarray:
    """(splitmix64(seed + i) >> 23) - 2^40: uniform in [-2^40, 2^40)."""
    return _chunked(n, lambda s0, c: (splitmix64_np(seed, s0, c) >> np.uint64(23)).astype(np.int64)
                    - (1 << 40))


def gather_indices(n: int, seed: int, table_rows: int) -> np.ndarray:
    """splitmix64(seed + i) mod table_rows (a power of two)."""
    return _chunked(n, lambda s0, c: (splitmix64_np(seed, s0, c) & np.uint64(table_rows - 1)).astype(np.int64))


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pmc_traffic_streaming_2p28.md")


def measured_traffic(label: str, log2n: int):
    """HBM bytes per launch of `label` at 2^28 rows from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated), or None when there is none for this label or size."""
    if log2n != 28 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    # gather_i64 runs the same kernel (8-byte rows) on the same sizes as gather_f64
    label = {"gather_i64": "gather_f64"}.get(label, label)
    for line in open(TRAFFIC_PROFILE):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic" and parts[1] == label:
            return int(parts[2])
    return None


def parse_trace(text: str):
    out = []
    for line in text.strip().splitlines():
        parts = line.split()
        if not parts or parts[0] != "trace":
            continue
        kv = dict(p.split("=", 1) for p in parts[2:] if "=" in p)
        out.append((parts[1], float(kv["time"])))
    return out


def gather_by_table(D, n: int, reps: int = 3) -> dict:
    """gather_f64 of the same n random indices into tables of 2^19 rows (4 MB: one L2), 2^24
    rows (128 MB: inside the 256 MB Infinity Cache) and 2^26 rows (512 MB: HBM), in us per
    launch. Equal times at 2^24 and 2^26 rows mean the bound is the rate of L2 misses (one
    128-byte line each), not the HBM bandwidth."""
    from verkehrt_amd import Array, Scalar as S
    out = {}
    for lg in (19, 24, 26):
        rows = 1 << lg
        tbl = D.call("rand_f64", [S("i64", rows), S("i64", 3)])
        idx = Array.from_numpy(D.ctx, gather_indices(n, 5, rows))
        D.call("gather_f64", [tbl, idx])  # warm-up
        D.ctx.sync()
        D.ctx.set_trace(True)
        D.ctx.take_trace()
        for _ in range(reps):
            D.call("gather_f64", [tbl, idx])
        D.ctx.sync()
        D.ctx.set_trace(False)
        t = [us for name, us in parse_trace(D.ctx.take_trace()) if name == "gather_f64"]
        out[f"2^{lg}"] = round(float(np.mean(t)), 1)
        del tbl, idx
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> int:
    """The host cores this process may use, at most 16 (a one-GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_streaming_steps(B, n: int, shard: int):
    """The streaming suite on host arrays of n rows (shard number `shard` picks the seeds), as
    (label, call, kept) with every result checked against the oracle first."""
    import ctypes as C
    from oracle import oracle as O
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    s0 = 16 * shard
    x1, x2 = O.rand_f64(n, 1 + s0), O.rand_f64(n, 2 + s0)
    i1 = i64_uniform_40(n, 4 + s0)
    tbl = O.rand_f64(min(1 << 26, n), 3 + s0)
    idx = gather_indices(n, 5 + s0, tbl.size)
    masks = {c: (x2 > c).astype(np.uint8) for c in SELECTIVITIES}
    fo = np.empty(n + 8, np.float64)
    io = np.empty(n, np.int64)
    bad = np.zeros(1, np.int64)
    # correctness gate before timing (benches/tier1.rs:403-433)
    assert B.vkb_dot_f64(p(x1), p(x2), n) == O.dot_f64(x1, x2)
    assert B.vkb_sum_f64(p(x1), n) == O.canonical_sum_f64(x1)
    assert B.vkb_count_gt_f64(p(x1), n, 0.5) == O.count_cmp(x1, "gt", 0.5)
    k = B.vkb_filter_gt_f64(p(x1), n, 0.5, p(fo))
    assert np.array_equal(fo[:k], O.filter_cmp(x1, "gt", 0.5))
    B.vkb_scan_add_i64(p(i1), n, p(io))
    assert np.array_equal(io, O.scan_add_i64(i1))
    kept = {c: int((x1 > c).sum()) for c in SELECTIVITIES}
    kept_m = {c: int(masks[c].sum()) for c in SELECTIVITIES}
    steps = [
        ("dot_f64", lambda: B.vkb_dot_f64(p(x1), p(x2), n), 0),
        ("reduce_add_f64", lambda: B.vkb_sum_f64(p(x1), n), 0),
        ("count_gt_f64", lambda: B.vkb_count_gt_f64(p(x1), n, 0.5), 0),
        ("reduce_min_f64", lambda: B.vkb_min_f64(p(x1), n), 0),
        ("argmax_f64", lambda: B.vkb_argmax_f64(p(x1), n), 0),
        ("scan_add_i64", lambda: B.vkb_scan_add_i64(p(i1), n, p(io)), 0),
    ]
    for c in SELECTIVITIES:
        steps.append((f"filter_gt_f64@{c}", lambda c=c: B.vkb_filter_gt_f64(p(x1), n, c, p(fo)), kept[c]))
    for c in SELECTIVITIES:
        steps.append((f"compress_f64@{c}", lambda c=c: B.vkb_compress_f64(p(x1), p(masks[c]), n, p(fo)), kept_m[c]))
    steps.append(("gather_f64", lambda: B.vkb_gather_f64(p(tbl), tbl.size, p(idx), n, p(fo), p(bad)), 0))
    keep = (x1, x2, i1, tbl, idx, masks, fo, io, bad)
    return steps, keep


def zipf_cdf(values: int = 1 << 20, s: float = 1.1) -> np.ndarray:
    """The CDF of Zipf(s) over 1..values: P(k) proportional to k^-s."""
    w = np.arange(1, values + 1, dtype=np.float64) ** -s
    c = np.cumsum(w)
    return c / c[-1]


def zipf_from_uniform(u: np.ndarray, cdf: np.ndarray) -> np.ndarray:
    """Inverse CDF: the least k with u < cdf[k - 1], as i64 in 1..len(cdf)."""
    return (np.searchsorted(cdf, u, side="right") + 1).astype(np.int64)


def zipf_device(D, n: int, seed: int, cdf: np.ndarray):
    """i64 Zipf keys on the device: for u = rand_f64(n, seed), the least k with u < cdf[k - 1]
    (zipf_from_uniform), by a binary search of the CDF built from the table's own gather,
    comparison, select and integer kernels (setup only, never timed)."""
    from verkehrt_amd import Array, Scalar as S
    dcdf = Array.from_numpy(D.ctx, cdf)
    u = D.call("rand_f64", [S("i64", n), S("i64", seed)])
    lo = D.call("fill_i64", [S("i64", n), S("i64", 0)])
    hi = D.call("fill_i64", [S("i64", n), S("i64", cdf.size - 1)])
    steps = int(cdf.size).bit_length() + 1
    for _ in range(steps):  # invariant: the answer index is in [lo, hi]
        mid = D.call("div_const_i64", [D.call("add_i64", [lo, hi]), S("i64", 2)])
        go = D.call("le_f64", [D.call("gather_f64", [dcdf, mid]), u])  # cdf[mid] <= u: answer > mid
        lo = D.call("select_i64", [go, D.call("add_i64", [mid, S("i64", 1)]), lo])
        hi = D.call("select_i64", [go, hi, mid])
    return D.call("add_i64", [lo, S("i64", 1)])


def cpu_sort_legs(B, n: int):
    """One core: argsort (u64 uniform, i64 Zipf(1.1) over 2^20 values, f64 normal), group_id +
    reduce_into_add_f64 over 1K, 1M and n/4 groups, and scan_by add, on n-row samples; each
    checked against the oracle at 2^16 rows first. Algorithmic GB/s as for the device."""
    import ctypes as C
    from oracle import oracle as O
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    cdf = zipf_cdf()
    rng = np.random.default_rng(3)

    def sort_keys(m):
        f = rng.normal(size=m)
        f[rng.integers(0, m, max(1, m // 100000))] = 0.0
        f[rng.integers(0, m, max(1, m // 100000))] = -0.0
        return [("u64", 0, splitmix64_np(7, 0, m)), ("i64_zipf", 1, zipf_from_uniform(O.rand_f64(m, 9), cdf)),
                ("f64_normal", 2, f)]

    def group_keys(m, g):
        pool = splitmix64_np(11, 0, g).view(np.int64)
        return pool[(splitmix64_np(12, 0, m) % np.uint64(g)).astype(np.int64)]

    small = 1 << 16
    for label, kind, k in sort_keys(small):
        out = np.empty(small, np.int64)
        assert B.vkb_argsort_64(kind, p(np.ascontiguousarray(k)), small, p(out)) == 0
        assert np.array_equal(out, O.argsort(k)), label
    kk = group_keys(small, 1000)
    ids = np.empty(small, np.int64)
    assert B.vkb_group_id_64(p(kk), small, p(ids)) == 0 and np.array_equal(ids, O.group_id(kk))
    res = {}
    out = np.empty(n, np.int64)
    for label, kind, k in sort_keys(n):
        k = np.ascontiguousarray(k)
        t0 = time.perf_counter()
        assert B.vkb_argsort_64(kind, p(k), n, p(out)) == 0
        dt = time.perf_counter() - t0
        b = argsort_alg_bytes(k, kind)
        res[f"argsort_{label}"] = {"rows": n, "ms": round(dt * 1e3, 1), "GBps": round(b / dt / 1e9, 3)}
    vals = O.rand_f64(n, 1)
    for g in (1 << 10, 1 << 20, n // 4):
        kk = group_keys(n, g)
        t0 = time.perf_counter()
        assert B.vkb_group_id_64(p(kk), n, p(out)) == 0
        t1 = time.perf_counter()
        ng = int(out.max()) + 1
        sums = np.empty(ng, np.float64)
        assert B.vkb_reduce_into_add_f64(p(out), p(vals), n, ng, p(sums)) == 0
        t2 = time.perf_counter()
        res[f"group_sum_{g}"] = {"rows": n, "group_id_ms": round((t1 - t0) * 1e3, 1),
                                 "reduce_into_ms": round((t2 - t1) * 1e3, 1),
                                 "GBps_min_io": round((16 * n + 8 * ng) / (t2 - t0) / 1e9, 3)}
    st = ((splitmix64_np(13, 0, n) % np.uint64(16)) == 0).astype(np.uint8)
    iv = i64_uniform_40(n, 4)
    t0 = time.perf_counter()
    B.vkb_scan_by_add_i64(p(iv), p(st), n, p(out))
    dt = time.perf_counter() - t0
    res["scan_by_add_i64"] = {"rows": n, "ms": round(dt * 1e3, 1), "GBps": round(17 * n / dt / 1e9, 3)}
    return res


def cpu_baseline(sample_rows: int, budget_s: float = 8.0):
    """The reference's best x86 CPU path, restated in C (oracle/cpu_best.c: x86-64-v4 auto
    loops, AVX-512 compress stores, vpgatherqq, and the scalar radix argsort, hash group_id and
    scan_by), on the host cores of this run over `sample_rows`-row samples of the same
    workloads. One core (the reference's design), then one independent instance per core on
    disjoint shards (the streaming suite). Every result is checked against the oracle (the
    scalar restatement) before timing."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    B = O.best()
    n = sample_rows
    steps, keep = cpu_streaming_steps(B, n, 0)
    total_bytes = sum(alg_bytes(nm.split("@")[0], n, k) for nm, _, k in steps)
    t_end = time.perf_counter() + budget_s
    passes, spent = 0, 0.0
    while time.perf_counter() < t_end or passes == 0:
        t0 = time.perf_counter()
        for _, f, _ in steps:
            f()
        spent += time.perf_counter() - t0
        passes += 1
    gbps = total_bytes * passes / spent / 1e9
    del keep, steps
    # all cores: ctypes releases the GIL during each foreign call, so threads run the C loops
    # in parallel, one shard each
    cores = host_cores()
    shards = [cpu_streaming_steps(B, n, s + 1) for s in range(cores)]
    all_passes = max(1, passes // 2)

    def run(sh):
        for _ in range(all_passes):
            for _, f, _ in sh[0]:
                f()

    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        list(ex.map(run, shards))
        wall = time.perf_counter() - t0
    all_bytes = sum(sum(alg_bytes(nm.split("@")[0], n, k) for nm, _, k in sh[0]) for sh in shards) * all_passes
    del shards
    sort = cpu_sort_legs(B, n)
    cpu = cpu_model()
    return {"value": round(gbps, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"the streaming suite (one gather table) at {n} rows, {passes} passes in {spent:.1f} s, "
                      f"C restatement of the reference's best x86-64-v4 path on one core of {cpu}",
            "all_cores": {"value": round(all_bytes / wall / 1e9, 3), "unit": "GB/s", "cores": cores,
                          "sample": f"{cores} independent instances of the same suite on disjoint {n}-row "
                                    f"shards, {all_passes} passes each in {wall:.1f} s"},
            "sort": sort}


def argsort_alg_bytes(keys: np.ndarray, kind: int) -> int:
    """2nw + 2(P-1)n(w+4) + 8n for 8-byte keys, P the non-constant byte digits of the images."""
    n = keys.size
    b = keys.view(np.uint64)
    img = b ^ np.uint64(1 << 63) if kind == 1 else (np.where(b >> np.uint64(63), ~b, b ^ np.uint64(1 << 63)) if kind == 2 else b)
    P = 0
    for d in range(8):
        dig = (img >> np.uint64(8 * d)) & np.uint64(255)
        P += int(not np.all(dig == dig[0]))
    return 16 * n + 2 * max(P - 1, 0) * n * 12 + 8 * n


def batched_suite(D, count: int):
    """The many-arrays form: `count` arrays with log-uniform lengths in [2^10, 2^20] rows, each
    running filter_gt(0.9)+sum, argsort and a group-by sum over 256 keys; one call per array
    and statement (plain) against one batched call per statement. Results are compared bit for
    bit (all sums, every 16th array's permutation and group sums)."""
    from verkehrt_amd import batch as B
    wl = B.make_workload(D, count=count)
    B.batched(D, wl)  # warm-up
    tb, rb = B.timed(B.batched, D, wl, repeat=3)
    tp, rp = B.timed(B.plain, D, wl, repeat=1)
    sample = B.Results(rb.sums, rb.perms[::16], rb.groups[::16])
    ref = B.Results(rp.sums, rp.perms[::16], rp.groups[::16])
    same = B.same(sample, ref)
    if not same:
        sys.exit("batched results differ from plain dispatch")
    # the whole program as one batch plan: one hipGraph replay per run, no host read inside
    del rb
    prog = B.program_plan(D, wl, tune=True)
    B.program_plan_run(D, prog)  # warm-up replay
    tg, _ = B.timed(B.program_plan_run, D, prog, repeat=3)
    rg = B.program_plan_results(prog, every=16)
    prog_tuning = prog.tuning().strip()
    prog.close()
    if not B.same(rg, ref):
        sys.exit("the program plan's results differ from plain dispatch")
    del rg
    # the first statement pair alone, three ways: one call per array, two batched calls, and a
    # batch plan (tiles planned on the device, tile shapes autotuned, one hipGraph replay)
    tfp, rfp = B.timed(B.filter_sum_plain, D, wl, repeat=1)
    tfb, rfb = B.timed(B.filter_sum_batched, D, wl, repeat=3)
    plan = B.filter_sum_plan(D, wl, tune=True)
    B.plan_sums(D, plan)  # warm-up replay
    tfg, rfg = B.timed(B.plan_sums, D, plan, repeat=5)
    tuning = plan.tuning().strip()
    plan.close()
    if not all(a.bits() == b.bits() == int(c) for a, b, c in zip(rfp, rfb, rfg)):
        sys.exit("filter+sum results differ between plain, batched and the batch plan")
    filter_sum = {"statements": "s = reduce_add_f64(filter_gt_f64(x, 0.9)) per array",
                  "plain_ms": round(tfp * 1e3, 1), "batched_ms": round(tfb * 1e3, 1),
                  "plan_ms": round(tfg * 1e3, 2), "plan_speedup_vs_plain": round(tfp / tfg, 1),
                  "autotuned": tuning, "identical": True}
    return {"arrays": count, "rows": wl.rows, "lengths": "log-uniform in [2^10, 2^20] rows",
            "program": "filter_gt_f64(x, 0.9) then reduce_add_f64; argsort_i64(k); group_id_i64(k), "
                       "reduce_max_i64 + 1, reduce_into_add_f64(g, x, ng); keys from 256 values",
            "plain_ms": round(tp * 1e3, 1), "batched_ms": round(tb * 1e3, 1),
            "speedup": round(tp / tb, 2), "identical": same,
            "plan": {"what": "the whole program as one batch plan (argsort, group_id and reduce_into_add_f64 "
                             "included), one hipGraph replay per run",
                     "ms": round(tg * 1e3, 1), "speedup_vs_plain": round(tp / tg, 2),
                     "vs_batched_calls": round(tb / tg, 2), "autotuned": prog_tuning, "identical": True},
            "filter_sum": filter_sum}


def program_plan_suite(D, count: int):
    """The many-arrays program as one batch plan alone (for profiling its graph replay): compile
    with autotuning, one warm-up replay, then three timed replays; results of every 16th item
    checked against the batched calls."""
    from verkehrt_amd import batch as B
    wl = B.make_workload(D, count=count)
    plan = B.program_plan(D, wl, tune=True)
    B.program_plan_run(D, plan)
    t, _ = B.timed(B.program_plan_run, D, plan, repeat=3)
    rb = B.batched(D, wl)
    same = B.same(B.program_plan_results(plan, every=16), B.Results(rb.sums, rb.perms[::16], rb.groups[::16]))
    plan.close()
    return {"arrays": count, "rows": wl.rows, "plan_ms": round(t * 1e3, 1), "identical": same}


def key_image(b: np.ndarray, kind: int) -> np.ndarray:
    """The u64 sort images of 8-byte keys as bits: 0 u64, 1 i64 (sign flipped), 2 f64 (negative
    values flip every bit, others the sign bit)."""
    b = b.view(np.uint64)
    if kind == 1:
        return b ^ np.uint64(1 << 63)
    if kind == 2:
        return np.where(b >> np.uint64(63), ~b, b ^ np.uint64(1 << 63))
    return b


def check_argsort(keys: np.ndarray, kind: int, perm: np.ndarray) -> bool:
    """Exact check of a stable ascending argsort: the positions are a permutation, the keys'
    images are non-decreasing along it, and equal images keep their positions' order. The
    stable ascending permutation is unique, so this equals a comparison with the oracle."""
    n = keys.size
    if perm.size != n or (n and (perm.min() < 0 or perm.max() >= n)):
        return False
    seen = np.zeros(n, np.bool_)
    seen[perm] = True
    if not seen.all():
        return False
    img = key_image(keys, kind)[perm]
    if not np.all(img[1:] >= img[:-1]):
        return False
    eq = img[1:] == img[:-1]
    return bool(np.all(perm[1:][eq] > perm[:-1][eq]))


def check_group_ids(keys: np.ndarray, ids: np.ndarray) -> bool:
    """Exact check of first-occurrence ids: rows with one id hold one key, distinct ids hold
    distinct keys, and the running maximum of the ids grows by at most 1 from 0 (every new key
    takes the next id). These fix the ids uniquely, as group.rs:1-5 defines them."""
    n = keys.size
    if ids.size != n:
        return False
    if n == 0:
        return True
    if ids[0] != 0 or ids.min() < 0:
        return False
    run = np.maximum.accumulate(ids)
    if np.any(np.diff(run) > 1):
        return False
    g = int(run[-1]) + 1
    key_of = np.empty(g, keys.dtype)
    key_of[ids] = keys
    if not np.array_equal(key_of[ids], keys):
        return False
    return np.unique(key_of).size == g


def batched_literal_suite(D, count: int = 4096, lo: int = 1 << 10, hi: int = 1 << 24):
    """The many-arrays form at the configuration's literal lengths, log-uniform in [1K, 16M]
    rows (about 7 G rows, 113 GB of values and keys): run in chunks of consecutive arrays of at
    most 2^29 rows, each chunk once as one call per array and statement and once as one batched
    call per statement. Every sum and every 64th array's permutation and group sums are compared
    bit for bit."""
    from verkehrt_amd import batch as B
    tp = tb = 0.0
    rows = 0
    same = True
    for first, take in B.chunks(count, lo, hi):
        wl = B.make_workload(D, count=count, lo=lo, hi=hi, first=first, take=take)
        rows += wl.rows
        t, rp = B.timed(B.plain, D, wl, repeat=1)
        tp += t
        t, rb = B.timed(B.batched, D, wl, repeat=1)
        tb += t
        step = 64
        pick = [i for i in range(take) if (first + i) % step == 0]
        same = same and B.same(B.Results(rp.sums, [rp.perms[i] for i in pick], [rp.groups[i] for i in pick]),
                               B.Results(rb.sums, [rb.perms[i] for i in pick], [rb.groups[i] for i in pick]))
        del wl, rp, rb
        D.ctx.trim()
    if not same:
        sys.exit("batched results differ from plain dispatch at the literal lengths")
    return {"arrays": count, "rows": rows, "lengths": "log-uniform in [2^10, 2^24] rows (chunks of <= 2^29 rows)",
            "plain_ms": round(tp * 1e3, 1), "batched_ms": round(tb * 1e3, 1), "speedup": round(tp / tb, 2),
            "identical": same}


def sort_suite(D, verify: bool = True):
    """Sort and grouping (the configuration's third workload): argsort of 2^27 keys (u64 uniform
    = splitmix64(7 + i), i64 Zipf(1.1) over 2^20 values by inverse CDF of rand_f64(2^27, 9),
    f64 normal(0,1) with +0 and -0 injected), group-by sum of rand_f64(2^28, 1) values over i64
    keys pool[splitmix64(12 + i) mod G], pool = splitmix64(11 + j), for G = 1K, 1M and 64M
    (group_id, reduce_max + 1, reduce_into_add_f64), and scan_by add over 2^28 i64 rows with
    starts where splitmix64(13 + i) mod 16 == 0. Inputs are generated on the device where the
    table's kernels allow it. Times are medians of HIP-event times per call; GB/s counts the
    algorithmic bytes (argsort 2nw + 2(P-1)n(w+4) + 8n, group-by 16n + 8G, scan_by 17n). Every
    result is checked exactly: argsort and group_id ids by the properties that fix them,
    reduce_into and scan_by against the oracle."""
    from oracle import oracle as O
    from verkehrt_amd import Array, Scalar as S
    ctx = D.ctx

    def timed(name, a, reps=3):
        D.call(name, a)
        ctx.sync()
        ctx.set_trace(True)
        ctx.take_trace()
        for _ in range(reps):
            out = D.call(name, a)
            del out
        ctx.sync()
        tr = parse_trace(ctx.take_trace())
        ctx.set_trace(False)
        return float(np.median([us for _, us in tr]))

    checks = {}
    out = {}
    n = 1 << 27
    cdf = zipf_cdf()
    zipf = zipf_device(D, n, 9, cdf)
    rng = np.random.default_rng(3)
    f = rng.normal(size=n)
    f[rng.integers(0, n, 1000)] = 0.0
    f[rng.integers(0, n, 1000)] = -0.0
    cases = [("u64", "argsort_u64", 0, D.call("cast_i64_u64", [D.call("rand_i64", [S("i64", n), S("i64", 7)])])),
             ("i64_zipf", "argsort_i64", 1, zipf), ("f64_normal", "argsort_f64", 2, Array.from_numpy(ctx, f))]
    del f, zipf
    for label, name, kind, a in cases:
        us = timed(name, [a])
        host = a.to_numpy()
        b = argsort_alg_bytes(host, kind)
        out[f"argsort_{label}"] = {"rows": n, "us": round(us, 1), "GBps": round(b / us / 1e3, 1),
                                   "frac": round(b / us / 1e3 / HBM_PEAK_GBPS, 4), "bytes": b}
        if verify:
            checks[f"argsort_{label}"] = check_argsort(host, kind, D.call(name, [a]).to_numpy())
        if label == "i64_zipf" and verify:
            ref = zipf_from_uniform(O.rand_f64(1 << 16, 9), cdf)  # the device's inverse CDF
            checks["zipf_inputs"] = bool(np.array_equal(host[:1 << 16], ref))
        del host
    del cases
    n = 1 << 28
    vals = D.call("rand_f64", [S("i64", n), S("i64", 1)])
    vals_h = O.rand_f64(n, 1) if verify else None
    for g in (1 << 10, 1 << 20, 1 << 26):
        pool = D.call("rand_i64", [S("i64", g), S("i64", 11)])
        sel = D.call("and_i64", [D.call("rand_i64", [S("i64", n), S("i64", 12)]), S("i64", g - 1)])
        k = D.call("gather_i64", [pool, sel])
        del pool, sel
        t_id = timed("group_id_i64", [k], reps=2)
        ids = D.call("group_id_i64", [k])
        ng = D.call("reduce_max_i64", [ids]).value + 1
        t_ri = timed("reduce_into_add_f64", [ids, vals, S("i64", ng)], reps=2)
        b = 16 * n + 8 * ng
        out[f"group_sum_{g}"] = {"rows": n, "groups": ng, "group_id_us": round(t_id, 1), "reduce_into_us": round(t_ri, 1),
                                 "GBps_min_io": round(b / (t_id + t_ri) / 1e3, 1),
                                 "frac": round(b / (t_id + t_ri) / 1e3 / HBM_PEAK_GBPS, 4)}
        if verify:
            kh, ih = k.to_numpy(), ids.to_numpy()
            checks[f"group_id_{g}"] = check_group_ids(kh, ih)
            del kh
            got = D.call("reduce_into_add_f64", [ids, vals, S("i64", ng)]).to_numpy()
            checks[f"reduce_into_add_f64_{g}"] = bool(np.array_equal(got.view(np.uint64),
                                                                     O.reduce_into(ih, vals_h, ng).view(np.uint64)))
            del ih, got
        del k, ids
    del vals, vals_h
    st = D.call("eq_i64", [D.call("and_i64", [D.call("rand_i64", [S("i64", n), S("i64", 13)]), S("i64", 15)]),
                           S("i64", 0)])
    iv_h = i64_uniform_40(n, 4)
    iv = Array.from_numpy(ctx, iv_h)
    us = timed("scan_by_add_i64", [iv, st])
    out["scan_by_add_i64"] = {"rows": n, "us": round(us, 1), "GBps": round(17 * n / us / 1e3, 1),
                              "frac": round(17 * n / us / 1e3 / HBM_PEAK_GBPS, 4)}
    if verify:
        got = D.call("scan_by_add_i64", [iv, st]).to_numpy()
        checks["scan_by_add_i64"] = bool(np.array_equal(got, O.scan_by_add_i64(iv_h, st.to_numpy())))
    out["verified"] = bool(checks) and all(checks.values())
    out["checks"] = checks
    return out


def region_suite(D, x1, x2, n):
    """Fused regions at n rows: sum((x * 2.0) + 1.0), sum((x - y) * (x - y)) and
    count((x * 2.0) > 1.0), each as one device pass (vk_region_reduce) against the same plan run
    kernel by kernel (elementwise results materialised in HBM). Results must be identical; times
    are HIP-event medians of three calls; GB/s counts the fused pass's input bytes."""
    from verkehrt_amd import Scalar as S
    ctx = D.ctx
    inputs = [x1, x2, S("f64", 2.0), S("f64", 1.0)]
    shapes = {
        "sum((x*2.0)+1.0)": ([("mul_f64", [("in", 0), ("in", 2)]), ("add_f64", [("k", 0), ("in", 3)])],
                             "reduce_add_f64", [("k", 1)], 8),
        "sum((x-y)*(x-y))": ([("sub_f64", [("in", 0), ("in", 1)]), ("mul_f64", [("k", 0), ("k", 0)])],
                             "reduce_add_f64", [("k", 1)], 16),
        "count((x*2.0)>1.0)": ([("mul_f64", [("in", 0), ("in", 2)]), ("gt_f64", [("k", 0), ("in", 3)])],
                               "count", [("k", 1)], 8),
    }
    out = {}
    for label, (kernels, root, rargs, bpr) in shapes.items():
        def fused():
            return D.region_reduce(kernels, root, rargs, inputs)

        def unfused():
            vals = []
            get = lambda s_: vals[s_[1]] if s_[0] == "k" else inputs[s_[1]]
            for name, srcs in kernels:
                vals.append(D.call(name, [get(s_) for s_ in srcs]))
            return D.call(root, [get(s_) for s_ in rargs])

        times = {}
        for mode, f in (("fused", fused), ("unfused", unfused)):
            f()
            ctx.sync()
            ctx.set_trace(True)
            ctx.take_trace()
            per = []
            for _ in range(3):
                res = f()
                ctx.sync()
                per.append(sum(us for _, us in parse_trace(ctx.take_trace())))
            ctx.set_trace(False)
            times[mode] = (float(np.median(per)), res)
        (rf, flag), ru = times["fused"][1], times["unfused"][1]
        out[label] = {"fused": flag, "identical": rf.bits() == ru.bits(), "fused_us": round(times["fused"][0], 1),
                      "unfused_us": round(times["unfused"][0], 1),
                      "speedup": round(times["unfused"][0] / times["fused"][0], 2),
                      "GBps": round(bpr * n / times["fused"][0] / 1e3, 1)}
    return out


def _timed_calls(D, name, args, reps=3):
    """One warm-up call, then `reps` traced calls: (results, mean microseconds per call)."""
    D.call(name, args)
    D.ctx.sync()
    D.ctx.set_trace(True)
    D.ctx.take_trace()
    res = [D.call(name, args) for _ in range(reps)]
    D.ctx.sync()
    D.ctx.set_trace(False)
    return res, float(np.mean([t for _, t in parse_trace(D.ctx.take_trace())]))


def zero_mean_suite(D, x1, n: int, verify: bool = True):
    """reduce_add_f64 and dot_f64 over normal(0,1) rows (numpy's generator, seeds 5 and 6, made on
    the host and copied in): every chain's running sum keeps returning near zero, so thousands
    of tiles per chain cross binades, the canonical form's hardest input. Also scan_add_f64
    (out[i] = the canonical sum of x[0..=i]) over the same rows and over the uniform rows x1.
    Timed per call over 3 calls after one warm-up, then checked bit for bit against the oracle."""
    from verkehrt_amd import Array
    rng5, rng6 = np.random.default_rng(5), np.random.default_rng(6)
    h, g = rng5.normal(size=n), rng6.normal(size=n)
    x, y = Array.from_numpy(D.ctx, h), Array.from_numpy(D.ctx, g)
    out = {"data": "normal(0,1) rows, numpy default_rng seeds 5 and 6", "rows": n}

    def entry(us, nbytes):
        return {"us": round(us, 1), "GBps": round(nbytes / us / 1e3, 1),
                "frac": round(nbytes / us / 1e3 / HBM_PEAK_GBPS, 4)}

    for name, args, nbytes in (("reduce_add_f64", [x], 8 * n), ("dot_f64", [x, y], 16 * n)):
        res, us = _timed_calls(D, name, args)
        out[name] = entry(us, nbytes)
        if verify:
            from oracle import oracle as O
            want = O.canonical_sum_f64(h) if name == "reduce_add_f64" else O.dot_f64(h, g)
            out[name]["verified"] = all(r.value == want for r in res)
    del y
    scans = {}
    for label, arr in (("normal", x), ("uniform", x1)):
        res, us = _timed_calls(D, "scan_add_f64", [arr], reps=2)
        scans[label] = entry(us, 16 * n)
        if verify:
            from oracle import oracle as O
            host = h if label == "normal" else O.rand_f64(n, 1)
            want = O.scan_add_f64(host).view(np.uint64)
            scans[label]["verified"] = all(np.array_equal(r.to_numpy().view(np.uint64), want) for r in res)
            del host, want
        del res
    out["scan_add_f64"] = scans
    del x
    return out


def f32_int_suite(D, n: int, verify: bool = True):
    """The f32 canonical sums (reduce_add_f32 over uniform and normal(0,1) rows, dot_f32) and an
    integer dot (dot_i64 over uniform i64 rows in [-2^40, 2^40)) at n rows, made on the host
    (numpy default_rng seed 7) and copied in; timed per call over 3 calls after one warm-up,
    then checked bit for bit against the oracle."""
    from verkehrt_amd import Array
    rng = np.random.default_rng(7)
    u = rng.random(n, dtype=np.float32)
    z = rng.standard_normal(n, dtype=np.float32)
    xu, xz = Array.from_numpy(D.ctx, u), Array.from_numpy(D.ctx, z)
    out = {"data": "f32 uniform [0, 1) and normal(0,1) rows, i64 uniform [-2^40, 2^40), numpy default_rng seed 7",
           "rows": n}

    def entry(us, nbytes):
        return {"us": round(us, 1), "GBps": round(nbytes / us / 1e3, 1),
                "frac": round(nbytes / us / 1e3 / HBM_PEAK_GBPS, 4)}

    for label, name, args, nbytes, want in (
            ("reduce_add_f32", "reduce_add_f32", [xu], 4 * n, lambda O: O.canonical_fold_f32("sum", u)),
            ("reduce_add_f32_normal", "reduce_add_f32", [xz], 4 * n, lambda O: O.canonical_fold_f32("sum", z)),
            ("dot_f32", "dot_f32", [xu, xz], 8 * n, lambda O: O.canonical_fold_f32("dot", u, z))):
        res, us = _timed_calls(D, name, args)
        out[label] = entry(us, nbytes)
        if verify:
            from oracle import oracle as O
            w = want(O)
            out[label]["verified"] = all(np.float32(r.value).tobytes() == np.float32(w).tobytes() for r in res)
    del xu, xz, u, z
    a = rng.integers(-(1 << 40), 1 << 40, n)
    b = rng.integers(-(1 << 40), 1 << 40, n)
    xa, xb = Array.from_numpy(D.ctx, a), Array.from_numpy(D.ctx, b)
    res, us = _timed_calls(D, "dot_i64", [xa, xb])
    out["dot_i64"] = entry(us, 16 * n)
    if verify:
        from oracle import oracle as O
        w = O.fused_sum_int("dot", a, b)
        out["dot_i64"]["verified"] = all(r.value == w for r in res)
    del xa, xb
    return out


def verify_streaming(D, calls, n, table_rows):
    """Each streaming primitive's 2^28-row result, computed once more after the timed region,
    against the oracle on the same inputs (as benches/tier1.rs:403-433 asserts before timing)."""
    from oracle import oracle as O
    x1, x2 = O.rand_f64(n, 1), O.rand_f64(n, 2)
    inputs = {"x1": x1, "x2": x2}
    want = {}
    res = {}
    for label, name, a in calls:
        got = D.call(name, a)
        if label == "dot_f64":
            ok = got.value == O.dot_f64(x1, x2)
        elif label == "reduce_add_f64":
            ok = got.value == O.canonical_sum_f64(x1)
        elif label == "count_gt_f64":
            ok = got.value == O.count_cmp(x1, "gt", 0.5)
        elif label == "reduce_min_f64":
            ok = got.value == float(O.reduce_extreme(x1, False))
        elif label == "argmax_f64":
            ok = got.value == O.arg_extreme(x1, True)
        elif label == "scan_add_i64":
            ok = np.array_equal(got.to_numpy(), O.scan_add_i64(i64_uniform_40(n, 4)))
        elif label.startswith("filter_gt_f64@"):
            c = float(label.split("@")[1])
            ok = np.array_equal(got.to_numpy().view(np.uint64), O.filter_cmp(x1, "gt", c).view(np.uint64))
        elif label.startswith("compress_f64@"):
            c = float(label.split("@")[1])
            ok = np.array_equal(got.to_numpy().view(np.uint64), O.compress(x1, (x2 > c).astype(np.uint8)).view(np.uint64))
        elif label == "gather_f64":
            if "idx" not in want:
                want["idx"] = gather_indices(n, 5, table_rows)
            ok = np.array_equal(got.to_numpy().view(np.uint64),
                                O.gather(O.rand_f64(table_rows, 3), want["idx"]).view(np.uint64))
        elif label == "gather_i64":
            if "idx" not in want:
                want["idx"] = gather_indices(n, 5, table_rows)
            ok = np.array_equal(got.to_numpy(), O.gather(i64_uniform_40(table_rows, 6), want["idx"]))
        else:
            continue
        res[label] = bool(ok)
        del got
    del inputs
    return res


def serve_suite(requests: int, qps: float):
    """The serving loop: a seeded Poisson stream of `requests` programs at `qps`, sizes
    log-uniform in [10K, 64M] rows, 40% reusing resident columns, 4 workers, a 200 GB HBM
    budget; latency percentiles at that rate, and the throughput of the same requests all
    arriving at once."""
    from verkehrt_amd import serve as SV
    reqs = SV.make_requests(requests, qps=qps, seed=11)
    # the warm-ups cover the whole size range, so the paced runs start with the workers' pools
    # holding blocks of every size class (steady state)
    srv = SV.Server(workers=4, budget=200 * 10**9, key_rows=64 << 20)
    srv.run(SV.make_requests(50, qps=1e9, seed=99), paced=False)  # warm-up
    srv.close()
    paced = srv.run(reqs).summary()
    srv.close()
    sat = srv.run(reqs, paced=False).summary()
    srv.close()
    # continuous batching: the same stream with arrived same-program requests coalesced into
    # batched calls (up to 32 per group)
    srv_c = SV.Server(workers=4, budget=200 * 10**9, key_rows=64 << 20, coalesce=32)
    srv_c.run(SV.make_requests(50, qps=1e9, seed=99), paced=False)  # warm-up
    srv_c.close()
    paced_c = srv_c.run(reqs).summary()
    srv_c.close()
    sat_c = srv_c.run(reqs, paced=False).summary()
    srv_c.close()
    return {"requests": requests, "qps": qps, "sizes": "log-uniform in [10K, 64M] rows",
            "programs": "sum_filter, dot, count, scan, argsort, group_sum (uniform)",
            "reuse": 0.4, "workers": 4, "budget_GB": 200,
            "p50_ms": paced["p50_ms"], "p99_ms": paced["p99_ms"], "GBps_at_qps": paced["GBps"],
            "max_concurrent": paced["max_concurrent"], "columns_created": paced["columns_created"],
            "saturated_GBps": sat["GBps"], "saturated_wall_s": sat["wall_s"],
            "coalesced": {"max_group": 32, "p50_ms": paced_c["p50_ms"], "p99_ms": paced_c["p99_ms"],
                          "mean_group_at_qps": paced_c["mean_group"], "saturated_GBps": sat_c["GBps"],
                          "saturated_wall_s": sat_c["wall_s"], "saturated_mean_group": sat_c["mean_group"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--cpu-rows", type=int, default=1 << 24)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the many-arrays batched comparison")
    ap.add_argument("--batch-arrays", type=int, default=4096)
    ap.add_argument("--no-literal", action="store_true",
                    help="skip the many-arrays form at the literal [1K, 16M] lengths (about 7 G rows)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-sort", action="store_true",
                    help="skip the sort and grouping cases (argsort of 2^27 keys, group-by sum over "
                         "1K/1M/64M groups at 2^28 rows, scan_by add)")
    ap.add_argument("--no-verify", action="store_true",
                    help="skip checking the results against the oracle after the timed region")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-qps", type=float, default=1000.0)
    ap.add_argument("--only", default="",
                    help="comma-separated primitive labels to run (for profiling); 'copy_f64' adds a "
                         "plain copy (add_f64 x + 0.0) used to calibrate HBM counters; 'zero_mean' "
                         "alone runs only the zero-mean canonical and f64 scan leg, 'sort' alone only the "
                         "sort and grouping suite, 'batched' alone only the many-arrays comparison, "
                         "'program_plan' alone only the many-arrays program as one batch plan, 'f32_int' "
                         "alone only the f32 canonical and integer dot leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this backend has no multi-GPU path (asked for {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    from verkehrt_amd import Array, Registry, Scalar
    D = Registry().resolve(0)
    ctx = D.ctx
    n = 1 << args.log2n
    table_rows = min(1 << 26, n)
    S = Scalar

    if args.only == "sort":
        print(json.dumps({"sort": sort_suite(D, verify=not args.no_verify)}))
        return
    if args.only == "batched":
        print(json.dumps({"batched": batched_suite(D, args.batch_arrays)}))
        return
    if args.only == "program_plan":
        print(json.dumps({"program_plan": program_plan_suite(D, args.batch_arrays)}))
        return
    x1 = D.call("rand_f64", [S("i64", n), S("i64", 1)])
    if args.only == "zero_mean":
        print(json.dumps({"zero_mean": zero_mean_suite(D, x1, n, verify=not args.no_verify)}))
        return
    if args.only == "f32_int":
        print(json.dumps({"f32_int": f32_int_suite(D, n, verify=not args.no_verify)}))
        return
    x2 = D.call("rand_f64", [S("i64", n), S("i64", 2)])
    i1 = Array.from_numpy(ctx, i64_uniform_40(n, 4))
    tbl = D.call("rand_f64", [S("i64", table_rows), S("i64", 3)])
    tbli = Array.from_numpy(ctx, i64_uniform_40(table_rows, 6))
    idx = Array.from_numpy(ctx, gather_indices(n, 5, table_rows))
    masks = {c: D.call("gt_f64", [x2, S("f64", c)]) for c in SELECTIVITIES}
    ctx.sync()

    calls = [
        ("dot_f64", "dot_f64", [x1, x2]),
        ("reduce_add_f64", "reduce_add_f64", [x1]),
        ("count_gt_f64", "count_gt_f64", [x1, S("f64", 0.5)]),
        ("reduce_min_f64", "reduce_min_f64", [x1]),
        ("argmax_f64", "argmax_f64", [x1]),
        ("scan_add_i64", "scan_add_i64", [i1]),
    ]
    for c in SELECTIVITIES:
        calls.append((f"filter_gt_f64@{c}", "filter_gt_f64", [x1, S("f64", c)]))
    for c in SELECTIVITIES:
        calls.append((f"compress_f64@{c}", "compress_f64", [x1, masks[c]]))
    calls.append(("gather_f64", "gather_f64", [tbl, idx]))
    calls.append(("gather_i64", "gather_i64", [tbli, idx]))
    if args.only:
        want = set(args.only.split(","))
        extra = [("copy_f64", "add_f64", [x1, S("f64", 0.0)])] if "copy_f64" in want else []
        calls = [c for c in calls if c[0] in want] + extra
        unknown = want - {c[0] for c in calls}
        if unknown:
            sys.exit(f"unknown primitive labels: {sorted(unknown)}")

    kept = {}

    def step():
        for label, name, a in calls:
            out = D.call(name, a)
            if label not in kept:
                kept[label] = len(out) if hasattr(out, "to_numpy") and name.startswith(("filter", "compress")) else 0
            del out

    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.set_trace(True)
    ctx.take_trace()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    wall = time.perf_counter() - t0
    ctx.set_trace(False)
    trace = parse_trace(ctx.take_trace())

    per = {}
    for (label, name, _), (tname, us) in zip(calls * args.steps, trace):
        assert name == tname, (name, tname)
        per.setdefault(label, []).append(us)
    step_bytes = sum(alg_bytes(label.split("@")[0], n, kept[label]) for label, _, _ in calls)
    ms_per_step = wall / args.steps * 1e3
    value = step_bytes / (wall / args.steps) / 1e9
    prims = {}
    for label, _, _ in calls:
        us = float(np.mean(per[label]))
        b = alg_bytes(label.split("@")[0], n, kept[label])
        prims[label] = {"us": round(us, 1), "GBps": round(b / us / 1e3, 1),
                        "frac": round(b / us / 1e3 / HBM_PEAK_GBPS, 4), "bytes": b}
    dom = max(prims, key=lambda k: prims[k]["us"])
    traffic = measured_traffic(dom, args.log2n)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": prims[dom]["GBps"], "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": prims[dom]["frac"], "traffic": traffic,
                "bytes_per_launch": prims[dom]["bytes"], "launch_us": prims[dom]["us"]}
    if traffic:
        # the HBM bytes the kernel actually moves per second (rocprofv3 FETCH_SIZE + WRITE_SIZE)
        roofline["traffic_GBps"] = round(traffic / prims[dom]["us"] / 1e3, 1)
        roofline["traffic_frac"] = round(traffic / prims[dom]["us"] / 1e3 / HBM_PEAK_GBPS, 4)
    if dom.startswith("gather"):
        # random 8-byte rows: each costs a 128-byte line, so the bound is the random-row rate
        roofline["random_rows_G_per_s"] = round(n / prims[dom]["us"] / 1e3, 1)
        if not args.only:
            roofline["us_by_table_rows"] = gather_by_table(D, n)
    regions = None if args.only else region_suite(D, x1, x2, n)
    zero_mean = None if args.only else zero_mean_suite(D, x1, n, verify=not args.no_verify)
    f32_int = None if args.only else f32_int_suite(D, n, verify=not args.no_verify)
    checks = None if (args.no_verify or args.only) else verify_streaming(D, calls, n, table_rows)
    sort = None if (args.no_sort or args.only) else sort_suite(D, verify=not args.no_verify)
    batched = None if (args.no_batched or args.only) else batched_suite(D, args.batch_arrays)
    if batched is not None and not args.no_literal:
        batched["literal_range"] = batched_literal_suite(D, args.batch_arrays)
    if not (args.no_serve or args.only):
        # the serving loop's workers allocate from pools of their own: give this context's
        # cached blocks (the sort and batched legs' buffers) back to the device first
        import gc
        gc.collect()
        ctx.sync()
        ctx.trim()
    serve = None if (args.no_serve or args.only) else serve_suite(args.serve_requests, args.serve_qps)
    cpu = None if (args.no_cpu or args.only) else cpu_baseline(args.cpu_rows, args.cpu_seconds)
    verified = (checks is not None and all(checks.values()) and (sort is None or sort["verified"])
                and (zero_mean is None or (all(zero_mean[k]["verified"] for k in ("reduce_add_f64", "dot_f64"))
                                           and all(v["verified"] for v in zero_mean["scan_add_f64"].values())))
                and (f32_int is None or all(v["verified"] for k, v in f32_int.items() if isinstance(v, dict))))
    line = {
        "metric": "HBM GB/s per primitive (% of peak); batched-vs-plain speedup; serve p99 ms",
        "value": round(value, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64/i64", "data": "synthetic",
        "config": {"workload": f"streaming primitives at 2^{args.log2n} rows: dot, sum, count_gt, min, "
                               f"argmax, cumsum, filter_gt and compress at 1/50/99% kept, gather of "
                               f"2^{args.log2n} random indices into a {table_rows}-row f64/i64 table",
                   "rows": n, "table_rows": table_rows},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "verified": verified,
        "checks": checks,
        "primitives": prims,
        "regions": regions,
        "zero_mean": zero_mean,
        "f32_int": f32_int,
        "batched": batched,
        "serve": serve,
        "sort": sort,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
