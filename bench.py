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


def cpu_baseline(sample_rows: int, budget_s: float = 12.0):
    """The reference's best x86 CPU path, restated in C (oracle/cpu_best.c: x86-64-v4 auto
    loops, AVX-512 compress stores and vpgatherqq), on one host core over the same suite at
    `sample_rows` rows, repeated for about `budget_s` seconds. Every result is checked against
    the oracle (the scalar restatement) before timing."""
    import ctypes as C
    from oracle import oracle as O
    B = O.best()
    n = sample_rows
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    x1, x2 = O.rand_f64(n, 1), O.rand_f64(n, 2)
    i1 = i64_uniform_40(n, 4)
    tbl = O.rand_f64(min(1 << 26, n), 3)
    idx = gather_indices(n, 5, tbl.size)
    masks = {c: (x2 > c).astype(np.uint8) for c in SELECTIVITIES}
    fo = np.empty(n + 8, np.float64)
    io = np.empty(n, np.int64)
    bad = np.zeros(1, np.int64)
    # correctness gate before timing
    assert B.vkb_dot_f64(p(x1), p(x2), n) == O.dot_f64(x1, x2)
    assert B.vkb_sum_f64(p(x1), n) == O.canonical_sum_f64(x1)
    assert B.vkb_count_gt_f64(p(x1), n, 0.5) == O.count_cmp(x1, "gt", 0.5)
    k = B.vkb_filter_gt_f64(p(x1), n, 0.5, p(fo))
    assert np.array_equal(fo[:k], O.filter_cmp(x1, "gt", 0.5))
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
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(gbps, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"the same suite (one gather table) at {n} rows, {passes} passes in {spent:.1f} s, "
                      f"C restatement of the reference's best x86-64-v4 path on one core of {cpu}"}


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
            "speedup": round(tp / tb, 2), "identical": same, "filter_sum": filter_sum}


def sort_suite(D):
    """Sort and grouping cases: argsort of 2^27 keys (u64 uniform splitmix64, i64 Zipf(1.1)
    over 2^20 values, f64 normal(0,1) with +-0), group-by sum of rand_f64 values over i64 keys
    with 1K, 1M and 64M distinct values at 2^28 rows (group_id, reduce_max + 1,
    reduce_into_add_f64), and scan_by add over 2^28 rows with starts of probability 1/16.
    Times are the median of three calls from HIP events; GB/s counts the algorithmic bytes."""
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

    out = {}
    n = 1 << 27
    rng = np.random.default_rng(3)
    keys = {"u64": Array.from_numpy(ctx, splitmix64_np(7, 0, n), "u64"),
            "i64_zipf": Array.from_numpy(ctx, np.minimum(rng.zipf(1.1, n), 1 << 20).astype(np.int64)),
            "f64_normal": None}
    f = rng.normal(size=n)
    f[rng.integers(0, n, 1000)] = 0.0
    f[rng.integers(0, n, 1000)] = -0.0
    keys["f64_normal"] = Array.from_numpy(ctx, f)
    del f
    for label, a in keys.items():
        name = {"u64": "argsort_u64", "i64_zipf": "argsort_i64", "f64_normal": "argsort_f64"}[label]
        us = timed(name, [a])
        out[f"argsort_{label}"] = {"rows": n, "us": round(us, 1)}
    del keys
    n = 1 << 28
    vals = D.call("rand_f64", [S("i64", n), S("i64", 1)])
    for g in (1 << 10, 1 << 20, 1 << 26):
        pool = splitmix64_np(11, 0, g).view(np.int64)
        sel = (splitmix64_np(12, 0, n) % np.uint64(g)).astype(np.int64)
        k = Array.from_numpy(ctx, pool[sel])
        del sel
        t_id = timed("group_id_i64", [k], reps=2)
        ids = D.call("group_id_i64", [k])
        ng = D.call("reduce_max_i64", [ids]).value + 1
        t_ri = timed("reduce_into_add_f64", [ids, vals, S("i64", ng)], reps=2)
        out[f"group_sum_{g}"] = {"rows": n, "group_id_us": round(t_id, 1), "reduce_into_us": round(t_ri, 1),
                                 "GBps_min_io": round((16 * n + 8 * g) / (t_id + t_ri) / 1e3, 1)}
        del k, ids
    st = np.zeros(n, np.uint8)
    st[(splitmix64_np(13, 0, n) % np.uint64(16)) == 0] = 1
    sa = Array.from_numpy(ctx, st, "bool")
    del st
    iv = Array.from_numpy(ctx, i64_uniform_40(n, 4))
    us = timed("scan_by_add_i64", [iv, sa])
    out["scan_by_add_i64"] = {"rows": n, "us": round(us, 1), "GBps": round(17 * n / us / 1e3, 1)}
    return out


def serve_suite(requests: int, qps: float):
    """The serving loop: a seeded Poisson stream of `requests` programs at `qps`, sizes
    log-uniform in [10K, 64M] rows, 40% reusing resident columns, 4 workers, a 200 GB HBM
    budget; latency percentiles at that rate, and the throughput of the same requests all
    arriving at once."""
    from verkehrt_amd import serve as SV
    reqs = SV.make_requests(requests, qps=qps, seed=11)
    srv = SV.Server(workers=4, budget=200 * 10**9, key_rows=64 << 20)
    srv.run(SV.make_requests(50, qps=1e9, seed=99, hi=1 << 20), paced=False)  # warm-up
    srv.close()
    paced = srv.run(reqs).summary()
    srv.close()
    sat = srv.run(reqs, paced=False).summary()
    srv.close()
    return {"requests": requests, "qps": qps, "sizes": "log-uniform in [10K, 64M] rows",
            "programs": "sum_filter, dot, count, scan, argsort, group_sum (uniform)",
            "reuse": 0.4, "workers": 4, "budget_GB": 200,
            "p50_ms": paced["p50_ms"], "p99_ms": paced["p99_ms"], "GBps_at_qps": paced["GBps"],
            "max_concurrent": paced["max_concurrent"], "columns_created": paced["columns_created"],
            "saturated_GBps": sat["GBps"], "saturated_wall_s": sat["wall_s"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--cpu-rows", type=int, default=1 << 24)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the many-arrays batched comparison")
    ap.add_argument("--batch-arrays", type=int, default=4096)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--sort", action="store_true",
                    help="also time the sort and grouping cases (argsort of 2^27 keys, group-by sum "
                         "over 1K/1M/64M groups at 2^28 rows, scan_by add)")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-qps", type=float, default=1000.0)
    ap.add_argument("--only", default="",
                    help="comma-separated primitive labels to run (for profiling); 'copy_f64' adds a "
                         "plain copy (add_f64 x + 0.0) used to calibrate HBM counters")
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

    x1 = D.call("rand_f64", [S("i64", n), S("i64", 1)])
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
    roofline = {"bound": "hbm", "kernel": dom, "achieved": prims[dom]["GBps"], "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": prims[dom]["frac"], "traffic": measured_traffic(dom, args.log2n),
                "bytes_per_launch": prims[dom]["bytes"], "launch_us": prims[dom]["us"]}
    batched = None if (args.no_batched or args.only) else batched_suite(D, args.batch_arrays)
    serve = None if (args.no_serve or args.only) else serve_suite(args.serve_requests, args.serve_qps)
    sort = sort_suite(D) if args.sort else None
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_rows, args.cpu_seconds)
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
        "primitives": prims,
        "batched": batched,
        "serve": serve,
        "sort": sort,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
