"""Serving loop: a request stream over device-resident columns, with continuous batching.

Requests arrive as a seeded Poisson stream at a fixed rate. Each runs one program from the
statement set over its columns, values x (f64) and keys k (i64):

    sum_filter:  reduce_add_f64(filter_gt_f64(x, 0.9))
    dot:         dot_f64(x, x2)
    count:       count_gt_f64(x, 0.5)
    scan:        scan_add_i64(k)
    argsort:     argsort_i64(k)
    group_sum:   g = group_id_i64(k); reduce_into_add_f64(g, x, reduce_max_i64(g) + 1)

A share of the requests (40% by default) reuses a column that is already resident in HBM; the
others create theirs on the device (rand_f64). This replaces the reference's compile-once
Prepared::execute (src/interp.rs:441-516) and columns::load_inputs (src/columns.rs:94-134):
columns stay on the device between requests instead of being loaded per program.

Scheduling: the scheduler admits a request as soon as a worker slot is free and the HBM
budget covers the request's columns, outputs and scratch (estimated per program), evicting
idle columns least recently used first. Admitted requests run at once on their workers' own
contexts (own HIP streams), so requests that arrive while others run overlap on the device;
nothing waits for a batch to fill or drain. With coalescing on, a worker that takes a request
also takes the arrived requests of the same program behind it (up to `coalesce` in all, as
the budget allows) and runs each statement once for all of them with vk_batch_call, so a
backlog of small requests becomes few launches. Latency is completion minus arrival time.

Every result is reduced to a digest on the device (wrapping sum and xor of its 64-bit words,
or the scalar's bits), so a result can be compared between a run in the mix and a run alone.
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import native as N
from .batch import MASK64, splitmix64, splitmix64_np
from .kernels import Array, Dispatch, Registry, Scalar

PROGRAMS = ("sum_filter", "dot", "count", "scan", "argsort", "group_sum")

# Device bytes per row a program needs beyond its input columns (outputs and scratch), an
# upper estimate used for admission.
EXTRA_BYTES_PER_ROW = {"sum_filter": 24, "dot": 0, "count": 0, "scan": 16, "argsort": 64,
                       "group_sum": 112}


@dataclass
class Request:
    rid: int
    t_arrive: float  # seconds after the stream starts
    program: str
    column: int      # column id
    rows: int
    reuse: bool


def make_requests(count: int, qps: float, seed: int = 5, lo: int = 10_000, hi: int = 64 << 20,
                  reuse: float = 0.4) -> List[Request]:
    """A seeded Poisson stream: exponential gaps of mean 1/qps, programs uniform over
    PROGRAMS, new columns with log-uniform row counts in [lo, hi], and a `reuse` share of
    requests taking a column created by an earlier request (uniformly among them)."""
    def u(i, k):
        return (splitmix64((seed * 0x100000001B3 + i * 8 + k) & MASK64) >> 11) * 2.0 ** -53
    reqs, t, sizes = [], 0.0, []
    for i in range(count):
        t += -math.log(1.0 - u(i, 0)) / qps
        prog = PROGRAMS[int(u(i, 1) * len(PROGRAMS))]
        if sizes and u(i, 2) < reuse:
            col = int(u(i, 3) * len(sizes))
            reqs.append(Request(i, t, prog, col, sizes[col], True))
        else:
            n = int(round(2.0 ** (math.log2(lo) + u(i, 4) * (math.log2(hi) - math.log2(lo)))))
            sizes.append(n)
            reqs.append(Request(i, t, prog, len(sizes) - 1, n, False))
    return reqs


@dataclass
class Column:
    x: Optional[Array]
    rows: int
    nbytes: int
    pins: int = 0
    last_use: float = 0.0
    ready: bool = True  # False while the request that creates it runs


class ColumnStore:
    """Device-resident value columns (f64, rand_f64 with the column id as seed), created on one
    loader context and read by every worker; keys are views of one shared key column."""

    def __init__(self, loader: Dispatch, budget: int, key_rows: int, seed: int = 9):
        self.D = loader
        self.budget = budget
        self.cols: Dict[int, Column] = {}
        self.resident = 0
        self.evictions = 0
        pool = splitmix64_np(seed, 0, 1024).view(np.int64)
        self.keys = Array.empty(loader.ctx, "i64", key_rows)
        step = 1 << 24
        for s0 in range(0, key_rows, step):
            c = min(step, key_rows - s0)
            sel = (splitmix64_np(seed + 1, s0, c) & np.uint64(1023)).astype(np.int64)
            chunk = Array.from_numpy(loader.ctx, pool[sel])
            loader.elementwise_tile("add_i64", [chunk, Scalar("i64", 0)], self.keys.slice(s0, c))
            del chunk
        self.x2 = loader.call("rand_f64", [Scalar("i64", key_rows), Scalar("i64", seed + 2)])
        loader.sync()
        self.base_bytes = key_rows * 16

    def fits(self, need: int, in_flight: int, keep: int) -> bool:
        """Whether `need` more bytes fit beside the resident columns and the in-flight
        requests, after evicting idle columns other than `keep` (which this does)."""
        while self.base_bytes + self.resident + in_flight + need > self.budget:
            idle = [(c.last_use, cid) for cid, c in self.cols.items() if c.pins == 0 and c.ready and cid != keep]
            if not idle:
                return False
            _, cid = min(idle)
            col = self.cols.pop(cid)
            self.resident -= col.nbytes
            self.evictions += 1
            col.x.release()
        return True

    def usable(self, req: Request) -> bool:
        """Whether the request's value column is resident and complete, or not resident at all
        (the request then creates it); a column that another request is creating is neither."""
        col = self.cols.get(req.column)
        return col is None or col.ready

    def acquire(self, req: Request, now: float) -> Tuple[Optional[Array], Array, Array, bool]:
        """The request's columns, pinned. A value column that is not resident is reserved and
        returned as None: the worker that runs the request creates it on its own stream, ahead of
        the request's kernels, and `created` hands it over when the request completes, so the
        scheduler never waits for a column."""
        col = self.cols.get(req.column)
        created = col is None
        if created:
            col = Column(None, req.rows, req.rows * 8, ready=False)
            self.cols[req.column] = col
            self.resident += col.nbytes
        col.pins += 1
        col.last_use = now
        return col.x, self.keys.slice(0, req.rows), self.x2.slice(0, req.rows), created

    def created(self, req: Request, x: Optional[Array]):
        """The column the request created, complete now that the request has finished."""
        col = self.cols[req.column]
        if x is not None:
            col.x = x
            col.ready = True

    def release(self, req: Request):
        self.cols[req.column].pins -= 1


def _digest_words(D: Dispatch, a: Array) -> Tuple[int, int]:
    """(wrapping sum, xor) of an array's 64-bit words, on the device."""
    w = Array.wrap(D.ctx, "i64", a.device_ptr(), len(a)) if a.type != "i64" else a
    s = D.call("reduce_add_i64", [w]).value & MASK64 if len(w) else 0
    x = D.call("reduce_xor_i64", [w]).value & MASK64 if len(w) else 0
    return s, x


def run_program(D: Dispatch, prog: str, x: Array, k: Array, x2: Array):
    """Runs one program; returns (digest, algorithmic bytes)."""
    S = Scalar
    n = len(x)
    if prog == "sum_filter":
        f = D.call("filter_gt_f64", [x, S("f64", 0.9)])
        s = D.call("reduce_add_f64", [f])
        return ("f64", s.bits()), 8 * n + 16 * len(f)
    if prog == "dot":
        return ("f64", D.call("dot_f64", [x, x2]).bits()), 16 * n
    if prog == "count":
        return ("i64", D.call("count_gt_f64", [x, S("f64", 0.5)]).value), 8 * n
    if prog == "scan":
        return ("array", _digest_words(D, D.call("scan_add_i64", [k]))), 16 * n
    if prog == "argsort":
        return ("array", _digest_words(D, D.call("argsort_i64", [k]))), 16 * n
    if prog == "group_sum":
        g = D.call("group_id_i64", [k])
        ng = D.call("reduce_max_i64", [g]).value + 1
        r = D.call("reduce_into_add_f64", [g, x, S("i64", ng)])
        return ("array", _digest_words(D, r)), 16 * n + 8 * ng
    raise ValueError(prog)


def _digest_group(D: Dispatch, arrays: List[Array]) -> List[Tuple[int, int]]:
    """_digest_words of every array, with one batched launch sequence per reduction."""
    views = [Array.wrap(D.ctx, "i64", a.device_ptr(), len(a)) if a.type != "i64" else a for a in arrays]
    live = [i for i, w in enumerate(views) if len(w)]
    out = [(0, 0)] * len(arrays)
    if live:
        sums = D.batch_call("reduce_add_i64", [[views[i]] for i in live])
        xors = D.batch_call("reduce_xor_i64", [[views[i]] for i in live])
        for i, a, b in zip(live, sums, xors):
            out[i] = (a.value & MASK64, b.value & MASK64)
    return out


def run_program_group(D: Dispatch, prog: str, cols: List[Tuple[Array, Array, Array]]):
    """Runs one program for a group of requests with one batched call per statement
    (vk_batch_call); returns (digest, algorithmic bytes) per request, each equal to
    run_program's for that request alone."""
    if len(cols) == 1:
        return [run_program(D, prog, *cols[0])]
    S = Scalar
    xs = [c[0] for c in cols]
    ns = [len(x) for x in xs]
    if prog == "sum_filter":
        fs = D.batch_call("filter_gt_f64", [[x, S("f64", 0.9)] for x in xs])
        ss = D.batch_call("reduce_add_f64", [[f] for f in fs])
        return [(("f64", s_.bits()), 8 * n + 16 * len(f)) for s_, f, n in zip(ss, fs, ns)]
    if prog == "dot":
        ds = D.batch_call("dot_f64", [[x, c[2]] for x, c in zip(xs, cols)])
        return [(("f64", d.bits()), 16 * n) for d, n in zip(ds, ns)]
    if prog == "count":
        cs = D.batch_call("count_gt_f64", [[x, S("f64", 0.5)] for x in xs])
        return [(("i64", c_.value), 8 * n) for c_, n in zip(cs, ns)]
    if prog == "scan":
        rs = D.batch_call("scan_add_i64", [[c[1]] for c in cols])
        return [(("array", d), 16 * n) for d, n in zip(_digest_group(D, rs), ns)]
    if prog == "argsort":
        rs = D.batch_call("argsort_i64", [[c[1]] for c in cols])
        return [(("array", d), 16 * n) for d, n in zip(_digest_group(D, rs), ns)]
    if prog == "group_sum":
        gs = D.batch_call("group_id_i64", [[c[1]] for c in cols])
        ngs = [m.value + 1 for m in D.batch_call("reduce_max_i64", [[g] for g in gs])]
        rs = D.batch_call("reduce_into_add_f64", [[g, x, S("i64", ng)] for g, x, ng in zip(gs, xs, ngs)])
        return [(("array", d), 16 * n + 8 * ng) for d, n, ng in zip(_digest_group(D, rs), ns, ngs)]
    raise ValueError(prog)


@dataclass
class Outcome:
    rid: int
    program: str
    rows: int
    reuse: bool
    t_arrive: float
    t_admit: float
    t_done: float
    digest: tuple
    nbytes: int
    concurrent: int  # requests running when this one was admitted (itself included)
    group: int = 1   # requests run together with this one (coalesced), itself included

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrive


@dataclass
class Report:
    outcomes: List[Outcome]
    wall: float
    workers: int
    budget: int
    evictions: int
    created: int

    def pct(self, q: float) -> float:
        lat = sorted(o.latency for o in self.outcomes)
        return lat[min(len(lat) - 1, int(math.ceil(q / 100.0 * len(lat))) - 1)] if lat else 0.0

    def summary(self) -> dict:
        nbytes = sum(o.nbytes for o in self.outcomes)
        return {"requests": len(self.outcomes), "workers": self.workers,
                "budget_GB": round(self.budget / 1e9, 1),
                "p50_ms": round(self.pct(50) * 1e3, 3), "p99_ms": round(self.pct(99) * 1e3, 3),
                "GBps": round(nbytes / self.wall / 1e9, 1), "wall_s": round(self.wall, 3),
                "columns_created": self.created, "evictions": self.evictions,
                "max_concurrent": max((o.concurrent for o in self.outcomes), default=0),
                "mean_group": round(sum(o.group for o in self.outcomes) / max(1, len(self.outcomes)), 2)}


class Server:
    """The request loop: one scheduler (this thread) and `workers` worker threads, each with
    its own context."""

    def __init__(self, workers: int = 4, budget: int = 200 * 10**9, key_rows: int = 64 << 20,
                 coalesce: int = 1, window: int = 256):
        self.R = Registry()
        self.loader = self.R.resolve(0)
        self.store = ColumnStore(self.loader, budget, key_rows)
        self.nworkers = workers
        self.workers = [self.R.resolve(0) for _ in range(workers)]
        # continuous batching: a worker that takes a request also takes up to coalesce - 1 more
        # requests of the same program that have arrived and fit the budget (looking `window`
        # pending requests ahead), and runs each statement once for all of them
        self.coalesce = max(1, coalesce)
        self.window = window

    def close(self):
        for c in list(self.store.cols.values()):
            if c.x is not None:
                c.x.release()
        self.store.cols.clear()

    def run(self, requests: List[Request], paced: bool = True) -> Report:
        """Serves `requests` (arrival times honoured when `paced`), returns the outcomes."""
        done_q: "queue.Queue" = queue.Queue()
        work_qs = [queue.Queue() for _ in range(self.nworkers)]
        start = time.perf_counter()
        clock = lambda: time.perf_counter() - start

        def worker(i: int):
            D = self.workers[i]
            while True:
                job = work_qs[i].get()
                if job is None:
                    return
                group, t_admit, conc = job
                made = [None] * len(group)
                try:
                    # columns this group creates: on this worker's stream, ahead of its kernels
                    cols = []
                    for g, (req, x, k, x2) in enumerate(group):
                        if x is None:
                            x = D.call("rand_f64", [Scalar("i64", req.rows), Scalar("i64", 1000 + req.column)])
                            made[g] = x
                        cols.append((x, k, x2))
                    res = run_program_group(D, group[0][0].program, cols)
                    D.sync()
                    err = None
                except Exception as e:  # reported by the scheduler
                    res, err = [(None, 0)] * len(group), e
                t_done = clock()
                done_q.put((i, [(req, Outcome(req.rid, req.program, req.rows, req.reuse, req.t_arrive, t_admit,
                                              t_done, dg, nb, conc, len(group)), m)
                                for (req, _, _, _), (dg, nb), m in zip(group, res, made)], err))

        threads = [threading.Thread(target=worker, args=(i,), daemon=True) for i in range(self.nworkers)]
        for t in threads:
            t.start()
        idle = list(range(self.nworkers))
        in_flight: Dict[int, int] = {}  # rid -> admitted bytes
        pending = list(requests)
        outcomes: List[Outcome] = []
        created = 0
        error = None

        def reap(block: bool):
            nonlocal error
            try:
                i, done, err = done_q.get(block=block, timeout=5.0 if block else None)
            except queue.Empty:
                return False
            idle.append(i)
            for req, out, made in done:
                in_flight.pop(req.rid, None)
                self.store.created(req, made)
                self.store.release(req)
                outcomes.append(out)
            if err is not None and error is None:
                error = err
            return True

        while pending or in_flight:
            while reap(False):
                pass
            if error is not None:
                break
            if not pending:
                reap(True)
                continue
            # the first pending request whose column is not being created by another request
            at = next((j for j in range(min(len(pending), self.window)) if self.store.usable(pending[j])), None)
            if at is None or (paced and clock() < pending[at].t_arrive and at > 0):
                reap(True)  # wait for a creating request to finish
                continue
            req = pending[at]
            if paced and clock() < req.t_arrive:
                if in_flight:
                    # wait for a completion or the next arrival, whichever is first
                    try:
                        got = done_q.get(timeout=max(0.0, req.t_arrive - clock()))
                        done_q.put(got)
                    except queue.Empty:
                        pass
                else:
                    time.sleep(max(0.0, req.t_arrive - clock()))
                continue
            extra = EXTRA_BYTES_PER_ROW[req.program] * req.rows
            need = extra + (0 if req.column in self.store.cols else 8 * req.rows)
            if not idle or not self.store.fits(need, sum(in_flight.values()), req.column):
                if not in_flight:
                    raise RuntimeError(f"request {req.rid} does not fit the HBM budget")
                reap(True)
                continue
            pending.pop(at)
            now = clock()
            x, k, x2, made = self.store.acquire(req, now)
            created += made
            in_flight[req.rid] = extra
            group = [(req, x, k, x2)]
            # coalesce arrived requests of the same program that fit the budget
            j = 0
            while len(group) < self.coalesce and j < min(len(pending), self.window):
                r2 = pending[j]
                if paced and r2.t_arrive > now:
                    break
                if r2.program != req.program or not self.store.usable(r2):
                    j += 1
                    continue
                extra2 = EXTRA_BYTES_PER_ROW[r2.program] * r2.rows
                need2 = extra2 + (0 if r2.column in self.store.cols else 8 * r2.rows)
                if not self.store.fits(need2, sum(in_flight.values()), r2.column):
                    break
                pending.pop(j)
                x_, k_, x2_, made = self.store.acquire(r2, now)
                created += made
                in_flight[r2.rid] = extra2
                group.append((r2, x_, k_, x2_))
            i = idle.pop()
            work_qs[i].put((group, clock(), len(in_flight)))
        for q in work_qs:
            q.put(None)
        for t in threads:
            t.join()
        if error is not None:
            raise error
        wall = time.perf_counter() - start
        outcomes.sort(key=lambda o: o.rid)
        return Report(outcomes, wall, self.nworkers, self.store.budget, self.store.evictions, created)
