"""The serving loop: requests admitted while others run overlap on the device, and every
result is the same alone and in the mix and equals the oracle's on the same columns (bit for
bit: scalar bits, or the wrapping sum and xor of an array result's 64-bit words). The serving
form has no counterpart in the reference; its parity bar is the single entries'."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

MASK64 = (1 << 64) - 1


def _alone(SV, reqs, key_rows):
    srv = SV.Server(workers=1, budget=40 * 10**9, key_rows=key_rows)
    out = {}
    for r in reqs:
        out[r.rid] = srv.run([r], paced=False).outcomes[0].digest
    srv.close()
    return out


class OracleColumns:
    """The server's columns restated on the host: value column c = rand_f64(rows, 1000 + c),
    keys = pool[splitmix64(10 + i) & 1023] with pool = splitmix64(9 + j), x2 = rand_f64(., 11)."""

    def __init__(self, SV, key_rows):
        pool = SV.splitmix64_np(9, 0, 1024).view(np.int64)
        self.keys = pool[(SV.splitmix64_np(10, 0, key_rows) & np.uint64(1023)).astype(np.int64)]
        self.x2 = O.rand_f64(key_rows, 11)


def _words(a):
    w = np.ascontiguousarray(a).view(np.uint64)
    return (int(np.sum(w, dtype=np.uint64)) & MASK64, int(np.bitwise_xor.reduce(w)) if w.size else 0)


def oracle_digest(cols, req):
    import struct
    n = req.rows
    x = O.rand_f64(n, 1000 + req.column)
    k = cols.keys[:n]
    fb = lambda v: ("f64", struct.unpack("<Q", struct.pack("<d", v))[0])
    if req.program == "sum_filter":
        return fb(O.canonical_sum_f64(O.filter_cmp(x, "gt", 0.9)))
    if req.program == "dot":
        return fb(O.dot_f64(x, cols.x2[:n]))
    if req.program == "count":
        return ("i64", O.count_cmp(x, "gt", 0.5))
    if req.program == "scan":
        return ("array", _words(O.scan_add_i64(k)))
    if req.program == "argsort":
        return ("array", _words(O.argsort(k)))
    ids = O.group_id(k)
    return ("array", _words(O.reduce_into(ids, x, int(ids.max()) + 1 if n else 0)))


def test_results_same_alone_and_in_the_mix():
    from verkehrt_amd import serve as SV
    key_rows = 1 << 22
    reqs = SV.make_requests(60, qps=20000.0, seed=3, lo=10_000, hi=key_rows, reuse=0.4)
    assert any(r.reuse for r in reqs) and any(not r.reuse for r in reqs)
    assert {r.program for r in reqs} == set(SV.PROGRAMS)
    srv = SV.Server(workers=4, budget=40 * 10**9, key_rows=key_rows)
    mix = srv.run(reqs)
    srv.close()
    s = mix.summary()
    assert s["requests"] == len(reqs) and s["max_concurrent"] >= 2
    assert 0 < s["p50_ms"] <= s["p99_ms"]
    alone = _alone(SV, reqs, key_rows)
    cols = OracleColumns(SV, key_rows)
    for o, r in zip(mix.outcomes, reqs):
        assert o.digest == alone[o.rid], (o.rid, o.program, o.rows)
        assert o.digest == oracle_digest(cols, r), (o.rid, o.program, o.rows)


def test_coalesced_requests_match_alone_and_the_oracle():
    # continuous batching: a saturated stream of small requests, each worker taking up to 16
    # arrived requests of one program and running every statement once for the group
    from verkehrt_amd import serve as SV
    key_rows = 1 << 21
    reqs = SV.make_requests(120, qps=1e9, seed=7, lo=1_000, hi=1 << 20, reuse=0.4)
    srv = SV.Server(workers=4, budget=40 * 10**9, key_rows=key_rows, coalesce=16)
    mix = srv.run(reqs, paced=False)
    srv.close()
    s = mix.summary()
    assert s["requests"] == len(reqs) and s["mean_group"] > 2, s
    alone = _alone(SV, reqs, key_rows)
    cols = OracleColumns(SV, key_rows)
    for o, r in zip(mix.outcomes, reqs):
        assert o.digest == alone[o.rid], (o.rid, o.program, o.rows, o.group)
        assert o.digest == oracle_digest(cols, r), (o.rid, o.program, o.rows, o.group)


def test_configured_stream_matches_the_oracle():
    # the configured serving form: sizes log-uniform in [10K, 64M] rows, 40% reuse, a 200 GB
    # budget, four workers, a paced 1000 QPS stream; every result against the oracle
    from verkehrt_amd import serve as SV
    key_rows = 64 << 20
    reqs = SV.make_requests(24, qps=1000.0, seed=11, lo=10_000, hi=key_rows, reuse=0.4)
    srv = SV.Server(workers=4, budget=200 * 10**9, key_rows=key_rows, coalesce=8)
    mix = srv.run(reqs)
    srv.close()
    s = mix.summary()
    assert s["requests"] == len(reqs) and s["max_concurrent"] >= 2
    cols = OracleColumns(SV, key_rows)
    for o, r in zip(mix.outcomes, reqs):
        assert o.digest == oracle_digest(cols, r), (o.rid, o.program, o.rows)


def test_budget_evicts_idle_columns_and_keeps_results():
    from verkehrt_amd import serve as SV
    key_rows = 1 << 21
    reqs = SV.make_requests(40, qps=20000.0, seed=4, lo=100_000, hi=key_rows, reuse=0.5)
    # room for the shared columns, the largest request with its column, and about six more
    # value columns: the 22 columns (168 MB) the stream creates cannot all stay resident
    budget = key_rows * 16 + key_rows * (8 + max(SV.EXTRA_BYTES_PER_ROW.values())) + 6 * key_rows * 8
    srv = SV.Server(workers=3, budget=budget, key_rows=key_rows)
    mix = srv.run(reqs)
    srv.close()
    assert mix.evictions > 0
    alone = _alone(SV, reqs, key_rows)
    cols = OracleColumns(SV, key_rows)
    for o, r in zip(mix.outcomes, reqs):
        assert o.digest == alone[o.rid], (o.rid, o.program)
        assert o.digest == oracle_digest(cols, r), (o.rid, o.program)
