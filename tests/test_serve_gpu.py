"""The serving loop: requests admitted while others run overlap on the device, and every
result is the same alone and in the mix (bit for bit: scalar bits, or the wrapping sum and xor
of an array result's 64-bit words). The serving form has no counterpart in the reference."""
import pytest

pytestmark = pytest.mark.gpu


def _alone(SV, reqs, key_rows):
    srv = SV.Server(workers=1, budget=40 * 10**9, key_rows=key_rows)
    out = {}
    for r in reqs:
        out[r.rid] = srv.run([r], paced=False).outcomes[0].digest
    srv.close()
    return out


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
    for o in mix.outcomes:
        assert o.digest == alone[o.rid], (o.rid, o.program, o.rows)


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
    for o in mix.outcomes:
        assert o.digest == alone[o.rid], (o.rid, o.program)
