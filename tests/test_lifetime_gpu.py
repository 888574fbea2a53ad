"""Context lifetime and look-back epoch regression tests.

- A context destroyed while arrays, a reducer, a scanner, a builder and a batch plan still
  hold it must stay alive until the last of them is released, in any order, and the process
  must exit normally. Runs in a child process with a time limit, so a regression shows as a
  failed exit instead of a hang of the test session.
- Look-back kernels (selections, scans) tag their status words with a 16-bit epoch taken per
  launch and clear the status buffers when it wraps; more than 2^17 launches on one context
  cross the wrap twice, and every result across it must stay exact.
"""
import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = textwrap.dedent("""
    import sys
    sys.path.insert(0, {root!r})
    import numpy as np
    from verkehrt_amd import Registry, Scalar, Input, Stmt
    order = {order!r}
    D = Registry().resolve(0)
    x = D.array(np.arange(100_000, dtype=np.float64))
    y = D.call("filter_gt_f64", [x, Scalar("f64", 10.0)])
    r = D.reducer("reduce_add_f64")
    r.update([x.slice(0, 4096)])
    sc = D.scanner("scan_add_i64")
    b = D.builder("f64")
    D.select_tile("filter_gt_f64", [x, Scalar("f64", 50_000.0)], 0, b)
    plan = D.batch_plan([[x], [y]])
    s0 = plan.add("filter_gt_f64", [Input(0), Scalar("f64", 5.0)])
    plan.add("reduce_add_f64", [Stmt(s0)])
    plan.compile()
    plan.run()
    D.ctx.sync()
    print("built", flush=True)
    D.ctx.close()  # vk_ctx_destroy while everything above still holds the context
    objs = {{"x": x, "y": y, "r": r, "sc": sc, "b": b, "plan": plan}}
    for name in order:
        o = objs.pop(name)
        if name == "r":
            o.free()
        elif name == "sc":
            o.free()
        elif name == "plan":
            o.close()
        elif name == "b":
            o.__del__()
            o.handle = None
        else:
            o.release()
        print("released", name, flush=True)
    print("done", flush=True)
""")


@pytest.mark.parametrize("order", [
    ["x", "y", "r", "sc", "b", "plan"],
    ["plan", "b", "sc", "r", "y", "x"],
    ["r", "x", "plan", "y", "sc", "b"],
])
def test_destroy_context_before_its_holders(order):
    script = CHILD.format(root=ROOT, order=order)
    p = subprocess.run([sys.executable, "-c", script], capture_output=True, text=True, timeout=90)
    assert p.returncode == 0, (p.returncode, p.stdout[-2000:], p.stderr[-2000:])
    assert p.stdout.strip().endswith("done"), p.stdout


def test_lookback_epoch_wraps():
    from verkehrt_amd import Registry, Scalar
    D = Registry().resolve(0)
    rng = np.random.default_rng(50)
    n = 40_000  # several look-back tiles per launch
    x = rng.random(n)
    xi = rng.integers(-2**40, 2**40, n).astype(np.int64)
    dx, di = D.array(x), D.array(xi)
    cs = [0.1, 0.5, 0.9, 0.99]
    want_f = {c: O.filter_cmp(x, "gt", c) for c in cs}
    want_s = O.scan_add_i64(xi)
    # two launches per round (a selection and a scan), each taking an epoch: 70,000 rounds
    # take 140,000 epochs, two wraps of the 16-bit epoch
    for k in range(70_000):
        c = cs[k % 4]
        f = D.call("filter_gt_f64", [dx, Scalar("f64", c)])
        assert len(f) == want_f[c].size, k
        s = D.call("scan_add_i64", [di])
        if k % 2048 == 0 or 32_000 <= k <= 33_500 or 65_000 <= k <= 66_000 or k >= 69_990:
            assert np.array_equal(f.to_numpy(), want_f[c]), k
            assert np.array_equal(s.to_numpy(), want_s), k
        del f, s
