"""The HIP step on the configurations that pin the oracle's control flow
(tests/test_oracle_control_flow.py): in-place ridging seen by the next pair,
nearest-first visiting, the per-floe summation order and stale-force stress
diagnostics.  FP64 through the C ABI, against the oracle: pair flags exact,
the order-sensitive force sums bitwise."""
import numpy as np
import pytest

from godar_amd.params import GodarParams
from test_oracle_control_flow import BASE, _field, _star

pytestmark = pytest.mark.gpu


def _run(p, f, steps=1):
    from godar_amd.sim import Context
    from oracle_py import OracleSim
    ctx = Context(f.n, 1, "fp64", record_forces=True)
    m = ctx.add_field(p, f)
    O = OracleSim(p, f)
    out = []
    for _ in range(steps):
        ctx.step(1)
        O.step()
        li, lj, lf = O.log()
        cand = sorted(zip(li.tolist(), lj.tolist()))
        cont = sorted((a, b) for a, b, c in zip(li.tolist(), lj.tolist(), lf.tolist()) if c & 1)
        ci, cj = ctx.pairs(m, "candidate")
        ki, kj = ctx.pairs(m, "contact")
        assert list(zip(ci.tolist(), cj.tolist())) == cand
        assert list(zip(ki.tolist(), kj.tolist())) == cont
        out.append((ctx.floes(m), ctx.forces(m), O.state(), O.forces(), cont))
    ctx.close()
    return out


def test_gpu_ridging_in_place_and_nearest_first():
    x0 = 5e4
    chain = _field([x0, x0 + 1800.0, x0 + 3795.0], [5e4] * 3, [1000.0] * 3, [2.0, 0.5, 2.0])
    star3 = _field([x0, x0 - 1995.0, x0 + 1800.0], [5e4] * 3, [1000.0] * 3, [0.5, 2.0, 2.0])
    p = GodarParams(**BASE, sigmanc_crit=1.0)
    for f, absent in ((chain, (1, 2)), (star3, (0, 1))):
        (a, fa, b, fb, cont), = _run(p, f)
        assert absent not in cont
        for k in ("r", "h", "x", "y", "u", "v"):
            np.testing.assert_allclose(a[k], b[k], rtol=1e-13, atol=0, err_msg=k)


def test_gpu_per_floe_summation_order_bitwise():
    """No transcendental on this path: floe 2's four-term contact sums equal
    the oracle's bit for bit (other orders differ in the last bits)."""
    pts, hs = _star()
    f = _field([q[0] for q in pts], [q[1] for q in pts], [1000.0] * 5, hs)
    (a, fa, b, fb, cont), = _run(GodarParams(**BASE, ridging=0), f)
    assert len(cont) == 4
    for k in ("fcx", "fcy", "mc", "tfx", "tfy"):
        assert np.array_equal(fa[k], fb[k]), k
    for k in ("x", "y", "u", "v"):
        assert np.array_equal(a[k], b[k]), k


def test_gpu_separated_pair_keeps_last_contact_stress():
    p = GodarParams(**BASE, ridging=0, Cdair=0.0, Cdwater=0.0, Csair=0.0, Cswater=0.0, lat=0.0)
    f = _field([5e4, 5e4 + 2205.0], [5e4, 5e4 + 40.0], [1000.0, 1200.0], [1.0, 1.0], us=[0.5, -0.5])
    stale = 0
    for a, fa, b, fb, cont in _run(p, f, steps=60):
        for k in ("tsigxx", "tsigyy", "tsigxy", "tsigyx", "tfx", "tfy"):
            np.testing.assert_allclose(fa[k], fb[k], rtol=1e-13, atol=0, err_msg=k)
        if not cont and abs(fb["tsigxx"][0]) > 0:
            stale += 1
    assert stale > 10
