"""The oracle against the committed golden fixtures (made by GODAR's own
routines compiled from the reference, see tests/golden/make_golden.py).
Bit-for-bit: runs without the reference present."""
import numpy as np

from fixtures import load, pairs_at

STATE = ["x", "y", "u", "v", "omega", "theta", "r", "h"]


def _check(name, steps, pair_steps):
    from oracle_py import OracleSim
    f, p, d = load(name)
    O = OracleSim(p, f)
    for t in range(1, max(steps | pair_steps) + 1):
        O.step()
        if t in pair_steps:
            li, lj, lf = O.log()
            assert sorted(zip(li.tolist(), lj.tolist())) == pairs_at(d, "cand", t)
            c = (lf & 1).astype(bool)
            assert sorted(zip(li[c].tolist(), lj[c].tolist())) == pairs_at(d, "cont", t)
        if t in steps:
            st = O.state()
            for k in STATE:
                assert np.array_equal(st[k], d[f"{k}_{t}"]), f"{name}: {k} at step {t}"


def test_config0_golden(oracle_lib):
    _check("config0_seed0.npz", {1, 100, 1000}, {1, 100, 1000})


def test_dense64_golden(oracle_lib):
    _check("dense64.npz", {1, 10, 50}, set(range(1, 51)))
