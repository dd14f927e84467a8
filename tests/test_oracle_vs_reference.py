"""Pin the C oracle against GODAR's own routines compiled from the reference.

oracle/_ref/<cfg>/libgodar_ref.so holds the reference's physics files,
compiled unmodified, plus the harness in oracle/ref_harness.f90.  Every
comparison here is bit-for-bit: same inputs, same step ordering, so the
oracle's restatement must reproduce every rounding of the compiled
reference (state, all force/diagnostic columns, pair sequence, pair state,
wall histories).
"""
import numpy as np
import pytest

import ref_harness as rh
from godar_amd import generators as g
from godar_amd.params import GodarParams
from workloads import config0_params, kick, packed_box, squeeze

pytestmark = []


def _need(cfg):
    if not rh.available(cfg):
        pytest.skip(f"oracle/_ref/{cfg} not built (needs /root/reference at build time)")


def _run_pair(cfg, field, params, nsteps, floe_wind=None, check_every=1):
    from oracle_py import OracleSim
    H = rh.RefHarness(cfg)
    assert H.n == field.n
    H.setup(params, field, bondmask=rh.dense_bondmask(field.n, params.bond_seed, params.bond_frac))
    O = OracleSim(params, field)
    seen = dict(contact=0, ridge=0, bonds=0)
    for t in range(1, nsteps + 1):
        if floe_wind is not None:
            st = H.state()
            ua, va = floe_wind(st["x"], st["y"])
            n = field.n
            H.set_floe_forcing(ua, va, np.full(n, params.uw), np.full(n, params.vw))
        H.step(t)
        O.step()
        if t % check_every and t != nsteps:
            continue
        a, b = H.state(), O.state()
        for k in a:
            assert np.array_equal(a[k], b[k]), f"state {k} differs at step {t}"
        fa, fb = H.forces(), O.forces()
        for k in fa:
            assert np.array_equal(fa[k], fb[k]), f"force column {k} differs at step {t}"
        li, lj, lc = H.log()
        oi, oj, of = O.log()
        assert np.array_equal(li, oi) and np.array_equal(lj, oj), f"pair sequence differs at {t}"
        assert np.array_equal(lc, (of & 1).astype(bool)), f"contact flags differ at {t}"
        seen["contact"] += int(lc.sum())
        seen["ridge"] += O.stats()["ridge_pair"] + O.stats()["ridge_wall"]
    bi, bj = H.bonds()
    obi, obj = O.bonds()
    assert np.array_equal(bi, obi) and np.array_equal(bj, obj)
    seen["bonds"] = int(bi.size)
    if bi.size:
        pa, pb = H.pairs(bi, bj), O.pairs(bi, bj)
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), f"bond pair state {k} differs"
    li, lj, lc = H.log()
    if lc.any():
        ci, cj = li[lc], lj[lc]
        pa, pb = H.pairs(ci, cj), O.pairs(ci, cj)
        for k in pa:
            assert np.array_equal(pa[k], pb[k]), f"contact pair state {k} differs"
    wa, wb = H.walls(), O.walls()
    for k in wa:
        assert np.array_equal(wa[k], wb[k]), f"wall history {k} differs"
    return seen


def test_defaults_match_get_default():
    _need("own")
    H = rh.RefHarness("own")
    prm, opt = H.defaults()
    p = GodarParams()
    vec, o = p.harness_vectors()
    assert np.array_equal(prm[:31], vec)
    assert np.array_equal(opt, o)
    # derived constants (par_get.f90:64-68)
    assert prm[31] == p.ec and prm[32] == p.gc and prm[33] == p.beta


def test_bond_mask_hash_matches_python(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(0)
    for frac in (0.0, 0.3, 0.5, 1.0):
        thr = L.or_bond_threshold(frac)
        assert thr == g.bond_threshold(frac)
        i = rng.integers(0, 5000, 400)
        j = i + rng.integers(1, 5000, 400)
        seed = int(rng.integers(0, 2 ** 63))
        py = g.bond_keep(i, j, seed, frac)
        c = np.array([L.or_bond_keep(int(a), int(b), seed, thr) for a, b in zip(i, j)], dtype=bool)
        assert np.array_equal(py, c)


def test_config0_bitexact_300_steps():
    """1,000 floes, 100 km box, wind 10 m/s, bonds on 50% of touching pairs."""
    _need("n1000")
    f = g.config0(0)
    p = config0_params(f)
    seen = _run_pair("n1000", f, p, 300, check_every=25)
    assert seen["bonds"] > 200


def test_dense_box_ridging_bitexact():
    """64 floes squeezed in a 20 km box: contacts, walls and heavy ridging."""
    _need("n64")
    f = kick(squeeze(packed_box(64, 3)), 1)
    p = config0_params(f, sigmanc_crit=1e5)
    seen = _run_pair("n64", f, p, 200)
    assert seen["contact"] > 1000 and seen["ridge"] > 1000


def test_dense_box_stiff_bond_breaking_bitexact():
    """Reference default stiffness: bonds break, fast collisions."""
    _need("n64")
    f = kick(squeeze(packed_box(64, 5), 0.995), 2, speed=0.5)
    p = config0_params(f, sigmanc_crit=1e5, e_modul=6e9, eb=6e9, bond_lim=50.0, bond_frac=1.0)
    seen = _run_pair("n64", f, p, 150)
    assert seen["contact"] > 100


def test_switches_off_bitexact():
    """shelter/ridging/cohesion/slipping off, including the zero-drag quirk of shelter off."""
    _need("n64")
    f = kick(squeeze(packed_box(64, 4)), 3)
    p = config0_params(f, shelter=0, ridging=0, slipping=0, cohesion=0)
    _run_pair("n64", f, p, 120, check_every=10)


def test_wind_grid_extension_bitexact():
    """Per-floe wind from a gridded field, sampled at the floe centres."""
    _need("n64")
    from oracle_py import OracleSim  # noqa: F401
    f = kick(squeeze(packed_box(64, 6)), 4)
    f.wind_grid = g.gaussian_wind_grid(f.box, 8.0, 3.0, 5e3, seed=9, spacing=1e3)
    p = config0_params(f, sigmanc_crit=1e5)
    _run_pair("n64", f, p, 120, floe_wind=f.wind_grid.sample, check_every=10)
