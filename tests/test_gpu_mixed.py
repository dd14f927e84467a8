"""Mixed precision (GODAR_MIXED) against the FP64 oracle.

In mixed precision the contact, bond, sheltering and drag force laws run in
FP32 (contact_forces.f90:1-87, bonds.f90:74-212, forcing.f90:32-186), while
positions, distances and overlaps, the contact and bond strain histories,
damage, ridging volume bookkeeping and the integration stay FP64.  Bars:

* one step from identical inputs: candidate and contact pair sets exact
  (geometry is FP64); every force column within MIXED_FORCE_TOL of that
  column's largest magnitude; the sheltering minima within
  MIXED_SHELTER_TOL absolute; the velocity and position increments of the
  step within MIXED_FORCE_TOL of the largest increment;
* free-running lockstep over 40 steps: pair sets equal at every step and the
  velocity error within MIXED_VEL_TOL of the field's velocity scale (the
  FP32 error of one step, ~1e-6 relative, compounds through the contacts);
* 10^4 steps of the 1,000-floe field: the relative L2 errors of positions
  and of velocities stay within the FP64 reference's own one-ulp
  sensitivity envelope (the deviation that moving every initial centre by
  one ulp causes; MIXED_ENV_FACTOR = 1), and total ice volume is kept.

Measured on an MI355X: one step 4.4e-7 (worst column), sheltering 1.5e-7;
40 steps 1.8e-7 of the velocity scale; 10^4 steps position L2 5.2e-4
(envelope 6.3e-4), velocity L2 0.235 (envelope 0.311).
"""
import numpy as np
import pytest

from fixtures import load, load_long
from workloads import config0_params, kick, packed_box, squeeze

pytestmark = pytest.mark.gpu

MIXED_FORCE_TOL = 1e-5      # of the column's max |value|, one step
MIXED_SHELTER_TOL = 1e-6    # absolute (minima lie in [0, 1])
MIXED_VEL_TOL = 1e-5        # of max |velocity|, over 40 free-running steps
MIXED_ENV_FACTOR = 1.0      # 10^4 steps: L2 errors vs the one-ulp envelope

FORCE_COLS = ["fcx", "fcy", "mc", "fbx", "fby", "mb", "fax", "fay", "ma", "fwx", "fwy", "mw",
              "fcorx", "fcory", "fx_bc", "fy_bc", "m_bc", "tfx", "tfy", "m", "tac", "pc", "tab",
              "pb", "ta_bc", "p_bc", "tsigxx", "tsigyy", "tsigxy", "tsigyx"]


def _cases():
    f0, p0, _ = load("config0_seed0.npz")
    fd, pd, _ = load("dense64.npz")
    fs = kick(squeeze(packed_box(64, 5), 0.995), 2, speed=0.5)
    ps = config0_params(fs, sigmanc_crit=1e5, e_modul=6e9, eb=6e9, bond_lim=50.0, bond_frac=1.0)
    return [("config0", f0, p0), ("dense64", fd, pd), ("stiff_bonds", fs, ps)]


def _pairsets(ctx, m, O):
    li, lj, lf = O.log()
    cand = sorted(zip(li.tolist(), lj.tolist()))
    cont = sorted((a, b) for a, b, c in zip(li.tolist(), lj.tolist(), lf.tolist()) if c & 1)
    ci, cj = ctx.pairs(m, "candidate")
    ki, kj = ctx.pairs(m, "contact")
    return (list(zip(ci.tolist(), cj.tolist())) == cand, list(zip(ki.tolist(), kj.tolist())) == cont,
            len(cont))


def _colerr(a, b):
    scale = float(np.max(np.abs(b)))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b))) / scale


@pytest.mark.parametrize("case", ["config0", "dense64", "stiff_bonds"])
def test_mixed_one_step_against_fp64_oracle(case):
    from godar_amd.sim import Context
    from oracle_py import OracleSim
    name, f, p = next(c for c in _cases() if c[0] == case)
    ctx = Context(f.n, 1, "mixed", record_forces=True)
    m = ctx.add_field(p, f)
    O = OracleSim(p, f)
    ctx.step(1)
    O.step()
    cand_ok, cont_ok, ncont = _pairsets(ctx, m, O)
    assert cand_ok and cont_ok, name
    fa, fb = ctx.forces(m), O.forces()
    errs = {k: _colerr(fa[k], fb[k]) for k in FORCE_COLS}
    sh = max(float(np.max(np.abs(fa[k] - fb[k]))) for k in ("hsfa_min", "hsfw_min"))
    a, b = ctx.floes(m), O.state()
    inc = {}
    for k, k0 in (("u", f.u), ("v", f.v), ("omega", f.omega), ("x", f.x), ("y", f.y)):
        da, db = a[k] - k0, b[k] - k0
        inc[k] = _colerr(da, db)
    ctx.close()
    print(f"[mixed 1 step] {name}: contacts {ncont}, worst force column "
          f"{max(errs, key=errs.get)} {max(errs.values()):.2e}, shelter {sh:.2e}, "
          f"increments {max(inc.values()):.2e}", flush=True)
    for k, e in errs.items():
        assert e <= MIXED_FORCE_TOL, (name, k, e)
    assert sh <= MIXED_SHELTER_TOL, (name, sh)
    for k, e in inc.items():
        assert e <= MIXED_FORCE_TOL, (name, k, e)


@pytest.mark.parametrize("case", ["config0", "dense64", "stiff_bonds"])
def test_mixed_lockstep_against_fp64_oracle(case):
    from godar_amd.sim import Context
    from oracle_py import OracleSim
    name, f, p = next(c for c in _cases() if c[0] == case)
    ctx = Context(f.n, 1, "mixed", record_forces=True)
    m = ctx.add_field(p, f)
    O = OracleSim(p, f)
    worst_v = worst_x = 0.0
    for t in range(1, 41):
        ctx.step(1)
        O.step()
        cand_ok, cont_ok, _ = _pairsets(ctx, m, O)
        assert cand_ok and cont_ok, (name, t)
        a, b = ctx.floes(m), O.state()
        vs = float(np.max(np.hypot(b["u"], b["v"]))) + 1e-300
        worst_v = max(worst_v, float(np.max(np.hypot(a["u"] - b["u"], a["v"] - b["v"]))) / vs)
        worst_x = max(worst_x, float(np.max(np.abs(np.concatenate([a["x"] - b["x"], a["y"] - b["y"]])))))
        vol = np.sum(a["r"] ** 2 * a["h"]) / np.sum(b["r"] ** 2 * b["h"])
        assert abs(vol - 1) < 1e-12, (name, t, vol)
    ctx.close()
    print(f"[mixed 40 steps] {name}: velocity error {worst_v:.2e} of max speed, "
          f"position error {worst_x:.2e} m", flush=True)
    assert worst_v <= MIXED_VEL_TOL, (name, worst_v)


def test_mixed_1e4_steps_within_envelope():
    """10,000 mixed steps of the 1,000-floe field against the reference-made
    trajectory: position and velocity L2 errors within MIXED_ENV_FACTOR x the
    one-ulp envelope, median position error within the envelope's max."""
    from godar_amd.sim import Context
    f, p, _ = load("config0_seed0.npz")
    L = load_long()
    ctx = Context(f.n, 1, "mixed")
    m = ctx.add_field(p, f)
    ctx.step(10000)
    a = ctx.floes(m)
    ctx.close()
    xy = np.concatenate([a["x"], a["y"]])
    ref = np.concatenate([L["x_10000"], L["y_10000"]])
    uv = np.concatenate([a["u"], a["v"]])
    uref = np.concatenate([L["u_10000"], L["v_10000"]])
    pos_l2 = float(np.linalg.norm(xy - ref) / np.linalg.norm(ref))
    vel_l2 = float(np.linalg.norm(uv - uref) / np.linalg.norm(uref))
    med = float(np.median(np.abs(xy - ref) / np.abs(ref)))
    env_pos = max(float(L[f"ulp_env_pos_l2_10000_{k}"]) for k in (0, 1))
    env_vel = max(float(L[f"ulp_env_vel_l2_10000_{k}"]) for k in (0, 1))
    env_max = max(float(L[f"ulp_env_max_10000_{k}"]) for k in (0, 1))
    print(f"[mixed 1e4] position L2 {pos_l2:.2e} (envelope {env_pos:.2e}), velocity L2 "
          f"{vel_l2:.2e} (envelope {env_vel:.2e}), median {med:.2e}", flush=True)
    assert pos_l2 <= MIXED_ENV_FACTOR * env_pos
    assert vel_l2 <= MIXED_ENV_FACTOR * env_vel
    assert med <= env_max
    vol = np.sum(a["r"] ** 2 * a["h"]) / np.sum(L["r_10000"] ** 2 * L["h_10000"])
    assert abs(vol - 1) < 1e-9
