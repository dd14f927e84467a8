"""The Fortran side of the boundary (godar_amd/fortran).

* host: the iso_c_binding module sees struct godar_params exactly as the C
  ABI and Python do (godar_params_default through the Fortran interface);
* GPU: stepper_gpu, the drop-in for stepper (stepper.f90:1), driven by a
  stand-in of the reference's time loop against the reference's COMMON
  blocks, gives bitwise the same floe state and bond set as the same run
  through the C ABI from Python.

The binaries are built by __graft_entry__.build() (make -C godar_amd/fortran
all demo).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from godar_amd import generators as g
from godar_amd.params import _DOUBLE_FIELDS, _INT_FIELDS, GodarParams
from workloads import config0_params, kick

HERE = os.path.dirname(os.path.abspath(__file__))
FDIR = os.path.join(os.path.dirname(HERE), "godar_amd", "fortran", "_build")
ABI = os.path.join(FDIR, "abi_check")
DEMO = os.path.join(FDIR, "n1000", "demo_driver")


def test_fortran_params_layout_matches_c():
    assert os.path.exists(ABI), "Fortran binding not built (run __graft_entry__.build())"
    out = subprocess.run([ABI], check=True, capture_output=True, text=True).stdout.split("\n")
    rows = {l.split()[0]: l.split()[1:] for l in out if l.strip()}
    assert rows["rc"] == ["0"]
    assert int(rows["sizeof"][0]) == ctypes.sizeof(GodarParams)
    p = GodarParams()
    doubles = [float(v) for v in rows["doubles"]]
    assert doubles == [getattr(p, k) for k, _ in _DOUBLE_FIELDS]
    ints = [int(v) for v in rows["ints"]]
    assert ints == [getattr(p, k) for k, _ in _INT_FIELDS] + [p.bond_seed]


# order of the 33 doubles of the demo input (demo_driver.f90 header)
_DEMO_DOUBLES = ["dt", "rtree", "rhoair", "rhoice", "rhowater", "Cdair", "Cdwater", "Csair",
                 "Cswater", "z0w", "lat", "e_modul", "poiss_ratio", "friction_coeff",
                 "rest_coeff", "sigmanc_crit", "eb", "lambda_rb", "lambda_lb", "sigmatb_crit",
                 "sigmacb_crit", "tau_crit", "bond_lim", "dmax", "dtd", "dth", "ua", "va", "uw",
                 "vw"]


@pytest.mark.gpu
def test_stepper_gpu_matches_c_abi(tmp_path):
    """stepper_gpu (Fortran COMMON blocks) == C ABI from Python, bitwise."""
    from godar_amd.sim import Context
    assert os.path.exists(DEMO), "Fortran demo driver not built (run __graft_entry__.build())"
    field = g.config0(0)
    assert field.n == 1000
    # the reference's own bonding rule (every pair within bond_lim bonds)
    p = config0_params(field, bond_frac=1.0)
    assert p.nx == 1e5 and p.ny == 1e5
    nt = 40
    for prec_name, prec in (("fp64", 0), ("mixed", 1)):
        fin, fout = tmp_path / f"in_{prec_name}.bin", tmp_path / f"out_{prec_name}.bin"
        with open(fin, "wb") as fh:
            fh.write(np.array([nt, 0, nt, prec], dtype=np.int32).tobytes())
            pv = [getattr(p, k) for k in _DEMO_DOUBLES] + [float(p.ntree), 0.0, 0.0]
            fh.write(np.array(pv, dtype=np.float64).tobytes())
            fh.write(np.array([p.slipping, p.cohesion, p.ridging, p.shelter], dtype=np.int32).tobytes())
            for a in (field.x, field.y, field.u, field.v, field.theta, field.omega, field.r, field.h):
                fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())
        print(f"[fortran demo] {prec_name} n=1000 steps={nt}", flush=True)
        subprocess.run([DEMO, str(fin), str(fout)], check=True, timeout=120)
        raw = open(fout, "rb").read()
        n = field.n
        arr = np.frombuffer(raw[:13 * n * 8], dtype=np.float64).reshape(13, n)
        nb = int(np.frombuffer(raw[13 * n * 8:13 * n * 8 + 4], dtype=np.int32)[0])
        pairs = np.frombuffer(raw[13 * n * 8 + 4:], dtype=np.int32).reshape(nb, 2)

        ctx = Context(n, 1, prec_name, record_forces=True)
        m = ctx.add_field(p, field)
        ctx.step(nt)
        st, fo = ctx.floes(m), ctx.forces(m)
        bi, bj = ctx.pairs(m, "bond")
        ctx.close()
        for k, key in enumerate(["x", "y", "u", "v", "theta", "omega", "r", "h", "mass", "inertia"]):
            assert np.array_equal(arr[k], st[key]), (prec_name, key)
        for k, key in zip((10, 11, 12), ("tfx", "tfy", "m")):
            assert np.array_equal(arr[k], fo[key]), (prec_name, key)
        assert nb == len(bi) and nb > 0
        assert np.array_equal(pairs[:, 0] - 1, bi) and np.array_equal(pairs[:, 1] - 1, bj)


def _demo_input(path, p, field, nt, comp, prec):
    with open(path, "wb") as fh:
        fh.write(np.array([nt, 0, comp, prec], dtype=np.int32).tobytes())
        pv = [getattr(p, k) for k in _DEMO_DOUBLES] + [float(p.ntree), 0.0, 0.0]
        fh.write(np.array(pv, dtype=np.float64).tobytes())
        fh.write(np.array([p.slipping, p.cohesion, p.ridging, p.shelter], dtype=np.int32).tobytes())
        for a in (field.x, field.y, field.u, field.v, field.theta, field.omega, field.r, field.h):
            fh.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def _tokens(path):
    with open(path) as fh:
        return fh.read().split()


@pytest.mark.gpu
def test_output_files_match_reference_writer(tmp_path):
    """godar_amd.output writes, from the device state, the same output files
    as the reference's own sea_ice_post (diagnostics_outputs.f90, compiled
    from the reference sources into the demo driver) writes from the COMMON
    blocks that stepper_gpu refreshes: per-floe state, forces, moments,
    stresses and pressure on every output step, bonded pairs with damage,
    and contact angles.  Values are compared as parsed numbers (the two
    writers differ in spacing and digits printed)."""
    from godar_amd.output import is_output_step, write_outputs
    from godar_amd.sim import Context
    assert os.path.exists(DEMO), "Fortran demo driver not built (run __graft_entry__.build())"
    field = kick(g.config0(0), 7, speed=2.0)      # contacts within the run
    p = config0_params(field, bond_frac=1.0)
    nt, comp, n = 30, 10, field.n
    ref_dir, our_dir = tmp_path / "ref", tmp_path / "ours"
    (ref_dir / "output").mkdir(parents=True)
    _demo_input(tmp_path / "in.bin", p, field, nt, comp, 0)
    print(f"[fortran demo + sea_ice_post] n={n} steps={nt} comp={comp}", flush=True)
    subprocess.run([DEMO, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "T001"], check=True,
                   timeout=120, cwd=ref_dir)
    ctx = Context(n, 1, "fp64", record_forces=True)
    m = ctx.add_field(p, field)
    steps = 0
    for t in range(1, nt + 1):
        ctx.step(1)
        if is_output_step(t, comp):
            write_outputs(ctx, m, t, comp, str(our_dir / "output"), "T001")
            steps += 1
    ctx.close()
    names = ["x", "y", "u", "v", "r", "h", "theta", "omega", "tfx", "tfy", "fcx", "fcy", "fbx",
             "fby", "mom", "mc", "mb", "tsigxx", "tsigyy", "tsigxy", "tsigyx", "tp"]
    for name in names:
        a = np.array(_tokens(ref_dir / "output" / f"{name}.T001"), dtype=np.float64).reshape(steps, n)
        b = np.array(_tokens(our_dir / "output" / f"{name}.T001"), dtype=np.float64).reshape(steps, n)
        assert np.allclose(a, b, rtol=1e-15, atol=0.0), name
    for name, width in (("bond", 3), ("damage", 4), ("angle", 4)):
        a = np.array(_tokens(ref_dir / "output" / f"{name}.T001"), dtype=np.float64).reshape(-1, width)
        b = np.array(_tokens(our_dir / "output" / f"{name}.T001"), dtype=np.float64).reshape(-1, width)
        assert a.shape == b.shape and len(a) > 0, (name, a.shape, b.shape)
        assert np.array_equal(a[:, :3], b[:, :3]), name
        if width == 4:
            assert np.allclose(a[:, 3], b[:, 3], rtol=1e-14, atol=1e-12), name
    # one empty line per output step closes each step's bond list
    with open(ref_dir / "output" / "bond.T001") as fa, open(our_dir / "output" / "bond.T001") as fb:
        assert sum(1 for l in fa if not l.strip()) == sum(1 for l in fb if not l.strip()) == steps
