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
from workloads import config0_params

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
