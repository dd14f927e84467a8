"""Spatio-temporal gridded wind averaged over the floe disc (the plan of the
reference's winds_currents hook, forcing.f90:1-10; README.md:133).

Oracle checks fixed by the mathematics (CPU):
* an affine field a + b x + c y per frame is reproduced exactly by bilinear
  interpolation, and the 7-point disc rule integrates it exactly: every
  floe takes the affine value at its centre, interpolated linearly in time
  between the frames around the step's start (t = (tstep - 1) dt), and
  held at the first and last frame outside them;
* a field x y inside one grid cell: the disc mean is cx cy (the first
  moments vanish by symmetry), where the centre value of a field with a
  curvature across cells would not be;
GPU (marked gpu): the HIP path against the oracle over 60 FP64 and mixed
steps of a colliding box with two-frame, disc-averaged wind.
"""
import numpy as np
import pytest

from godar_amd import generators as g
from godar_amd.params import GodarParams
from workloads import config0_params, kick, packed_box, squeeze


def _frames(nx, ny, dx, fns):
    xs = np.arange(nx) * dx
    ys = np.arange(ny) * dx
    X, Y = np.meshgrid(xs, ys)
    ua = np.stack([fu(X, Y) for fu, _ in fns])
    va = np.stack([fv(X, Y) for _, fv in fns])
    return ua, va


def _field():
    f = g.make_field(40, 11, radius="uniform", r_lo=300, r_hi=900, h_median=1.0, h_sigma=0.3,
                     packing=0.3, box=2.0e4, wind_speed=0.0, current_speed=0.0)
    return f


def test_affine_wind_is_exact_in_space_and_time(oracle_lib):
    from oracle_py import OracleSim
    f = _field()
    p = config0_params(f, dt=30.0, cohesion=0, ridging=0)
    fns = [(lambda X, Y: 2.0 + 1e-4 * X - 2e-4 * Y, lambda X, Y: -1.0 + 3e-4 * X + 1e-4 * Y),
           (lambda X, Y: 6.0 - 2e-4 * X + 1e-4 * Y, lambda X, Y: 4.0 + 0.0 * X - 3e-4 * Y)]
    ua, va = _frames(21, 21, 1000.0, fns)
    t0, fdt = 60.0, 300.0
    O = OracleSim(p, f)
    O.set_wind_frames(ua, va, 0.0, 0.0, 1000.0, 1000.0, t0=t0, frame_dt=fdt, disc_average=True)
    for tstep in range(1, 16):
        x, y = O.state()["x"], O.state()["y"]
        O.step()
        wu, wv = O.winds()
        t = (tstep - 1) * p.dt
        w = min(max((t - t0) / fdt, 0.0), 1.0)
        eu = (1 - w) * fns[0][0](x, y) + w * fns[1][0](x, y)
        ev = (1 - w) * fns[0][1](x, y) + w * fns[1][1](x, y)
        assert np.max(np.abs(wu - eu)) < 1e-12 * 10, tstep
        assert np.max(np.abs(wv - ev)) < 1e-12 * 10, tstep


def test_disc_mean_of_bilinear_field(oracle_lib):
    """x y (in km) inside one 50 km cell: the mean over a disc is the value at
    its centre; a field x^2 over the same disc has mean cx^2 + r^2/4 (so
    the rule integrates a curvature the centre value misses)."""
    from oracle_py import OracleSim
    f = _field()
    # every floe well inside the one cell [0, 50 km]^2 of a 2 x 2 node grid
    f.x = 5e3 + (f.x / f.box) * 1e4
    f.y = 5e3 + (f.y / f.box) * 1e4
    p = config0_params(f, dt=1.0, cohesion=0, ridging=0, nx=5e4, ny=5e4)
    xs = np.array([0.0, 5e4])
    X, Y = np.meshgrid(xs, xs)
    ua = (X * Y / 1e6)[None]
    va = np.zeros_like(ua)
    O = OracleSim(p, f)
    O.set_wind_frames(ua, va, 0.0, 0.0, 5e4, 5e4, disc_average=True)
    x, y = O.state()["x"], O.state()["y"]
    O.step()
    wu, _ = O.winds()
    np.testing.assert_allclose(wu, x * y / 1e6, rtol=1e-12)
    # the 7-point rule of degree 5 on x^2 (evaluated directly, the quadrature
    # points the oracle uses): mean over the disc of radius r is cx^2 + r^2/4
    r = 700.0
    rho = r * np.sqrt(2.0 / 3.0)
    ang = np.arange(6) * np.pi / 3
    cx = 1234.0
    m = 0.25 * cx ** 2 + 0.125 * np.sum((cx + rho * np.cos(ang)) ** 2)
    assert abs(m - (cx ** 2 + r * r / 4)) < 1e-9 * cx ** 2


@pytest.mark.gpu
def test_wind_frames_disc_average_gpu_matches_oracle():
    from godar_amd.sim import Context
    from oracle_py import OracleSim
    f = kick(squeeze(packed_box(64, 6)), 4)
    p = config0_params(f, sigmanc_crit=1e5)
    rng = np.random.default_rng(3)
    ua = 8.0 + 3.0 * rng.standard_normal((3, 6, 6))
    va = -2.0 + 3.0 * rng.standard_normal((3, 6, 6))
    args = (0.0, 0.0, f.box / 5, f.box / 5)
    for prec in ("fp64", "mixed"):
        ctx = Context(f.n, 1, prec, record_forces=True)
        m = ctx.add_field(p, f)
        ctx.set_wind_frames(m, ua, va, *args, t0=5.0, frame_dt=20.0, disc_average=True)
        O = OracleSim(p, f)
        O.set_wind_frames(ua, va, *args, t0=5.0, frame_dt=20.0, disc_average=True)
        worst = 0.0
        for t in range(1, 61):
            ctx.step(1)
            O.step()
            a, b = ctx.floes(m), O.state()
            for k in ("x", "y", "u", "v"):
                sc = float(np.max(np.abs(b[k]))) + 1e-300
                worst = max(worst, float(np.max(np.abs(a[k] - b[k]))) / sc)
        ctx.close()
        print(f"[wind frames] {prec}: worst relative deviation {worst:.3e}", flush=True)
        assert worst == 0.0 if prec == "fp64" else worst < 1e-5, (prec, worst)
