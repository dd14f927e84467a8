"""The correctly rounded exp, log, asin and acos of the FP64 force laws
(godar_amd/csrc/cr_math.h, built here for the CPU from the same source the
device compiles) against values computed to 60 significant digits with
Python's decimal module: every result must be the double nearest to the
exact value.  Beside it, how often glibc's libm (which the reference's
physics call) rounds differently: that rate, not a device libm's few
percent, is how often the FP64 path can still part from the reference."""
import ctypes
import math
import os
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
getcontext().prec = 60


@pytest.fixture(scope="module")
def crm(tmp_path_factory):
    out = tmp_path_factory.mktemp("crm") / "libcrm.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                           "-I", os.path.join(ROOT, "godar_amd", "csrc"), "-o", str(out),
                           os.path.join(ROOT, "tests", "cr_math_check.cpp")])
    L = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    fns = {}
    for name in ("exp", "log", "asin", "acos"):
        f = getattr(L, f"crm_{name}_n")
        f.argtypes = [dp, dp, ctypes.c_int]

        def call(x, f=f):
            x = np.ascontiguousarray(x, dtype=np.float64)
            y = np.empty_like(x)
            f(x.ctypes.data_as(dp), y.ctypes.data_as(dp), x.size)
            return y
        fns[name] = call
    return fns


def _sin(x: Decimal) -> Decimal:
    s = t = x
    k = 1
    while True:
        t = -t * x * x / ((2 * k) * (2 * k + 1))
        if abs(t) < Decimal(10) ** -62:
            return s + t
        s += t
        k += 1


def _cos(x: Decimal) -> Decimal:
    s = t = Decimal(1)
    k = 1
    while True:
        t = -t * x * x / ((2 * k - 1) * (2 * k))
        if abs(t) < Decimal(10) ** -62:
            return s + t
        s += t
        k += 1


def _asin(x: float) -> float:
    if x == 0.0:
        return x
    y = Decimal(math.asin(x))
    for _ in range(3):     # Newton on sin y = x, quadratic from a 1-ulp start
        y = y - (_sin(y) - Decimal(x)) / _cos(y)
    return float(y)


def _acos(x: float) -> float:
    if x == 1.0:
        return 0.0
    y = Decimal(math.acos(x))
    for _ in range(3):
        y = y + (_cos(y) - Decimal(x)) / _sin(y)
    return float(y)


EXACT = {
    "exp": lambda x: float(Decimal(x).exp()),
    "log": lambda x: float(Decimal(x).ln()),
    "asin": _asin,
    "acos": _acos,
}
GLIBC = {"exp": math.exp, "log": math.log, "asin": math.asin, "acos": math.acos}


def _inputs(name, rng):
    if name == "exp":
        return np.concatenate([
            rng.uniform(-745.0, 709.7, 1500),           # the whole finite range
            -rng.uniform(0.0, 60.0, 1500),              # sheltering: exp(-0.18 |dn/cos| / hf)
            rng.uniform(-1e-6, 1e-6, 300),
            [0.0, -0.0, 2.0 ** -53, -(2.0 ** -54), 1e-300, 0.5, -0.5, 1.0, 709.78, -708.39]])
    if name == "log":
        return np.concatenate([
            np.exp(rng.uniform(math.log(1e-300), math.log(1e300), 1500)),
            rng.uniform(100.0, 1.5e4, 1500),            # log profile: log(hf / z0w)
            1.0 + rng.uniform(-0.02, 0.02, 500),
            [1.0, 2.0, 0.5, 10.0, 1.0 + 2.0 ** -52, 1.0 - 2.0 ** -53, 5e-324, 1.7e308]])
    base = [0.0, 0.5, -0.5, 1.0, -1.0, 1e-300, 0.999999999999, -0.9999999999999999]
    return np.concatenate([rng.uniform(-1.0, 1.0, 600), 1.0 - 10.0 ** -rng.uniform(1, 15, 150),
                           -(1.0 - 10.0 ** -rng.uniform(1, 15, 150)), base])


@pytest.mark.parametrize("name", ["exp", "log", "asin", "acos"])
def test_correctly_rounded(crm, name):
    rng = np.random.default_rng({"exp": 1, "log": 2, "asin": 3, "acos": 4}[name])
    x = _inputs(name, rng)
    got = crm[name](x)
    bad = []
    glibc_differs = 0
    for xi, gi in zip(x.tolist(), got.tolist()):
        want = EXACT[name](xi)
        if gi != want and not (gi == 0.0 and want == 0.0):
            bad.append((xi, gi, want))
        glibc_differs += GLIBC[name](xi) != want
    print(f"[cr_math] {name}: {x.size} inputs, glibc rounds {glibc_differs} of them "
          f"differently ({glibc_differs / x.size:.2%})")
    assert not bad, bad[:5]
    assert glibc_differs / x.size < 0.01


def test_special_values(crm):
    inf, nan = float("inf"), float("nan")
    e = crm["exp"](np.array([inf, -inf, nan, 710.0, -746.0]))
    assert e[0] == inf and e[1] == 0.0 and math.isnan(e[2]) and e[3] == inf and e[4] == 0.0
    lg = crm["log"](np.array([0.0, -1.0, inf, nan]))
    assert lg[0] == -inf and math.isnan(lg[1]) and lg[2] == inf and math.isnan(lg[3])
    for nm in ("asin", "acos"):
        v = crm[nm](np.array([1.5, -1.0000001, nan]))
        assert all(math.isnan(t) for t in v)
    assert crm["acos"](np.array([-1.0]))[0] == math.pi
    assert crm["asin"](np.array([1.0]))[0] == math.pi / 2
