"""Physical and numerical parameters of one GODAR simulation.

Mirrors the reference's parameter COMMON blocks (``model/inc/CB_const.h``,
``CB_bond.h``, ``CB_options.h``, ``CB_forcings.h``) and the compile-time
box of ``model/inc/parameter.h``.  Defaults are ``get_default``
(``model/src/par_get.f90:1-111``); the derived constants ``ec``, ``gc`` and
``beta`` follow ``par_get.f90:263-266``.  ``GodarParams`` is laid out exactly
like ``struct godar_params`` in ``include/godar_gpu.h``.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

# par_get.f90:47 assigns the default-real literal 1.4 to a double: the value
# the reference runs with is the float32 rounding of 1.4.
LAT_DEFAULT = float(np.float32(1.4))

_DOUBLE_FIELDS = [
    ("dt", 1e-3), ("rtree", 1e3), ("nx", 5e3), ("ny", 5e3),
    ("rhoair", 1.3), ("rhoice", 9e2), ("rhowater", 1026.0),
    ("Cdair", 3e-1), ("Cdwater", 3e-1), ("Csair", 5e-4), ("Cswater", 2e-3),
    ("z0w", 3.0e-4), ("lat", LAT_DEFAULT),
    ("e_modul", 6e9), ("poiss_ratio", 33e-2), ("friction_coeff", 7e-1),
    ("rest_coeff", 88e-2), ("sigmanc_crit", 1e6),
    ("eb", 6e9), ("lambda_rb", 8e-1), ("lambda_lb", 1.0),
    ("sigmatb_crit", 1e5), ("sigmacb_crit", 1e6), ("tau_crit", 1e6),
    ("bond_lim", 1e-2), ("dmax", 9e-1), ("dtd", 1.0), ("dth", 1e5),
    ("ua", 0.0), ("va", 0.0), ("uw", 0.0), ("vw", 0.0),
    ("bond_frac", 1.0),
]
_INT_FIELDS = [
    ("ntree", 1), ("slipping", 1), ("cohesion", 1), ("ridging", 1), ("shelter", 1),
    ("restart", 0),
]


class GodarParams(ctypes.Structure):
    _fields_ = ([(name, ctypes.c_double) for name, _ in _DOUBLE_FIELDS]
                + [(name, ctypes.c_int32) for name, _ in _INT_FIELDS]
                + [("bond_seed", ctypes.c_uint64)])

    def __init__(self, **kw):
        super().__init__()
        for name, val in _DOUBLE_FIELDS + _INT_FIELDS:
            setattr(self, name, val)
        self.bond_seed = 0
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)

    # ---- derived constants (par_get.f90:64-68, recomputed at :263-266)
    @property
    def pi(self) -> float:
        return 4.0 * math.atan(1.0)

    @property
    def ec(self) -> float:
        return self.e_modul / (2.0 * (1.0 - self.poiss_ratio * self.poiss_ratio))

    @property
    def gc(self) -> float:
        return self.e_modul / (4.0 * (1.0 - self.poiss_ratio) * (2.0 + self.poiss_ratio))

    @property
    def beta(self) -> float:
        lr = math.log(self.rest_coeff)
        return 10.0 * lr / math.sqrt(lr * lr + self.pi * self.pi)

    def copy(self, **kw) -> "GodarParams":
        p = GodarParams()
        ctypes.memmove(ctypes.byref(p), ctypes.byref(self), ctypes.sizeof(self))
        for k, v in kw.items():
            setattr(p, k, v)
        return p

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

    # ---- vectors for the reference harness (oracle/ref_harness.f90 order)
    def harness_vectors(self):
        names = ["dt", "rtree", None, "rhoair", "rhoice", "rhowater", "Cdair",
                 "Cdwater", "Csair", "Cswater", "z0w", "lat", "e_modul",
                 "poiss_ratio", "friction_coeff", "rest_coeff", "sigmanc_crit",
                 "eb", "lambda_rb", "lambda_lb", "sigmatb_crit", "sigmacb_crit",
                 "tau_crit", "bond_lim", "dmax", "dtd", "dth", "ua", "va", "uw", "vw"]
        prm = np.array([float(self.ntree) if n is None else getattr(self, n) for n in names],
                       dtype=np.float64)
        opt = np.array([self.slipping, self.cohesion, self.ridging, self.shelter],
                       dtype=np.int32)
        return prm, opt


def workload_params(field, *, dt: float = 1.0, rtree=None, **kw) -> GodarParams:
    """Parameters for a generated floe field: its box, its uniform forcing and
    rtree = ceil(max r) (the reference's generator advice,
    tools/init/gridded_placement.py:273-275); other values from ``kw``."""
    if rtree is None:
        rtree = float(math.ceil(float(np.max(field.r))))
    p = GodarParams(dt=dt, rtree=rtree, nx=field.box, ny=field.box,
                    ua=field.wind[0], va=field.wind[1],
                    uw=field.current[0], vw=field.current[1])
    for k, v in kw.items():
        setattr(p, k, v)
    return p
