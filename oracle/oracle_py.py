"""ctypes wrapper for oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only as the checker.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

FORCE_KEYS = ["tfx", "tfy", "m", "fcx", "fcy", "mc", "fbx", "fby", "mb", "fax", "fay",
              "ma", "fwx", "fwy", "mw", "fcorx", "fcory", "fx_bc", "fy_bc", "m_bc",
              "hsfa_min", "hsfw_min", "tp", "tsigxx", "tsigyy", "tsigxy", "tsigyx",
              "tac", "pc", "tab", "pb", "ta_bc", "p_bc"]
PAIR_KEYS = ["deltat", "thetarelc", "bond", "damage", "deltanb", "deltatb",
             "thetarelb_ji", "thetarelb_ij", "rb", "hb", "lb", "sb", "ib",
             "fcn", "fct", "fbn", "fbt", "mcc"]
STATE_KEYS = ["x", "y", "r", "h", "theta", "omega", "u", "v", "mass", "inertia", "hfa", "hfw"]


def build() -> str:
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
    return LIB_PATH


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.or_create.restype = ctypes.c_void_p
        L.or_create.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_destroy.argtypes = [ctypes.c_void_p]
        L.or_set_state.argtypes = [ctypes.c_void_p] + [_dp] * 8
        L.or_set_wind_grid.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, _dp, _dp]
        L.or_set_wind_frames.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.or_get_winds.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_get_state.argtypes = [ctypes.c_void_p] + [_dp] * 12
        L.or_get_forces.argtypes = [ctypes.c_void_p, _dp]
        L.or_get_walls.argtypes = [ctypes.c_void_p, _dp]
        L.or_get_log.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _ip]
        L.or_get_log.restype = ctypes.c_int
        L.or_get_bonds.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip]
        L.or_get_bonds.restype = ctypes.c_int
        L.or_get_pairs.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _dp]
        L.or_get_stats.argtypes = [ctypes.c_void_p, _ip]
        L.or_bond_threshold.argtypes = [ctypes.c_double]
        L.or_bond_threshold.restype = ctypes.c_uint32
        L.or_bond_keep.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint32]
        L.or_bond_keep.restype = ctypes.c_int
        L.or_fit.argtypes = [ctypes.c_double]
        L.or_fit.restype = ctypes.c_double
        _lib = L
    return _lib


class OracleSim:
    """One simulation on the CPU oracle (0-based indices)."""

    def __init__(self, params, field=None, n=None):
        self.L = lib()
        self.params = params
        self.n = int(field.n if field is not None else n)
        self.h = self.L.or_create(ctypes.byref(params), self.n)
        self.tstep = 0
        if field is not None:
            self.set_state(field.x, field.y, field.r, field.h, field.theta, field.omega,
                           field.u, field.v)
            if field.wind_grid is not None:
                self.set_wind_grid(field.wind_grid)

    def __del__(self):
        try:
            if self.h:
                self.L.or_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_state(self, x, y, r, h, theta, omega, u, v):
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, r, h, theta, omega, u, v)]
        self.L.or_set_state(self.h, *[_ptr(a) for a in self._keep])

    def set_wind_grid(self, wg):
        ua = np.ascontiguousarray(wg.ua, dtype=np.float64)
        va = np.ascontiguousarray(wg.va, dtype=np.float64)
        self.L.or_set_wind_grid(self.h, ua.shape[1], ua.shape[0], wg.x0, wg.y0, wg.dx, wg.dy,
                                _ptr(ua), _ptr(va))

    def set_wind_frames(self, ua, va, x0, y0, dx, dy, t0=0.0, frame_dt=1.0, disc_average=True):
        ua = np.ascontiguousarray(ua, dtype=np.float64)
        va = np.ascontiguousarray(va, dtype=np.float64)
        nf, nyn, nxn = ua.shape
        self._keep_wind = (ua, va)
        self.L.or_set_wind_frames(self.h, nxn, nyn, x0, y0, dx, dy, nf, t0, frame_dt,
                                  1 if disc_average else 0, _ptr(ua), _ptr(va))

    def winds(self):
        """(ua_i, va_i) of the last step."""
        a, b = np.zeros(self.n), np.zeros(self.n)
        self.L.or_get_winds(self.h, _ptr(a), _ptr(b))
        return a, b

    def step(self, k: int = 1):
        for _ in range(k):
            self.tstep += 1
            self.L.or_step(self.h, self.tstep)

    def state(self):
        outs = [np.zeros(self.n) for _ in range(12)]
        self.L.or_get_state(self.h, *[_ptr(a) for a in outs])
        return dict(zip(STATE_KEYS, outs))

    def forces(self):
        f = np.zeros((33, self.n))
        self.L.or_get_forces(self.h, _ptr(f))
        return dict(zip(FORCE_KEYS, f))

    def walls(self):
        w = np.zeros((4, self.n))
        self.L.or_get_walls(self.h, _ptr(w))
        return dict(zip(["deltat_bc1", "deltat_bc2", "theta_bc1", "theta_bc2"], w))

    def log(self):
        cap = 1 << 16
        while True:
            li = np.zeros(cap, dtype=np.int32)
            lj = np.zeros(cap, dtype=np.int32)
            lf = np.zeros(cap, dtype=np.int32)
            c = self.L.or_get_log(self.h, cap, _ptr(li, _ip), _ptr(lj, _ip), _ptr(lf, _ip))
            if c <= cap:
                return li[:c], lj[:c], lf[:c]
            cap = c

    def bonds(self):
        cap = 1 << 16
        while True:
            bi = np.zeros(cap, dtype=np.int32)
            bj = np.zeros(cap, dtype=np.int32)
            c = self.L.or_get_bonds(self.h, cap, _ptr(bi, _ip), _ptr(bj, _ip))
            if c <= cap:
                return bi[:c], bj[:c]
            cap = c

    def pairs(self, i, j):
        i = np.ascontiguousarray(i, dtype=np.int32)
        j = np.ascontiguousarray(j, dtype=np.int32)
        out = np.zeros((18, i.size))
        self.L.or_get_pairs(self.h, i.size, _ptr(i, _ip), _ptr(j, _ip), _ptr(out))
        return dict(zip(PAIR_KEYS, out))

    def stats(self):
        s = np.zeros(4, dtype=np.int32)
        self.L.or_get_stats(self.h, _ptr(s, _ip))
        return dict(ridge_pair=int(s[0]), ridge_wall=int(s[1]), ncand=int(s[2]),
                    ncontact=int(s[3]))
