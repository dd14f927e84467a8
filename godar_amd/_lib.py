"""ctypes binding of libgodar_amd.so (the C ABI in include/godar_gpu.h).

There is no fallback: if the library is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes
import os

from .params import GodarParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgodar_amd.so")

GODAR_OK = 0
GODAR_ERR_ARG = 1
GODAR_ERR_HIP = 2
GODAR_ERR_CAPACITY = 3
GODAR_ERR_NOMEM = 4
GODAR_ERR_STATE = 5
GODAR_FP64 = 0
GODAR_MIXED = 1
PAIRS_CANDIDATE = 0
PAIRS_CONTACT = 1
PAIRS_BOND = 2

EXPORTS = [
    "godar_params_default", "godar_create", "godar_destroy", "godar_last_error",
    "godar_member_add", "godar_member_remove", "godar_member_footprint",
    "godar_capacity_left", "godar_set_forcing", "godar_set_wind_grid", "godar_set_wind_frames",
    "godar_set_bonds",
    "godar_step", "godar_step_async", "godar_sync", "godar_member_tstep", "godar_reserve",
    "godar_member_run_until", "godar_member_tsteps",
    "godar_get_floes", "godar_record_forces", "godar_get_forces", "godar_get_walls",
    "godar_get_pairs", "godar_get_bond_state", "godar_get_stats", "godar_set_block", "godar_use_graphs", "godar_graphs_on",
    "godar_set_stage_block", "godar_get_stage_blocks", "godar_autotune",
    "godar_set_dense_search", "godar_set_repair", "godar_set_skin", "godar_set_search", "godar_get_search", "godar_checkpoint_size", "godar_checkpoint_save",
    "godar_member_restore",
    "godar_profile", "godar_get_profile", "godar_stage_name", "godar_get_totals",
    "godar_get_totals_noalloc",
]
NSTAGES = 12


class GodarError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"godar error {code}: {msg}")
        self.code = code


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_d = ctypes.c_double

_lib = None


def load(path: str = LIB_PATH):
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m godar_amd.build` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    sig = {
        "godar_params_default": (_i32, [ctypes.POINTER(GodarParams)]),
        "godar_create": (_i32, [_i64, _i32, _i32, ctypes.POINTER(_vp)]),
        "godar_destroy": (_i32, [_vp]),
        "godar_last_error": (ctypes.c_char_p, [_vp]),
        "godar_member_add": (_i32, [_vp, ctypes.POINTER(GodarParams), _i32] + [_dp] * 8 + [_i32, _ip]),
        "godar_member_remove": (_i32, [_vp, _i32]),
        "godar_member_footprint": (_i32, [_vp, _i32, _lp]),
        "godar_capacity_left": (_i32, [_vp, _lp, _ip]),
        "godar_set_forcing": (_i32, [_vp, _i32, _d, _d, _d, _d]),
        "godar_set_wind_grid": (_i32, [_vp, _i32, _i32, _i32, _d, _d, _d, _d, _dp, _dp]),
        "godar_set_wind_frames": (_i32, [_vp, _i32, _i32, _i32, _d, _d, _d, _d, _i32, _d, _d, _i32,
                                         _dp, _dp]),
        "godar_set_bonds": (_i32, [_vp, _i32, _i32, _ip, _ip, _dp]),
        "godar_step": (_i32, [_vp, _i32]),
        "godar_step_async": (_i32, [_vp, _i32]),
        "godar_sync": (_i32, [_vp]),
        "godar_member_tstep": (_i32, [_vp, _i32, _ip]),
        "godar_member_run_until": (_i32, [_vp, _i32, _i32]),
        "godar_member_tsteps": (_i32, [_vp, _ip]),
        "godar_reserve": (_i32, [_vp, _i32, _i32, _i32, _i32]),
        "godar_get_floes": (_i32, [_vp, _i32] + [_dp] * 12),
        "godar_record_forces": (_i32, [_vp, _i32]),
        "godar_get_forces": (_i32, [_vp, _i32, _dp]),
        "godar_get_walls": (_i32, [_vp, _i32, _dp]),
        "godar_get_pairs": (_i32, [_vp, _i32, _i32, _i32, _ip, _ip, _ip]),
        "godar_get_bond_state": (_i32, [_vp, _i32, _i32, _dp, _ip]),
        "godar_get_stats": (_i32, [_vp, _i32, _lp]),
        "godar_set_block": (_i32, [_vp, _i32]),
        "godar_use_graphs": (_i32, [_vp, _i32]),
        "godar_graphs_on": (_i32, [_vp, _ip]),
        "godar_set_stage_block": (_i32, [_vp, _i32, _i32]),
        "godar_get_stage_blocks": (_i32, [_vp, _ip]),
        "godar_autotune": (_i32, [_vp, _i32]),
        "godar_set_dense_search": (_i32, [_vp, _i32]),
        "godar_set_repair": (_i32, [_vp, _i32, _i32]),
        "godar_set_skin": (_i32, [_vp, _d]),
        "godar_set_search": (_i32, [_vp, _d, _i32, _i32]),
        "godar_get_search": (_i32, [_vp, _dp, _ip, _ip]),
        "godar_checkpoint_size": (_i32, [_vp, _i32, _lp]),
        "godar_checkpoint_save": (_i32, [_vp, _i32, ctypes.POINTER(GodarParams), _vp, _i64]),
        "godar_member_restore": (_i32, [_vp, _vp, _i64, _ip]),
        "godar_profile": (_i32, [_vp, _i32]),
        "godar_get_profile": (_i32, [_vp, _dp, _lp]),
        "godar_stage_name": (ctypes.c_char_p, [_i32]),
        "godar_get_totals": (_i32, [_vp, _lp]),
        "godar_get_totals_noalloc": (_i32, [_vp, _lp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(ctx, code):
    if code != GODAR_OK:
        msg = "?"
        try:
            if ctx:
                msg = load().godar_last_error(ctx).decode()
        except Exception:
            pass
        raise GodarError(code, msg)
