"""Python host side over the C ABI: contexts, members, stepping, read-back.

The reference's host interface for this path is the single call
``call stepper(tstep, restart)`` (``model/src/godar.f90:193``) over COMMON
blocks; here a :class:`Context` holds device state for one or many
independent floe fields ("members") and :meth:`Context.step` is the
replacement of ``stepper`` for all of them at once.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .params import GodarParams

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

STATE_KEYS = ["x", "y", "u", "v", "theta", "omega", "r", "h", "mass", "inertia", "hfa", "hfw"]
FORCE_KEYS = ["tfx", "tfy", "m", "fcx", "fcy", "mc", "fbx", "fby", "mb", "fax", "fay",
              "ma", "fwx", "fwy", "mw", "fcorx", "fcory", "fx_bc", "fy_bc", "m_bc",
              "hsfa_min", "hsfw_min", "tp", "tsigxx", "tsigyy", "tsigxy", "tsigyx",
              "tac", "pc", "tab", "pb", "ta_bc", "p_bc"]
BOND_KEYS = ["bond", "damage", "deltanb", "deltatb", "thetarelb_ji", "thetarelb_ij",
             "rb", "hb", "lb", "sb", "ib"]


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


def _c64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.shape != (n,):
        raise ValueError(f"expected shape ({n},), got {a.shape}")
    return a


class Context:
    """A device context: up to ``max_floes`` floes in up to ``max_members`` fields."""

    def __init__(self, max_floes: int, max_members: int = 1, precision: str = "fp64",
                 outer: int = 0, inner: int = 0, hist: int = 0, bonds: int = 0,
                 record_forces: bool = False):
        self.L = _lib.load()
        prec = {"fp64": _lib.GODAR_FP64, "mixed": _lib.GODAR_MIXED}[precision]
        self.precision = precision
        h = ctypes.c_void_p()
        code = self.L.godar_create(int(max_floes), int(max_members), prec, ctypes.byref(h))
        if code != _lib.GODAR_OK:
            raise _lib.GodarError(code, "godar_create failed")
        self.h = h
        self.max_floes = int(max_floes)
        self.max_members = int(max_members)
        self.sizes: dict[int, int] = {}
        if outer or inner or hist or bonds:
            self.reserve(outer, inner, hist, bonds)
        if record_forces:
            self.record_forces(True)

    def close(self):
        if getattr(self, "h", None):
            self.L.godar_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, code):
        _lib.check(self.h, code)

    # ------------------------------------------------------------ members
    def add(self, params: GodarParams, x, y, r, h, theta=None, omega=None, u=None, v=None,
            tstep0: int = 1) -> int:
        n = int(np.asarray(x).size)
        z = np.zeros(n)
        arrs = [_c64(a if a is not None else z, n)
                for a in (x, y, u, v, theta, omega, r, h)]
        m = ctypes.c_int32()
        self._chk(self.L.godar_member_add(self.h, ctypes.byref(params), n,
                                          *[_ptr(a) for a in arrs], int(tstep0), ctypes.byref(m)))
        self.sizes[m.value] = n
        return m.value

    def add_field(self, params: GodarParams, field, tstep0: int = 1) -> int:
        m = self.add(params, field.x, field.y, field.r, field.h, field.theta, field.omega,
                     field.u, field.v, tstep0=tstep0)
        if getattr(field, "wind_grid", None) is not None:
            self.set_wind_grid(m, field.wind_grid)
        return m

    def checkpoint(self, m: int, params: GodarParams) -> bytes:
        """Complete state of member m (floes, walls, contact and bond pair
        state, step number, parameters) as a byte block."""
        n = ctypes.c_int64()
        self._chk(self.L.godar_checkpoint_size(self.h, int(m), ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._chk(self.L.godar_checkpoint_save(self.h, int(m), ctypes.byref(params), buf, n.value))
        return buf.raw

    def restore(self, block: bytes) -> int:
        """Add a member from a checkpoint block; it continues exactly."""
        m = ctypes.c_int32()
        buf = ctypes.create_string_buffer(block, len(block))
        self._chk(self.L.godar_member_restore(self.h, buf, len(block), ctypes.byref(m)))
        n = int(np.frombuffer(block[12:16], dtype=np.int32)[0])   # floe count in the header
        self.sizes[m.value] = n
        return m.value

    def remove(self, m: int):
        self._chk(self.L.godar_member_remove(self.h, int(m)))
        self.sizes.pop(m, None)

    def footprint(self, n: int) -> int:
        b = ctypes.c_int64()
        self._chk(self.L.godar_member_footprint(self.h, int(n), ctypes.byref(b)))
        return b.value

    def capacity_left(self):
        f = ctypes.c_int64()
        m = ctypes.c_int32()
        self._chk(self.L.godar_capacity_left(self.h, ctypes.byref(f), ctypes.byref(m)))
        return f.value, m.value

    def set_forcing(self, m, ua, va, uw, vw):
        self._chk(self.L.godar_set_forcing(self.h, int(m), ua, va, uw, vw))

    def set_wind_grid(self, m, wg):
        ua = np.ascontiguousarray(wg.ua, dtype=np.float64)
        va = np.ascontiguousarray(wg.va, dtype=np.float64)
        self._chk(self.L.godar_set_wind_grid(self.h, int(m), ua.shape[1], ua.shape[0],
                                             wg.x0, wg.y0, wg.dx, wg.dy, _ptr(ua), _ptr(va)))

    def set_wind_frames(self, m, ua, va, x0, y0, dx, dy, t0=0.0, frame_dt=1.0,
                        disc_average=True):
        """Spatio-temporal wind: ua, va of shape (frames, nyn, nxn), frame f
        at model time t0 + f frame_dt, averaged over each floe's disc (or
        taken at its centre)."""
        ua = np.ascontiguousarray(ua, dtype=np.float64)
        va = np.ascontiguousarray(va, dtype=np.float64)
        if ua.ndim != 3 or ua.shape != va.shape:
            raise ValueError("ua and va must have the shape (frames, nyn, nxn)")
        nf, nyn, nxn = ua.shape
        self._chk(self.L.godar_set_wind_frames(self.h, int(m), nxn, nyn, float(x0), float(y0),
                                               float(dx), float(dy), nf, float(t0),
                                               float(frame_dt), 1 if disc_average else 0,
                                               _ptr(ua), _ptr(va)))

    def set_bonds(self, m, bi, bj, state):
        bi = np.ascontiguousarray(bi, dtype=np.int32)
        bj = np.ascontiguousarray(bj, dtype=np.int32)
        st = np.ascontiguousarray(state, dtype=np.float64).reshape(-1)
        self._chk(self.L.godar_set_bonds(self.h, int(m), bi.size, _ptr(bi, _ip), _ptr(bj, _ip),
                                         _ptr(st)))

    def reserve(self, outer=0, inner=0, hist=0, bonds=0):
        self._chk(self.L.godar_reserve(self.h, int(outer), int(inner), int(hist), int(bonds)))

    def record_forces(self, on=True):
        self._chk(self.L.godar_record_forces(self.h, 1 if on else 0))

    def set_block(self, threads: int):
        self._chk(self.L.godar_set_block(self.h, int(threads)))

    def set_dense_search(self, floes: int):
        """Members added afterwards examining >= floes per search use one wavefront per floe."""
        self._chk(self.L.godar_set_dense_search(self.h, int(floes)))

    def set_skin(self, frac: float = 0.1):
        """Superset skin of members added afterwards, x (max r + rtree).
        Results do not depend on it."""
        self._chk(self.L.godar_set_skin(self.h, float(frac)))

    def set_search(self, skin_frac: float = 0.1, thread_span: int = 1, dense_scan: int = 96):
        """Search shape of the current members (and later ones): superset
        skin, cells the thread search scans on each side, dense-search
        threshold.  The next step builds the supersets; results do not depend
        on any of them."""
        self._chk(self.L.godar_set_search(self.h, float(skin_frac), int(thread_span), int(dense_scan)))

    def search(self) -> dict:
        f = ctypes.c_double(0.0)
        sp = ctypes.c_int32(0)
        ds = ctypes.c_int32(0)
        self._chk(self.L.godar_get_search(self.h, ctypes.byref(f), ctypes.byref(sp), ctypes.byref(ds)))
        return dict(skin_frac=f.value, thread_span=sp.value, dense_scan=ds.value)

    def set_repair(self, lanes: int = 64, pending: int = 1024):
        """Ridging repair: events evaluated at once (1: one at a time) and
        pending events kept on chip.  Results do not depend on either."""
        self._chk(self.L.godar_set_repair(self.h, int(lanes), int(pending)))

    def set_stage_block(self, stage: int, threads: int):
        self._chk(self.L.godar_set_stage_block(self.h, int(stage), int(threads)))

    def stage_blocks(self) -> dict:
        t = np.zeros(_lib.NSTAGES, dtype=np.int32)
        self._chk(self.L.godar_get_stage_blocks(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return {self.L.godar_stage_name(k).decode(): int(t[k]) for k in range(_lib.NSTAGES)}

    def autotune(self, trial: int = 3) -> dict:
        """Pick each stage's block size (the bond and wall stages, which
        overlap the contact stage, take 64 threads), then the search shape
        (skin, thread span, dense threshold), then the launch mode (graph
        replay or one-by-one launches), by timing real steps (members
        advance).
        Returns the stage block sizes; `search()` holds the search shape."""
        self._chk(self.L.godar_autotune(self.h, int(trial)))
        return self.stage_blocks()

    def use_graphs(self, on: bool = True):
        """Replay a captured HIP graph per step (default) or launch kernels one by one."""
        self._chk(self.L.godar_use_graphs(self.h, int(bool(on))))

    def graphs_on(self) -> bool:
        """The current launch mode (godar_autotune may change it)."""
        v = ctypes.c_int32(0)
        self._chk(self.L.godar_graphs_on(self.h, ctypes.byref(v)))
        return bool(v.value)

    def profile(self, on: bool = True):
        self._chk(self.L.godar_profile(self.h, 1 if on else 0))

    def stage_times(self) -> dict:
        """{stage: (total ms, timed launches)} since the last profile(True)."""
        ms = np.zeros(_lib.NSTAGES)
        n = np.zeros(_lib.NSTAGES, dtype=np.int64)
        self._chk(self.L.godar_get_profile(self.h, _ptr(ms),
                                           n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return {self.L.godar_stage_name(k).decode(): (float(ms[k]), int(n[k]))
                for k in range(_lib.NSTAGES)}

    # ------------------------------------------------------------ stepping
    def step(self, nsteps: int = 1):
        self._chk(self.L.godar_step(self.h, int(nsteps)))

    def step_async(self, nsteps: int = 1):
        self._chk(self.L.godar_step_async(self.h, int(nsteps)))

    def sync(self):
        self._chk(self.L.godar_sync(self.h))

    def run_until(self, m: int, tlast: int):
        """Member m stops advancing after step tlast (0: no limit)."""
        self._chk(self.L.godar_member_run_until(self.h, int(m), int(tlast)))

    def tsteps(self) -> np.ndarray:
        """Next step number of every member slot (-1: free slot)."""
        t = np.zeros(self.max_members, dtype=np.int32)
        self._chk(self.L.godar_member_tsteps(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return t

    def tstep(self, m: int) -> int:
        t = ctypes.c_int32()
        self._chk(self.L.godar_member_tstep(self.h, int(m), ctypes.byref(t)))
        return t.value

    # ------------------------------------------------------------ read-back
    def floes(self, m: int) -> dict:
        n = self.sizes[m]
        outs = [np.zeros(n) for _ in STATE_KEYS]
        self._chk(self.L.godar_get_floes(self.h, int(m), *[_ptr(a) for a in outs]))
        return dict(zip(STATE_KEYS, outs))

    def forces(self, m: int) -> dict:
        n = self.sizes[m]
        f = np.zeros((33, n))
        self._chk(self.L.godar_get_forces(self.h, int(m), _ptr(f)))
        return dict(zip(FORCE_KEYS, f))

    def walls(self, m: int) -> dict:
        n = self.sizes[m]
        w = np.zeros((4, n))
        self._chk(self.L.godar_get_walls(self.h, int(m), _ptr(w)))
        return dict(zip(["deltat_bc1", "deltat_bc2", "theta_bc1", "theta_bc2"], w))

    def pairs(self, m: int, kind: str = "candidate"):
        k = {"candidate": _lib.PAIRS_CANDIDATE, "contact": _lib.PAIRS_CONTACT,
             "bond": _lib.PAIRS_BOND}[kind]
        cnt = ctypes.c_int32()
        self._chk(self.L.godar_get_pairs(self.h, int(m), k, 0, None, None, ctypes.byref(cnt)))
        pi = np.zeros(max(cnt.value, 1), dtype=np.int32)
        pj = np.zeros(max(cnt.value, 1), dtype=np.int32)
        self._chk(self.L.godar_get_pairs(self.h, int(m), k, pi.size, _ptr(pi, _ip), _ptr(pj, _ip),
                                         ctypes.byref(cnt)))
        return pi[:cnt.value], pj[:cnt.value]

    def bond_state(self, m: int) -> dict:
        cnt = ctypes.c_int32()
        self._chk(self.L.godar_get_bond_state(self.h, int(m), 0, None, ctypes.byref(cnt)))
        st = np.zeros((max(cnt.value, 1), 11))
        self._chk(self.L.godar_get_bond_state(self.h, int(m), cnt.value, _ptr(st), ctypes.byref(cnt)))
        st = st[:cnt.value]
        return dict(zip(BOND_KEYS, st.T))

    def totals(self) -> dict:
        """Candidate pairs, contacts (last step) and bonded pairs over all
        members, the floes the last step queued for contact, bond and wall
        work, and the ridging seeds and events the repair re-evaluated."""
        t = np.zeros(8, dtype=np.int64)
        self._chk(self.L.godar_get_totals(self.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(candidates=int(t[0]), contacts=int(t[1]), bonds=int(t[2]),
                    queued_contact=int(t[3]), queued_bond=int(t[4]), queued_wall=int(t[5]),
                    ridging_seeds=int(t[6]), repaired_events=int(t[7]))

    def stats(self, m: int) -> dict:
        s = np.zeros(4, dtype=np.int64)
        self._chk(self.L.godar_get_stats(self.h, int(m), s.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return dict(candidates=int(s[0]), contacts=int(s[1]), seeds=int(s[2]), repaired=int(s[3]))


class Simulation:
    """One floe field on its own context (the plain single-simulation path)."""

    def __init__(self, params: GodarParams, field, precision: str = "fp64",
                 record_forces: bool = False, headroom: float = 1.0, **reserve):
        self.ctx = Context(max(field.n, 1), 1, precision, record_forces=record_forces, **reserve)
        self.m = self.ctx.add_field(params, field)
        self.n = field.n

    def step(self, k: int = 1):
        self.ctx.step(k)

    def state(self):
        return self.ctx.floes(self.m)

    def forces(self):
        return self.ctx.forces(self.m)

    def pairs(self, kind="candidate"):
        return self.ctx.pairs(self.m, kind)

    def close(self):
        self.ctx.close()
