"""Serving loop: continuous batching of independent simulations on one GPU.

No counterpart in the reference (it runs one simulation per program run).
Requests (a floe field, its parameters and a number of steps) arrive over
time.  One context holds every admitted request as a member: they advance
together in the batched step, their state stays on the device, and a
request leaves as soon as its own steps are done (its slot and floe range are
reused by later requests).

Admission is by device footprint: the context is sized for an HBM budget
(``godar_member_footprint`` bytes per floe), and a request is admitted, in
arrival order, as soon as a contiguous floe range and a member slot are free.
A request that does not fit waits; later, smaller requests may be admitted
past it (no head-of-line blocking), and it is admitted once room frees up.

The loop advances all members ``chunk`` steps at a time (fewer when a member
would finish earlier); each member stops exactly at its own last step
(``godar_member_run_until``), so results do not depend on the chunking or on
the batch mates.
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque
from typing import Callable, Optional

import numpy as np

from . import _lib
from .sim import Context


@dataclasses.dataclass
class Request:
    rid: int
    field: object          # generators.FloeField
    params: object         # params.GodarParams
    steps: int
    arrival: float = 0.0   # seconds after the start of the run


@dataclasses.dataclass
class Result:
    rid: int
    n: int
    steps: int
    arrival: float
    admitted: float
    finished: float
    state: Optional[dict] = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class Server:
    """Continuous-batching request loop over one context.

    hbm_bytes: device memory the context may use for floe state (admission
    budget); max_members: concurrent requests; chunk: steps per batched
    launch; clock: seconds since an arbitrary origin (time.perf_counter by
    default; a test passes a virtual clock)."""

    def __init__(self, hbm_bytes: int, max_members: int = 256, precision: str = "mixed",
                 chunk: int = 25, keep_state: bool = False,
                 clock: Callable[[], float] = time.perf_counter,
                 sleep: Callable[[float], None] = time.sleep):
        probe = Context(1, 1, precision)
        per_floe = probe.footprint(1)
        probe.close()
        self.max_floes = max(1, int(hbm_bytes // per_floe))
        self.ctx = Context(self.max_floes, max_members, precision)
        self.per_floe = per_floe
        self.chunk = int(chunk)
        self.keep_state = keep_state
        self.clock = clock
        self.sleep = sleep

    def close(self):
        self.ctx.close()

    def _try_admit(self, req: Request) -> Optional[int]:
        try:
            m = self.ctx.add_field(req.params, req.field)
        except _lib.GodarError as e:
            if e.code == _lib.GODAR_ERR_NOMEM:
                return None
            raise
        self.ctx.run_until(m, req.steps)
        return m

    def run(self, requests: list[Request]) -> dict:
        """Serve `requests` (arrival times relative to the start); returns
        per-request results and the run's aggregate figures."""
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting: list[Request] = []
        active: dict[int, tuple[Request, float]] = {}
        results: list[Result] = []
        t0 = self.clock()
        busy = 0.0
        peak_members = peak_floes = 0
        while pending or waiting or active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            still = []
            for req in waiting:
                if req.field.n > self.max_floes:
                    raise ValueError(f"request {req.rid}: {req.field.n} floes exceed the context")
                m = self._try_admit(req)
                if m is None:
                    still.append(req)
                else:
                    active[m] = (req, self.clock() - t0)
            waiting = still
            if not active:
                if pending:
                    self.sleep(max(0.0, pending[0].arrival - (self.clock() - t0)))
                continue
            peak_members = max(peak_members, len(active))
            peak_floes = max(peak_floes, sum(r.field.n for r, _ in active.values()))
            ts = self.ctx.tsteps()
            left = min(r.steps + 1 - int(ts[m]) for m, (r, _) in active.items())
            k = max(1, min(self.chunk, left))
            b0 = self.clock()
            self.ctx.step(k)
            busy += self.clock() - b0
            ts = self.ctx.tsteps()
            for m in list(active):
                req, adm = active[m]
                if int(ts[m]) > req.steps:
                    st = self.ctx.floes(m) if self.keep_state else None
                    self.ctx.remove(m)
                    del active[m]
                    results.append(Result(req.rid, req.field.n, req.steps, req.arrival, adm,
                                          self.clock() - t0, st))
        makespan = self.clock() - t0
        lat = np.array([r.latency for r in results]) if results else np.zeros(1)
        work = float(sum(r.n * r.steps for r in results))
        return dict(
            results=sorted(results, key=lambda r: r.rid),
            requests=len(results),
            floe_steps=work,
            makespan_s=makespan,
            busy_s=busy,
            floe_steps_per_s=work / makespan if makespan > 0 else 0.0,
            latency_p50_s=float(np.percentile(lat, 50)),
            latency_p99_s=float(np.percentile(lat, 99)),
            peak_members=peak_members,
            peak_floes=peak_floes,
            capacity_floes=self.max_floes,
        )


def poisson_arrivals(count: int, rate: float, seed: int) -> np.ndarray:
    """Arrival times of a Poisson process of the given rate (per second)."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=count))
