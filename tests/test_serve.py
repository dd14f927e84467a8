"""Serving loop (godar_amd/serve.py): continuous batching with footprint
admission.

* host: the scheduling logic against a stand-in context -- admission in
  arrival order without head-of-line blocking, no member ever stepped past
  its last step, every request completed once;
* GPU: each request served in the mix (admitted while others run, into
  reused floe ranges and member slots) ends bitwise where it ends alone.
"""
import numpy as np
import pytest

from godar_amd import _lib
from godar_amd import generators as g
from godar_amd.serve import Request, Server, poisson_arrivals
from workloads import config0_params


class _VirtualClock:
    def __init__(self, tick=0.01):
        self.t = 0.0
        self.tick = tick

    def __call__(self):
        self.t += self.tick
        return self.t

    def sleep(self, dt):
        self.t += dt


class _FakeField:
    def __init__(self, n):
        self.n = n


class _FakeCtx:
    """Members with floe ranges in [0, cap); step advances live members."""

    def __init__(self, cap, slots):
        self.cap, self.slots = cap, slots
        self.mem = {}          # m -> [first, n, tstep, tlast]
        self.steps_of = {}     # m -> steps actually taken

    def add_field(self, params, field):
        free = [m for m in range(self.slots) if m not in self.mem]
        used = sorted((v[0], v[1]) for v in self.mem.values())
        pos, first = 0, None
        for a, n in used:
            if a - pos >= field.n:
                first = pos
                break
            pos = max(pos, a + n)
        if first is None and self.cap - pos >= field.n:
            first = pos
        if not free or first is None:
            raise _lib.GodarError(_lib.GODAR_ERR_NOMEM, "no room")
        m = free[0]
        self.mem[m] = [first, field.n, 1, 0]
        self.steps_of[m] = 0
        return m

    def run_until(self, m, tlast):
        self.mem[m][3] = tlast

    def tsteps(self):
        t = np.full(self.slots, -1, dtype=np.int32)
        for m, v in self.mem.items():
            t[m] = v[2]
        return t

    def step(self, k):
        for m, v in self.mem.items():
            for _ in range(k):
                if v[3] and v[2] > v[3]:
                    break
                v[2] += 1
                self.steps_of[m] += 1

    def floes(self, m):
        return {}

    def remove(self, m):
        assert self.steps_of[m] == self.mem[m][3], "member stepped past its last step"
        del self.mem[m]


def _fake_server(cap, slots, chunk):
    s = Server.__new__(Server)
    clk = _VirtualClock()
    s.ctx, s.max_floes, s.per_floe = _FakeCtx(cap, slots), cap, 1
    s.chunk, s.keep_state, s.clock, s.sleep = chunk, False, clk, clk.sleep
    return s


def test_scheduler_admission_and_stop_steps():
    rng = np.random.default_rng(5)
    arr = poisson_arrivals(40, 20.0, seed=5)
    assert np.all(np.diff(arr) > 0) and abs(arr[-1] / 40 - 1 / 20.0) < 0.03
    reqs = [Request(k, _FakeField(int(rng.integers(100, 900))), None,
                    int(rng.integers(5, 60)), float(arr[k])) for k in range(40)]
    reqs[3].field.n = 1900          # large request: must wait for room, others pass it
    srv = _fake_server(cap=2000, slots=4, chunk=7)
    out = srv.run(reqs)
    assert out["requests"] == 40
    res = {r.rid: r for r in out["results"]}
    assert sorted(res) == list(range(40))
    assert all(r.admitted >= r.arrival and r.finished > r.admitted for r in res.values())
    # requests that arrived after the large one were admitted before it
    assert any(res[k].admitted < res[3].admitted for k in range(4, 40))
    assert out["peak_floes"] <= 2000 and out["peak_members"] <= 4
    assert out["latency_p99_s"] >= out["latency_p50_s"] > 0


@pytest.mark.gpu
def test_served_equals_alone_bitwise():
    """Requests served together (footprint-limited, slots reused) end
    bitwise where each ends alone, mixed precision and FP64."""
    from godar_amd.sim import Context
    sizes, steps = [700, 400, 900, 500, 650, 300], [40, 25, 30, 45, 20, 35]
    fields = [g.config0_member(100 + k, n) for k, n in enumerate(sizes)]
    params = [config0_params(f, bond_seed=7 + k) for k, f in enumerate(fields)]
    for prec in ("mixed", "fp64"):
        reqs = [Request(k, f, p, s, 0.02 * k) for k, (f, p, s) in
                enumerate(zip(fields, params, steps))]
        probe = Context(1, 1, prec)
        per = probe.footprint(1)
        probe.close()
        clk = _VirtualClock(0.005)
        srv = Server(hbm_bytes=per * 1700, max_members=3, precision=prec, chunk=8,
                     keep_state=True, clock=clk, sleep=clk.sleep)
        print(f"[serve] {prec}: {len(reqs)} requests, room for {srv.max_floes} floes", flush=True)
        out = srv.run(reqs)
        srv.close()
        assert out["requests"] == len(reqs)
        assert out["peak_floes"] <= 1700
        for r in out["results"]:
            alone = Context(fields[r.rid].n, 1, prec)
            m = alone.add_field(params[r.rid], fields[r.rid])
            alone.step(steps[r.rid])
            a = alone.floes(m)
            alone.close()
            for key in ("x", "y", "u", "v", "omega", "r", "h", "theta"):
                assert np.array_equal(r.state[key], a[key]), (prec, r.rid, key)
