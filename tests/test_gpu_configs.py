"""The batched and served workloads at their real shapes.

* The batched ensemble: 512 independent fields of 500-8,000 floes
  (1.38 M floes), mixed precision, 60 batched steps.  Every member keeps
  each floe's r^2 h (ridging conserves it, ridging.f90:75-77); sampled
  members give the FP64 oracle's candidate and contact pair sets at step 1
  and end bitwise where they end when run alone.
* Served requests: ten requests of 1,000-50,000 floes from the config-1
  generator (power-law radii, Gaussian wind field), Poisson arrivals,
  admitted by footprint into a context too small to hold them all at once:
  each ends bitwise where it ends alone, and the largest request's first
  step gives the oracle's pair sets.
"""
import numpy as np
import pytest

from godar_amd import workloads as W

pytestmark = pytest.mark.gpu

STATE = ["x", "y", "u", "v", "omega", "r", "h", "theta"]


def _oracle_pairsets(f, p):
    from oracle_py import OracleSim
    O = OracleSim(p, f)
    O.step()
    li, lj, lf = O.log()
    cand = sorted(zip(li.tolist(), lj.tolist()))
    cont = sorted((a, b) for a, b, c in zip(li.tolist(), lj.tolist(), lf.tolist()) if c & 1)
    return cand, cont


def _pairsets(ctx, m):
    ci, cj = ctx.pairs(m, "candidate")
    ki, kj = ctx.pairs(m, "contact")
    return list(zip(ci.tolist(), cj.tolist())), list(zip(ki.tolist(), kj.tolist()))


def test_batched_ensemble_512_fields():
    from godar_amd.sim import Context
    members = W.config3()
    total = sum(f.n for f, _ in members)
    assert len(members) == 512 and total > 1_300_000
    print(f"[config 3] 512 fields, {total} floes", flush=True)
    ctx = Context(total, len(members), "mixed")
    ms = [ctx.add_field(p, f) for f, p in members]
    sample = [0, 101, 257, 511]
    ctx.step(1)
    for k in sample:
        f, p = members[k]
        assert _pairsets(ctx, ms[k]) == _oracle_pairsets(f, p), k
    ctx.step(59)
    tot = ctx.totals()
    print(f"[config 3] after 60 steps: {tot}", flush=True)
    for (f, p), m in zip(members, ms):
        a = ctx.floes(m)
        assert np.max(np.abs(a["r"] ** 2 * a["h"] / (f.r ** 2 * f.h) - 1)) < 1e-12, m
    for k in sample:
        f, p = members[k]
        alone = Context(f.n, 1, "mixed")
        ma = alone.add_field(p, f)
        alone.step(60)
        a, b = ctx.floes(ms[k]), alone.floes(ma)
        alone.close()
        for key in STATE:
            assert np.array_equal(a[key], b[key]), (k, key)
    ctx.close()


def test_served_requests_config1_generator():
    from godar_amd.serve import Request, Server
    from godar_amd.sim import Context
    reqs = W.config4_requests(10, seed=44)
    rng = np.random.default_rng(5)
    steps = [int(s) for s in rng.integers(20, 81, len(reqs))]
    sizes = [f.n for f, _, _, _ in reqs]
    assert min(sizes) >= 1000 and max(sizes) <= 50000
    assert all(f.wind_grid is not None for f, _, _, _ in reqs)
    print(f"[config 4] requests of {sizes} floes", flush=True)
    probe = Context(1, 1, "mixed")
    per = probe.footprint(1)
    probe.close()
    room = max(sizes) + sum(sorted(sizes)[:2])      # never all ten at once
    srv = Server(hbm_bytes=per * room, max_members=6, precision="mixed", chunk=16, keep_state=True)
    out = srv.run([Request(k, f, p, s, arr) for k, ((f, p, _, arr), s) in enumerate(zip(reqs, steps))])
    srv.close()
    assert out["requests"] == len(reqs) and out["peak_floes"] <= room
    print(f"[config 4] served: p50 {out['latency_p50_s']:.2f} s, p99 {out['latency_p99_s']:.2f} s, "
          f"peak {out['peak_members']} members", flush=True)
    for r in out["results"]:
        f, p, _, _ = reqs[r.rid]
        alone = Context(f.n, 1, "mixed")
        m = alone.add_field(p, f)
        alone.step(steps[r.rid])
        a = alone.floes(m)
        alone.close()
        for key in STATE:
            assert np.array_equal(r.state[key], a[key]), (r.rid, key)
    big = int(np.argmax(sizes))
    f, p, _, _ = reqs[big]
    ctx = Context(f.n, 1, "mixed")
    m = ctx.add_field(p, f)
    ctx.step(1)
    assert _pairsets(ctx, m) == _oracle_pairsets(f, p)
    ctx.close()
