"""The benchmark workloads (configs 0-4) as (field(s), parameters).

Namelist values shared by every workload: the reference's defaults
(par_get.f90:1-111) except e_modul = eb = 1e8 Pa (the defaults, 6e9 Pa, make
the explicit contact/bond oscillator unstable at dt = 1 s: bonds break within
ten steps) and bond_lim = 10 m (the generators leave >= 1 m clearance between
floes, so "touching" means a gap of at most 10 m).
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import generators as g
from .params import GodarParams, workload_params

SOFT = dict(e_modul=1e8, eb=1e8, bond_lim=10.0)


def params_for(field, **kw) -> GodarParams:
    d = dict(SOFT, bond_seed=12345, bond_frac=0.5, dt=1.0)
    d.update(kw)
    return workload_params(field, **d)


def config0(seed: int = 0):
    """1,000 floes, 100 km box, r ~ U[500,2000], 50% of touching pairs bonded,
    wind 10 m/s, current 0.1 m/s, dt = 1 s (1,000 steps in the reference run)."""
    f = g.config0(seed)
    return f, params_for(f)


def config1(seed: int = 1, n: int = 262144):
    """262,144 floes, power-law radii on [200, 5000] m, packing 0.7 (754 km box),
    30% bonded, Gaussian wind field, dt = 0.5 s (2,000 steps)."""
    f = g.config1(seed, n=n)
    return f, params_for(f, dt=0.5, bond_frac=0.3, bond_seed=seed * 7919 + 1)


def _member(args):
    s, n = args
    f = g.config0_member(s, n)
    return f, params_for(f, bond_seed=s)


def config3(seed: int = 3, members: int = 512, lo: int = 500, hi: int = 8000, workers: int = 0):
    """512 independent fields, floe counts log-uniform in [500, 8000], each the
    config-0 generator with its own seed and wind direction (box scaled to keep
    the packing); mixed precision, 5,000 steps.  Members are generated in
    parallel processes (deterministic: each depends on its own seed only)."""
    sizes = g.member_sizes(members, lo, hi, seed)
    jobs = [(seed * 100003 + k, int(n)) for k, n in enumerate(sizes)]
    if workers <= 0:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and len(jobs) > 8:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            return pool.map(_member, jobs, chunksize=4)
    return [_member(j) for j in jobs]


def config4_request(rng: np.random.Generator, seed: int):
    """One served request: 1,000-50,000 floes (log-uniform) from the config-1
    generator (power-law radii, packing 0.7, Gaussian wind field), 500-5,000
    steps, dt = 1 s."""
    n = int(math.floor(math.exp(rng.uniform(math.log(1e3), math.log(5e4)))))
    steps = int(rng.integers(500, 5001))
    f, p = config4_field((n, seed))
    return f, p, steps


def config4_field(args):
    """Field and parameters of a served request of n floes (seeded): the
    config-1 generator, Gaussian wind field included."""
    n, seed = args
    f = g.config1(seed, n=n)
    return f, params_for(f, bond_frac=0.3, bond_seed=seed)


def config4_requests(count: int, seed: int = 4, rate: float = 20.0, workers: int = 0):
    """`count` served requests: Poisson arrivals at `rate` per second, sizes
    and step counts as in config4_request, fields generated in parallel
    processes (each depends on its own seed only).  Returns
    [(field, params, steps, arrival)]."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        n = int(math.floor(math.exp(rng.uniform(math.log(1e3), math.log(5e4)))))
        steps = int(rng.integers(500, 5001))
        specs.append((n, steps, seed * 1000003 + k))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    jobs = [(n, s) for n, _, s in specs]
    if workers <= 0:
        workers = min(16, os.cpu_count() or 1)
    if workers > 1 and count > 4:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(workers) as pool:
            fps = pool.map(config4_field, jobs, chunksize=1)
    else:
        fps = [config4_field(j) for j in jobs]
    return [(f, p, st, float(a)) for (f, p), (_, st, _), a in zip(fps, specs, arrivals)]


def algorithmic_bytes(n_floes: int, ncand: int, ncont: int, nbond: int) -> float:
    """Bytes per floe-step a step must move at minimum (FP64):
    144 per floe (read 12 state words, write 6) + 8 per candidate pair
    + 32 per contact (history read + write) + 120 per bond (5 evolving words
    read + write, 5 constants read)."""
    return 144.0 * n_floes + 8.0 * ncand + 32.0 * ncont + 120.0 * nbond
