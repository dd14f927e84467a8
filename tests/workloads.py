"""Small seeded workloads shared by the tests (sizes the oracle finishes in seconds)."""
from __future__ import annotations

import numpy as np

from godar_amd import generators as g
from godar_amd.workloads import SOFT, params_for  # noqa: F401

config0_params = params_for


def packed_box(n=64, seed=3, box=2.0e4, r_lo=500.0, r_hi=1500.0):
    """A dense walled box: contacts, walls and ridging from the first steps."""
    return g.make_field(n, seed, radius="uniform", r_lo=r_lo, r_hi=r_hi, h_median=1.0,
                        h_sigma=0.3, packing=0.6, box=box, wind_speed=15.0, current_speed=0.2,
                        gap=0.5)


def squeeze(field, factor=0.997):
    """Shrink positions toward the box centre so neighbours overlap slightly."""
    c = field.box / 2
    field.x = c + (field.x - c) * factor
    field.y = c + (field.y - c) * factor
    return field


def kick(field, seed=0, speed=0.3, spin=2e-4):
    rng = np.random.default_rng(seed)
    field.u = rng.normal(0, speed, field.n)
    field.v = rng.normal(0, speed, field.n)
    field.omega = rng.normal(0, spin, field.n)
    return field


def clusters(rows=3, per=64, seed=5, box=2.0e4, gap=6.0e3):
    """rows x rows kicked, squeezed packed boxes side by side (gap metres
    apart) in one field: separate contact clusters, each ridging from the
    first steps."""
    parts = []
    for a in range(rows * rows):
        f = kick(squeeze(packed_box(per, seed + a, box=box)), seed + a)
        parts.append((f, (a % rows) * (box + gap), (a // rows) * (box + gap)))
    f0 = parts[0][0]
    cat = lambda k: np.concatenate([getattr(f, k) + (dx if k == "x" else dy if k == "y" else 0.0)
                                    for f, dx, dy in parts])
    out = g.FloeField(x=cat("x"), y=cat("y"), r=cat("r"), h=cat("h"), theta=cat("theta"),
                      omega=cat("omega"), u=cat("u"), v=cat("v"),
                      box=rows * box + (rows - 1) * gap, seed=seed, wind=f0.wind,
                      current=f0.current)
    return out
