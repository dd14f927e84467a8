"""Load the committed golden fixtures (tests/golden/*.npz) as (field, params, data)."""
from __future__ import annotations

import os

import numpy as np

from godar_amd.generators import FloeField
from godar_amd.params import GodarParams

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name: str):
    d = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    f = FloeField(d["x0"].copy(), d["y0"].copy(), d["r0"].copy(), d["h0"].copy(),
                  d["theta0"].copy(), d["omega0"].copy(), d["u0"].copy(), d["v0"].copy(),
                  float(d["box"]), 0, (0.0, 0.0), (0.0, 0.0))
    p = GodarParams()
    for k, v in zip(d["pnames"].tolist(), d["pvalues"].tolist()):
        cur = getattr(p, k)
        setattr(p, k, type(cur)(v) if not isinstance(cur, float) else float(v))
    return f, p, d


def pairs_at(d, kind: str, t: int):
    a = d[f"{kind}_{t}"]
    return sorted(zip(a[0].tolist(), a[1].tolist()))


def load_long():
    return np.load(os.path.join(GOLDEN, "config0_1e4.npz"), allow_pickle=False)
