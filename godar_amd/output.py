"""Output files of a member in the reference's text layout, written from the
device state (no (n, n) arrays).

The reference writes them in `sea_ice_post` (diagnostics_outputs.f90:1-189)
on output steps (godar.f90:195: `modulo(tstep, int(comp)) == 0 .or.
tstep + restart == 1`), appending to `output/<name>.<expno>`:

* one line per output step with the n values of x, y, u, v, r, h, theta,
  omega, tfx, tfy, fcx, fcy, fbx, fby, m (file `mom`), mc, mb, tsigxx,
  tsigyy, tsigxy, tsigyx, tp (diagnostics_outputs.f90:149-156, 169-182);
* per bonded pair (j > i, visited i then j ascending): `step j i` in `bond`
  and `step j i damage` in `damage`, then an empty line in `bond`
  (:157-168); per contact pair of the step, `step j i angle` in `angle`,
  the angle of the line from i to j in degrees (:165-167); step is
  int(tstep / comp), indices 1-based.

Numbers are written with 17 significant digits (they read back as the same
doubles); the reference's list-directed layout differs only in spacing, and
its tools read both alike (whitespace-separated values, `np.loadtxt`).
The member must record forces (`Context(record_forces=True)`) for the force
and stress files.
"""
from __future__ import annotations

import math
import os

import numpy as np

# file name -> key of Context.floes / Context.forces
_FLOE_FILES = [("x", "x"), ("y", "y"), ("u", "u"), ("v", "v"), ("r", "r"), ("h", "h"),
               ("theta", "theta"), ("omega", "omega")]
_FORCE_FILES = [("tfx", "tfx"), ("tfy", "tfy"), ("fcx", "fcx"), ("fcy", "fcy"), ("fbx", "fbx"),
                ("fby", "fby"), ("mom", "m"), ("mc", "mc"), ("mb", "mb"), ("tsigxx", "tsigxx"),
                ("tsigyy", "tsigyy"), ("tsigxy", "tsigxy"), ("tsigyx", "tsigyx"), ("tp", "tp")]
PI = 4.0 * math.atan(1.0)          # par_get.f90:35


def is_output_step(tstep: int, comp: float, restart: int = 0) -> bool:
    """godar.f90:195."""
    return tstep % int(comp) == 0 or tstep + restart == 1


def _line(vals) -> str:
    return " ".join(f"{v:.17g}" for v in vals) + "\n"


def write_outputs(ctx, m: int, tstep: int, comp: float, outdir: str = "output",
                  expno: str = "0001") -> None:
    """Append member m's state after step `tstep` to the output files."""
    os.makedirs(outdir, exist_ok=True)
    step = int(tstep / comp)
    fl = ctx.floes(m)
    fo = ctx.forces(m)

    def path(name):
        return os.path.join(outdir, f"{name}.{expno.strip()}")

    for name, key in _FLOE_FILES:
        with open(path(name), "a") as fh:
            fh.write(_line(fl[key]))
    # bonded pairs, visited by column i then row j > i (the pair lists are
    # sorted by (i, j))
    bi, bj = ctx.pairs(m, "bond")
    st = ctx.bond_state(m)
    keep = st["bond"] == 1.0 if len(bi) else np.zeros(0, dtype=bool)
    with open(path("bond"), "a") as fb, open(path("damage"), "a") as fd:
        for i, j, dmg in zip(bi[keep], bj[keep], st["damage"][keep] if len(bi) else []):
            fb.write(f"{step} {j + 1} {i + 1}\n")
            fd.write(f"{step} {j + 1} {i + 1} {dmg:.17g}\n")
        fb.write("\n")
    ci, cj = ctx.pairs(m, "contact")
    x, y = fl["x"], fl["y"]
    with open(path("angle"), "a") as fa:
        for i, j in zip(ci, cj):
            ang = math.atan2(y[j] - y[i], x[j] - x[i]) * 180.0 / PI
            fa.write(f"{step} {j + 1} {i + 1} {ang:.17g}\n")
    for name, key in _FORCE_FILES:
        with open(path(name), "a") as fh:
            fh.write(_line(fo[key]))
