"""The oracle's step control flow, pinned by what geometry and arithmetic fix.

The reference step (stepper.f90:71-280) is Gauss-Seidel: for each floe i in
index order it visits its candidates j > i in ascending distance, a ridging
event reshapes the thinner floe in place (ridging.f90:36-99) before the next
pair is evaluated, and each per-floe accumulator is summed in visiting
order.  These tests build small configurations where each of those choices
changes an observable, so that a Jacobi restatement, an index-ordered visit
or another summation order fails:

* a chain 0-1-2 where pair (0,1) ridges floe 1, whose shrunken radius then
  leaves pair (1,2) apart -- evaluated with the start-of-step radius the pair
  would still touch;
* a thin floe whose nearer partner (the higher index) ridges it first, so its
  farther partner no longer touches it;
* a floe with four contacts whose forces, added in the reference's order
  (inner pairs by ascending outer index, then outer pairs by distance), give
  a sum that differs in the last bits from other orders;
* a pair that has separated but is still a candidate keeps adding its last
  contact force to the stress diagnostics (stepper.f90:238-271), with the
  stress trace fixed at -|F| r_i for both floes.
"""
import math
from fractions import Fraction
from itertools import permutations

import numpy as np

from godar_amd import generators as g
from godar_amd.params import GodarParams

BOX = 1e5
BASE = dict(nx=BOX, ny=BOX, rtree=2000.0, dt=1.0, cohesion=0, shelter=0)


def _field(xs, ys, rs, hs, us=None):
    n = len(xs)
    z = np.zeros(n)
    u = np.zeros(n) if us is None else np.asarray(us, float)
    return g.FloeField(np.asarray(xs, float), np.asarray(ys, float), np.asarray(rs, float),
                       np.asarray(hs, float), z.copy(), z.copy(), u, z.copy(), BOX, 0,
                       (0.0, 0.0), (0.0, 0.0))


def _lens_area(d, r1, r2):
    """Area of the intersection of two discs (textbook form with the kite
    area from the side lengths, not the chord construction the step uses)."""
    a1 = r1 * r1 * math.acos((d * d + r1 * r1 - r2 * r2) / (2 * d * r1))
    a2 = r2 * r2 * math.acos((d * d + r2 * r2 - r1 * r1) / (2 * d * r2))
    kite = 0.5 * math.sqrt((-d + r1 + r2) * (d + r1 - r2) * (d - r1 + r2) * (d + r1 + r2))
    return a1 + a2 - kite


def _reshaped_radius(d, r_thin, h_thin, r_other):
    """Radius of the thinner floe after one ridging event: the ridged volume
    lens area x h_thin / 2 (ridging.f90:70-77) thickens it at constant r^2 h."""
    vol = _lens_area(d, r_thin, r_other) * h_thin / 2
    dh = vol / (math.pi * r_thin * r_thin)
    return r_thin * math.sqrt(h_thin / (h_thin + dh)), h_thin + dh


def _flags(O):
    li, lj, lf = O.log()
    return {(int(a), int(b)): int(f) for a, b, f in zip(li, lj, lf)}


def test_ridging_reshapes_in_place_before_the_next_pair(oracle_lib):
    """Chain 0-1-2 on a line: floe 1 (thin) overlaps floe 0 by 200 m and
    floe 2 by 5 m.  Pair (0,1) comes first (i = 0) and ridges floe 1, which
    shrinks by about 9 m; pair (1,2) is then 4 m apart: a candidate, no
    contact.  A restatement that evaluated (1,2) with the start-of-step
    radius would log a contact there."""
    from oracle_py import OracleSim
    x0 = 5e4
    xs = [x0, x0 + 1800.0, x0 + 1800.0 + 1995.0]
    f = _field(xs, [5e4] * 3, [1000.0] * 3, [2.0, 0.5, 2.0])
    O = OracleSim(GodarParams(**BASE, sigmanc_crit=1.0), f)
    O.step()
    fl = _flags(O)
    assert fl[(0, 1)] & 1 and fl[(0, 1)] & 4, "pair (0,1) must touch and ridge"
    assert (1, 2) in fl, "pair (1,2) stays a candidate"
    assert not fl[(1, 2)] & 1, "pair (1,2) must see the reshaped floe 1"
    r1, h1 = _reshaped_radius(1800.0, 1000.0, 0.5, 1000.0)
    assert r1 + 1000.0 < 1995.0 < 2000.0
    st = O.state()
    assert abs(st["r"][1] / r1 - 1) < 1e-12 and abs(st["h"][1] / h1 - 1) < 1e-12
    assert st["r"][0] == 1000.0 and st["r"][2] == 1000.0
    # without floe 0 the same pair (then (0,1)) touches
    O2 = OracleSim(GodarParams(**BASE, sigmanc_crit=1.0), _field(xs[1:], [5e4] * 2, [1000.0] * 2, [0.5, 2.0]))
    O2.step()
    assert _flags(O2)[(0, 1)] & 1


def test_candidates_visited_nearest_first(oracle_lib):
    """Floe 0 (thin) has two partners: index 1 at 1995 m (5 m overlap) and
    index 2 at 1800 m (200 m overlap).  Nearest first, pair (0,2) ridges
    floe 0 and pair (0,1) no longer touches; visiting by index would make
    (0,1) a contact."""
    from oracle_py import OracleSim
    x0 = 5e4
    f = _field([x0, x0 - 1995.0, x0 + 1800.0], [5e4] * 3, [1000.0] * 3, [0.5, 2.0, 2.0])
    O = OracleSim(GodarParams(**BASE, sigmanc_crit=1.0), f)
    O.step()
    li, lj, lf = O.log()
    assert list(zip(li.tolist(), lj.tolist()))[:2] == [(0, 2), (0, 1)], "ascending distance"
    fl = _flags(O)
    assert fl[(0, 2)] & 5 == 5
    assert not fl[(0, 1)] & 1
    r0, _ = _reshaped_radius(1800.0, 1000.0, 0.5, 1000.0)
    assert abs(O.state()["r"][0] / r0 - 1) < 1e-12


def _star(seed=20):
    """Floe 2 at the centre, partners 0, 1 (lower index) and 3, 4 (higher)
    around it, each overlapping it by 0.5-60 m and clear of each other."""
    rng = np.random.default_rng(seed)
    c = (5e4, 5e4)
    angs = np.deg2rad(np.array([20, 110, 200, 290]) + rng.uniform(-15, 15, 4))
    ds = 2000.0 - rng.uniform(0.5, 60, 4)
    nb = [(c[0] + d * np.cos(a), c[1] + d * np.sin(a)) for a, d in zip(angs, ds)]
    hs = list(rng.uniform(0.5, 2.5, 5))
    pts = [nb[0], nb[1], c, nb[2], nb[3]]
    return pts, hs


def test_per_floe_summation_order(oracle_lib):
    """Floe 2's contact force is the floating-point sum of its four pair
    terms, first those where it is the inner index (pairs (0,2), (1,2)),
    then its own pairs in ascending distance.  Each term is the same as in a
    run of that pair alone (no tangential force, no ridging), so the sum can
    be formed here in any order: the oracle's equals the reference order to
    the last bit, and the index order, the all-by-distance order and 20 other
    orders give different last bits in this configuration."""
    from oracle_py import OracleSim
    p = GodarParams(**BASE, ridging=0)
    pts, hs = _star()
    O = OracleSim(p, _field([q[0] for q in pts], [q[1] for q in pts], [1000.0] * 5, hs))
    O.step()
    fo = O.forces()
    terms = {}
    for k in (0, 1, 3, 4):
        idx = [k, 2] if k < 2 else [2, k]
        O2 = OracleSim(p, _field([pts[q][0] for q in idx], [pts[q][1] for q in idx], [1000.0] * 2,
                                 [hs[q] for q in idx]))
        O2.step()
        f2 = O2.forces()
        at = idx.index(2)
        terms[k] = (float(f2["fcx"][at]), float(f2["fcy"][at]))
        assert _flags(O2)[(0, 1)] & 1

    def total(order, c):
        acc = 0.0
        for k in order:
            acc = acc + terms[k][c]
        return acc

    dist = {k: math.hypot(pts[k][0] - pts[2][0], pts[k][1] - pts[2][1]) for k in (0, 1, 3, 4)}
    ref = [0, 1] + sorted((3, 4), key=lambda k: dist[k])
    assert ref == [0, 1, 4, 3], "the configuration puts the nearer outer partner at the higher index"
    for c, key in ((0, "fcx"), (1, "fcy")):
        assert total(ref, c) == fo[key][2], key
    differ = lambda o: any(total(list(o), c) != total(ref, c) for c in (0, 1))
    assert differ([0, 1, 3, 4]), "index order"
    assert differ(sorted((0, 1, 3, 4), key=lambda k: dist[k])), "distance order over all partners"
    assert sum(differ(o) for o in permutations((0, 1, 3, 4)) if list(o) != ref) >= 20


def test_separated_pair_keeps_last_contact_stress(oracle_lib):
    """Two floes collide and part.  While they remain candidates, the stress
    diagnostics keep adding the pair's last contact force (stepper.f90:238-271
    read fcn/fct, which reset_contact does not clear): with no drag and no
    walls, the trace sigxx + sigyy of both floes is -|F_last| r_0, where
    r_0 is the lower-index floe's radius (the reference uses r(i) for both)."""
    from oracle_py import OracleSim
    p = GodarParams(**BASE, ridging=0, Cdair=0.0, Cdwater=0.0, Csair=0.0, Cswater=0.0, lat=0.0)
    f = _field([5e4, 5e4 + 2205.0], [5e4, 5e4 + 40.0], [1000.0, 1200.0], [1.0, 1.0],
               us=[0.5, -0.5])
    O = OracleSim(p, f)
    last, contact_steps, stale_steps = None, 0, 0
    for _ in range(200):
        O.step()
        fl = _flags(O)
        if (0, 1) not in fl:
            break
        ps = O.pairs([0], [1])
        if fl[(0, 1)] & 1:
            contact_steps += 1
            last = math.hypot(ps["fcn"][0], ps["fct"][0])
            continue
        if last is None:
            continue
        stale_steps += 1
        fo = O.forces()
        for k in (0, 1):
            tr = fo["tsigxx"][k] + fo["tsigyy"][k]
            assert abs(tr - (-last * 1000.0)) <= 1e-12 * last * 1000.0, (k, tr, last)
        # the history the next contact would start from is cleared (reset.f90:1-15)
        assert ps["deltat"][0] == 0.0 and ps["thetarelc"][0] == 0.0
    # a contact that ends gently ends at zero force; this one parts within a
    # step of 1 s while the force is still large, so the check has teeth
    assert contact_steps >= 1 and last > 1e8 and stale_steps > 10, (contact_steps, last, stale_steps)


def test_fit_function_rational(oracle_lib):
    """The oracle's fit (contact_forces.f90:247-265) against the rational
    function evaluated exactly: within 3 ulp everywhere, and its limits
    fit(0) = p3/q2, fit(inf) = p1."""
    L = oracle_lib.lib()
    p1, p2, p3 = Fraction(9117, 10000), Fraction(2722, 10000), Fraction(3324, 1000000)
    q1, q2 = Fraction(1524, 1000), Fraction(3159, 100000)
    for xi in (0.0, 1e-9, 1e-3, 0.037, 0.1, 0.5, 1.0, 2.5, 40.0, 1e3, 1e8):
        X = Fraction(xi)
        exact = float((p1 * X * X + p2 * X + p3) / (X * X + q1 * X + q2))
        got = L.or_fit(xi)
        assert abs(got - exact) <= 3 * math.ulp(exact), (xi, got, exact)
    assert abs(L.or_fit(0.0) - float(p3 / q2)) <= math.ulp(float(p3 / q2))
    assert abs(L.or_fit(1e20) - 0.9117) <= math.ulp(0.9117)
