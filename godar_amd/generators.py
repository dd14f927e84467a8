"""Seeded floe-field generators for the benchmark workloads.

GODAR reads its initial state from one-value-per-line text files
(``files/{x,y,r,h,theta,omega}.dat``, read at ``model/src/ini_get.f90:146-163``)
that its own Python tools produce (``tools/init/gridded_placement.py:98-313``,
which draws radii/thicknesses from a distribution and jitters centres on a
grid).  The workloads here need the same kind of field at sizes the reference
tools do not reach, with no initial overlap (an initial overlap of a few
percent of ``r`` makes the first step explode), so this module:

* draws radii and thicknesses from the requested distributions,
* puts the centres on a jittered grid inside the walled box,
* removes overlaps with a deterministic push-apart relaxation,
* optionally writes the reference's text input files.

Everything is a pure function of the seed.  Two workload extensions that the
reference lacks are generated here as well:

* a bond mask: the reference bonds *every* pair with ``deltan >= -bond_lim`` at
  step 1 (``model/src/stepper.f90:93-101``); the workloads bond a seeded
  fraction of them.  ``bond_keep`` is the integer hash both the HIP path and
  the oracle evaluate (fraction 1 reproduces the reference rule exactly);
* a gridded wind field (Gaussian random field) sampled bilinearly at floe
  centres, replacing the uniform ``winds_currents`` (``model/src/forcing.f90:1-29``).
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

MASK64 = (1 << 64) - 1


# --------------------------------------------------------------------------
# bond mask hash (shared bit-for-bit with the HIP kernels and the oracle)
# --------------------------------------------------------------------------
def splitmix64(z):
    """splitmix64 finaliser on numpy uint64 arrays (wrapping arithmetic)."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def bond_threshold(frac: float) -> int:
    """24-bit acceptance threshold for ``bond_keep`` (frac=1 keeps all)."""
    if frac >= 1.0:
        return 1 << 24
    if frac <= 0.0:
        return 0
    return int(math.floor(frac * (1 << 24)))


def bond_keep(i, j, seed: int, frac: float):
    """True where the 0-based pair (i<j) may bond at step 1.

    key = splitmix64(seed ^ (i << 32 | j)); keep iff (key >> 40) < thr24.
    """
    i = np.asarray(i, dtype=np.uint64)
    j = np.asarray(j, dtype=np.uint64)
    key = splitmix64(np.uint64(seed & MASK64) ^ ((i << np.uint64(32)) | j))
    return (key >> np.uint64(40)) < np.uint64(bond_threshold(frac))


# --------------------------------------------------------------------------
# floe fields
# --------------------------------------------------------------------------
@dataclasses.dataclass
class FloeField:
    x: np.ndarray
    y: np.ndarray
    r: np.ndarray
    h: np.ndarray
    theta: np.ndarray
    omega: np.ndarray
    u: np.ndarray
    v: np.ndarray
    box: float               # walls at 0 and box on both axes [m]
    seed: int
    wind: tuple              # (ua, va) uniform wind [m/s]
    current: tuple           # (uw, vw) uniform current [m/s]
    meta: dict = dataclasses.field(default_factory=dict)
    wind_grid: Optional["WindGrid"] = None

    @property
    def n(self) -> int:
        return int(self.x.size)

    @property
    def packing(self) -> float:
        return float(np.sum(np.pi * self.r ** 2) / self.box ** 2)

    def write_reference_inputs(self, folder: str, suffix: str = "") -> None:
        """Write ``{x,y,r,h,theta,omega}<suffix>.dat`` as ini_get.f90:146-163 reads them."""
        os.makedirs(folder, exist_ok=True)
        for name in ("x", "y", "r", "h", "theta", "omega"):
            arr = getattr(self, name)
            with open(os.path.join(folder, f"{name}{suffix}.dat"), "w") as f:
                f.write("\n".join(repr(float(a)) for a in arr))
                f.write("\n")


@dataclasses.dataclass
class WindGrid:
    """Regular cell-centred-free grid: node (a,b) sits at (x0 + a*dx, y0 + b*dy)."""
    ua: np.ndarray   # (ny_nodes, nx_nodes) float64
    va: np.ndarray
    x0: float
    y0: float
    dx: float
    dy: float

    def sample(self, x, y):
        """Bilinear sample; the HIP kernels and the oracle evaluate the same expression."""
        return sample_grid(self.ua, self.va, self.x0, self.y0, self.dx, self.dy, x, y)


def sample_grid(ua, va, x0, y0, dx, dy, x, y):
    nyn, nxn = ua.shape
    fx = (np.asarray(x, dtype=np.float64) - x0) / dx
    fy = (np.asarray(y, dtype=np.float64) - y0) / dy
    fx = np.minimum(np.maximum(fx, 0.0), nxn - 1.000001)
    fy = np.minimum(np.maximum(fy, 0.0), nyn - 1.000001)
    ia = np.floor(fx).astype(np.int64)
    ib = np.floor(fy).astype(np.int64)
    tx = fx - ia
    ty = fy - ib
    out = []
    for g in (ua, va):
        g00 = g[ib, ia]
        g10 = g[ib, ia + 1]
        g01 = g[ib + 1, ia]
        g11 = g[ib + 1, ia + 1]
        a = g00 + (g10 - g00) * tx
        b = g01 + (g11 - g01) * tx
        out.append(a + (b - a) * ty)
    return out[0], out[1]


def gaussian_wind_grid(box: float, mean_speed: float, sigma: float, corr: float,
                       seed: int, spacing: Optional[float] = None) -> WindGrid:
    """Seeded 2-D Gaussian random field by spectral synthesis (Gaussian covariance)."""
    rng = np.random.default_rng([seed, 7])
    if spacing is None:
        spacing = corr / 5.0
    m = int(math.ceil(box / spacing)) + 1
    m = max(m, 8)
    # generate on a periodic grid twice the size to avoid wrap correlation
    M = 2 * m
    k = np.fft.fftfreq(M, d=spacing) * 2 * np.pi
    kx, ky = np.meshgrid(k, k, indexing="xy")
    # Gaussian covariance exp(-r^2 / (2 l^2)) -> spectrum ~ exp(-k^2 l^2 / 2)
    spec = np.exp(-0.25 * (kx ** 2 + ky ** 2) * corr ** 2)
    theta = rng.uniform(0.0, 2 * np.pi)
    comps = []
    for _ in range(2):
        noise = rng.standard_normal((M, M))
        f = np.real(np.fft.ifft2(np.fft.fft2(noise) * spec))
        f = f[:m, :m]
        f = (f - f.mean()) / (f.std() + 1e-300)
        comps.append(f * sigma)
    ua = mean_speed * math.cos(theta) + comps[0]
    va = mean_speed * math.sin(theta) + comps[1]
    return WindGrid(np.ascontiguousarray(ua), np.ascontiguousarray(va),
                    0.0, 0.0, spacing, spacing)


def _relax_overlaps(x, y, r, box, gap, max_iter=20000):
    """Push overlapping disks apart until every pair has >= gap clearance.

    Verlet-list relaxation: candidate pairs within r_i + r_j + skin are found
    with a k-d tree (scipy), then each sweep moves both floes of every
    violating pair apart by half the violation and clamps them into the box.
    """
    from scipy.spatial import cKDTree

    n = x.size
    pos = np.stack([x, y], axis=1)
    skin = 0.25 * float(np.median(r))
    rmax = float(r.max())
    lo = r + gap
    hi = box - r - gap
    if np.any(hi < lo):
        raise ValueError("box too small for the largest floe")

    def build_pairs(p):
        tree = cKDTree(p)
        out = []
        # size classes keep the pair query bounded for wide radius ranges
        edges = np.quantile(r, [0.0, 0.5, 0.9, 0.99, 1.0])
        for a in range(len(edges) - 1):
            sel = np.nonzero((r >= edges[a]) & ((r < edges[a + 1]) | (a == len(edges) - 2)))[0]
            if sel.size == 0:
                continue
            rad = float(r[sel].max()) + rmax + gap + skin
            lists = tree.query_ball_point(p[sel], rad, return_sorted=False)
            lens = np.fromiter((len(l) for l in lists), dtype=np.int64, count=len(lists))
            if lens.sum() == 0:
                continue
            ii = np.repeat(sel, lens)
            jj = np.fromiter((j for l in lists for j in l), dtype=np.int64, count=int(lens.sum()))
            keep = ii < jj
            ii, jj = ii[keep], jj[keep]
            d = np.hypot(p[ii, 0] - p[jj, 0], p[ii, 1] - p[jj, 1])
            keep = d < r[ii] + r[jj] + gap + skin
            out.append(np.stack([ii[keep], jj[keep]], axis=1))
        if not out:
            return np.zeros((0, 2), dtype=np.int64)
        pr = np.concatenate(out, axis=0)
        pr = np.unique(pr, axis=0)
        return pr

    pairs = build_pairs(pos)
    ref = pos.copy()
    for it in range(max_iter):
        if np.max(np.hypot(*(pos - ref).T)) > 0.4 * skin:
            pairs = build_pairs(pos)
            ref = pos.copy()
        if pairs.shape[0] == 0:
            break
        ii, jj = pairs[:, 0], pairs[:, 1]
        dxy = pos[jj] - pos[ii]
        d = np.hypot(dxy[:, 0], dxy[:, 1])
        need = r[ii] + r[jj] + gap
        viol = need - d
        bad = viol > 0
        if not np.any(bad):
            break
        ii, jj, dxy, d, viol = ii[bad], jj[bad], dxy[bad], d[bad], viol[bad]
        d = np.where(d > 0, d, 1e-9)
        # under-relaxed push (a floe may sit in several overlaps at once) plus a
        # small fixed clearance so the last violations close in finite sweeps
        push = ((0.25 * viol + 0.25 * gap) / d)[:, None] * dxy
        delta = np.zeros_like(pos)
        np.add.at(delta, jj, push)
        np.add.at(delta, ii, -push)
        pos += delta
        pos[:, 0] = np.clip(pos[:, 0], lo, hi)
        pos[:, 1] = np.clip(pos[:, 1], lo, hi)
    else:
        raise RuntimeError("overlap relaxation did not converge; lower the packing")
    return pos[:, 0].copy(), pos[:, 1].copy(), it


def _place_lib():
    import ctypes
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "libgodar_place.so")
    if not os.path.exists(path):
        from .build import build_place
        build_place()
    L = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    L.godar_place_disks.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_uint64, ctypes.c_int, ctypes.c_int, dp, dp,
                                    ctypes.POINTER(ctypes.c_int)]
    L.godar_place_disks.restype = ctypes.c_int
    return L


def place_disks(r, box, seed, gap=1.0, max_tries=400, max_relax=20000):
    """Non-overlapping centres for radii r in a [0, box]^2 walled box (compiled
    largest-first random sequential addition + relaxation, csrc/place.c)."""
    import ctypes
    L = _place_lib()
    r = np.ascontiguousarray(r, dtype=np.float64)
    x = np.zeros(r.size)
    y = np.zeros(r.size)
    st = (ctypes.c_int * 2)()
    dp = ctypes.POINTER(ctypes.c_double)
    rc = L.godar_place_disks(r.size, r.ctypes.data_as(dp), float(box), float(gap),
                             int(seed) & MASK64, int(max_tries), int(max_relax),
                             x.ctypes.data_as(dp), y.ctypes.data_as(dp), st)
    if rc != 0:
        raise RuntimeError(f"placement failed (code {rc}); lower the packing")
    return x, y, dict(rsa_failed=int(st[0]), relax_iters=int(st[1]))


def relax_disks(x, y, r, box, gap=1.0, max_relax=20000):
    """Compiled overlap removal from given centres (csrc/place.c)."""
    import ctypes
    L = _place_lib()
    dp = ctypes.POINTER(ctypes.c_double)
    L.godar_relax_disks.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, dp, dp, ctypes.POINTER(ctypes.c_int)]
    L.godar_relax_disks.restype = ctypes.c_int
    r = np.ascontiguousarray(r, dtype=np.float64)
    x = np.array(x, dtype=np.float64)
    y = np.array(y, dtype=np.float64)
    sw = ctypes.c_int()
    rc = L.godar_relax_disks(r.size, r.ctypes.data_as(dp), float(box), float(gap), int(max_relax),
                             x.ctypes.data_as(dp), y.ctypes.data_as(dp), ctypes.byref(sw))
    if rc != 0:
        raise RuntimeError(f"overlap relaxation failed (code {rc}); lower the packing")
    return x, y, sw.value


def make_field(n: int, seed: int, *, radius: str, r_lo: float, r_hi: float,
               h_median: float, h_sigma: float, packing: float,
               box: Optional[float] = None, power: float = -2.5,
               wind_speed: float = 10.0, current_speed: float = 0.1,
               gap: float = 1.0, placement: str = "relax") -> FloeField:
    """Generic seeded field: radii, lognormal thickness, non-overlapping centres.

    placement="relax": jittered grid + numpy relaxation; "grid": the same with
    the compiled relaxation; "rsa": compiled largest-first random sequential
    addition (needed for large, dense fields).
    """
    rng = np.random.default_rng(seed)
    if radius == "uniform":
        r = rng.uniform(r_lo, r_hi, n)
    elif radius == "powerlaw":
        # inverse CDF of p(r) ~ r^power on [r_lo, r_hi]
        a = power + 1.0
        u = rng.uniform(0.0, 1.0, n)
        r = (r_lo ** a + u * (r_hi ** a - r_lo ** a)) ** (1.0 / a)
    else:
        raise ValueError(radius)
    h = h_median * np.exp(h_sigma * rng.standard_normal(n))
    area = float(np.sum(np.pi * r ** 2))
    if box is None:
        box = math.sqrt(area / packing)
    if placement in ("relax", "grid"):
        # jittered grid + push-apart relaxation ("relax": numpy, "grid": compiled)
        m = int(math.ceil(math.sqrt(n)))
        spacing = box / m
        cells = rng.permutation(m * m)[:n]
        cx = (cells % m + 0.5) * spacing + rng.uniform(-0.25, 0.25, n) * spacing
        cy = (cells // m + 0.5) * spacing + rng.uniform(-0.25, 0.25, n) * spacing
        cx = np.clip(cx, r + gap, box - r - gap)
        cy = np.clip(cy, r + gap, box - r - gap)
        if placement == "relax":
            x, y, iters = _relax_overlaps(cx, cy, r, box, gap)
        else:
            x, y, iters = relax_disks(cx, cy, r, box, gap)
        pstats = dict(relax_iters=int(iters))
    else:
        x, y, pstats = place_disks(r, box, int(rng.integers(0, 2 ** 63)), gap)
        iters = pstats["relax_iters"]
    ang = rng.uniform(0.0, 2 * np.pi)
    cang = rng.uniform(0.0, 2 * np.pi)
    wind = (wind_speed * math.cos(ang), wind_speed * math.sin(ang))
    current = (current_speed * math.cos(cang), current_speed * math.sin(cang))
    z = np.zeros(n)
    meta = dict(seed=seed, n=n, radius=radius, r_lo=r_lo, r_hi=r_hi,
                h_median=h_median, h_sigma=h_sigma, packing_target=packing,
                packing=float(area / box ** 2), box=box, relax_iters=int(iters),
                placement=placement, placement_stats=pstats,
                wind_angle=ang, current_angle=cang)
    return FloeField(x, y, r, h, z.copy(), z.copy(), z.copy(), z.copy(), box, seed,
                     wind, current, meta)


def config0(seed: int = 0, n: int = 1000, box: Optional[float] = 1e5) -> FloeField:
    """1,000 floes, r ~ U[500, 2000] m, h lognormal (median 1 m, sigma 0.3),
    100 km walled box (realised packing is recorded in ``meta``), wind 10 m/s
    in a seeded direction, current 0.1 m/s."""
    return make_field(n, seed, radius="uniform", r_lo=500.0, r_hi=2000.0,
                      h_median=1.0, h_sigma=0.3, packing=0.6, box=box,
                      wind_speed=10.0, current_speed=0.1)


def config0_member(seed: int, n: int) -> FloeField:
    """Config-0 generator at another floe count: the box scales to keep the
    packing of the 1,000-floe / 100 km field."""
    return make_field(n, seed, radius="uniform", r_lo=500.0, r_hi=2000.0,
                      h_median=1.0, h_sigma=0.3, packing=0.55, box=None,
                      wind_speed=10.0, current_speed=0.1, placement="grid")


def config1(seed: int = 1, n: int = 262144, packing: float = 0.7,
            with_wind_grid: bool = True) -> FloeField:
    """Power-law radii (exponent -2.5) on [200, 5000] m, h lognormal median 1.5 m,
    packing 0.7 (box derived from the packing), Gaussian wind field
    (mean 8 m/s, sigma 3 m/s, 50 km correlation)."""
    f = make_field(n, seed, radius="powerlaw", r_lo=200.0, r_hi=5000.0,
                   h_median=1.5, h_sigma=0.3, packing=packing, box=None,
                   wind_speed=8.0, current_speed=0.1, placement="rsa")
    if with_wind_grid:
        f.wind_grid = gaussian_wind_grid(f.box, 8.0, 3.0, 5e4, seed)
        f.meta["wind_grid"] = dict(mean=8.0, sigma=3.0, corr=5e4,
                                   spacing=f.wind_grid.dx, shape=f.wind_grid.ua.shape)
    return f


def member_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Floe counts drawn log-uniform in [lo, hi]."""
    rng = np.random.default_rng([seed, 11])
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), count))).astype(np.int64)
