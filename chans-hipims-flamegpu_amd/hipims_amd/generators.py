"""Seeded synthetic inputs in the reference's raster format.

Every generator returns rasters with one ring of nodata (mask -9999) around
the domain, the outermost valid ring carrying boundary codes as the
reference's preProcessing.importDEMData_And_BC writes them (leading digit =
type, rest = id; preProcessing.py:282-337): 40 = slip wall (closed on all
four sides), 30 = the reference's "rigid" wall, 50 = open outflow, 1 =
interior.  All randomness comes from numpy PCG64 with explicit seeds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

NODATA = -9999
MANNING_6 = np.array([0.02, 0.03, 0.035, 0.05, 0.06, 0.08])
SEED_DEM, SEED_CLASSES, SEED_RAIN, SEED_GAUGES = 1001, 1002, 1003, 1004


@dataclass
class Case:
    name: str
    mask: np.ndarray
    h: np.ndarray
    qx: np.ndarray
    qy: np.ndarray
    z: np.ndarray
    dx: float
    landuse: np.ndarray
    manning: np.ndarray
    cfl: float = 0.5
    first_dt: float = 1.0e-4
    rain_table: np.ndarray | None = None      # [rows, 2] uniform rain, m/s
    rain_field: np.ndarray | None = None      # per-cell rain, m/s
    extra: dict = field(default_factory=dict)

    @property
    def n_valid(self) -> int:
        return int((self.mask > 0).sum())


def ring_mask(ny: int, nx: int, wall: int = 40, east: int | None = None) -> np.ndarray:
    """(ny+2, nx+2) raster: nodata ring, boundary ring of `wall`, interior 1;
    the east boundary column gets `east` when given."""
    m = np.full((ny + 2, nx + 2), NODATA, dtype=np.int32)
    m[1:-1, 1:-1] = 1
    m[1, 1:-1] = wall
    m[-2, 1:-1] = wall
    m[1:-1, 1] = wall
    m[1:-1, -2] = wall if east is None else east
    if east is not None:
        m[1, -2] = east
        m[-2, -2] = east
    return m


def pad(a: np.ndarray, fill=0.0) -> np.ndarray:
    return np.pad(a, 1, constant_values=fill)


def dam_break(n: int = 256, dx: float = 1.0, h_left: float = 2.0, h_right: float = 0.5, manning: float = 0.03,
              wall: int = 30, cfl: float = 0.5) -> Case:
    """Flat-bed dam break: h = h_left for columns < n/2, h_right elsewhere."""
    mask = ring_mask(n, n, wall=wall)
    h = np.full((n, n), h_right)
    h[:, : n // 2] = h_left
    z = np.zeros((n, n))
    first_dt = cfl * dx / math.sqrt(9.81 * h_left)
    return Case("dam_break", mask, pad(h), pad(np.zeros((n, n))), pad(np.zeros((n, n))), pad(z), dx,
                pad(np.zeros((n, n), dtype=np.int64)).astype(np.int64), np.array([manning]), cfl, first_dt)


def diamond_square(n: int, seed: int, roughness: float = 0.6) -> np.ndarray:
    """Fractal terrain on an n x n crop of a (2^k + 1)^2 diamond-square grid;
    the displacement amplitude shrinks by `roughness` per octave."""
    k = max(1, int(math.ceil(math.log2(max(n - 1, 1)))))
    size = 2 ** k + 1
    rng = np.random.Generator(np.random.PCG64(seed))
    g = np.zeros((size, size), dtype=np.float64)
    g[0, 0], g[0, -1], g[-1, 0], g[-1, -1] = rng.standard_normal(4)
    step, amp = size - 1, 1.0
    while step > 1:
        half = step // 2
        # diamond: centres of squares
        c = (g[:-1:step, :-1:step] + g[:-1:step, step::step] + g[step::step, :-1:step] + g[step::step, step::step]) / 4
        g[half::step, half::step] = c + amp * rng.standard_normal(c.shape)
        # square: edge midpoints, the points (i, j) of the lattice of
        # spacing `half` with i + j odd; each takes the mean of its lattice
        # neighbours (north, south, west, east, in that order of summation;
        # three on the border) plus noise drawn in row-major order
        lat = g[::half, ::half]
        m = lat.shape[0]
        acc = np.zeros((m, m))
        cnt = np.zeros((m, m))
        acc[1:, :] += lat[:-1, :]
        cnt[1:, :] += 1
        acc[:-1, :] += lat[1:, :]
        cnt[:-1, :] += 1
        acc[:, 1:] += lat[:, :-1]
        cnt[:, 1:] += 1
        acc[:, :-1] += lat[:, 1:]
        cnt[:, :-1] += 1
        odd = np.zeros((m, m), dtype=bool)
        odd[0::2, 1::2] = True
        odd[1::2, 0::2] = True
        lat[odd] = acc[odd] / cnt[odd] + amp * rng.standard_normal(int(odd.sum()))
        step, amp = half, amp * roughness
    return g[:n, :n]


def scale(a: np.ndarray, lo: float, hi: float) -> np.ndarray:
    a = a - a.min()
    m = a.max()
    return lo + (hi - lo) * (a / m if m > 0 else a)


def carve_channel(z: np.ndarray, dx: float, width_m: float = 10.0, depth_m: float = 2.0) -> np.ndarray:
    """Channel along the middle row from west to east: bed = running minimum
    of the terrain along the path minus depth_m, so it drains to the east."""
    z = z.copy()
    ny, nx = z.shape
    w = max(1, int(round(width_m / dx)))
    r0 = ny // 2 - w // 2
    path = z[r0: r0 + w, :].min(axis=0)
    bed = np.minimum.accumulate(path) - depth_m
    z[r0: r0 + w, :] = bed[None, :]
    return z


def voronoi_classes(ny: int, nx: int, seed: int, n_sites: int = 64, n_classes: int = 6, block: int = 4) -> np.ndarray:
    """Landcover classes from a seeded Voronoi tessellation (labels evaluated
    on `block` x `block` cell blocks)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sites = rng.uniform(0, 1, size=(n_sites, 2)) * np.array([ny, nx])
    cls = rng.integers(0, n_classes, size=n_sites)
    by, bx = int(math.ceil(ny / block)), int(math.ceil(nx / block))
    yc = (np.arange(by) + 0.5) * block
    xc = (np.arange(bx) + 0.5) * block
    best = np.full((by, bx), np.inf)
    lab = np.zeros((by, bx), dtype=np.int64)
    for s in range(n_sites):
        d = (yc[:, None] - sites[s, 0]) ** 2 + (xc[None, :] - sites[s, 1]) ** 2
        better = d < best
        best[better] = d[better]
        lab[better] = cls[s]
    return np.repeat(np.repeat(lab, block, axis=0), block, axis=1)[:ny, :nx]


def gaussian_random_field(ny: int, nx: int, dx: float, corr_m: float, mean: float, std: float, seed: int,
                          coarse: int = 8) -> np.ndarray:
    """Stationary field with a Gaussian covariance of length corr_m (FFT
    filtering of white noise on a grid `coarse` times coarser, bilinearly
    refined), normalised to mean/std and clipped at 0."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cy, cx = int(math.ceil(ny / coarse)) + 1, int(math.ceil(nx / coarse)) + 1
    cdx = dx * coarse
    noise = rng.standard_normal((cy, cx))
    ky = np.fft.fftfreq(cy, d=cdx)
    kx = np.fft.fftfreq(cx, d=cdx)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    filt = np.exp(-(math.pi ** 2) * corr_m ** 2 * k2 / 2.0)
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * filt))
    f = (f - f.mean()) / (f.std() + 1e-30)
    # bilinear refinement to (ny, nx)
    yi = np.arange(ny) / coarse
    xi = np.arange(nx) / coarse
    y0 = np.floor(yi).astype(int)
    x0 = np.floor(xi).astype(int)
    wy = (yi - y0)[:, None]
    wx = (xi - x0)[None, :]
    f = ((1 - wy) * (1 - wx) * f[np.ix_(y0, x0)] + (1 - wy) * wx * f[np.ix_(y0, x0 + 1)]
         + wy * (1 - wx) * f[np.ix_(y0 + 1, x0)] + wy * wx * f[np.ix_(y0 + 1, x0 + 1)])
    f = (f - f.mean()) / (f.std() + 1e-30)
    return np.clip(mean + std * f, 0.0, None)


def mmh(v: float) -> float:
    """mm/h -> m/s."""
    return v / 1000.0 / 3600.0


def uniform_rain_table(rate_ms: float, duration_s: float) -> np.ndarray:
    """[t, r] table: constant rate until duration_s, then 0."""
    return np.array([[0.0, rate_ms], [duration_s, rate_ms], [duration_s * (1 + 1e-12) + 1e-9, 0.0]])


def triangular_hyetograph(peak_ms: float, duration_s: float) -> np.ndarray:
    return np.array([[0.0, 0.0], [duration_s / 2.0, peak_ms], [duration_s, 0.0]])


def rainfall_runoff(n: int = 2048, dx: float = 5.0, seed: int = SEED_DEM, class_seed: int = SEED_CLASSES,
                    rain_mmh: float = 50.0, duration_s: float = 3600.0, relief: float = 50.0,
                    wall: int = 40) -> Case:
    """Seeded fractal DEM (0..relief m) with a carved channel, dry start,
    uniform rain, 6 Voronoi landcover classes, open east edge."""
    z = scale(diamond_square(n + 1, seed, 0.6), 0.0, relief)[:n, :n]
    z = carve_channel(z, dx)
    cls = voronoi_classes(n, n, class_seed)
    mask = ring_mask(n, n, wall=wall, east=50)
    zero = np.zeros((n, n))
    c = Case(f"rainfall_runoff_{n}", mask, pad(zero), pad(zero), pad(zero), pad(z), dx,
             pad(cls, 0).astype(np.int64), MANNING_6.copy(), 0.5, 1.0e-4,
             rain_table=uniform_rain_table(mmh(rain_mmh), duration_s))
    return c


def roofline_case(n: int = 8192, dx: float = 2.0) -> Case:
    """rainfall_runoff terrain at n x n, dx, with a static Gaussian random
    rain field (mean 30 mm/h, std 10 mm/h, correlation 500 m)."""
    c = rainfall_runoff(n, dx)
    field_ = gaussian_random_field(n, n, dx, 500.0, mmh(30.0), mmh(10.0), SEED_RAIN)
    c.rain_table = None
    c.rain_field = pad(field_, 0.0).astype(np.float32)
    c.name = f"roofline_{n}"
    return c


ENSEMBLE_SIZES = (512, 768, 1024, 1536, 2048)


def ensemble_specs(n_scen: int = 64, seed: int = 2000):
    """(size, seed, peak_ms, duration_s) of each ensemble scenario."""
    rng = np.random.Generator(np.random.PCG64(seed))
    sizes = rng.choice(ENSEMBLE_SIZES, size=n_scen)
    peaks = rng.uniform(20.0, 120.0, size=n_scen)
    durs = rng.uniform(0.5, 3.0, size=n_scen) * 3600.0
    return [(int(sizes[k]), seed + k, mmh(float(peaks[k])), float(durs[k])) for k in range(n_scen)]


def ensemble_case(size: int, seed: int, peak_ms: float, duration_s: float, dx: float = 5.0) -> Case:
    c = rainfall_runoff(size, dx, seed=seed, class_seed=seed + 500)
    c.rain_table = triangular_hyetograph(peak_ms, duration_s)
    c.name = f"scenario_{seed}_{size}"
    return c
