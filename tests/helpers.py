"""Small builders shared by the tests."""
import numpy as np

from hipims_amd import generators as gen


def box(ny, nx, wall=40, east=None, z=None, h=None, qx=None, qy=None):
    """(ny+2, nx+2) rasters with a nodata ring and a wall ring."""
    mask = gen.ring_mask(ny, nx, wall=wall, east=east)
    z = np.zeros((ny, nx)) if z is None else z
    h = np.zeros((ny, nx)) if h is None else h
    qx = np.zeros((ny, nx)) if qx is None else qx
    qy = np.zeros((ny, nx)) if qy is None else qy
    return mask, gen.pad(h), gen.pad(qx), gen.pad(qy), gen.pad(z)


def compact(mask, raster, dtype=np.float64):
    return np.ascontiguousarray(np.asarray(raster)[mask > 0], dtype=dtype)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(initial=0.0), 1e-300)
    return float(np.abs(a - b).max(initial=0.0) / scale)


def random_state(ny, nx, seed, wet=True, rough=0.5):
    rng = np.random.Generator(np.random.PCG64(seed))
    z = rough * rng.standard_normal((ny, nx))
    h = (0.5 + rng.uniform(0, 1, (ny, nx))) if wet else np.where(rng.uniform(0, 1, (ny, nx)) < 0.5, 0.0,
                                                                    rng.uniform(0, 1, (ny, nx)))
    qx = 0.3 * rng.standard_normal((ny, nx)) * (h > 0)
    qy = 0.3 * rng.standard_normal((ny, nx)) * (h > 0)
    return z, h, qx, qy
