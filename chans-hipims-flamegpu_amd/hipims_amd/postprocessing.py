"""Output formats of the reference driver, without rasterio.

  * snapshots: ``Godunov.exportField`` writes the compact arrays (row-major
    over ``mask > 0``) with torch.save as ``<name>_<str(t + dt)>.pt`` for h,
    qx, qy, wl and h_max (SWE_CUDA.py:836-868); ``load_snapshot`` reads one
    back (weights_only, nothing executed from the file);
  * rasters: ``exportRaster_tiff`` (postProcessing.py:26-155) scatters each
    snapshot onto the DEM grid with -9999 at nodata and names the output
    ``<path up to the last '_'>_<text between '[' and ']'>tif``; here the
    same names end in ``asc`` or ``npy`` (GeoTIFF needs rasterio);
  * time series: ``t.txt`` = columns (t, dt) per step and ``gauges.txt`` =
    rows (t, h at each gauge) every 10th step (CatchFlood_main.py:195-202,
    SWE_CUDA.py:290-301), both np.savetxt with its default format.
"""
from __future__ import annotations

import glob
import os

import numpy as np
import torch

from .preprocessing import NODATA, write_raster


def load_snapshot(path: str) -> torch.Tensor:
    return torch.load(path, map_location="cpu", weights_only=True)


def raster_from_compact(mask: np.ndarray, values, nodata: float = NODATA) -> np.ndarray:
    """Scatter compact values onto the raster grid (``mask > 0`` row-major)."""
    mask = np.asarray(mask)
    vals = values.detach().cpu().numpy() if torch.is_tensor(values) else np.asarray(values)
    out = np.full(mask.shape, nodata, dtype=np.float32)
    out[mask > 0] = vals.astype(np.float32)
    return out


def raster_name(snapshot_path: str, ext: str) -> str:
    """The reference's output name for a snapshot file."""
    top = snapshot_path[:snapshot_path.rfind("_")]
    tag = snapshot_path[snapshot_path.find("[") + 1:snapshot_path.find("]")]
    return top + "_" + tag + ext


def export_rasters(mask: np.ndarray, out_path: str, ext: str = "asc",
                   transform=(1.0, 0.0, 0.0, 0.0, -1.0, 0.0)) -> list[str]:
    """Convert every ``*.pt`` snapshot under out_path to a raster file;
    returns the files written."""
    written = []
    for f in sorted(glob.glob(os.path.join(out_path, "*.pt"))):
        name = raster_name(f, ext)
        write_raster(name, raster_from_compact(mask, load_snapshot(f)), transform, NODATA)
        written.append(name)
    return written


def write_time_series(out_path: str, t_list, dt_list, gauge_rows=None, wall_seconds: float | None = None):
    """t.txt (t, dt per step; the reference appends the run's wall time as a
    last row of both columns) and gauges.txt (rows of t then gauge depths)."""
    t = list(t_list)
    dt = list(dt_list)
    if wall_seconds is not None:
        t.append(wall_seconds)
        dt.append(wall_seconds)
    np.savetxt(os.path.join(out_path, "t.txt"), np.column_stack((np.array(t), np.array(dt))))
    if gauge_rows is not None:
        np.savetxt(os.path.join(out_path, "gauges.txt"), np.array(gauge_rows))
