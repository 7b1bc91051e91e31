"""Case pre-processing without rasterio (CHANS-HiPIMS/hipims/pythonHipims/preProcessing.py).

The reference reads GeoTIFFs with rasterio and builds the solver inputs on
the GPU with torch; here the same inputs are built with numpy from rasters
held as arrays plus an affine transform, read from ESRI ASCII grids (.asc)
or .npy files with a JSON sidecar.  What is restated, with its rules:

  * boundary cells (preProcessing.py:236-251): valid cells with a nodata
    neighbour E, N, W or S, found with wrap-around rolls of the nodata mask;
  * mask codes (:223-337): 1 interior, 0 nodata, ``default_BC`` on every
    boundary cell, then per box ``int(str(type) + str(n))`` with n counted
    per boundary type from 0 (the default BC's type already used 0);
  * gauge compact indices in the caller's gauge order (:255-281);
  * landuse re-indexing (:345-369) and the rain-station mask (:372-380);
  * the station rain table (:393-395, a whitespace text matrix).
Quirks kept: the per-type counters cover RIGID..WL_GIVEN only, so a FALL
box (type 9) raises IndexError as in the reference; landuse level 0 leaves
the smallest class value unchanged and maps the i-th larger one to i - 1
(values can merge); gauges on nodata cells shift the later gauge indices.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import dataclass

import numpy as np

BC_TYPES = {"RIGID": 3, "WALL_SLIP": 4, "OPEN": 5, "H_GIVEN": 6, "Q_GIVEN": 7, "WL_GIVEN": 8, "FALL": 9}
NODATA = -9999.0


@dataclass
class Raster:
    """A single-band raster: data[row, col], a nodata mask and the affine
    transform (a, b, c, d, e, f) with x = a*col + b*row + c and
    y = d*col + e*row + f at a cell's upper-left corner (GDAL / rasterio
    order)."""
    data: np.ndarray
    nodata_mask: np.ndarray
    transform: tuple = (1.0, 0.0, 0.0, 0.0, -1.0, 0.0)

    @property
    def shape(self):
        return self.data.shape

    @property
    def cell_size(self) -> float:
        return float(self.transform[0])

    @property
    def bounds(self):
        """(left, bottom, right, top) for a north-up transform."""
        a, _, c, _, e, f = self.transform
        ny, nx = self.data.shape
        return (c, f + e * ny, c + a * nx, f)

    def rowcol(self, x, y):
        """Row and column of map coordinates, floored like rasterio's
        transform.rowcol (the inverse affine map, then floor)."""
        a, b, c, d, e, f = self.transform
        det = a * e - b * d
        x = np.asarray(x, dtype=np.float64) - c
        y = np.asarray(y, dtype=np.float64) - f
        col = (e * x - b * y) / det
        row = (-d * x + a * y) / det
        return np.floor(row).astype(np.int64), np.floor(col).astype(np.int64)

    def filled(self, value=NODATA) -> np.ndarray:
        out = np.array(self.data, dtype=np.float64, copy=True)
        out[self.nodata_mask] = value
        return out


def read_raster(path: str, nodata=None) -> Raster:
    """ESRI ASCII grid (.asc) or .npy (+ optional ``<path>.json`` holding
    ``{"transform": [a, b, c, d, e, f], "nodata": v}``)."""
    if path.endswith(".asc"):
        header = {}
        with open(path) as fh:
            for _ in range(6):
                pos = fh.tell()
                parts = fh.readline().split()
                if len(parts) != 2 or parts[0][0].isdigit() or parts[0][0] == "-":
                    fh.seek(pos)
                    break
                header[parts[0].lower()] = float(parts[1])
            data = np.loadtxt(fh, dtype=np.float64, ndmin=2)
        nrows, ncols = int(header["nrows"]), int(header["ncols"])
        if data.shape != (nrows, ncols):
            raise ValueError(f"{path}: header says {nrows}x{ncols}, data is {data.shape}")
        cs = header["cellsize"]
        x0 = header.get("xllcorner", header.get("xllcenter", 0.0) - cs / 2)
        y0 = header.get("yllcorner", header.get("yllcenter", 0.0) - cs / 2)
        nd = header.get("nodata_value", nodata)
        transform = (cs, 0.0, x0, 0.0, -cs, y0 + cs * nrows)
    elif path.endswith(".npy"):
        data = np.load(path, allow_pickle=False)
        meta = {}
        if os.path.exists(path + ".json"):
            with open(path + ".json") as fh:
                meta = json.load(fh)
        transform = tuple(meta.get("transform", (1.0, 0.0, 0.0, 0.0, -1.0, 0.0)))
        nd = meta.get("nodata", nodata)
    else:
        raise ValueError(f"unsupported raster format: {path} (use .asc or .npy)")
    data = np.asarray(data)
    mask = ~np.isfinite(data) if data.dtype.kind == "f" else np.zeros(data.shape, bool)
    if nd is not None:
        mask |= data == nd
    return Raster(data, mask, tuple(float(v) for v in transform))


def write_raster(path: str, data: np.ndarray, transform=(1.0, 0.0, 0.0, 0.0, -1.0, 0.0), nodata=NODATA):
    """Write .asc (north-up transforms only) or .npy + JSON sidecar."""
    data = np.asarray(data)
    if path.endswith(".asc"):
        a, b, c, d, e, f = transform
        if b != 0.0 or d != 0.0 or abs(a) != abs(e):
            raise ValueError("ESRI ASCII grids need a north-up transform with square cells")
        ny, nx = data.shape
        with open(path, "w") as fh:
            fh.write(f"ncols {nx}\nnrows {ny}\nxllcorner {c!r}\nyllcorner {f + e * ny!r}\n"
                     f"cellsize {a!r}\nNODATA_value {nodata!r}\n")
            np.savetxt(fh, data, fmt="%.17g")
    elif path.endswith(".npy"):
        np.save(path, data)
        with open(path + ".json", "w") as fh:
            json.dump({"transform": list(transform), "nodata": nodata}, fh)
    else:
        raise ValueError(f"unsupported raster format: {path} (use .asc or .npy)")


def degree_to_meter(degree_unit: float) -> float:
    """preProcessing.py:111-113 (mean Earth radius 6371004 m)."""
    return degree_unit * (2.0 * math.pi * 6371004.0) / 360.0


def boundary_cells(nodata: np.ndarray) -> np.ndarray:
    """Valid cells with a nodata neighbour E, N, W or S; the neighbours are
    taken with wrap-around like torch.roll (preProcessing.py:241-251)."""
    nd = nodata.astype(np.int32)
    count = (np.roll(nd, -1, axis=1) + np.roll(nd, 1, axis=0) + np.roll(nd, 1, axis=1) + np.roll(nd, -1, axis=0))
    count[nodata] = 0
    return count > 0


def dem_and_bc(dem: Raster, gauges_position=None, boundBox=None, default_BC: int = 50, bc_type=5):
    """importDEMData_And_BC (preProcessing.py:221-337) on a Raster.

    Returns (dem filled with -9999 at nodata, int32 mask codes, int64 gauge
    compact indices in the order of ``gauges_position`` rows (x, y))."""
    nodata = np.asarray(dem.nodata_mask, bool)
    z = dem.filled(NODATA)
    bnd = boundary_cells(nodata)
    mask = (~nodata).astype(np.int32)

    gauge_index = np.zeros(0, np.int64)
    gp = np.zeros((0, 2)) if gauges_position is None else np.asarray(gauges_position, dtype=np.float64).reshape(-1, 2)
    if gp.shape[0] > 0:
        rows, cols = dem.rowcol(gp[:, 0], gp[:, 1])
        mg = mask.copy()
        mg[rows, cols] = 100
        gauge_index = np.flatnonzero(mg[mg > 0] >= 99).astype(np.int64)
        ranks = np.argsort(np.argsort(rows * z.shape[1] + cols))
        gauge_index = gauge_index[ranks]

    types = list(BC_TYPES.values())
    counts = [-1] * 6  # the reference's counter list: no slot for FALL
    mask[bnd] = default_BC
    counts[types.index(int(str(default_BC)[0]))] += 1
    boxes = np.zeros((0, 4)) if boundBox is None else np.array(boundBox, dtype=np.float64).reshape(-1, 4)
    if boxes.shape[0] > 0:
        kinds = bc_type if isinstance(bc_type, (list, tuple, np.ndarray)) else [bc_type] * len(boxes)
        left, bottom, right, top = dem.bounds
        for i, b in enumerate(boxes):
            t = BC_TYPES[kinds[i]] if isinstance(kinds[i], str) else int(kinds[i])
            x0, y0 = max(left, b[0]), max(bottom, b[1])
            x1, y1 = min(right, b[2]), min(top, b[3])
            r_min, c_min = dem.rowcol(x0, y0)
            r_max, c_max = dem.rowcol(x1, y1)
            box = np.zeros_like(bnd)
            box[int(r_max):int(r_min), int(c_min):int(c_max)] = True  # rows run top to bottom
            counts[types.index(t)] += 1
            mask[bnd & box] = int(str(t) + str(counts[types.index(t)]))
    return z, mask, gauge_index


def landuse_classes(landuse: Raster, level: int = 0):
    """importLanduseData (preProcessing.py:345-369): returns (class raster,
    class index array)."""
    if level == 1:
        lu = np.where(landuse.nodata_mask, 11, landuse.data).astype(np.float64)
        lu = (lu / 10).astype(np.int64) - 1
        n_cls = len(np.unique(lu)) - 1
    else:
        lu = np.where(landuse.nodata_mask, 1, landuse.data)
        n_cls = len(np.unique(lu)) - 1
        values = np.unique(lu)
        lu = lu.copy()
        for i in range(1, len(values)):
            lu[lu == values[i]] = i - 1
    return lu, np.arange(n_cls)


def rain_station_mask(r: Raster) -> np.ndarray:
    """importRainStationMask (preProcessing.py:372-380): nodata -> station 0."""
    return np.where(r.nodata_mask, 0, r.data).astype(np.int16)


def rain_source(path: str) -> np.ndarray:
    """voronoiDiagramGauge_rainfall_source (preProcessing.py:393-395): rows
    [t, rate_0 .. rate_S-1] in m/s."""
    return np.genfromtxt(path)


def neighbour_index(mask: np.ndarray) -> np.ndarray:
    """The compact-layout table the reference's op surface takes
    (SWE_CUDA.py:173-222): int32 [5, N]; row 0 the mask code of each valid
    cell (row-major ``mask > 0`` order), rows 1..4 the compact id of its E,
    N (row - 1), W and S neighbour or -1, with wrap-around neighbours as the
    reference's torch.roll gives them."""
    mask = np.asarray(mask)
    valid = mask > 0
    n = int(valid.sum())
    ids = np.full(mask.shape, -1, np.int64)
    ids[valid] = np.arange(n)
    out = np.empty((5, n), np.int32)
    out[0] = mask[valid]
    for k, (shift, axis) in enumerate(((-1, 1), (1, 0), (1, 1), (-1, 0))):
        out[k + 1] = np.roll(ids, shift, axis=axis)[valid]
    return out
