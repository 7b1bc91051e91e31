"""Case pre-processing (preProcessing.py restated without rasterio): mask
codes, boundary boxes, gauge order, landuse classes, station mask, raster
files, and the compact neighbour table against the oracle's own builder and
the table shape the reference's backCase notebook prints."""
import math
import os

import numpy as np
import pytest

from hipims_amd import preprocessing as pre

ND = -9999.0


def ring_raster(ny, nx, value=1.0):
    data = np.full((ny, nx), value)
    data[0, :] = data[-1, :] = data[:, 0] = data[:, -1] = ND
    return pre.Raster(data, data == ND, (1.0, 0.0, 0.0, 0.0, -1.0, float(ny)))


def test_boundary_cells_and_default_code():
    r = ring_raster(5, 6)
    z, mask, g = pre.dem_and_bc(r, default_BC=50)
    expect = np.zeros((5, 6), np.int32)
    expect[1:4, 1:5] = 50
    expect[2, 2:4] = 1  # the only valid cells with no nodata neighbour
    assert np.array_equal(mask, expect)
    assert (z[r.nodata_mask] == ND).all() and g.size == 0


def test_boundary_wraps_like_roll():
    data = np.ones((4, 4))
    data[0, 0] = ND
    r = pre.Raster(data, data == ND)
    b = pre.boundary_cells(r.nodata_mask)
    # E/W/N/S neighbours of (0, 0) with wrap-around: (0, 1), (0, 3), (1, 0), (3, 0)
    assert set(map(tuple, np.argwhere(b))) == {(0, 1), (0, 3), (1, 0), (3, 0)}


def test_boxes_count_codes_per_type():
    r = ring_raster(6, 8)  # x = col, y = 6 - row
    boxes = [[6.0, 0.0, 8.0, 6.0],   # east column of valid cells (col 6)
             [0.0, 0.0, 2.0, 6.0],   # west column (col 1)
             [2.0, 4.0, 6.0, 6.0]]   # top row, cols 2..5 (row 1)
    z, mask, _ = pre.dem_and_bc(r, boundBox=np.array(boxes), default_BC=50, bc_type=["OPEN", "H_GIVEN", "H_GIVEN"])
    assert set(mask[1:5, 6]) == {51}          # second OPEN code
    assert set(mask[1:5, 1]) == {60}          # first H_GIVEN
    assert set(mask[1, 2:6]) == {61}          # second H_GIVEN
    assert set(mask[4, 2:6]) == {50}          # untouched default
    assert (mask[2:4, 2:6] == 1).all()


def test_fall_box_raises_like_the_reference():
    r = ring_raster(6, 8)
    with pytest.raises(IndexError):
        pre.dem_and_bc(r, boundBox=np.array([[0.0, 0.0, 8.0, 6.0]]), bc_type=["FALL"])


def test_gauges_in_caller_order():
    r = ring_raster(6, 8)
    # gauges at (row 3, col 5), (row 1, col 1), (row 2, col 2): x = col + .5, y = 6 - row - .5
    pos = np.array([[5.5, 2.5], [1.5, 4.5], [2.5, 3.5]])
    _, mask, g = pre.dem_and_bc(r, gauges_position=pos)
    valid = np.argwhere(mask > 0)
    cid = {tuple(v): k for k, v in enumerate(valid)}
    assert g.tolist() == [cid[(3, 5)], cid[(1, 1)], cid[(2, 2)]]


def test_landuse_classes_and_station_mask():
    data = np.array([[ND, 5, 2], [7, 2, 5]])
    r = pre.Raster(data, data == ND)
    lu, idx = pre.landuse_classes(r, 0)
    # nodata -> 1; unique {1, 2, 5, 7}: 2 -> 0, 5 -> 1, 7 -> 2, 1 stays 1 (merges with 5)
    assert lu.tolist() == [[1, 1, 0], [2, 0, 1]] and idx.tolist() == [0, 1, 2]
    d1 = np.array([[ND, 10, 20], [30, 20, 10]])
    lu1, idx1 = pre.landuse_classes(pre.Raster(d1, d1 == ND), 1)
    assert lu1.tolist() == [[0, 0, 1], [2, 1, 0]] and idx1.tolist() == [0, 1]
    st = pre.rain_station_mask(pre.Raster(np.array([[ND, 3.0], [1.0, 2.0]]), np.array([[True, False], [False, False]])))
    assert st.dtype == np.int16 and st.tolist() == [[0, 3], [1, 2]]


def test_degree_to_meter():
    assert pre.degree_to_meter(1.0) == pytest.approx(2 * math.pi * 6371004.0 / 360.0, rel=1e-15)


@pytest.mark.parametrize("ext", [".asc", ".npy"])
def test_raster_files_round_trip(tmp_path, ext):
    rng = np.random.Generator(np.random.PCG64(5))
    data = rng.uniform(0, 50, (7, 9))
    data[0, :] = ND
    tf = (2.0, 0.0, 100.0, 0.0, -2.0, 500.0)
    path = os.path.join(tmp_path, "dem" + ext)
    pre.write_raster(path, data, tf, ND)
    r = pre.read_raster(path)
    assert np.array_equal(r.data, data) and r.nodata_mask[0].all() and not r.nodata_mask[1:].any()
    assert np.allclose(r.transform, tf)
    assert r.bounds == (100.0, 486.0, 118.0, 500.0)
    row, col = r.rowcol(103.9, 497.5)
    assert (int(row), int(col)) == (1, 1)


def test_neighbour_index_matches_oracle_and_backcase_shape(oracle_lib):
    # backCase.ipynb: a 27 x 227 raster with a nodata ring; the reference
    # prints the neighbour table as torch.Size([5, 5625])
    mask = np.full((27, 227), 10.0)
    mask[1, :] = mask[-2, :] = 40
    mask[:, 1] = 60
    mask[:, -2] = 50
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = -9999
    idx = pre.neighbour_index(mask)
    assert idx.shape == (5, 5625) and idx.dtype == np.int32
    ref = oracle_lib.build_index(mask).reshape(5, -1)
    assert np.array_equal(idx, ref)
    rng = np.random.Generator(np.random.PCG64(9))
    m = np.where(rng.uniform(size=(13, 17)) < 0.7, 1, -9999)
    got = pre.neighbour_index(m)
    assert np.array_equal(got, oracle_lib.build_index(m).reshape(5, -1))
    # brute force: E, N (row - 1), W, S neighbours with wrap-around
    ny, nx = m.shape
    cells = [(r, c) for r in range(ny) for c in range(nx) if m[r, c] > 0]
    cid = {rc: k for k, rc in enumerate(cells)}
    for k, (r, c) in enumerate(cells):
        for j, (dr, dc) in enumerate(((0, 1), (-1, 0), (0, -1), (1, 0))):
            assert got[j + 1, k] == cid.get(((r + dr) % ny, (c + dc) % nx), -1)
