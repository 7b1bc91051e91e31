"""Output formats: snapshot names, raster reconstruction (mask > 0 row-major
scatter, -9999 elsewhere), the reference's output file names, and the
t.txt / gauges.txt layouts."""
import os

import numpy as np
import torch

from hipims_amd import postprocessing as post
from hipims_amd import preprocessing as pre


def test_snapshot_to_raster_round_trip(tmp_path):
    rng = np.random.Generator(np.random.PCG64(3))
    mask = np.where(rng.uniform(size=(6, 9)) < 0.6, 1, -9999)
    n = int((mask > 0).sum())
    h = torch.from_numpy(rng.uniform(0, 2, n))
    tag = str(torch.tensor([120.0], dtype=torch.float64))  # the reference's str(t + dt)
    for name in ("h", "h_max"):
        torch.save(h * (2 if name == "h_max" else 1), os.path.join(tmp_path, f"{name}_{tag}.pt"))
    files = post.export_rasters(mask, str(tmp_path), "asc", (5.0, 0.0, 0.0, 0.0, -5.0, 30.0))
    assert sorted(os.path.basename(f) for f in files) == ["h_120.asc", "h_max_120.asc"]
    r = pre.read_raster(os.path.join(tmp_path, "h_max_120.asc"))
    expect = np.full(mask.shape, -9999.0)
    expect[mask > 0] = (2 * h.numpy()).astype(np.float32)
    assert np.array_equal(r.data, expect) and r.cell_size == 5.0
    assert np.array_equal(r.nodata_mask, mask <= 0)


def test_raster_name_rule():
    p = "/out/qx_tensor([3600.], device='cuda:0', dtype=torch.float64).pt"
    assert post.raster_name(p, "tif") == "/out/qx_3600.tif"


def test_time_series_files(tmp_path):
    post.write_time_series(str(tmp_path), [1.0, 2.5], [1.0, 1.5], [[0.0, 0.1, 0.2], [10.0, 0.3, 0.4]],
                           wall_seconds=9.0)
    t = np.loadtxt(os.path.join(tmp_path, "t.txt"))
    g = np.loadtxt(os.path.join(tmp_path, "gauges.txt"))
    assert t.tolist() == [[1.0, 1.0], [2.5, 1.5], [9.0, 9.0]]
    assert g.shape == (2, 3) and g[1, 0] == 10.0
