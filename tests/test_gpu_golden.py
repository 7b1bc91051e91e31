"""The HIP structured path (FP64) against the committed golden fixtures
(tests/golden/*.npz): the same inputs run free (not in lockstep) for the
fixture's step count through the two-pass and the single-pass step, with
the same time schedule emerging on the device; depths, discharges, h_max,
and every step's t and dt agree with the oracle's committed outputs to 1e-10
relative."""
import glob
import os

import numpy as np
import pytest
import torch

from helpers import rel_err
from hipims_amd.grid import GridSolver

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
TOL = 1e-10


def load(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def solver(d, cuda):
    opt = lambda k: d[k] if k in d else None
    g = GridSolver(d["mask"], d["h"], d["qx"], d["qy"], d["z"], float(d["dx"]), cfl=float(d["cfl"]),
                   order=int(d["order"]), dtype=torch.float64, device=cuda, first_dt=float(d["first_dt"]),
                   manning=d["manning"], landuse=d["landuse"], given_depth=opt("given_depth"),
                   given_wl=opt("given_wl"), given_q=opt("given_q"),
                   export_T=float(d["export_T"]) if "export_T" in d else 1.0e30)
    if "rain_table" in d:
        g.set_station_rain(d["rain_table"], d["station"] if "station" in d else None)
    return g


@pytest.mark.parametrize("fused", [False, True], ids=["two_pass", "single_pass"])
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_hip_matches_golden(cuda, path, fused):
    d = load(path)
    print(f"[golden {os.path.basename(path)} {'single-pass' if fused else 'two-pass'}]", flush=True)
    g = solver(d, cuda)
    sched = []
    for _ in range(int(d["steps"])):
        sched.append((g.t, g.dt))
        g.step_fused(1) if fused else g.step(1)
    torch.cuda.synchronize()
    errs = {k: rel_err(g.compact(getattr(g, k)).cpu().numpy(), d[f"{k}_out"]) for k in ("h", "qx", "qy", "h_max")}
    errs["sched"] = float(np.abs(np.array(sched) - d["sched"]).max() / np.abs(d["sched"]).max())
    errs["t"] = abs(g.t - float(d["t_end"])) / float(d["t_end"])
    errs["dt"] = abs(g.dt - float(d["dt_end"])) / float(d["dt_end"])
    print("  " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    assert all(v <= TOL for v in errs.values()), errs
