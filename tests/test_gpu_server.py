"""The RPC server and the output formats over the HIP Godunov drop-in (GPU):
depths served to the agent-model client equal a direct run of the same
server step sequence, and exported snapshots convert to rasters equal to
the state at the export time."""
import os
import sys

import numpy as np
import pytest
import torch

from hipims_amd import generators as gen
from hipims_amd import postprocessing as post
from hipims_amd import preprocessing as pre
from hipims_amd.server import TIME_STEP, FloodServer, RPCClient, Simulation, serve
from hipims_amd.SWE_CUDA import Godunov

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def godunov(c, dev, out, export_dt=1.0e9, dtype=torch.float64):
    T = lambda a: torch.as_tensor(np.asarray(a))
    num = Godunov(dev, c.dx, c.cfl, export_dt, t=0.0, export_n=0, firstTimeStep=1.0, secondOrder=False,
                  tensorType=dtype)
    num.setOutPutPath(str(out))
    num.init__fluidField_tensor(T(c.mask), T(c.h), T(c.qx), T(c.qy), T(c.h + c.z), T(c.z), dev)
    num.set__frictionField_tensor(c.manning, dev)
    num.set_landuse(T(c.mask), T(c.landuse), dev)
    num.set_distributed_rainfall_station_Mask(T(c.mask), T(np.zeros_like(c.mask)), dev)
    num.set_boundary_tensor({}, dev)
    return num


def test_server_serves_the_direct_run(cuda, tmp_path):
    say("[rpc server over Godunov]")
    c = gen.rainfall_runoff(48, 5.0, rain_mmh=200.0, relief=4.0)
    abm = np.zeros(c.mask.shape, bool)
    abm[10:20, 8:30] = True
    abm &= c.mask > 0
    idx = Simulation.abm_compact_index(c.mask, abm)
    sim = Simulation(godunov(c, cuda, tmp_path), c.rain_table, idx, cuda)
    srv = serve(FloodServer(sim), "127.0.0.1", 0)
    direct = godunov(c, cuda, tmp_path)
    idx_t = torch.from_numpy(idx).to(cuda)
    try:
        client = RPCClient("127.0.0.1", srv.server_address[1])
        for k in range(1, 4):
            t = k * TIME_STEP
            data = client.call("request_data", t)
            while direct.t.item() < t:
                direct.addFlux()
                direct.addStation_PrecipitationSource(c.rain_table, cuda)
                direct.time_friction_euler_update_cuda(cuda)
            ref = direct.get_h()[idx_t].cpu().numpy()
            say(f"  t {t}: {len(data)} values, max {max(data):.4f} m")
            assert len(data) == len(idx) and np.array_equal(np.array(data), ref)
            client.call("set_target", t + TIME_STEP)
        client.close()
    finally:
        srv.shutdown()
        srv.server_close()


def test_exported_snapshots_become_rasters(cuda, tmp_path):
    say("[exports -> rasters]")
    c = gen.rainfall_runoff(40, 5.0, rain_mmh=300.0, relief=3.0)
    num = godunov(c, cuda, tmp_path, export_dt=60.0, dtype=torch.float32)
    has = lambda prefix: any(f.startswith(prefix) for f in os.listdir(tmp_path))
    num.exportField()  # at t = 0 the files are named t + dt = 1.0
    assert has("h_tensor([1.]")
    while not has("h_tensor([120.]"):
        num.addFlux()
        num.addStation_PrecipitationSource(c.rain_table, cuda)
        num.time_friction_euler_update_cuda(cuda)
        assert num.t.item() < 200
    assert num.t.item() == 120.0
    files = post.export_rasters(c.mask, str(tmp_path), "npy")
    assert os.path.join(tmp_path, "h_120.npy") in files and os.path.join(tmp_path, "h_max_60.npy") in files
    r = pre.read_raster(os.path.join(tmp_path, "h_120.npy"))
    expect = np.full(c.mask.shape, -9999.0, np.float32)
    expect[c.mask > 0] = num.get_h().cpu().numpy()
    assert np.array_equal(r.data, expect)
