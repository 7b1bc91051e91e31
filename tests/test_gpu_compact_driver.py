"""The reference's compact per-step sequence over the HIP compact ops
(CompactGodunov) against the oracle and the structured Godunov drop-in
(GPU): same step sequence, FP64, 1e-10."""
import sys

import numpy as np
import pytest
import torch

from helpers import rel_err
from hipims_amd import generators as gen
from hipims_amd.compact_driver import CompactGodunov
from hipims_amd.SWE_CUDA import Godunov

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(cls, c, dev, order):
    T = lambda a: torch.as_tensor(np.asarray(a))
    num = cls(dev, c.dx, c.cfl, 1.0e9, t=0.0, export_n=0, firstTimeStep=c.first_dt, secondOrder=order == 2,
              tensorType=torch.float64)
    num.init__fluidField_tensor(T(c.mask), T(c.h), T(c.qx), T(c.qy), T(c.h + c.z), T(c.z), dev)
    num.set__frictionField_tensor(c.manning, dev)
    num.set_landuse(T(c.mask), T(c.landuse), dev)
    num.set_distributed_rainfall_station_Mask(T(c.mask), T(np.zeros_like(c.mask)), dev)
    num.set_boundary_tensor({}, dev)
    return num


@pytest.mark.parametrize("order", [1, 2])
def test_compact_sequence_matches_oracle_and_structured(oracle_lib, cuda, order):
    say(f"[compact driver order {order}]")
    c = gen.rainfall_runoff(64, 5.0, rain_mmh=300.0, relief=5.0)
    a = setup(CompactGodunov, c, cuda, order)
    b = setup(Godunov, c, cuda, order)
    o = oracle_lib.OracleSim(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, first_dt=c.first_dt,
                             manning=c.manning, landuse=c.landuse, rain_table=c.rain_table,
                             station=np.zeros_like(c.mask))
    for k in range(60):
        for num in (a, b):
            num.addFlux()
            num.addStation_PrecipitationSource(c.rain_table, cuda)
            num.time_friction_euler_update_cuda(cuda)
    o.run(60)
    ha, hb = a.get_h().cpu().numpy(), b.get_h().cpu().numpy()
    e_o, e_b = rel_err(ha, o.h_c), rel_err(ha, hb)
    say(f"  vs oracle {e_o:.2e}, vs structured {e_b:.2e}, t {a.t.item():.2f} / {o.time:.2f}")
    assert e_o <= 1e-10 and e_b <= 1e-10
    assert abs(a.t.item() - o.time) <= 1e-10 * o.time
