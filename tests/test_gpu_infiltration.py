"""Green-Ampt infiltration (GPU): the compact op against the oracle op, the
structured two-pass step with its infiltration pass against the oracle's
whole step in lockstep (FP64, 1e-10), the single-pass step against the
two-pass step bitwise, the Godunov call sequence of CatchFlood_main.py
(addFlux, station rain, addInfiltrationSource, update) against the oracle,
and the mass balance with the infiltrated volume."""
import sys

import numpy as np
import pytest
import torch

from helpers import compact, rel_err
from hipims_amd import ops
from hipims_amd import generators as gen
from hipims_amd.grid import GridSolver

pytestmark = pytest.mark.gpu
KS = np.array([2e-6, 8e-6, 1.5e-5, 3e-6, 6e-6, 1e-5])
PHI = np.array([0.05, 0.09, 0.11, 0.2, 0.15, 0.07])
DTH = np.array([0.15, 0.25, 0.35, 0.3, 0.2, 0.4])


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def case(n=72):
    return gen.rainfall_runoff(n, 5.0, rain_mmh=250.0, relief=6.0)


def pair(orc, c, dtype=torch.float64, order=2):
    g = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, dtype=dtype, first_dt=c.first_dt,
                   manning=c.manning, landuse=c.landuse)
    g.set_station_rain(c.rain_table)
    g.set_infiltration(KS, PHI, DTH)
    o = orc.OracleSim(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, first_dt=c.first_dt,
                      manning=c.manning, landuse=c.landuse, rain_table=c.rain_table, station=np.zeros_like(c.mask),
                      hydraulic_conductivity=KS, capillary_head=PHI, water_content_diff=DTH)
    return g, o


def test_compact_op_matches_oracle(oracle_lib, cuda):
    say("[infiltration compact op]")
    rng = np.random.Generator(np.random.PCG64(21))
    n = 5000
    for dt_ in (np.float64, np.float32):
        h = rng.uniform(0, 0.5, n).astype(dt_)
        h[::7] = 1e-7
        F0 = rng.uniform(0, 0.2, n).astype(dt_)
        cls = rng.integers(0, 6, n).astype(np.uint8)
        wet = np.sort(rng.choice(n, 3000, replace=False)).astype(np.int32)
        dh0 = rng.standard_normal(n).astype(dt_) * 1e-3
        # the oracle in FP64 on the same (float-representable) inputs: the
        # FP32 path is held to its exact result
        dh_o, F_o = dh0.astype(np.float64), F0.astype(np.float64)
        oracle_lib.infiltration(wet, dh_o, cls, h.astype(np.float64), KS, PHI, DTH, F_o, 3.0)
        T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
        dh_g, F_g = T(dh0), T(F0)
        tt = torch.float64 if dt_ == np.float64 else torch.float32
        ops.add_infiltration(T(wet), dh_g, T(cls), T(h), T(KS).to(tt), T(PHI).to(tt), T(DTH).to(tt), F_g,
                             torch.tensor([3.0], dtype=tt, device=cuda))
        torch.cuda.synchronize()
        tol = 1e-14 if dt_ == np.float64 else 1e-6
        e = max(rel_err(dh_g.cpu().numpy(), dh_o), rel_err(F_g.cpu().numpy(), F_o))
        say(f"  {dt_.__name__}: rel err {e:.2e}")
        assert e <= tol


def test_two_pass_lockstep_fp64(oracle_lib, cuda):
    say("[infiltration two-pass vs oracle, fp64 lockstep]")
    c = case()
    g, o = pair(oracle_lib, c)
    o.run(30)
    worst = 0.0
    for k in range(25):
        for name, arr in (("h", o.h_c), ("qx", o.qx_c), ("qy", o.qy_c), ("h_max", o.h_max), ("F", o.F)):
            g.load_compact(getattr(g, name), torch.from_numpy(arr))
        g.ctrl[0], g.ctrl[1] = o.time, o.dt
        g.flux()
        g.infiltration()
        g.update()
        o.run(1)
        torch.cuda.synchronize()
        for name, ref in (("h", o.h_c), ("qx", o.qx_c), ("qy", o.qy_c), ("F", o.F)):
            e = rel_err(g.compact(getattr(g, name)).cpu().numpy(), ref)
            assert e <= 1e-10, (k, name, e)
            worst = max(worst, e)
        assert abs(g.dt - o.dt) <= 1e-10 * o.dt
    say(f"  worst rel err {worst:.2e}, infiltrated depth sum {o.F.sum():.4e} m, t {o.time:.1f} s")
    assert o.F.sum() > 0


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_and_step_match_two_pass(oracle_lib, cuda, dtype):
    say(f"[infiltration: step(), step_fused() vs flux/infiltration/update, {dtype}]")
    c = case(64)
    a, _ = pair(oracle_lib, c, dtype)
    b, _ = pair(oracle_lib, c, dtype)
    f, _ = pair(oracle_lib, c, dtype)
    for _ in range(40):
        a.flux()
        a.infiltration()
        a.update()
    b.step(40)
    f.step_fused(40)
    torch.cuda.synchronize()
    for s in (b, f):
        for name in ("h", "qx", "qy", "h_max", "F"):
            assert torch.equal(getattr(a, name), getattr(s, name)), name
        assert torch.equal(a.ctrl[:2], s.ctrl[:2])


def test_godunov_sequence_and_mass_balance(oracle_lib, cuda, tmp_path):
    say("[infiltration through Godunov]")
    from hipims_amd.SWE_CUDA import Godunov
    c = case(56)
    T = lambda a: torch.as_tensor(np.asarray(a))
    num = Godunov(cuda, c.dx, c.cfl, 1.0e9, t=0.0, export_n=0, firstTimeStep=c.first_dt, secondOrder=True,
                  tensorType=torch.float64)
    num.setOutPutPath(str(tmp_path))
    num.init__fluidField_tensor(T(c.mask), T(c.h), T(c.qx), T(c.qy), T(c.h + c.z), T(c.z), cuda)
    num.set__frictionField_tensor(c.manning, cuda)
    num.set_landuse(T(c.mask), T(c.landuse), cuda)
    num.set_distributed_rainfall_station_Mask(T(c.mask), T(np.zeros_like(c.mask)), cuda)
    num.set_boundary_tensor({}, cuda)
    num.set__infiltrationField_tensor(torch.tensor(KS), torch.tensor(PHI), torch.tensor(DTH), cuda)
    _, o = pair(oracle_lib, c)
    for k in range(80):
        num.addFlux()
        num.addStation_PrecipitationSource(c.rain_table, cuda)
        num.addInfiltrationSource()
        num.time_friction_euler_update_cuda(cuda)
    o.run(80)
    h = num.get_h().cpu().numpy()
    F = num._cumulativeWaterDepth.cpu().numpy()
    e = max(rel_err(h, o.h_c), rel_err(F, o.F))
    assert e <= 1e-10, e
    ctrl = num.solver.ctrl_dict()
    area = c.dx * c.dx
    balance = ctrl["volume"] - (compact(c.mask, c.h).sum() * area + ctrl["rain_volume"] + ctrl["flux_volume"]
                                - F.sum() * area)
    say(f"  rel err vs oracle {e:.2e}; infiltrated {F.sum() * area:.3e} m3 of rain {ctrl['rain_volume']:.3e} m3; "
        f"balance {balance:.2e} m3")
    assert abs(balance) <= 1e-9 * ctrl["rain_volume"]
