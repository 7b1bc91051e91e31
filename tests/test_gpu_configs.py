"""The configured workloads, each run once on the GPU at its configured size.

config1  flat 256 x 256 dam break, dx 1 m, FP64, order 2, rigid walls (code 30),
         1000 steps run free on the HIP path and on the oracle, compared at
         four checkpoints (h, qx, qy, h_max, t, dt within 1e-7 relative run
         free, 1e-10 for one step from the oracle's state).
config2  2048 x 2048 rainfall-runoff DEM (dx 5 m, 50 mm/h), FP64: ten steps in
         lockstep with the oracle from a state 300 steps into the storm.
config3  8192 x 8192 DEM at dx 2 m with the Gaussian random rain field,
         FP32 state with FP64 mass sums: a few hundred single-pass steps,
         finite state and the mass balance closed to 1e-5.
config4  64 scenarios at the configured sizes (512^2 .. 2048^2), FP32, one
         graph-captured ensemble: one scenario of each size against its
         standalone single-pass run bit for bit, the mass balance of every
         scenario.
config5  1024 x 1024 calibration, 2000 steps, checkpoints every 100, 32
         gauges, 7 parameters: the FP64 adjoint gradient against FP64 central
         differences at eps 1e-6 (bar 1e-4) and eps 1e-3 (reported).
"""
import concurrent.futures as cf
import multiprocessing as mp
import os
import sys

import numpy as np
import pytest
import torch

from helpers import rel_err
from hipims_amd import generators as gen
from hipims_amd.grid import GridSolver

pytestmark = pytest.mark.gpu
TOL64 = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def solver(c, dtype, order=2, **kw):
    s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, dtype=dtype, first_dt=c.first_dt,
                   manning=c.manning, landuse=c.landuse, **kw)
    if c.rain_table is not None:
        s.set_station_rain(c.rain_table)
    elif c.rain_field is not None:
        s.set_rain_field(c.rain_field)
    return s


def oracle_of(orc, c, order=2):
    st = np.zeros_like(c.mask) if c.rain_table is not None else None
    return orc.OracleSim(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, first_dt=c.first_dt,
                         manning=c.manning, landuse=c.landuse, rain_table=c.rain_table, station=st)


def errors(g, o):
    torch.cuda.synchronize()
    e = {name: rel_err(g.compact(getattr(g, name)).cpu().numpy(), ref)
         for name, ref in (("h", o.h_c), ("qx", o.qx_c), ("qy", o.qy_c), ("h_max", o.h_max))}
    e["t"] = abs(g.t - o.time) / max(o.time, 1e-300)
    e["dt"] = abs(g.dt - o.dt) / o.dt
    return e


def balance(s):
    c = s.ctrl_dict()
    return (c["volume"] - c["rain_volume"] - c["flux_volume"]) / c["rain_volume"]


@pytest.mark.parametrize("path", ["two_pass", "single_pass"])
def test_config1_dam_break_256_fp64_order2_1000_steps_free_running(oracle_lib, cuda, path):
    """Run free for 1000 steps beside the oracle.  Order 2 amplifies
    rounding-level differences (about threefold per 250 steps on this
    case: 3.7e-11 at step 250, 1.1e-10 at 500), so the free-running bar is
    1e-7; at each checkpoint one step from the oracle's state is also taken
    in lockstep, which must agree to 1e-10."""
    say(f"[config1: dam break 256^2 FP64 order 2, 1000 steps free running, {path}]")
    c = gen.dam_break(256, wall=30)
    g = solver(c, torch.float64)
    o = oracle_of(oracle_lib, c)
    lock = solver(c, torch.float64)
    run = lambda s, n: (s.step if path == "two_pass" else s.step_fused)(n)
    for k in range(4):
        run(g, 250)
        o.run(250)
        e = errors(g, o)
        say(f"  step {250 * (k + 1)}: t {o.time:.3f} s, free: " + ", ".join(f"{n} {v:.1e}" for n, v in e.items()))
        bad = {n: v for n, v in e.items() if not v <= 1e-7}
        assert not bad, f"step {250 * (k + 1)}: {bad}"
        # one step from the oracle's state (the oracle is advanced by one step
        # here, so its later checkpoints are one step further on each time)
        for name, arr in (("h", o.h_c), ("qx", o.qx_c), ("qy", o.qy_c), ("h_max", o.h_max)):
            lock.load_compact(getattr(lock, name), torch.from_numpy(arr))
        lock.ctrl[0] = o.time
        lock.ctrl[1] = o.dt
        run(lock, 1)
        o.run(1)
        run(g, 1)
        e1 = errors(lock, o)
        say("    lockstep: " + ", ".join(f"{n} {v:.1e}" for n, v in e1.items()))
        bad = {n: v for n, v in e1.items() if not v <= TOL64}
        assert not bad, f"lockstep after step {250 * (k + 1)}: {bad}"
    assert g.ctrl_dict()["steps"] == 1004


@pytest.mark.parametrize("path", ["two_pass", "single_pass"])
def test_config2_runoff_2048_fp64_lockstep_10_steps(oracle_lib, cuda, path):
    say(f"[config2: rainfall-runoff 2048^2 FP64 order 2, 10 lockstep steps after 300, {path}]")
    c = gen.rainfall_runoff(2048, 5.0)
    g = solver(c, torch.float64)
    g.step(300)
    torch.cuda.synchronize()
    o = oracle_of(oracle_lib, c)
    o.h_c[:] = g.compact(g.h).cpu().numpy()
    o.qx_c[:] = g.compact(g.qx).cpu().numpy()
    o.qy_c[:] = g.compact(g.qy).cpu().numpy()
    o.h_max[:] = g.compact(g.h_max).cpu().numpy()
    o.wl_c[:] = o.h_c + o.z_c
    o.ctrl[0], o.ctrl[1] = g.t, g.dt
    wet = int((o.h_c > 1e-6).sum())
    say(f"  t {g.t:.2f} s, {wet} of {o.N} cells wet, max h {o.h_c.max():.4f} m")
    assert wet > 0
    worst = 0.0
    for k in range(10):
        for name, arr in (("h", o.h_c), ("qx", o.qx_c), ("qy", o.qy_c), ("h_max", o.h_max)):
            g.load_compact(getattr(g, name), torch.from_numpy(arr))
        g.ctrl[0] = o.time
        g.ctrl[1] = o.dt
        (g.step if path == "two_pass" else g.step_fused)(1)
        o.run(1)
        e = errors(g, o)
        bad = {n: v for n, v in e.items() if not v <= TOL64}
        assert not bad, f"lockstep step {k}: {bad}"
        worst = max(worst, max(e.values()))
    say(f"  worst relative difference over the 10 steps {worst:.1e}")


def test_config3_roofline_8192_fp32_rain_field_mass_sums_300_steps(cuda):
    say("[config3: 8192^2 dx 2 m FP32, Gaussian rain field, FP64 mass sums, 300 single-pass steps]")
    c = gen.roofline_case(8192, 2.0)
    say("  generated")
    g = solver(c, torch.float32)
    g.set_mass_sums(True)
    g.step_fused(300)
    torch.cuda.synchronize()
    ctrl = g.ctrl_dict()
    nx = g.layout.nx
    fin = all(bool(torch.isfinite(t[:, :nx]).all()) for t in (g.h, g.qx, g.qy))
    b = balance(g)
    say(f"  t {ctrl['t']:.2f} s, steps {ctrl['steps']:.0f}, volume {ctrl['volume']:.6e} m3, rain "
        f"{ctrl['rain_volume']:.6e} m3, boundary {ctrl['flux_volume']:.3e} m3, balance {b:.2e}")
    # the reference clamps no negative depth, and on this steep terrain
    # (relief 50 m at dx 2 m) its own FP64 update overdraws a few hundred
    # cells by millimetres (the oracle at 512^2: min -6 mm after 600 steps)
    neg = g.h[:, :nx] < 0
    say(f"  min h {float(g.h[:, :nx].min()):.2e} m in {int(neg.sum())} of {g.layout.n_valid} cells")
    assert ctrl["steps"] == 300 and fin
    assert abs(b) < 1e-5


def test_config4_ensemble_64_scenarios_configured_sizes(cuda):
    from hipims_amd.ensemble import Ensemble
    specs = gen.ensemble_specs(64, 2000)
    say(f"[config4: 64 scenarios, sizes {sorted(set(s[0] for s in specs))}, FP32, graph-replayed, 100 steps]")
    workers = max(1, min(8, len(os.sched_getaffinity(0))))
    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        cases = list(ex.map(gen.ensemble_case, *zip(*specs)))
    say("  generated")
    ens_s = [solver(c, torch.float32) for c in cases]
    sample = {}
    for k, sp in enumerate(specs):
        sample.setdefault(sp[0], k)
    alone = {k: solver(cases[k], torch.float32) for k in sample.values()}
    e = Ensemble(ens_s)
    e.capture(50)
    e.replay()
    e.replay()
    for s in alone.values():
        s.step_fused(100)
    torch.cuda.synchronize()
    for k, s in alone.items():
        a = ens_s[k]
        for name in ("h", "qx", "qy", "h_max"):
            assert torch.equal(getattr(a, name), getattr(s, name)), f"scenario {k} ({specs[k][0]}^2): {name}"
        ca, cs = a.ctrl_dict(), s.ctrl_dict()
        assert ca["t"] == cs["t"] and ca["dt"] == cs["dt"] and ca["steps"] == cs["steps"] == 100, k
    bal = [balance(s) for s in ens_s]
    say(f"  bitwise equal: scenarios {sorted(alone)}; worst mass balance {max(abs(b) for b in bal):.2e}")
    assert all(s.ctrl_dict()["steps"] == 100 for s in ens_s)
    assert max(abs(b) for b in bal) < 1e-5


def test_config5_calibration_1024_2000_steps_fp64_adjoint_vs_fd(cuda):
    from hipims_amd.adjoint import Calibration
    say("[config5: calibration 1024^2, 2000 steps, checkpoint 100, 32 gauges, FP64 adjoint vs FD]")
    c = gen.rainfall_runoff(1024, 5.0)
    pilot = solver(c, torch.float64, order=1)
    pilot.step(2000)
    h = pilot.raster(pilot.h)
    rng = np.random.Generator(np.random.PCG64(1004))
    wet = np.argwhere((h > 1e-3) & (c.mask > 0))
    gauges = wet[rng.choice(len(wet), size=min(32, len(wet)), replace=False)]
    del pilot
    cal = Calibration(solver(c, torch.float64, order=1), gauges, 2000, 100)
    th0 = cal.theta0()
    assert th0.numel() == 7
    pert = torch.tensor(np.random.Generator(np.random.PCG64(7)).choice([0.8, 1.2], size=th0.numel() - 1))
    th_true = th0.clone()
    th_true[:-1] *= pert
    cal.obs = cal.simulate(th_true)
    loss = cal.run(th0)
    grad = cal.gradient()
    errs = {}
    for eps in (1e-6, 1e-3):
        fd = torch.zeros_like(grad)
        for j in range(th0.numel()):
            d = eps * abs(float(th0[j]))
            tp, tm = th0.clone(), th0.clone()
            tp[j] += d
            tm[j] -= d
            fd[j] = (cal.run(tp).item() - cal.run(tm).item()) / (2 * d)
        errs[eps] = ((grad - fd).abs().max() / grad.abs().max()).item()
        say(f"  eps {eps:g}: fd {fd.numpy()}")
    say(f"  loss {loss.item():.6e}, {len(gauges)} gauges\n  adjoint {grad.numpy()}\n"
        f"  max rel err vs FD: eps 1e-6 {errs[1e-6]:.2e}, eps 1e-3 {errs[1e-3]:.2e}")
    assert errs[1e-6] <= 1e-4
