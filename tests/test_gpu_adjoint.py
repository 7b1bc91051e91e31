"""Adjoint gradients (GPU): the hand-written reverse of the step against
central finite differences of the forward model with the same frozen
time-step schedule, FP32 against FP64, and the autograd.Function wrapper."""
import sys

import numpy as np
import pytest
import torch

from hipims_amd import generators as gen
from hipims_amd.adjoint import Calibration, CalibrationLoss
from hipims_amd.grid import GridSolver

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(dtype, n=64, steps=80, ck=25, seed=4242, obs=None, overlap=True):
    c = gen.rainfall_runoff(n, 5.0, seed=seed, rain_mmh=150.0, relief=8.0)
    s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=0.5, order=2, dtype=torch.float64, first_dt=0.5,
                   manning=c.manning, landuse=c.landuse)
    s.set_station_rain(c.rain_table)
    # gauges: seeded choice among cells wet after a pilot run
    s.step(steps)
    h = s.raster(s.h)
    rng = np.random.Generator(np.random.PCG64(1004))
    wet = np.argwhere((h > 1e-3) & (c.mask > 0))
    gauges = wet[rng.choice(len(wet), size=min(16, len(wet)), replace=False)]
    s2 = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=0.5, order=2, dtype=dtype, first_dt=0.5,
                    manning=c.manning, landuse=c.landuse)
    s2.set_station_rain(c.rain_table)
    cal = Calibration(s2, gauges, steps, ck, overlap=overlap)
    th0 = cal.theta0()
    pert = torch.tensor(np.random.Generator(np.random.PCG64(7)).choice([0.8, 1.2], size=th0.numel() - 1))
    th_true = th0.clone()
    th_true[:-1] *= pert
    cal.obs = cal.simulate(th_true) if obs is None else obs
    return cal, th0


def fd_grad(cal, th0, eps=1e-3):
    g = torch.zeros_like(th0)
    for j in range(th0.numel()):
        d = eps * abs(float(th0[j]))
        tp, tm = th0.clone(), th0.clone()
        tp[j] += d
        tm[j] -= d
        g[j] = (cal.run(tp).item() - cal.run(tm).item()) / (2 * d)
    return g


def test_fp64_adjoint_matches_finite_differences(cuda):
    say("[adjoint fp64 vs FD]")
    cal, th0 = make(torch.float64)
    loss = cal.run(th0)
    grad = cal.gradient()
    fd3 = fd_grad(cal, th0, 1e-3)
    fd5 = fd_grad(cal, th0, 1e-5)
    say(f"  loss {loss.item():.6e}\n  adjoint {grad.numpy()}\n  fd(1e-3) {fd3.numpy()}\n  fd(1e-5) {fd5.numpy()}")
    scale = grad.abs().max()
    e3 = ((grad - fd3).abs() / scale).max().item()
    e5 = ((grad - fd5).abs() / scale).max().item()
    say(f"  max rel err vs FD: eps 1e-3 {e3:.2e}, eps 1e-5 {e5:.2e}")
    assert min(e3, e5) < 1e-4


def test_fp32_adjoint_against_fp64(cuda):
    say("[adjoint fp32 vs fp64]")
    cal64, th0 = make(torch.float64)
    cal64.run(th0)
    g64 = cal64.gradient()
    cal32, th0_32 = make(torch.float32, obs=cal64.obs)
    cal32.run(th0_32)
    g32 = cal32.gradient()
    err = ((g32 - g64).abs() / g64.abs().max()).max().item()
    say(f"  fp32 {g32.numpy()}\n  fp64 {g64.numpy()}\n  max rel diff {err:.2e}")
    assert err < 2e-2


def test_autograd_function(cuda):
    say("[autograd]")
    cal, th0 = make(torch.float64, n=48, steps=40, ck=15)
    theta = th0.clone().requires_grad_(True)
    loss = CalibrationLoss.apply(theta, cal)
    loss.backward()
    ref = cal.gradient()
    assert torch.allclose(theta.grad, ref, rtol=1e-12, atol=0.0)
    assert torch.isfinite(theta.grad).all() and theta.grad.abs().max() > 0


def test_store_all_matches_recompute(cuda):
    say("[adjoint: every state kept vs checkpoint + recompute]")
    cal1, th0 = make(torch.float64, n=48, steps=40, ck=1)
    cal2, _ = make(torch.float64, n=48, steps=40, ck=15, obs=cal1.obs)
    cal1.run(th0)
    cal2.run(th0)
    g1, g2 = cal1.gradient(), cal2.gradient()
    say(f"  {g1.numpy()}")
    assert torch.equal(g1, g2)


def test_order1_long_horizon_fp32_matches_fd(cuda):
    """Order 1 over a few hundred steps in FP32: the adjoint stays bounded
    and agrees with central differences on the frozen schedule (to 2e-2 of
    the largest component: the FP32 loss limits the differences' precision)."""
    say("[adjoint order 1, fp32, 400 steps]")
    c = gen.rainfall_runoff(96, 5.0, seed=4242, rain_mmh=150.0, relief=8.0)

    def solver(dtype):
        s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=0.5, order=1, dtype=dtype, first_dt=0.5,
                       manning=c.manning, landuse=c.landuse)
        s.set_station_rain(c.rain_table)
        return s
    pilot = solver(torch.float64)
    pilot.step(400)
    h = pilot.raster(pilot.h)
    rng = np.random.Generator(np.random.PCG64(1004))
    wet = np.argwhere((h > 1e-3) & (c.mask > 0))
    gauges = wet[rng.choice(len(wet), size=min(16, len(wet)), replace=False)]
    cal = Calibration(solver(torch.float32), gauges, 400, 50)
    th0 = cal.theta0()
    pert = torch.tensor(np.random.Generator(np.random.PCG64(7)).choice([0.8, 1.2], size=th0.numel() - 1))
    th_true = th0.clone()
    th_true[:-1] *= pert
    cal.obs = cal.simulate(th_true)
    cal.run(th0)
    grad = cal.gradient()
    fd = fd_grad(cal, th0, 1e-3)
    err = ((grad - fd).abs() / fd.abs().max()).max().item()
    say(f"  adjoint {grad.numpy()}\n  fd      {fd.numpy()}\n  max rel err {err:.2e}")
    assert torch.isfinite(grad).all() and err < 2e-2


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_order1_tiled_reverse_matches_separate_passes(cuda, dtype):
    """The single-pass tiled reverse step (order 1) against the separate
    update, edge and gather passes: the cotangent of the initial state and
    the gradient.  The domain is not a multiple of the 64 x 8 tile (partial
    tiles on the east and north sides), 37 steps run in segments of 10 (the
    odd last segment leaves the cotangent in the other set of planes, which
    is then copied back) and as one stored segment, with distinct gauges (the
    loss enters in the tiled pass) and with a gauge given twice (the loss
    enters through its own kernel)."""
    say(f"[adjoint order 1 tiled vs separate passes, {dtype}]")
    c = gen.rainfall_runoff(70, 5.0, seed=99, rain_mmh=150.0, relief=8.0)

    def solver():
        s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=0.5, order=1, dtype=dtype, first_dt=0.5,
                       manning=c.manning, landuse=c.landuse)
        s.set_station_rain(c.rain_table)
        return s
    pilot = solver()
    pilot.step(37)
    h = pilot.raster(pilot.h)
    say(f"  domain {pilot.layout.ny} x {pilot.layout.nx}, pitch {pilot.layout.pitch}")
    rng = np.random.Generator(np.random.PCG64(1004))
    wet = np.argwhere((h > 1e-3) & (c.mask > 0))
    uniq = wet[rng.choice(len(wet), size=min(12, len(wet)), replace=False)]
    dup = np.concatenate([uniq, uniq[:1]])
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    for gauges in (uniq, dup):
        for ck in (10, 1):
            ref = Calibration(solver(), gauges, 37, ck, reverse="separate")
            th0 = ref.theta0()
            th_true = th0.clone()
            th_true[:-1] *= 1.2
            ref.obs = ref.simulate(th_true)
            ref.run(th0)
            g_ref = ref.gradient()
            til = Calibration(solver(), gauges, 37, ck, obs=ref.obs, reverse="tiled")
            assert (til.gauge_of is None) == (gauges is dup)
            til.run(th0)
            g_til = til.gradient()
            e_g = ((g_til - g_ref).abs().max() / g_ref.abs().max()).item()
            e_l = max(((a - b).abs().max() / b.abs().max().clamp_min(1e-300)).item()
                      for a, b in zip(til.lam, ref.lam))
            same = all(torch.equal(a, b) for a, b in zip(til.lam, ref.lam))
            say(f"  gauges {len(gauges)} ck {ck}: grad rel {e_g:.2e}, lambda rel {e_l:.2e}, lambda bitwise {same}")
            assert torch.isfinite(g_til).all() and g_ref.abs().max() > 0
            assert e_g <= tol and e_l <= tol


def test_overlapped_recompute_matches_serial(cuda):
    """Recomputing the next segment on a second stream while the current one
    is reversed gives bitwise the gradient and cotangent of the serial
    schedule (37 steps in segments of 10, FP32)."""
    say("[adjoint: overlapped vs serial recomputation]")
    g, lam, obs = {}, {}, None
    for overlap in (False, True):
        cal, th0 = make(torch.float32, n=64, steps=37, ck=10, obs=obs, overlap=overlap)
        obs = cal.obs
        assert cal.overlap == overlap
        cal.run(th0)
        g[overlap] = cal.gradient()
        lam[overlap] = [t.clone() for t in cal.lam]
    say(f"  {g[True].numpy()}")
    assert torch.equal(g[False], g[True])
    assert all(torch.equal(a, b) for a, b in zip(lam[False], lam[True]))
