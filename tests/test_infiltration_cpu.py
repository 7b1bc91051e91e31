"""Green-Ampt infiltration in the oracle (infiltrationCUDA_Kernel.cu:20-46),
pinned by the mathematics: F1 is the positive root of
F^2 - (F0 + dt Ks) F - dt Ks (phi + h) dtheta = 0, dtheta = 0 reduces to
conductivity-limited infiltration, the increment is capped by the depth, and
the removed depth equals the infiltrated depth."""
import numpy as np
import pytest


def run(orc, h, F, cls, ks, phi, dth, dt, wet=None, dtype=np.float64):
    n = h.size
    wet = np.arange(n, dtype=np.int32) if wet is None else np.asarray(wet, np.int32)
    dh = np.zeros(n, dtype)
    F = F.astype(dtype).copy()
    orc.infiltration(wet, dh, cls.astype(np.uint8), h.astype(dtype), ks, phi, dth, F, dt)
    return dh, F


def test_root_of_the_green_ampt_quadratic(oracle_lib):
    rng = np.random.Generator(np.random.PCG64(11))
    n = 400
    h = rng.uniform(0.05, 2.0, n)          # deep enough that the cap is inactive
    F0 = rng.uniform(0.0, 0.3, n)
    cls = rng.integers(0, 3, n)
    ks = np.array([1e-6, 5e-6, 2e-5])
    phi = np.array([0.05, 0.11, 0.2])
    dth = np.array([0.1, 0.3, 0.45])
    dt = 2.5
    dh, F = run(oracle_lib, h, F0, cls, ks, phi, dth, dt)
    a = F0 + dt * ks[cls]
    resid = F * F - a * F - dt * ks[cls] * (phi[cls] + h) * dth[cls]
    assert np.abs(resid).max() < 1e-14 * np.maximum(1.0, F * F).max()
    assert (F > F0).all() and np.array_equal(dh, -(F - F0))


def test_no_deficit_is_conductivity_limited(oracle_lib):
    h = np.array([0.5, 1.0, 0.2])
    F0 = np.array([0.0, 0.1, 0.3])
    dh, F = run(oracle_lib, h, F0, np.zeros(3), np.array([3e-6]), np.array([0.1]), np.array([0.0]), 7.0)
    assert F - F0 == pytest.approx(np.full(3, 7.0 * 3e-6), rel=1e-12)


def test_increment_capped_by_depth_and_wet_list(oracle_lib):
    h = np.array([1e-5, 0.4, 0.3, 1e-7])
    F0 = np.zeros(4)
    dh, F = run(oracle_lib, h, F0, np.zeros(4), np.array([1e-3]), np.array([0.1]), np.array([0.4]), 10.0,
                wet=[0, 1, 3])
    assert F[0] == h[0] and F[3] == h[3]         # shallow cells lose all their water
    assert F[2] == 0.0 and dh[2] == 0.0          # not listed: untouched
    assert 0 < F[1] < h[1] and dh[1] == -F[1]
