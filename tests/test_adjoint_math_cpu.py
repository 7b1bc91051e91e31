"""The hand-written reverse derivatives (swe_adjoint.hpp) against central
finite differences of the forward physics (swe_physics.hpp), both compiled for
the host in FP64 by tests/native.  Random edge states cover wet, dry and
thin-film cells, both reconstruction orders, every boundary ghost type and
all four directions; samples whose finite-difference stencil straddles a
branch switch (one-sided differences disagree) are skipped, and the test
requires that most samples are usable."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
_P = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def ec():
    subprocess.run(["make", "-s", "-C", NATIVE], check=True)
    L = ctypes.CDLL(os.path.join(NATIVE, "libedgecheck.so"))
    L.ec_edge.argtypes = [ctypes.c_int, _P, ctypes.c_double, ctypes.c_double, _P]
    L.ec_edge_vjp.argtypes = [ctypes.c_int, _P, ctypes.c_double, ctypes.c_double, _P, _P]
    L.ec_ghost_edge.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_double, ctypes.c_double,
                                _P]
    L.ec_ghost_edge_vjp.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_double,
                                    ctypes.c_double, _P, _P]
    L.ec_edge_frozen.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double, _P]
    L.ec_edge_vjp_frozen.argtypes = [_P, ctypes.c_double, ctypes.c_double, _P, _P]
    L.ec_friction.argtypes = [ctypes.c_double] * 5 + [_P]
    L.ec_friction_vjp.argtypes = [ctypes.c_double] * 5 + [_P, _P]
    L.ec_rcbrt.argtypes = [ctypes.c_double]
    L.ec_rcbrt.restype = ctypes.c_double
    return L


def test_fp64_friction_cube_root_is_accurate_to_a_few_ulp(ec):
    """The FP64 friction's h^(-1/3) (single-precision seed, two Newton
    steps) against the double pow over every depth a wet cell can hold
    (1e-6 m .. 1e4 m, log-spaced and random): within 1e-15 relative, and
    h * y^3 = 1 to the same."""
    rng = np.random.Generator(np.random.PCG64(17))
    hs = np.concatenate([np.logspace(-6, 4, 4001), 10.0 ** rng.uniform(-6, 4, 20000)])
    y = np.array([ec.ec_rcbrt(float(h)) for h in hs])
    ref = np.power(hs, -1.0 / 3.0)
    assert np.abs(y / ref - 1.0).max() <= 1e-15
    assert np.abs(hs * y ** 3 - 1.0).max() <= 2e-15


def P(a):
    return a.ctypes.data_as(_P)


def fd_check(f, x, idx, v, grad, eps_rel=1e-6, tol=1e-6, depth_idx=(), kink=1e-5, floor=1.0):
    """Compare grad[k] with the central difference of v . f along x[idx[k]].
    Depth inputs (depth_idx) are not differenced through 0 or across the
    1e-6 wet/dry threshold.  Returns (compared, skipped)."""
    ok = skip = 0
    f0 = v @ f(x)
    for k, j in enumerate(idx):
        eps = eps_rel * max(floor, abs(x[j]))
        if j in depth_idx and (x[j] - eps < 0.0 or abs(x[j] - 1e-6) < 2 * eps):
            skip += 1
            continue
        xp, xm = x.copy(), x.copy()
        xp[j] += eps
        xm[j] -= eps
        fp, fm = v @ f(xp), v @ f(xm)
        fwd, bwd, cen = (fp - f0) / eps, (f0 - fm) / eps, (fp - fm) / (2 * eps)
        if abs(fwd - bwd) > kink * (abs(fwd) + abs(bwd)) + 1e-9:
            skip += 1  # a branch switches inside the stencil
            continue
        assert abs(grad[k] - cen) <= tol * max(abs(cen), abs(grad[k])) + 1e-9, (j, grad[k], cen)
        ok += 1
    return ok, skip


def random_cells(rng, kind):
    cells = np.zeros(16)
    for k in range(4):
        if kind == "wet":
            h = rng.uniform(0.05, 2.0)
        elif kind == "mixed":
            h = rng.choice([0.0, rng.uniform(0.0, 2e-6), rng.uniform(0.01, 1.5)])
        else:
            h = rng.uniform(0.0, 0.02)
        cells[4 * k: 4 * k + 4] = [h, rng.uniform(-1.5, 1.5) * h, rng.uniform(-1.0, 1.0) * h, rng.uniform(-0.4, 0.4)]
    return cells


@pytest.mark.parametrize("kind", ["wet", "mixed", "thin"])
def test_frozen_slope_edge_vjp_matches_finite_differences(ec, kind):
    """The order-2 reverse with frozen slopes against central differences of
    the edge whose faces take their limited slopes (and dry-cell momentum
    rule) from a fixed base state: only the two cells of the edge receive
    cotangents, and those are the derivative of that edge."""
    rng = np.random.Generator(np.random.PCG64(300 + len(kind)))
    dx, dt = 5.0, 0.7
    total_ok = total_skip = 0
    for _ in range(200):
        x = random_cells(rng, kind)
        base = x.copy()
        v = rng.standard_normal(5)

        def f(xx):
            out = np.zeros(5)
            ec.ec_edge_frozen(P(np.ascontiguousarray(xx)), P(base), dx, dt, P(out))
            return out
        grad = np.zeros(12)
        ec.ec_edge_vjp_frozen(P(x), dx, dt, P(v), P(grad))
        assert np.all(grad[[0, 1, 2, 9, 10, 11]] == 0.0)  # l and rr
        ok, sk = fd_check(f, x, [4, 5, 6, 8, 9, 10], v, grad[3:9], depth_idx=(4, 8))
        total_ok += ok
        total_skip += sk
    # thin films put more stencils across a branch switch (skipped)
    assert total_ok > 0.5 * (total_ok + total_skip), (total_ok, total_skip)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("kind", ["wet", "mixed", "thin"])
def test_interior_edge_vjp_matches_finite_differences(ec, order, kind):
    rng = np.random.Generator(np.random.PCG64(100 + order * 10 + len(kind)))
    dx, dt = 5.0, 0.7
    h_idx = [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14]  # (h, qn, qt) of l, c, r, rr; z is static
    total_ok = total_skip = active = 0
    for _ in range(200):
        x = random_cells(rng, kind)
        v = rng.standard_normal(5)

        def f(xx):
            out = np.zeros(5)
            ec.ec_edge(order, P(np.ascontiguousarray(xx)), dx, dt, P(out))
            return out
        grad = np.zeros(12)
        act = ec.ec_edge_vjp(order, P(x), dx, dt, P(v), P(grad))
        active += act
        ok, sk = fd_check(f, x, h_idx, v, grad, depth_idx=(0, 4, 8, 12))
        total_ok += ok
        total_skip += sk
    assert active > 50
    assert total_ok > 0.7 * (total_ok + total_skip), (total_ok, total_skip)


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("btype", [3, 4, 5, 6, 7, 8, 9])
def test_ghost_edge_vjp_matches_finite_differences(ec, order, btype):
    rng = np.random.Generator(np.random.PCG64(7 * btype + order))
    dx, dt = 2.0, 0.3
    total_ok = total_skip = 0
    for d in range(4):
        for _ in range(25):
            c = np.array([rng.uniform(0.02, 1.5), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(-0.3, 0.3)])
            given = np.array([rng.uniform(0.0, 1.0), rng.uniform(-0.5, 1.5), rng.uniform(-0.5, 0.5),
                              rng.uniform(-0.5, 0.5)])
            v = rng.standard_normal(5)

            def f(xx):
                out = np.zeros(5)
                ec.ec_ghost_edge(order, btype, d, P(np.ascontiguousarray(xx)), P(given), dx, dt, P(out))
                return out
            grad = np.zeros(3)
            ec.ec_ghost_edge_vjp(order, btype, d, P(c), P(given), dx, dt, P(v), P(grad))
            ok, sk = fd_check(f, c, [0, 1, 2], v, grad, depth_idx=(0,))
            total_ok += ok
            total_skip += sk
    assert total_ok > 0.7 * (total_ok + total_skip), (total_ok, total_skip)


def test_friction_vjp_matches_finite_differences(ec):
    rng = np.random.Generator(np.random.PCG64(3))
    total_ok = total_skip = 0
    for _ in range(300):
        h = rng.choice([rng.uniform(1e-4, 0.01), rng.uniform(0.01, 3.0)])
        n = rng.uniform(0.01, 0.1)
        dt = rng.uniform(0.05, 5.0)
        q = rng.uniform(-2, 2, 2) * h ** 1.5 * 3
        a = rng.standard_normal(2)
        x = np.array([q[0], q[1], h, n])

        def f(xx):
            out = np.zeros(2)
            ec.ec_friction(xx[2], xx[3], dt, xx[0], xx[1], P(out))
            return out
        grad = np.zeros(4)
        ec.ec_friction_vjp(h, n, dt, q[0], q[1], P(a), P(grad))
        ok, sk = fd_check(f, x, [0, 1, 2, 3], a, grad, eps_rel=1e-6, tol=1e-4, depth_idx=(2,), kink=1e-2, floor=1e-6)
        total_ok += ok
        total_skip += sk
    assert total_ok > 0.9 * (total_ok + total_skip), (total_ok, total_skip)
