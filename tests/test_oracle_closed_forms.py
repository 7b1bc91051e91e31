"""Pins the CPU oracle (oracle/swe_oracle.cpp) against what mathematics and
physics fix, independently of the oracle's own formulas.  The reference ships
no tests, fixtures or golden data for this path, so these checks are what make
the oracle trustworthy; each is chosen so that a plausible slip (a sign, an
index, a swapped operand, a dropped term) fails it.

Which oracle function each test pins:
  flux_compact (both orders) .. lake at rest, uniform-flow consistency, mass
                                conservation, 90-degree rotation symmetry,
                                Stoker dam break, order-2 accuracy gain
  bc_ghost / decode_code ...... slip wall zero mass flux, open-edge outflow,
                                H-given inflow column, rigid-wall N/S quirk
  flux_mask ................... cell-0 quirk, t == 0 rule
  friction_update ............. closed-form quadratic root (brentq), n = 0,
                                limiter invariants, dry cells
  time_control ................ exact-root CFL values, dry = 60
  station_rain + rain_interp .. volume identity, linear interpolation at the
                                driver's t (np.interp), zero after the table
  driver dt logic ............. 10x growth cap, export clamp
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from helpers import box, compact, rel_err, random_state
from hipims_amd import generators as gen

G = 9.81


def one_flux(orc, order, mask, h, qx, qy, z, t=1.0, dt=0.01, dx=1.0, gd=None, gw=None, gq=None, listed=None):
    h_c, qx_c, qy_c, z_c = (compact(mask, a) for a in (h, qx, qy, z))
    idx = orc.build_index(mask)
    wet = orc.flux_list(h_c, idx, t) if listed is None else listed
    dh, dqx, dqy = (np.zeros_like(h_c) for _ in range(3))
    orc.flux(order, wet, dh, dqx, dqy, h_c, h_c + z_c, z_c, qx_c, qy_c, idx,
             gd if gd is not None else [[0.0, 0.0]], gw if gw is not None else [[0.0, 0.0]],
             gq if gq is not None else [[0.0, 0.0, 0.0]], dx, t, dt)
    return dh, dqx, dqy, h_c


def to_raster(mask, v):
    out = np.zeros(mask.shape)
    out[mask > 0] = v
    return out


@pytest.mark.parametrize("order", [1, 2])
def test_lake_at_rest_rough_bed(oracle_lib, order):
    rng = np.random.Generator(np.random.PCG64(7))
    z = rng.uniform(-1.0, 1.0, (12, 14))
    eta = 2.0
    mask, h, qx, qy, zz = box(12, 14, z=z, h=eta - z)
    dh, dqx, dqy, _ = one_flux(oracle_lib, order, mask, h, qx, qy, zz)
    assert np.abs(dh).max() < 1e-15
    assert np.abs(dqx).max() < 1e-14
    assert np.abs(dqy).max() < 1e-14


@pytest.mark.parametrize("order", [1, 2])
def test_uniform_flow_is_steady_with_open_edges(oracle_lib, order):
    # a uniform state on a flat bed is an exact solution; with transmissive
    # (open, 50) edges every cell sees equal inflow and outflow
    n = 10
    mask = gen.ring_mask(n, n, wall=50)
    h = gen.pad(np.full((n, n), 1.3))
    qx = gen.pad(np.full((n, n), 0.7))
    qy = gen.pad(np.full((n, n), -0.4))
    z = gen.pad(np.zeros((n, n)))
    dh, dqx, dqy, _ = one_flux(oracle_lib, order, mask, h, qx, qy, z)
    assert np.abs(dh).max() < 1e-15
    assert np.abs(dqx).max() < 1e-14 and np.abs(dqy).max() < 1e-14


def test_hllc_supercritical_flux_equals_physical_flux(oracle_lib):
    # both states moving right faster than the wave speed: the edge flux is
    # the left physical flux F = (hu, hu^2 + g h^2 / 2, huv); a single open
    # east edge then shows dh = -(F_in - F_out)/dx*dt with F_out = F(ghost = cell)
    n = 6
    mask = gen.ring_mask(1, n, wall=50)
    h_in = np.array([[1.0, 1.0, 1.0, 0.8, 0.8, 0.8]])
    u = 6.0  # > sqrt(g h) ~ 3.1
    qx = h_in * u
    v = 0.25
    qy = h_in * v
    mask, h, qxr, qyr, z = box(1, n, wall=50, h=h_in, qx=qx, qy=qy)
    dt, dx = 0.01, 1.0
    dh, dqx, dqy, _ = one_flux(oracle_lib, 1, mask, h, qxr, qyr, z, dt=dt, dx=dx)
    # cell 3 receives F(h=1) from the west and emits F(h=0.8) to the east
    F = lambda hh: np.array([hh * u, hh * u * u + 0.5 * G * hh * hh, hh * u * v])
    expect = -(F(0.8) - F(1.0)) / dx * dt
    assert abs(dh[3] - expect[0]) < 1e-13
    assert abs(dqx[3] - expect[1]) < 1e-12
    assert abs(dqy[3] - expect[2]) < 1e-13


@pytest.mark.parametrize("order", [1, 2])
def test_mass_conservation_closed_box(oracle_lib, order):
    z, h, qx, qy = random_state(16, 18, seed=11)
    mask, h, qx, qy, z = box(16, 18, wall=40, z=z, h=h, qx=qx, qy=qy)
    dh, _, _, h_c = one_flux(oracle_lib, order, mask, h, qx, qy, z)
    assert abs(dh.sum()) < 1e-13 * np.abs(dh).sum()


@pytest.mark.parametrize("order", [1, 2])
def test_rotation_symmetry(oracle_lib, order):
    z, h, qx, qy = random_state(9, 13, seed=5)
    mask, H, QX, QY, Z = box(9, 13, wall=40, z=z, h=h, qx=qx, qy=qy)
    dh, dqx, dqy, _ = one_flux(oracle_lib, order, mask, H, QX, QY, Z)
    # rotate the picture 90 degrees counter-clockwise: (qx, qy) -> (-qy, qx)
    r = np.rot90
    m2, H2, Z2 = r(mask).copy(), r(H).copy(), r(Z).copy()
    QX2, QY2 = r(-QY).copy(), r(QX).copy()
    dh2, dqx2, dqy2, _ = one_flux(oracle_lib, order, m2, H2, QX2, QY2, Z2)
    R = lambda v: to_raster(mask, v)
    R2 = lambda v: to_raster(m2, v)
    assert rel_err(R2(dh2), r(R(dh))) < 1e-12
    assert rel_err(R2(dqx2), r(-R(dqy))) < 1e-12
    assert rel_err(R2(dqy2), r(R(dqx))) < 1e-12


def test_slip_wall_has_no_mass_flux_and_open_edge_only_drains(oracle_lib):
    n = 8
    h_in = np.full((n, n), 1.0)
    qx = np.full((n, n), 0.5)  # flow towards the east edge
    # closed (slip) east edge: interior flow, zero net mass change
    mask, h, q1, q2, z = box(n, n, wall=40, h=h_in, qx=qx)
    dh, _, _, _ = one_flux(oracle_lib, 1, mask, h, q1, q2, z)
    assert abs(dh.sum()) < 1e-14
    # open (transmissive) east edge: the ghost copies the cell, so the edge
    # passes the physical mass flux h*u = 0.5 of each of the n boundary rows
    mask, h, q1, q2, z = box(n, n, wall=40, east=50, h=h_in, qx=qx)
    dt, dx = 0.01, 1.0
    dh, _, _, _ = one_flux(oracle_lib, 1, mask, h, q1, q2, z, dt=dt, dx=dx)
    assert abs(dh.sum() - (-n * 0.5 * dt / dx)) < 1e-14


def test_rigid_wall_is_transmissive_on_north_south_edges(oracle_lib):
    # the reference's type-3 ghost reverses q only on E/W edges; on N/S edges
    # its expressions are the identity, so northward flow leaves the domain
    n = 6
    qy = np.full((n, n), 0.4)
    mask, h, qx, qyr, z = box(n, n, wall=30, h=np.ones((n, n)), qy=qy)
    dh30, _, _, _ = one_flux(oracle_lib, 1, mask, h, qx, qyr, z)
    mask4, *_ = box(n, n, wall=40)
    dh40, _, _, _ = one_flux(oracle_lib, 1, mask4, h, qx, qyr, z)
    rows = np.nonzero(mask > 0)[0]
    north = rows == rows.min()
    assert abs(dh40.sum()) < 1e-14           # slip walls close the box:
    assert dh40[north].sum() > 1e-6          # northward flow piles up at the N wall
    assert np.abs(dh30).max() < 1e-15        # "rigid" walls let the uniform flow through


def test_h_given_boundary_reads_its_own_column(oracle_lib):
    # code 61: type 6 (H given), series id 1 -> column 2 of given_depth
    n = 5
    mask = gen.ring_mask(n, n, wall=40)
    mask[1:-1, 1] = 61
    z = gen.pad(np.zeros((n, n)))
    zero = gen.pad(np.zeros((n, n)))
    gd_a = [[0.0, 0.0, 1.5], [1e9, 0.0, 1.5]]
    gd_b = [[0.0, 0.0, 0.0], [1e9, 1.5, 0.0]]
    dh_a, *_ = one_flux(oracle_lib, 1, mask, zero, zero, zero, z, t=0.0, gd=gd_a)
    dh_b, *_ = one_flux(oracle_lib, 1, mask, zero, zero, zero, z, t=0.0, gd=gd_b)
    west = compact(mask, mask == 61, dtype=bool)
    assert dh_a[west].min() > 1e-6          # 1.5 m given at the west edge flows in
    assert np.abs(dh_b).max() == 0.0        # column 1 is not this boundary's series


def test_flux_mask_cell0_and_t0_rules(oracle_lib):
    n = 5
    h_in = np.zeros((n, n))
    h_in[0, 0] = 1.0  # compact cell 0 (first valid cell in row-major order)
    mask, h, qx, qy, z = box(n, n, wall=40, h=h_in)
    idx = oracle_lib.build_index(mask)
    h_c = compact(mask, h)
    listed = oracle_lib.flux_list(h_c, idx, t=3.0)
    assert listed.tolist() == [0]           # its dry neighbours 1 and n are not listed
    assert oracle_lib.flux_list(h_c, idx, t=0.0).size == h_c.size
    h_in2 = np.zeros((n, n))
    h_in2[0, 1] = 1.0                       # compact cell 1 does count as a wet neighbour
    mask, h2, *_ = box(n, n, wall=40, h=h_in2)
    assert set(oracle_lib.flux_list(compact(mask, h2), idx, t=3.0).tolist()) == {0, 1, 2, n + 1}


def test_friction_matches_closed_form_root(oracle_lib):
    h, n, dt = 0.7, 0.05, 0.5
    qx0, qy0 = 0.9, -0.4
    q = math.hypot(qx0, qy0)
    a = dt * G * n * n * h ** (-7.0 / 3.0)
    # implicit Manning: |q'| (1 + a |q'|) = |q|, direction kept
    qn = brentq(lambda s: s * (1 + a * s) - q, 0.0, q)
    arrs = [np.array([v]) for v in (0.0, 0.0, 0.0, 0.0)]
    hh, wl, qx, qy, zz = (np.array([v]) for v in (h, h, qx0, qy0, 0.0))
    oracle_lib.friction_update(np.array([0], np.int32), *arrs, np.zeros(1, np.uint8), hh, wl, qx, qy, zz,
                               np.array([n]), dt)
    assert abs(qx[0] - qx0 * qn / q) < 1e-13
    assert abs(qy[0] - qy0 * qn / q) < 1e-13
    assert abs(math.hypot(qx[0], qy[0]) - qn) < 1e-13


def test_friction_zero_manning_dry_cell_and_limiter(oracle_lib):
    def run(h, qx0, qy0, n, dt=1.0, dh=0.0):
        hh, wl, qx, qy, zz = (np.array([v]) for v in (h, h, qx0, qy0, 0.0))
        oracle_lib.friction_update(np.array([0], np.int32), np.array([dh]), np.zeros(1), np.zeros(1), np.zeros(1),
                                   np.zeros(1, np.uint8), hh, wl, qx, qy, zz, np.array([n]), dt)
        return hh[0], qx[0], qy[0]
    assert run(0.5, 0.3, 0.2, 0.0)[1:] == (0.3, 0.2)            # n = 0: unchanged
    assert run(5e-7, 0.3, 0.2, 0.03)[1:] == (0.0, 0.0)           # h < 1e-6: momentum zeroed
    hn, qxn, _ = run(0.5, 0.3, 0.2, 0.03, dh=-0.5 + 1e-8)        # dries during the update
    assert qxn == 0.0 and hn < 1e-6
    # supercritical: the limiter branch returns 2*temp - q clipped at 0,
    # never reversing the flow and never exceeding the implicit value
    h, q0, n, dt = 0.01, 3.0, 0.2, 0.01
    a = dt * G * n * n * h ** (-7.0 / 3.0)
    temp = brentq(lambda s: s * (1 + a * s) - q0, 0.0, q0)
    assert temp > 5 * h * math.sqrt(G * h)
    _, qx_l, _ = run(h, q0, 0.0, n, dt)
    assert 0.0 <= qx_l <= temp


def test_cfl_candidates(oracle_lib):
    h = np.array([4.0 / G, 1.0 / G, 0.0, 1e-7])
    qx = np.array([-3.0 * h[0], 0.5 * h[1], 0.0, 0.0])
    qy = np.array([1.0 * h[0], -2.0 * h[1], 0.0, 0.0])
    wet = np.arange(4, dtype=np.int32)
    accel = np.zeros(4)
    hmax = np.array([0.1, 10.0, 0.0, 0.0])
    oracle_lib.time_control(wet, accel, hmax, h, qx, qy, dx=2.0, cfl=0.5)
    # dt = CFL dx / (sqrt(g h) + max|u|, |v|):  1/(2+3),  1/(1+2)
    assert abs(accel[0] - 1.0 / 5.0) < 1e-7 and abs(accel[1] - 1.0 / 3.0) < 1e-7
    assert accel[2] == 60.0 and accel[3] == 60.0
    assert hmax.tolist() == [h[0], 10.0, 0.0, 1e-7]


def test_station_rain_volume_and_interpolation(oracle_lib):
    # flat closed box, no water: each step adds r(t) * dt to every cell, with
    # t the driver time before the step (which already includes this dt)
    n = 6
    mask, h, qx, qy, z = box(n, n, wall=40)
    table = np.array([[0.0, 1e-3], [10.0, 3e-3], [20.0, 0.0]])
    sim = oracle_lib.OracleSim(mask, h, qx, qy, z, dx=1.0, order=1, first_dt=0.5, rain_table=table,
                               station=np.zeros_like(mask), manning=[0.0])
    expect = 0.0
    for _ in range(40):
        t_before, dt_before = sim.time, sim.dt
        sim.run(1)
        r = np.interp(t_before, table[:, 0], table[:, 1]) if t_before < table[-1, 0] else 0.0
        expect += r * dt_before
        assert abs(sim.h_c - expect).max() < 1e-14
    assert sim.time > 20.0  # ran past the table: rain stopped


def test_dt_growth_cap_and_export_clamp(oracle_lib):
    n = 4
    mask, h, qx, qy, z = box(n, n, wall=40, h=np.full((n, n), 1.0))
    sim = oracle_lib.OracleSim(mask, h, qx, qy, z, dx=10.0, order=1, first_dt=1e-3, export_T=1.0)
    sim.run(1)
    assert abs(sim.dt - 1e-2) < 1e-15          # CFL step ~1.6 s, capped at 10 x 1e-3
    sim.run(2)                                  # 1e-1, then clamped to land on t = 1.0
    assert abs(sim.time - 1.0) < 1e-15 and sim.exports == 1


def stoker(x, t, hl, hr, x0):
    """Exact wet-bed dam-break solution (Stoker), frictionless, flat bed."""
    cl, cr = math.sqrt(G * hl), math.sqrt(G * hr)
    f = lambda hm: 2 * (cl - math.sqrt(G * hm)) - (hm - hr) * math.sqrt(0.5 * G * (1 / hm + 1 / hr))
    hm = brentq(f, hr, hl)
    cm = math.sqrt(G * hm)
    um = 2 * (cl - cm)
    s = hm * um / (hm - hr)
    xi = (x - x0) / t
    h = np.where(xi < -cl, hl, np.where(xi < um - cm, (2 * cl - xi) ** 2 / (9 * G), np.where(xi < s, hm, hr)))
    return h, s


def stoker_run(orc, order):
    """Run the oracle on a 400 x 3 flume to ~18 s; return (L1 error, front error)."""
    nx, ny, x0, hl, hr = 400, 3, 200, 2.0, 0.5
    h_in = np.where(np.arange(nx)[None, :] < x0, hl, hr) * np.ones((ny, 1))
    mask, h, qx, qy, z = box(ny, nx, wall=40, h=h_in)
    first = 0.5 / math.sqrt(G * hl)
    sim = orc.OracleSim(mask, h, qx, qy, z, dx=1.0, order=order, first_dt=first, manning=[0.0])
    while sim.time - sim.dt + first < 18.0:  # t advances by the next dt (reference driver)
        sim.run(1)
    elapsed = sim.time - sim.dt + first
    xc = np.arange(nx) + 0.5
    exact, s = stoker(xc, elapsed, hl, hr, x0)
    got = sim.raster("h_c")[1:-1, 1:-1].mean(axis=0)
    l1 = np.abs(got - exact).sum() / np.abs(exact).sum()
    hm = exact[int(x0 + 0.5 * s * elapsed)]
    front = xc[np.nonzero(got > 0.5 * (hr + hm))[0].max()]
    return l1, abs(front - (x0 + s * elapsed))


@pytest.mark.parametrize("order,tol", [(1, 0.012), (2, 0.005)])
def test_stoker_dam_break(oracle_lib, order, tol):
    l1, front_err = stoker_run(oracle_lib, order)
    assert l1 < tol
    assert front_err < 2.0


def test_order2_more_accurate_than_order1(oracle_lib):
    assert stoker_run(oracle_lib, 2)[0] < stoker_run(oracle_lib, 1)[0]


def test_closed_box_volume_conserved_over_many_steps(oracle_lib):
    z, h, qx, qy = random_state(20, 24, seed=3, rough=0.2)
    mask, h, qx, qy, z = box(20, 24, wall=40, z=z, h=h, qx=qx, qy=qy)
    sim = oracle_lib.OracleSim(mask, h, qx, qy, z, dx=2.0, order=2, first_dt=1e-3, manning=[0.03])
    v0 = sim.h_c.sum()
    sim.run(200)
    assert abs(sim.h_c.sum() - v0) < 1e-11 * v0
    assert sim.h_c.min() >= 0.0
