"""The oracle's frozen-slope loss (a test construct for the order-2 adjoint
checks, not part of the reference's algorithm) against plain oracle runs.

With the perturbed run's parameters equal to the source's, reconstructing
with the source's slopes is reconstructing with its own: the trajectory and
the loss must equal a plain order-2 run's on the same schedule bit for bit
(a wrong neighbour, a swapped side or a dropped momentum rule in the
slope-source branch breaks the equality; the loss sums the same squares in
another order).  With different parameters the slopes stay the source's,
which makes the loss smooth in the parameters where the plain order-2 loss
is not."""
import numpy as np

from hipims_amd import generators as gen


def _sims(orc, c, n_scale=1.0, rain_scale=1.0):
    rt = c.rain_table.copy()
    rt[:, 1:] *= rain_scale
    return orc.OracleSim(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=2, first_dt=0.5,
                         manning=c.manning * n_scale, landuse=c.landuse, rain_table=rt,
                         station=np.zeros_like(c.mask))


def _plain(orc, c, sched, gauges, obs, **kw):
    o = _sims(orc, c, **kw)
    loss = 0.0
    for k in range(sched.shape[0]):
        o.ctrl[0:2] = sched[k]
        o.run(1)
        d = o.h_c[gauges] - obs[k]
        loss += float((d * d).sum())
    return loss, o


def _case():
    return gen.rainfall_runoff(40, 5.0, seed=77, rain_mmh=150.0, relief=6.0)


def _schedule(orc, c, steps):
    o = _sims(orc, c)
    sched = np.zeros((steps, 2))
    for k in range(steps):
        sched[k] = o.ctrl[0:2]
        o.run(1)
    return sched


def test_frozen_slope_loss_with_own_parameters_is_the_plain_run(oracle_lib):
    c = _case()
    sched = _schedule(oracle_lib, c, 150)
    rng = np.random.Generator(np.random.PCG64(3))
    n_valid = int((c.mask > 0).sum())
    gauges = rng.choice(n_valid, size=12, replace=False).astype(np.int32)
    obs = rng.uniform(0.0, 0.01, size=(150, 12))
    plain, o = _plain(oracle_lib, c, sched, gauges, obs)
    src, pert = _sims(oracle_lib, c), _sims(oracle_lib, c)
    frozen = oracle_lib.frozen_slope_loss(src, pert, sched, gauges, obs)
    for name in ("h_c", "qx_c", "qy_c"):
        assert np.array_equal(getattr(pert, name), getattr(o, name)), name
        assert np.array_equal(getattr(src, name), getattr(o, name)), name
    # the same squared misfits, summed in another order
    assert abs(frozen - plain) <= 1e-13 * plain
    assert (o.h_c > 1e-6).sum() > n_valid // 2


def test_frozen_slope_loss_is_smooth_in_the_parameters(oracle_lib):
    """The plain order-2 loss jumps as limiter branches switch (its one-sided
    differences at d = 1e-2, 2.5e-3, 6e-4 disagree by tens of percent on this
    case); with the slopes held at the source's, the loss is a smooth
    function of the Manning scale: its one-sided differences agree to 3 %."""
    c = _case()
    sched = _schedule(oracle_lib, c, 120)
    rng = np.random.Generator(np.random.PCG64(4))
    n_valid = int((c.mask > 0).sum())
    gauges = rng.choice(n_valid, size=10, replace=False).astype(np.int32)
    obs = np.zeros((120, 10))
    base, _ = _plain(oracle_lib, c, sched, gauges, obs)
    slopes_frozen, slopes_plain = [], []
    for d in (1e-2, 2.5e-3, 6e-4):
        plain, _ = _plain(oracle_lib, c, sched, gauges, obs, n_scale=1.0 + d)
        frozen = oracle_lib.frozen_slope_loss(_sims(oracle_lib, c), _sims(oracle_lib, c, n_scale=1.0 + d),
                                              sched, gauges, obs)
        slopes_frozen.append((frozen - base) / d)
        slopes_plain.append((plain - base) / d)
    sf, sp = np.array(slopes_frozen), np.array(slopes_plain)
    assert np.ptp(sf) <= 0.03 * np.abs(sf).max()
    assert np.ptp(sp) > 0.1 * np.abs(sp).max()


def test_order2_trajectory_amplifies_rounding_order1_does_not(oracle_lib):
    """Why the order-2 gradient is taken with frozen slopes: on the same
    schedule, a perturbation of 1e-15 m in every seventh depth grows to more
    than 1e-4 of the largest depth within 300 order-2 steps (MUSCL-minmod
    with forward Euler switches limiter branches under it), while order 1
    keeps it below 1e-12.  The exact derivative of an order-2 run therefore
    describes a trajectory no two platforms share beyond a few hundred
    steps."""
    c = _case()
    worst = {}
    for order in (1, 2):
        def sim():
            return oracle_lib.OracleSim(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=order, first_dt=0.5,
                                        manning=c.manning, landuse=c.landuse, rain_table=c.rain_table,
                                        station=np.zeros_like(c.mask))
        base, pert = sim(), sim()
        pert.h_c[:] += 1e-15 * (np.arange(pert.N) % 7 == 0)
        for _ in range(300):
            pert.ctrl[0:2] = base.ctrl[0:2]
            base.run(1)
            pert.run(1)
        worst[order] = float(np.abs(pert.h_c - base.h_c).max() / base.h_c.max())
    assert worst[1] < 1e-12 and worst[2] > 1e-4, worst


def test_frozen_slope_model_has_a_kink_at_the_base_state(oracle_lib):
    """From a wet state (300 oracle steps on a 96^2 case), three steps, the
    loss at one gauge: the forward and backward differences of the
    frozen-slope loss in the first Manning coefficient differ by 0.4 % at
    every eps from 1e-5 to 1e-7 (a kink exactly at the unperturbed
    parameters, where central differences return the mean of the two
    sides), while over two steps the difference shrinks with eps (smooth)."""
    c = gen.rainfall_runoff(96, 5.0, seed=4242, rain_mmh=150.0, relief=8.0)
    o = oracle_lib.OracleSim(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=0.5, order=2, first_dt=0.5,
                             manning=c.manning, landuse=c.landuse, rain_table=c.rain_table,
                             station=np.zeros_like(c.mask))
    o.run(300)
    h0, qx0, qy0, dt0 = o.raster("h_c", 0.0), o.raster("qx_c", 0.0), o.raster("qy_c", 0.0), o.dt
    n_classes = len(c.manning)
    th0 = np.concatenate([c.manning, [1.0]])

    def sim(theta):
        rt = c.rain_table.copy()
        rt[:, 1:] *= theta[n_classes]
        return oracle_lib.OracleSim(c.mask, h0, qx0, qy0, c.z, c.dx, cfl=0.5, order=2, first_dt=dt0,
                                    manning=theta[:n_classes], landuse=c.landuse, rain_table=rt,
                                    station=np.zeros_like(c.mask))
    compact_of = np.cumsum(c.mask > 0).reshape(c.mask.shape) - 1
    gauge = np.array([compact_of[71, 18]], dtype=np.int32)
    spreads = {}
    for K in (2, 3):
        base = sim(th0)
        sched = np.zeros((K, 2))
        for k in range(K):
            sched[k] = base.ctrl[0:2]
            base.run(1)
        obs = np.zeros((K, 1))
        L0 = oracle_lib.frozen_slope_loss(sim(th0), sim(th0), sched, gauge, obs)
        for eps in (1e-5, 1e-7):
            d = eps * th0[0]
            tp, tm = th0.copy(), th0.copy()
            tp[0] += d
            tm[0] -= d
            fwd = (oracle_lib.frozen_slope_loss(sim(th0), sim(tp), sched, gauge, obs) - L0) / d
            bwd = (L0 - oracle_lib.frozen_slope_loss(sim(th0), sim(tm), sched, gauge, obs)) / d
            spreads[K, eps] = abs(fwd - bwd) / abs(fwd)
    assert spreads[3, 1e-5] > 3e-3 and spreads[3, 1e-7] > 3e-3, spreads
    assert spreads[2, 1e-7] < 0.1 * spreads[2, 1e-5] and spreads[2, 1e-7] < 1e-5, spreads
