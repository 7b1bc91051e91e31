"""Writes the golden fixtures in this directory: small inputs of the hot path
and the oracle's FP64 outputs after a fixed number of server-sequence steps
(addFlux -> station rain -> friction/update/time control).

The reference ships no test vectors for this path, so the expected outputs
come from the oracle (its C++ restatement, pinned by the closed-form tests);
the fixtures freeze them so that later changes to the oracle or to the HIP
kernels are checked against the same numbers.  Each fixture holds its inputs
(rasters, tables, scalars) as plain arrays, so the tests rebuild every case
without the generators.

    python tests/golden/make_golden.py      (from the repository root)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "chans-hipims-flamegpu_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
from hipims_amd import generators as gen  # noqa: E402


def ring(ny, nx, wall=30):
    return gen.ring_mask(ny, nx, wall=wall)


def dam_break():
    c = gen.dam_break(32, wall=30)
    c.h[10:14, 5:9] += 0.3
    return dict(mask=c.mask, h=c.h, qx=c.qx, qy=c.qy, z=c.z, dx=c.dx, cfl=c.cfl, first_dt=c.first_dt, order=2,
                manning=c.manning, landuse=c.landuse, steps=40)


def runoff():
    c = gen.rainfall_runoff(40, 5.0, rain_mmh=300.0, relief=3.0)
    return dict(mask=c.mask, h=c.h, qx=c.qx, qy=c.qy, z=c.z, dx=c.dx, cfl=c.cfl, first_dt=1.0, order=1,
                manning=c.manning, landuse=c.landuse, rain_table=c.rain_table, station=np.zeros_like(c.mask),
                export_T=60.0, steps=60)


def boundary_mix():
    rng = np.random.Generator(np.random.PCG64(9))
    ny, nx = 24, 31
    z = 0.3 * rng.standard_normal((ny, nx))
    h = np.where(rng.uniform(0, 1, (ny, nx)) < 0.5, 0.0, rng.uniform(0, 1, (ny, nx)))
    qx = 0.3 * rng.standard_normal((ny, nx)) * (h > 0)
    qy = 0.3 * rng.standard_normal((ny, nx)) * (h > 0)
    mask = ring(ny, nx, wall=30)
    mask[1:-1, 1] = 61
    mask[1, 2:8] = 70
    mask[1, 8:14] = 80
    mask[1, 14:20] = 9
    mask[-2, 2:14] = 50
    mask[-2, 14:24] = 40
    return dict(mask=mask, h=gen.pad(h), qx=gen.pad(qx), qy=gen.pad(qy), z=gen.pad(z), dx=1.0, cfl=0.5,
                first_dt=0.01, order=2, manning=np.array([0.025]), landuse=np.zeros_like(mask, dtype=np.int64),
                given_depth=np.array([[0.0, 0.0, 0.8], [5.0, 0.0, 1.2], [1e9, 0.0, 0.4]]),
                given_wl=np.array([[0.0, 0.9], [1e9, 1.1]]),
                given_q=np.array([[0.0, 0.3, -0.6], [3.0, -0.2, -0.9], [1e9, 0.0, 0.0]]), steps=30)


def stations():
    n = 24
    h = np.zeros((n, n))
    h[0, 0] = 0.8  # compact cell 0 wet, its neighbours dry
    h[10:14, 10:14] = 0.5
    mask = ring(n, n, wall=40)
    st = np.zeros_like(mask)
    st[:, n // 2:] = 1
    st[n // 2:, :] += 2
    table = np.array([[0.0, 1e-4, 2e-4, 0.0, 5e-5], [30.0, 3e-4, 0.0, 1e-4, 5e-5], [60.0, 0.0, 1e-4, 2e-4, 0.0]])
    zero = np.zeros((n, n))
    return dict(mask=mask, h=gen.pad(h), qx=gen.pad(zero), qy=gen.pad(zero), z=gen.pad(zero), dx=1.0, cfl=0.5,
                first_dt=0.05, order=2, manning=np.array([0.03]), landuse=np.zeros_like(mask, dtype=np.int64),
                rain_table=table, station=st, steps=30)


CASES = {"dam_break_32_order2": dam_break, "rainfall_runoff_40_order1": runoff,
         "boundary_mix_24x31_order2": boundary_mix, "cell0_stations_24_order2": stations}


def oracle_run(inp):
    """OracleSim over the stored inputs; per-step (t, dt) and the final state."""
    o = oracle.OracleSim(inp["mask"], inp["h"], inp["qx"], inp["qy"], inp["z"], float(inp["dx"]),
                         cfl=float(inp["cfl"]), order=int(inp["order"]), first_dt=float(inp["first_dt"]),
                         export_T=float(inp.get("export_T", 1.0e30)), manning=inp["manning"],
                         landuse=inp["landuse"], rain_table=inp.get("rain_table"), station=inp.get("station"),
                         given_depth=inp.get("given_depth"), given_wl=inp.get("given_wl"),
                         given_q=inp.get("given_q"))
    sched = []
    for _ in range(int(inp["steps"])):
        sched.append((o.time, o.dt))
        o.run(1)
    return {"sched": np.array(sched), "t_end": np.float64(o.time), "dt_end": np.float64(o.dt),
            "h_out": o.h_c.copy(), "qx_out": o.qx_c.copy(), "qy_out": o.qy_c.copy(), "h_max_out": o.h_max.copy()}


def main():
    oracle.build()
    for name, make in CASES.items():
        inp = make()
        out = oracle_run(inp)
        arrays = {k: np.asarray(v) for k, v in inp.items()}
        arrays.update(out)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **arrays)
        print(f"{path}: {out['h_out'].size} cells, t {float(out['t_end']):.4f} s")


if __name__ == "__main__":
    main()
