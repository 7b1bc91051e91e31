"""Throughput of the HiPIMS shallow-water step on one MI355X.

Workload (default): rainfall-runoff on a seeded 2048 x 2048 fractal DEM,
dx = 5 m, FP32 state, MUSCL (order 2) + HLLC + implicit Manning friction,
6 Voronoi landcover classes, 50 mm/h uniform rain, open east edge.  Inputs are
synthetic (seeded generators), resident in HBM before timing starts.

One step = flux pass + update pass (rain, friction, update, h_max, CFL) +
device time control.  `value` = valid cells x steps / seconds (cell-updates/s).
Per-kernel durations come from HIP events recorded on the launch stream
between the kernels of the timed steps; the roofline uses the algorithmic
bytes per cell of the dominant kernel.

The CPU baseline times the oracle (the reference's update logic restated in
C++/OpenMP on the reference's compact layout) on the host cores for a bounded
number of steps of the same workload, starting from the GPU's warmed state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "chans-hipims-flamegpu_amd"))

METRIC = "cell-updates/s & % HBM peak, 2048² fp32; ensemble + adjoint speedup vs plain"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-update of each pass (state planes read/written
# once per pass; class bytes; no credit for cache reuse):
#   flux:   read h, qx, qy, z + type byte; write dh, dqx, dqy
#   update: read dh, dqx, dqy, h, qx, qy, h_max + class byte (+ rain field);
#           write h, qx, qy, h_max
#   fused:  read h, qx, qy, z, h_max + type and class bytes (+ rain field);
#           write h, qx, qy, h_max
def alg_bytes(esize: int, rain_field: bool) -> dict:
    flux = 4 * esize + 1 + 3 * esize
    update = 7 * esize + 1 + 4 * esize + (4 if rain_field else 0)
    fused = 5 * esize + 2 + 4 * esize + (4 if rain_field else 0)
    return {"k_flux": flux, "k_update": update, "k_step_fused": fused}


def time_two_pass(g, steps):
    """Two-pass step (flux kernel; update + finalize), HIP events between the
    kernels on the launch stream."""
    import torch
    stream = torch.cuda.current_stream(g.device)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall = time.perf_counter()
    start.record(stream)
    for k in range(steps):
        ev[k][0].record(stream)
        g.flux()
        ev[k][1].record(stream)
        g.update()
        ev[k][2].record(stream)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    return {"path": "two_pass", "steps": steps, "total_ms": start.elapsed_time(stop), "wall_s": wall,
            "kernels_ms": {"k_flux": sum(e[0].elapsed_time(e[1]) for e in ev) / steps,
                           "k_update": sum(e[1].elapsed_time(e[2]) for e in ev) / steps}}


def time_fused(g, steps):
    """Single-pass step (fused kernel + finalize), two steps per call so the
    state returns to its own buffers; events around each call."""
    import torch
    stream = torch.cuda.current_stream(g.device)
    pairs = max(1, steps // 2)
    g.step_fused(2)  # allocate the second buffers outside the timed region
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(pairs + 1)]
    torch.cuda.synchronize()
    wall = time.perf_counter()
    ev[0].record(stream)
    for k in range(pairs):
        g.step_fused(2)
        ev[k + 1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    total = ev[0].elapsed_time(ev[-1])
    return {"path": "fused", "steps": 2 * pairs, "total_ms": total, "wall_s": wall,
            "kernels_ms": {"k_step_fused": total / (2 * pairs)}}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--size", type=int, default=2048)
    p.add_argument("--dtype", choices=["fp32", "fp64"], default="fp32")
    p.add_argument("--order", type=int, choices=[1, 2], default=2)
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    p.add_argument("--cpu-threads", type=int, default=0, help="0: min(16, available cores)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--mode", choices=["step", "copy"], default="step",
                   help="copy: a plain device copy of known size, the reference point for HBM byte counters")
    p.add_argument("--copy-mib", type=int, default=1024)
    return p.parse_args()


def copy_calibration(a):
    """Plain float32 device-to-device copies of --copy-mib MiB: each copy
    reads and writes exactly that many bytes."""
    import torch
    dev = torch.device("cuda", 0)
    n = a.copy_mib * (1 << 20) // 4
    src = torch.rand(n, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(a.steps):
        dst.copy_(src)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / a.steps
    print(json.dumps({"mode": "copy", "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4,
                      "ms_per_copy": ms, "GBps_read_plus_write": 2 * n * 4 / (ms / 1e3) / 1e9}), flush=True)


def main():
    a = parse()
    if a.mode == "copy":
        return copy_calibration(a)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described, not implemented)"}))
        sys.exit(2)
    import numpy as np
    import torch

    from hipims_amd import generators as gen
    from hipims_amd.grid import GridSolver

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    dtype = torch.float32 if a.dtype == "fp32" else torch.float64
    esize = 4 if dtype == torch.float32 else 8
    case = gen.rainfall_runoff(a.size, 5.0)
    g = GridSolver(case.mask, case.h, case.qx, case.qy, case.z, case.dx, cfl=case.cfl, order=a.order, dtype=dtype,
                   device=dev, first_dt=case.first_dt, manning=case.manning, landuse=case.landuse)
    g.set_station_rain(case.rain_table)
    cells = case.n_valid

    g.step(a.warmup)
    torch.cuda.synchronize()
    two = time_two_pass(g, a.steps)
    fused = time_fused(g, a.steps)
    ctrl = g.ctrl_dict()
    assert np.isfinite(ctrl["t"]) and ctrl["dt"] > 0

    B = alg_bytes(esize, rain_field=False)
    best = fused if fused["total_ms"] < two["total_ms"] else two
    ms_step = best["total_ms"] / best["steps"]
    value = cells * best["steps"] / (best["total_ms"] / 1e3)
    per_kernel = best["kernels_ms"]
    dom = max(per_kernel, key=per_kernel.get)
    achieved = B[dom] * cells / (per_kernel[dom] / 1e3) / 1e9
    two_pass = (B["k_flux"] + B["k_update"]) * value / 1e9
    t_wall = best["wall_s"]

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a, g, case, dtype)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if esize == 4 else "f64",
        "data": "synthetic (seeded fractal DEM, Voronoi landcover, uniform rain)",
        "config": {"workload": f"rainfall_runoff_{a.size}x{a.size}_dx5_{a.dtype}_order{a.order}",
                   "cells": cells, "dx_m": 5.0, "order": a.order, "rain_mm_per_h": 50.0,
                   "sim_time_s": ctrl["t"]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "alg_bytes_per_cell": B[dom], "kernel_ms": per_kernel[dom]},
        "kernels_ms": per_kernel,
        "path": best["path"],
        "paths_ms_per_step": {"two_pass": two["total_ms"] / two["steps"], "fused": fused["total_ms"] / fused["steps"]},
        "two_pass_alg_GBps": two_pass,
        "two_pass_hbm_frac": two_pass / HBM_PEAK_GBS,
        "wall_s": t_wall,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def cpu_baseline(a, g, case, dtype):
    """Oracle (C++/OpenMP, the reference's compact layout and op sequence) on
    host cores, from the GPU's current state, for a bounded step count."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    orc.build()
    threads = a.cpu_threads or min(16, len(os.sched_getaffinity(0)))
    orc.set_threads(threads)
    np_dt = np.float32 if dtype == torch.float32 else np.float64
    sim = orc.OracleSim(case.mask, case.h, case.qx, case.qy, case.z, case.dx, cfl=case.cfl, order=a.order,
                        dtype=np_dt, manning=case.manning, landuse=case.landuse, rain_table=case.rain_table,
                        station=np.zeros_like(case.mask))
    sim.h_c[:] = g.compact(g.h).cpu().numpy()
    sim.qx_c[:] = g.compact(g.qx).cpu().numpy()
    sim.qy_c[:] = g.compact(g.qy).cpu().numpy()
    sim.wl_c[:] = sim.h_c + sim.z_c
    sim.ctrl[0], sim.ctrl[1] = g.t, g.dt
    sim.run(1)  # warm-up
    probe = sim.last_seconds
    n = max(2, int(a.cpu_seconds / max(probe, 1e-3)))
    sim.run(n)
    secs = sim.last_seconds
    return {"value": sim.N * n / secs, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the same {a.size}x{a.size} workload from the GPU's warmed state "
                      f"({secs:.1f} s, oracle C++/OpenMP, {np_dt.__name__})"}


if __name__ == "__main__":
    main()
