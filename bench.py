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


def kernel_names(esize: int) -> dict:
    """The kernel behind each pass: float domains run both the flux pass and
    the single-pass step as column-walk kernels (k_walk, its flux-pass and
    step instances), double domains the LDS-tile kernels."""
    if esize == 4:
        return {"k_flux": "k_walk_flux", "k_update": "k_update", "k_step_fused": "k_walk_step"}
    return {"k_flux": "k_flux", "k_update": "k_update", "k_step_fused": "k_step_fused"}


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` on `workload` as recorded from the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (copy-calibrated) in profiles/, or
    None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| measured traffic per launch |"):
            return float(line.split("|")[2].split()[0]) * 1e6
    return None


def profiled_valu(kernel: str, workload: str):
    """VALU lane-instructions per cell of `kernel` on `workload` from the
    rocprofv3 SQ_INSTS_VALU run summarised in profiles/, or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("| SQ_INSTS_VALU |") and "(" in line:
            return float(line.split("(")[1].split()[0])
    return None


# Non-packed VALU issue rate of MI355X at its peak clock: 256 CUs x 64 lanes
# per clock x 2.4 GHz; the rate measured by `--mode valu` (profiles/) is used
# when present.
VALU_LANE_OPS_PER_S = 256 * 64 * 2.4e9


def valu_peak():
    """(lane-instructions/s, source): the FMA probe's measured rate from
    profiles/valu_probe_fma.md, else the peak-clock figure."""
    path = os.path.join(ROOT, "profiles", "valu_probe_fma.md")
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("| FMA lane-instructions per second |"):
                return float(line.split("|")[2].split()[0]), "measured (profiles/valu_probe_fma.md)"
    return VALU_LANE_OPS_PER_S, "peak clock"


def cached_case(size: int):
    """The roofline case (slow to generate at 8192^2), kept in the temporary
    directory between runs of one session."""
    import tempfile

    import numpy as np

    from hipims_amd import generators as gen
    path = os.path.join(tempfile.gettempdir(), f"hipims_roofline_case_{size}.npz")
    if os.path.exists(path):
        with np.load(path, allow_pickle=False) as f:
            c = gen.Case(f"roofline_{size}", f["mask"], f["h"], f["qx"], f["qy"], f["z"], float(f["dx"]),
                         f["landuse"], f["manning"], float(f["cfl"]), float(f["first_dt"]), rain_table=None,
                         rain_field=f["rain_field"])
        return c
    c = gen.roofline_case(size, 2.0)
    np.savez(path, mask=c.mask, h=c.h, qx=c.qx, qy=c.qy, z=c.z, dx=c.dx, landuse=c.landuse, manning=c.manning,
             cfl=c.cfl, first_dt=c.first_dt, rain_field=c.rain_field)
    return c


def workload_name(a) -> str:
    if a.case == "runoff":
        return f"rainfall_runoff_{a.size}x{a.size}_dx5_{a.dtype}_order{a.order}"
    return f"roofline_{a.size}x{a.size}_dx2_rainfield_{a.dtype}_order{a.order}"


def workload_tag(a) -> str:
    """Workload part of the profiles/ file names."""
    if a.case == "runoff":
        return f"rainfall_runoff_{a.size}_{a.dtype}_order{a.order}"
    return f"roofline_{a.size}_{a.dtype}_order{a.order}"


# every tenth timed step carries HIP events around its kernels
KERNEL_SAMPLE_EVERY = 10


def time_two_pass(g, steps):
    """Two-pass step (flux kernel; update + finalize): HIP events on the
    launch stream around all the steps, and between the kernels of every
    tenth step (the per-kernel means)."""
    import torch
    stream = torch.cuda.current_stream(g.device)
    sampled = list(range(0, steps, KERNEL_SAMPLE_EVERY))
    ev = {k: [torch.cuda.Event(enable_timing=True) for _ in range(3)] for k in sampled}
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall = time.perf_counter()
    start.record(stream)
    for k in range(steps):
        e = ev.get(k)
        if e:
            e[0].record(stream)
        g.flux()
        if e:
            e[1].record(stream)
        g.update()
        if e:
            e[2].record(stream)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall
    n = len(sampled)
    return {"path": "two_pass", "steps": steps, "total_ms": start.elapsed_time(stop), "wall_s": wall,
            "kernels_ms": {"k_flux": sum(e[0].elapsed_time(e[1]) for e in ev.values()) / n,
                           "k_update": sum(e[1].elapsed_time(e[2]) for e in ev.values()) / n}}


def time_fused(g, steps):
    """Single-pass step (fused kernel + time control): `steps` steps in one
    call of hp_grid_step_fused_timed, which records HIP events on the launch
    stream around all the steps and around the single-pass kernel of every
    tenth step (an event between two kernels delays the next launch by a few
    microseconds, so only those steps carry it), so the kernel's own duration
    and the steps' total come from the same launches."""
    import ctypes

    import torch

    from hipims_amd import _lib
    g.ensure_out()
    g.step_fused(2)  # untimed: first launches on the second set of planes
    torch.cuda.synchronize()
    k_ms, tot_ms = ctypes.c_float(), ctypes.c_float()
    wall = time.perf_counter()
    _lib.check(_lib.lib().hp_grid_step_fused_timed(ctypes.byref(g.desc), int(steps), KERNEL_SAMPLE_EVERY,
                                                    ctypes.byref(k_ms), ctypes.byref(tot_ms),
                                                    _lib.stream_ptr(g.device)), "hp_grid_step_fused_timed")
    wall = time.perf_counter() - wall
    kern = k_ms.value
    return {"path": "fused", "steps": steps, "total_ms": tot_ms.value, "wall_s": wall,
            "kernels_ms": {"k_step_fused": kern, "k_finalize+gaps": tot_ms.value / steps - kern}}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--size", type=int, default=2048)
    p.add_argument("--case", choices=["runoff", "roofline"], default="runoff",
                   help="runoff: dx 5 m, uniform 50 mm/h station rain; roofline: dx 2 m, static Gaussian "
                        "random rain field (mean 30, std 10 mm/h)")
    p.add_argument("--dtype", choices=["fp32", "fp64"], default="fp32")
    p.add_argument("--order", type=int, choices=[1, 2], default=2)
    p.add_argument("--tile-rows", type=int, default=0, help="flux tile height (0: the fastest of the legal heights, timed per step path)")
    p.add_argument("--mass-sums", action="store_true",
                   help="also keep the FP64 volume / rain / boundary-flux sums each step (diagnostics the "
                        "reference step does not compute)")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="0: the physical cores (lscpu) among the CPUs this process may run on, at most the "
                        "host's per-GPU share (OMP_NUM_THREADS when set)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--mode", choices=["step", "copy", "valu", "ensemble", "adjoint"], default="step",
                   help="copy: a plain device copy of known size, the reference point for HBM byte counters; "
                        "valu: an FP32 FMA chain kernel, the reference point for the VALU issue rate; "
                        "ensemble: 64 uneven scenarios, graph-replayed ensemble vs the same run one by one "
                        "through the Godunov driver; adjoint: gradient cost against a plain forward run")
    p.add_argument("--copy-mib", type=int, default=1024)
    p.add_argument("--scenarios", type=int, default=64)
    p.add_argument("--sim-seconds", type=float, default=900.0, help="ensemble: simulated time of every scenario")
    p.add_argument("--replay-steps", type=int, default=50, help="ensemble: steps per graph replay (even)")
    p.add_argument("--ens-two-pass", action="store_true", help="ensemble: flux and update passes per step")
    p.add_argument("--ens-retune-below", type=float, default=0.25,
                   help="ensemble: time the unit height again once the active scenarios hold less than this "
                        "fraction of the cells it was chosen for (0: never)")
    p.add_argument("--no-baselines", action="store_true",
                   help="ensemble: of the sequential baseline runs keep only the standalone graphs (for long "
                        "simulated times)")
    p.add_argument("--ensemble-only", action="store_true", help="ensemble: no baseline runs at all")
    p.add_argument("--ensemble-seconds", type=float, default=900.0,
                   help="step mode: simulated time of every scenario in the ensemble run reported beside the "
                        "step (0: none; the configured 2 h takes minutes: --mode ensemble --sim-seconds 7200)")
    p.add_argument("--no-adjoint", action="store_true", help="step mode: no calibration gradient run beside the step")
    p.add_argument("--adj-size", type=int, default=1024)
    p.add_argument("--adj-steps", type=int, default=2000)
    p.add_argument("--adj-checkpoint", type=int, default=100)
    p.add_argument("--adj-gauges", type=int, default=32)
    p.add_argument("--adj-order", type=int, choices=[1, 2], default=1,
                   help="adjoint: scheme order (the server's default, order 1, has a bounded adjoint over long "
                        "horizons; order-2 runs amplify rounding differences, and beyond a few hundred steps "
                        "their derivative, with exact or frozen slopes, is not defined: costs only)")
    p.add_argument("--adj-slopes", choices=["exact", "frozen"], default="exact")
    p.add_argument("--no-fd", action="store_true", help="adjoint: skip the finite-difference check")
    p.add_argument("--adj-serial", action="store_true",
                   help="adjoint: recompute segments between the reverse ones instead of beside them")
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


def valu_calibration(a):
    """The FP32 FMA probe (hp_probe_fma): eight independent non-packed FMA
    chains per lane, 32 blocks of 256 lanes per CU; HIP events around
    --steps launches give FMA lane-instructions per second."""
    import torch

    from hipims_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda", 0)
    blocks = 256 * 32
    out = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    st = _lib.stream_ptr(dev)
    _lib.check(L.hp_probe_fma(blocks, out.data_ptr(), st), "hp_probe_fma")
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(a.steps):
        _lib.check(L.hp_probe_fma(blocks, out.data_ptr(), st), "hp_probe_fma")
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / a.steps
    lanes = blocks * 256 * int(L.hp_probe_fma_per_lane())
    print(json.dumps({"mode": "valu", "fma_lane_instructions_per_launch": lanes, "ms_per_launch": ms,
                      "fma_lane_instructions_per_s": lanes / (ms / 1e3),
                      "fp32_tflops_non_packed": 2 * lanes / (ms / 1e3) / 1e12}), flush=True)


def _ensemble_case(spec):
    from hipims_amd import generators as gen
    return gen.ensemble_case(*spec)


def ensemble_cases(n_scen):
    """The ensemble's scenarios, generated on host worker processes; call
    before this process touches the GPU (the pool forks)."""
    import concurrent.futures as cf

    from hipims_amd import generators as gen
    specs = gen.ensemble_specs(n_scen, 2000)
    t0 = time.perf_counter()
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    with cf.ProcessPoolExecutor(workers) as ex:
        cases = list(ex.map(_ensemble_case, specs))
    gen_s = time.perf_counter() - t0
    print(f"[ensemble] generated {len(cases)} scenarios in {gen_s:.1f} s", file=sys.stderr, flush=True)
    return cases, gen_s


def _ensemble_solver(c, dev):
    import torch

    from hipims_amd.grid import GridSolver
    s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=2, dtype=torch.float32, device=dev,
                   first_dt=c.first_dt, manning=c.manning, landuse=c.landuse)
    s.set_station_rain(c.rain_table)
    return s


def run_ensemble(a, cases, dev, sim_seconds, baselines):
    """Config: 64 scenarios, sizes drawn from {512, 768, 1024, 1536, 2048}^2,
    own DEM / landcover per scenario (seed 2000 + k), triangular hyetograph,
    each run to the same simulated time with its own CFL dt; FP32, order 2.
    Ensemble: every scenario advanced by one launch set per step, K steps per
    graph replay, one host read of the scenarios' times per replay.
    Baselines, the same scenarios one after another:
      standalone_graphs  each scenario alone through the single-pass step,
                         K steps per graph replay, one host read of t per
                         replay (the same kernels and host pattern as the
                         ensemble, one scenario per launch);
      godunov            the structured Godunov drop-in's per-step calls
                         (addFlux, addStation_PrecipitationSource,
                         time_friction_euler_update_cuda, a host read of t
                         each step);
      compact            CompactGodunov, the reference's own compact-layout
                         sequence (wet lists by nonzero, host dt logic) over
                         the HIP compact ops."""
    import torch

    from hipims_amd.ensemble import Ensemble
    from hipims_amd.compact_driver import CompactGodunov
    from hipims_amd.SWE_CUDA import Godunov

    # first launches of the step kernels outside the timed runs
    warm = _ensemble_solver(cases[0], dev)
    warm.step_fused(2)
    Ensemble([warm], single_pass=not a.ens_two_pass).step(2)
    del warm

    from hipims_amd import autotune
    members = [_ensemble_solver(c, dev) for c in cases]
    ens_rows = autotune.tune_ensemble(members, single_pass=not a.ens_two_pass)
    ens_tuning = dict(members[0].tune_timings)
    ens = Ensemble(members, t_stop=sim_seconds, single_pass=not a.ens_two_pass)
    ens.capture(a.replay_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    last = [t0]

    def progress(k, n_act):
        now = time.perf_counter()
        if now - last[0] > 10.0:
            last[0] = now
            print(f"[ensemble] {k} replays, {n_act} scenarios active, {now - t0:.1f} s", file=sys.stderr,
                  flush=True)

    replays = ens.run(a.replay_steps, progress=progress, retune_below=a.ens_retune_below)
    torch.cuda.synchronize()
    ens_s = time.perf_counter() - t0
    retunes = list(ens.retunes)
    steps = [int(s.ctrl_dict()["steps"]) for s in ens.solvers]
    cells = [s.layout.n_valid for s in ens.solvers]
    work = float(sum(n * k for n, k in zip(cells, steps)))
    print(f"[ensemble] {replays} replays, {ens_s:.2f} s; steps per scenario {min(steps)}..{max(steps)}",
          file=sys.stderr, flush=True)
    del ens
    torch.cuda.empty_cache()

    seq = {}
    if "standalone_graphs" in baselines:
        seq["standalone_graphs"] = standalone_graphs(cases, sim_seconds, a.replay_steps, dev)
    for name, cls in (("godunov", Godunov), ("compact", CompactGodunov)):
        if name in baselines:
            seq[name] = sequential(cls, cases, sim_seconds, dev, name)
    return {"scenarios": len(cases), "sizes": [c.mask.shape[0] - 2 for c in cases],
            "sim_seconds": sim_seconds, "dtype": "f32", "order": 2, "steps_per_replay": a.replay_steps,
            "tile_rows": ens_rows, "tile_rows_tuning_ms_per_step": {str(k): v for k, v in ens_tuning.items()},
            "retune_below": a.ens_retune_below,
            "retunes": [{"scenarios": n, "cells": c, "tile_rows": r, "seconds": t} for n, c, r, t in retunes],
            "step": "two-pass" if a.ens_two_pass else "single-pass",
            "cell_updates": work, "ensemble_s": ens_s, "ensemble_cell_updates_per_s": work / ens_s,
            "steps_min": min(steps), "steps_max": max(steps),
            "baselines": {name: {"seconds": v[0], "cell_updates_per_s": work / v[0], "speedup": v[0] / ens_s,
                                 "same_step_counts": v[1] == steps} for name, v in seq.items()}}


def standalone_graphs(cases, t_stop, steps, dev):
    """Each scenario alone: its single-pass steps captured `steps` to a HIP
    graph (captured before the clock starts) at the unit height tuned for
    its size, the graph replayed until the scenario reaches t_stop, one host
    read of t per replay."""
    import torch

    from hipims_amd import _lib, autotune
    total = 0.0
    counts = []
    last = time.perf_counter()
    for k, c in enumerate(cases):
        s = _ensemble_solver(c, dev)
        autotune.tune(s, path="fused")  # memoised per size class
        s.set_t_stop(t_stop)
        s.ensure_out()
        s.desc  # workspace allocated outside the capture
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            s.step_fused(steps)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while float(s.ctrl[_lib.CTRL_T]) < t_stop:
            g.replay()
        torch.cuda.synchronize()
        total += time.perf_counter() - t0
        counts.append(int(s.ctrl_dict()["steps"]))
        del g, s
        if time.perf_counter() - last > 10.0:
            last = time.perf_counter()
            print(f"[ensemble] standalone graphs: {k + 1} scenarios, {total:.1f} s", file=sys.stderr, flush=True)
    print(f"[ensemble] standalone graphs: {total:.2f} s", file=sys.stderr, flush=True)
    return total, counts


def ensemble_mode(a):
    import torch
    cases, gen_s = ensemble_cases(a.scenarios)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    bl = ("standalone_graphs",) if a.no_baselines else ("standalone_graphs", "godunov", "compact")
    if a.ensemble_only:
        bl = ()
    line = {"mode": "ensemble", **run_ensemble(a, cases, dev, a.sim_seconds, bl), "generation_s": gen_s}
    print(json.dumps(line), flush=True)


def sequential(cls, cases, t_stop, dev, name):
    """The scenarios one after another through a driver's per-step calls:
    addFlux, addStation_PrecipitationSource, time_friction_euler_update_cuda
    and a host read of t each step (CHANS-HiPIMS_server.py:246-251)."""
    import numpy as np
    import torch
    seq_s = 0.0
    seq_steps = []
    T = lambda x: torch.as_tensor(np.asarray(x))
    for k, c in enumerate(cases):
        num = cls(dev, c.dx, c.cfl, 1.0e30, t=0.0, export_n=0, firstTimeStep=c.first_dt, secondOrder=True,
                  tensorType=torch.float32)
        num.init__fluidField_tensor(T(c.mask), T(c.h), T(c.qx), T(c.qy), T(c.h + c.z), T(c.z), dev)
        num.set__frictionField_tensor(c.manning, dev)
        num.set_landuse(T(c.mask), T(c.landuse), dev)
        num.set_distributed_rainfall_station_Mask(T(c.mask), T(np.zeros_like(c.mask)), dev)
        num.set_boundary_tensor({}, dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = 0
        while num.t.item() < t_stop:
            num.addFlux()
            num.addStation_PrecipitationSource(c.rain_table, dev)
            num.time_friction_euler_update_cuda(dev)
            n += 1
        torch.cuda.synchronize()
        seq_s += time.perf_counter() - t0
        seq_steps.append(n)
        del num
        print(f"[ensemble] sequential {name}: scenario {k + 1} ({c.mask.shape[0] - 2}^2, {n} steps), "
              f"{seq_s:.1f} s so far", file=sys.stderr, flush=True)
    return seq_s, seq_steps


def adjoint_mode(a):
    print(json.dumps({"mode": "adjoint", **run_adjoint(a, fp64=not a.no_fd)}), flush=True)


def run_adjoint(a, fp64=True):
    """Config: 1024^2, dx = 5 m, FP32 forward and adjoint, FP64 gradient
    sums, 2000 steps, checkpoint every 100; 32 gauges among cells wet at the
    end of a pilot run; observations from a run with every class n scaled by
    0.8 or 1.2 (seeded); gradients w.r.t. the 6 class n and the rain
    multiplier.  Cost ratio = (forward with checkpoints + reverse with
    recomputation) / plain forward of the same steps; the gradient is checked
    against central differences (eps = 1e-3 relative) on the frozen schedule,
    and an FP64 run of the same calibration gives the reference gradient
    (checked against its own central differences).  Relative errors are
    max |difference| / max |gradient| over the seven parameters."""
    import numpy as np
    import torch

    from hipims_amd import generators as gen
    from hipims_amd.adjoint import Calibration
    from hipims_amd.grid import GridSolver

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    c = gen.rainfall_runoff(a.adj_size, 5.0)

    rows = 0

    def solver():
        s = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=a.adj_order, dtype=torch.float32,
                       device=dev, first_dt=c.first_dt, manning=c.manning, landuse=c.landuse, tile_rows=rows)
        s.set_station_rain(c.rain_table)
        return s

    # one tile height for the plain and the calibration runs: the fastest
    # single-pass one on this domain
    from hipims_amd import autotune
    rows = autotune.tune(solver(), path="fused", apply=False)
    pilot = solver()
    pilot.step(a.adj_steps)
    h = pilot.raster(pilot.h)
    rng = np.random.Generator(np.random.PCG64(1004))
    wet = np.argwhere((h > 1e-3) & (c.mask > 0))
    gauges = wet[rng.choice(len(wet), size=min(a.adj_gauges, len(wet)), replace=False)]
    del pilot

    # plain forward: the model's own CFL steps, one call of the library's
    # fastest step (the single-pass kernel, as the calibration runs use)
    plain = solver()
    plain.step_fused(6)
    torch.cuda.synchronize()
    plain = solver()
    plain.ensure_out()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain.step_fused(a.adj_steps)
    torch.cuda.synchronize()
    fwd_s = time.perf_counter() - t0

    cal = Calibration(solver(), gauges, a.adj_steps, a.adj_checkpoint, slopes=a.adj_slopes, overlap=not a.adj_serial)
    th0 = cal.theta0()
    pert = torch.tensor(np.random.Generator(np.random.PCG64(7)).choice([0.8, 1.2], size=th0.numel() - 1))
    th_true = th0.clone()
    th_true[:-1] *= pert
    cal.obs = cal.simulate(th_true)
    # untimed warm-up of the same calibration (first launches of the record,
    # recompute and reverse kernels), then the timed one on fresh state
    warm = Calibration(solver(), gauges, a.adj_steps, a.adj_checkpoint, obs=cal.obs, slopes=a.adj_slopes,
                       overlap=not a.adj_serial)
    warm.run(th0)
    warm.gradient()
    del warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loss = cal.run(th0)
    torch.cuda.synchronize()
    run_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad = cal.gradient()
    torch.cuda.synchronize()
    rev_s = time.perf_counter() - t0
    # the same gradient keeping every state and flux of the forward run
    # (no recomputation; K x 6 planes of HBM)
    cal_all = Calibration(solver(), gauges, a.adj_steps, 1, obs=cal.obs, slopes=a.adj_slopes)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cal_all.run(th0)
    torch.cuda.synchronize()
    run_all_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    grad_all = cal_all.gradient()
    torch.cuda.synchronize()
    rev_all_s = time.perf_counter() - t0
    store_all = {"recorded_forward_s": run_all_s, "reverse_s": rev_all_s,
                 "cost_ratio": (run_all_s + rev_all_s) / fwd_s,
                 "stored_bytes": 2 * cal_all.segU.numel() * cal_all.segU.element_size(),
                 "grad_equal": bool(torch.equal(grad_all, grad))}
    del cal_all
    torch.cuda.empty_cache()
    line = {"size": a.adj_size, "order": a.adj_order, "slopes": a.adj_slopes, "tile_rows": rows,
            "plain_path": "single-pass step (column-walk step kernel + k_finalize)",
            "steps": a.adj_steps, "checkpoint": a.adj_checkpoint,
            "recompute": "serial" if a.adj_serial else "overlapped on a second stream",
            "gauges": int(len(gauges)), "dtype": "f32 state, f64 gradient sums", "loss": float(loss),
            "grad": [float(v) for v in grad], "plain_forward_s": fwd_s, "recorded_forward_s": run_s,
            "reverse_s": rev_s, "cost_ratio": (run_s + rev_s) / fwd_s, "store_all": store_all}
    if not a.no_fd:
        fd = []
        for j in range(th0.numel()):
            d = 1e-3 * abs(float(th0[j]))
            tp, tm = th0.clone(), th0.clone()
            tp[j] += d
            tm[j] -= d
            fd.append((cal.run(tp).item() - cal.run(tm).item()) / (2 * d))
        fd = np.array(fd)
        g = grad.numpy()
        line["fd"] = fd.tolist()
        line["fd_max_rel_err"] = float(np.abs(g - fd).max() / np.abs(g).max())
    if fp64:
        g = grad.numpy()
        # FP64 reference: the same gauges, observations and loss on the FP64
        # solver (its own recorded schedule), its adjoint gradient against its
        # own central differences, and the FP32 gradient against it
        s64 = GridSolver(c.mask, c.h, c.qx, c.qy, c.z, c.dx, cfl=c.cfl, order=a.adj_order, dtype=torch.float64,
                         device=dev, first_dt=c.first_dt, manning=c.manning, landuse=c.landuse)
        s64.set_station_rain(c.rain_table)
        cal64 = Calibration(s64, gauges, a.adj_steps, a.adj_checkpoint, obs=cal.obs, slopes=a.adj_slopes)
        cal64.run(th0)
        g64 = cal64.gradient().numpy()
        line["fp64"] = {"grad": g64.tolist(), "fp32_vs_fp64_max_rel": float(np.abs(g - g64).max() / np.abs(g64).max())}
        # eps 1e-3 as for FP32; eps 1e-6 keeps the perturbed runs on the same
        # side of the wet/dry and limiter switches, where the discrete model
        # is differentiable
        for eps in (1e-3, 1e-6):
            fd64 = []
            for j in range(th0.numel()):
                d = eps * abs(float(th0[j]))
                tp, tm = th0.clone(), th0.clone()
                tp[j] += d
                tm[j] -= d
                fd64.append((cal64.run(tp).item() - cal64.run(tm).item()) / (2 * d))
            fd64 = np.array(fd64)
            line["fp64"][f"fd_eps{eps:g}"] = {
                "fd": fd64.tolist(), "max_rel_err": float(np.abs(g64 - fd64).max() / np.abs(g64).max()),
                "rel_err_per_param": (np.abs(g64 - fd64) / np.abs(fd64)).tolist()}
    return line


def main():
    a = parse()
    if a.mode == "copy":
        return copy_calibration(a)
    if a.mode == "valu":
        return valu_calibration(a)
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described, not implemented)"}))
        sys.exit(2)
    if a.mode == "ensemble":
        return ensemble_mode(a)
    if a.mode == "adjoint":
        return adjoint_mode(a)
    ens_cases = ensemble_cases(a.scenarios)[0] if a.ensemble_seconds > 0 else None
    import numpy as np
    import torch

    from hipims_amd import generators as gen
    from hipims_amd.grid import GridSolver

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    dtype = torch.float32 if a.dtype == "fp32" else torch.float64
    esize = 4 if dtype == torch.float32 else 8
    case = gen.rainfall_runoff(a.size, 5.0) if a.case == "runoff" else cached_case(a.size)
    g = GridSolver(case.mask, case.h, case.qx, case.qy, case.z, case.dx, cfl=case.cfl, order=a.order, dtype=dtype,
                   device=dev, first_dt=case.first_dt, manning=case.manning, landuse=case.landuse)
    if case.rain_table is not None:
        g.set_station_rain(case.rain_table)
    else:
        g.set_rain_field(case.rain_field)
    field = case.rain_table is None
    g.set_mass_sums(a.mass_sums)
    cells = case.n_valid

    g.step(a.warmup)
    torch.cuda.synchronize()
    # Tile height per path: the fixed --tile-rows, or the fastest legal one
    # timed on this domain (the tuner restores the state it steps).
    rows, tuned = {}, {}
    for path in ("flux", "fused"):
        if a.tile_rows:
            rows[path] = a.tile_rows
        else:
            from hipims_amd import autotune
            rows[path] = autotune.tune(g, path=path, apply=False)
            tuned[path] = dict(getattr(g, "tune_timings", {}))
    g.set_tile_rows(rows["flux"])
    two = time_two_pass(g, a.steps)
    g.set_tile_rows(rows["fused"])
    fused = time_fused(g, a.steps)
    ctrl = g.ctrl_dict()
    assert np.isfinite(ctrl["t"]) and ctrl["dt"] > 0

    B = alg_bytes(esize, rain_field=field)
    best = fused if fused["total_ms"] < two["total_ms"] else two
    ms_step = best["total_ms"] / best["steps"]
    value = cells * best["steps"] / (best["total_ms"] / 1e3)
    per_kernel, roofline = roofline_fields(best["kernels_ms"], ms_step, esize, field, cells, workload_tag(a))
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
        "data": "synthetic (seeded fractal DEM, Voronoi landcover, " + ("uniform rain)" if not field else
                                                                          "Gaussian random rain field)"),
        "config": {"workload": workload_name(a), "cells": cells, "dx_m": case.dx, "order": a.order,
                   "rain": "station 50 mm/h" if not field else "field 30 +- 10 mm/h",
                   "tile_rows": rows["fused" if best["path"] == "fused" else "flux"],
                   "tile_rows_per_path": {"two_pass": rows["flux"], "fused": rows["fused"]},
                   "tile_rows_tuning_ms": {p: {str(k): v for k, v in t.items()} for p, t in tuned.items()},
                   "mass_sums": a.mass_sums,
                   "sim_time_s": ctrl["t"]},
        "roofline": roofline,
        "kernels_ms": per_kernel,
        "path": best["path"],
        "paths_ms_per_step": {"two_pass": two["total_ms"] / two["steps"], "fused": fused["total_ms"] / fused["steps"]},
        "two_pass_alg_GBps": two_pass,
        "two_pass_hbm_frac": two_pass / HBM_PEAK_GBS,
        "wall_s": t_wall,
        "cpu_baseline": cpu,
    }
    if ens_cases is not None:
        line["ensemble"] = run_ensemble(a, ens_cases, dev, a.ensemble_seconds, ("standalone_graphs",))
        del line["ensemble"]["sizes"]
    if not a.no_adjoint:
        line["adjoint"] = run_adjoint(a, fp64=False)
    print(json.dumps(line), flush=True)


def roofline_fields(kernels_ms: dict, ms_step: float, esize: int, rain_field: bool, cells: int, tag: str):
    """(kernel times keyed by kernel name, the line's roofline object) for the
    timed path's kernels (keys as time_two_pass / time_fused give them).  The
    object is priced on the HBM roofline (no MFMA work on this path);
    `limiter` names the resource the dominant kernel is closer to
    saturating: its HBM fraction or its VALU issue fraction (the counted
    instructions per cell from profiles/ against the measured issue rate)."""
    B = alg_bytes(esize, rain_field=rain_field)
    dom = max((k for k in kernels_ms if k in B), key=kernels_ms.get)
    assert kernels_ms[dom] <= ms_step, "the dominant kernel's time exceeds the step it belongs to"
    achieved = B[dom] * cells / (kernels_ms[dom] / 1e3) / 1e9
    names = kernel_names(esize)
    kname = names[dom]
    per_kernel = {names.get(k, k): v for k, v in kernels_ms.items()}
    valu = profiled_valu(kname, tag)
    vpeak, vsrc = valu_peak()
    valu_frac = None if valu is None else valu * cells / (per_kernel[kname] / 1e3) / vpeak
    return per_kernel, {
        "bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(kname, tag),
        "alg_bytes_per_cell": B[dom], "kernel_ms": per_kernel[kname],
        "timed_with": "HIP events around the kernel of every tenth timed step",
        "valu_lane_instr_per_cell": valu, "valu_issue_frac": valu_frac,
        "limiter": "valu" if valu_frac is not None and valu_frac > achieved / HBM_PEAK_GBS else "hbm",
        "valu_peak_lane_instr_per_s": vpeak, "valu_peak_source": vsrc}


def physical_cores():
    """(physical cores among the CPUs this process may run on, physical cores
    of the host) from `lscpu -p=CPU,CORE,SOCKET`; logical CPU counts when
    lscpu is unavailable."""
    import subprocess
    allowed = os.sched_getaffinity(0)
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True, check=True).stdout
    except (OSError, subprocess.CalledProcessError):
        return len(allowed), os.cpu_count() or len(allowed)
    mine, every = set(), set()
    for line in out.splitlines():
        if not line or line.startswith("#"):
            continue
        cpu, core, sock = (int(v) if v else 0 for v in line.split(",")[:3])
        every.add((sock, core))
        if cpu in allowed:
            mine.add((sock, core))
    return len(mine) or len(allowed), len(every) or os.cpu_count() or len(allowed)


def cpu_baseline(a, g, case, dtype):
    """Oracle (C++/OpenMP, the reference's compact layout and op sequence) on
    host cores, from the GPU's current state, for a bounded step count."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    orc.build()
    mine, host = physical_cores()
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = a.cpu_threads or (min(mine, share) if share > 0 else mine)
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
            "cores_note": f"{threads} OpenMP threads, one per physical core: lscpu counts {mine} physical cores "
                          f"on the CPUs this process may use and {host} on the host"
                          + (f"; the host's per-GPU share is OMP_NUM_THREADS={share}" if share > 0 else ""),
            "sample": f"{n} steps of the same {a.size}x{a.size} workload from the GPU's warmed state "
                      f"({secs:.1f} s, oracle C++/OpenMP, {np_dt.__name__}"
                      + (")" if case.rain_table is not None else "; the oracle driver has no rain-field source, "
                         "so its steps run without rain)")}


if __name__ == "__main__":
    main()
