"""Flux unit-height autotuner.

Float domains run the flux pass and the single-pass step as column walks:
one wave per unit of 60 columns x R rows, R in {2, 4, 6, 8, 10, 12, 16}.  A taller
unit re-loads fewer halo rows and solves fewer duplicate y-edges (four
prologue rows and one edge per unit), a shorter one gives more waves, which
matters when the last round of waves on the 256 CUs would run part-empty.
Double domains stage 64 x TY cells plus a 2-cell halo in LDS, TY in {4, 8}: a
taller tile solves fewer halo edges but holds more LDS per block.  Which
wins depends on dtype, order, domain size and on the step path.  tune() times
each legal height on the caller's own domain with HIP events - the flux pass
alone, or whole single-pass steps - restores the domain's state afterwards
and keeps the fastest; results are memoised per (dtype, order, size class,
path) for the process.
"""
from __future__ import annotations

import torch

from .grid import GridSolver

CHOICES = {torch.float32: (2, 4, 6, 8, 10, 12, 16), torch.float64: (4, 8)}
_cache: dict = {}


def _key(s: GridSolver, path: str):
    cells = s.layout.ny * s.layout.pitch
    return (s.dtype, s.order, max(0, cells.bit_length() - 18), path)


def _time(s: GridSolver, path: str, reps: int) -> float:
    run = (lambda: s.step_fused(2)) if path == "fused" else s.flux
    run()  # warm-up (and the fused path's second buffers)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        run()
    b.record()
    torch.cuda.synchronize(s.device)
    return a.elapsed_time(b) / reps


def tune(s: GridSolver, reps: int = 20, apply: bool = True, path: str = "flux") -> int:
    """Best tile height for `path` ("flux": the two-pass flux pass,
    "fused": the single-pass step) on this domain."""
    if path not in ("flux", "fused"):
        raise ValueError("path must be 'flux' or 'fused'")
    key = _key(s, path)
    if key in _cache:
        best = _cache[key]
    else:
        saved = [s.U.clone(), s.D.clone(), s.h_max.clone(), s.ctrl.clone(), None if s.F is None else s.F.clone()]
        saved_rows = s.tile_rows
        timings = {}
        for ty in CHOICES[s.dtype]:
            s.set_tile_rows(ty)
            timings[ty] = _time(s, path, reps)
            s.U.copy_(saved[0])
            s.D.copy_(saved[1])
            s.h_max.copy_(saved[2])
            s.ctrl.copy_(saved[3])
            if saved[4] is not None:
                s.F.copy_(saved[4])
        s.set_tile_rows(saved_rows)
        best = min(timings, key=timings.get)
        _cache[key] = best
        s.tune_timings = timings
    if apply:
        s.set_tile_rows(best)
    return best


def tune_ensemble(solvers: list[GridSolver], single_pass: bool = True, steps: int = 10,
                  max_rows: int | None = None, restore: bool = True) -> int:
    """Best unit height for one ensemble launch over `solvers` (all must share
    dtype and order): each legal height (up to `max_rows` when given) is
    timed over `steps` ensemble steps (after two untimed ones).  With
    `restore` every domain's state is put back afterwards; without it the
    steps stand as the run's own (the states do not depend on the height),
    so a run that retunes on the way loses no work to it.  Sets the winner
    on every solver and returns it."""
    from .ensemble import Ensemble

    if not solvers:
        raise ValueError("no domains")
    dt = solvers[0].dtype
    steps += steps & 1  # the single-pass ensemble steps in pairs
    saved = [(s.U.clone(), s.h_max.clone(), s.ctrl.clone()) for s in solvers] if restore else []
    timings = {}
    for ty in CHOICES[dt]:
        if max_rows is not None and ty > max_rows and ty != min(CHOICES[dt]):
            continue
        for s in solvers:
            s.set_tile_rows(ty)
        e = Ensemble(solvers, single_pass=single_pass)
        e.step(2)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        e.step(steps)
        b.record()
        torch.cuda.synchronize(solvers[0].device)
        timings[ty] = a.elapsed_time(b) / steps
        del e
        for s, (u, hm, c) in zip(solvers, saved):
            s.U.copy_(u)
            s.h_max.copy_(hm)
            s.ctrl.copy_(c)
            s.state_written()
    best = min(timings, key=timings.get)
    for s in solvers:
        s.set_tile_rows(best)
    solvers[0].tune_timings = timings
    return best
