"""Ensemble of independent domains advanced together on one GPU.

The reference runs one domain per process and steps it from Python with
several host synchronisations per step (SWE_CUDA.py:780-831).  Here every
scenario keeps its own planes and device control block (its own dt, t, export
clamp and stop time), and one set of launches advances all of them by one
step each: by default two - the single-pass step of every tile of every
scenario (alternating each scenario's two sets of state planes), time
control of every scenario - or, with single_pass=False, three (flux of every
tile, update of every chunk, time control).  K such steps are captured
once in a HIP graph and replayed; a scenario that has reached its stop time is
skipped by every pass, so the ensemble runs until the slowest scenario is done
with one host check per replay.
"""
from __future__ import annotations

import ctypes
import time

import torch

from . import _lib, autotune
from ._lib import HpEnsemble, HpGridDesc, check, lib
from .grid import GridSolver


class Ensemble:
    def __init__(self, solvers: list[GridSolver], t_stop: float | None = None, single_pass: bool = True):
        """single_pass: advance with the single-pass step (every step count
        given to step(), capture() and run() must then be even, so that the
        state ends in the scenarios' own planes)."""
        if not solvers:
            raise ValueError("empty ensemble")
        dev = solvers[0].device
        if any(s.device != dev for s in solvers):
            raise ValueError("all domains of an ensemble must live on one device")
        self.solvers = solvers
        self.device = dev
        self.single_pass = bool(single_pass)
        if self.single_pass:
            for s in solvers:
                s.ensure_out()
        if t_stop is not None:
            for s in solvers:
                s.set_t_stop(t_stop)
        self._build(list(range(len(solvers))))

    def _build(self, members: list[int]):
        """(Re)build the launch maps over the scenarios `members`."""
        self.members = members
        n = len(members)
        self._descs = (HpGridDesc * n)(*[self.solvers[k].desc for k in members])
        need = int(lib().hp_ensemble_workspace_bytes(n, self._descs))
        self.work = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        self.handle = HpEnsemble()
        check(lib().hp_ensemble_init(n, self._descs, int(self.single_pass), self.work.data_ptr(), self.work.numel(),
                                     ctypes.byref(self.handle), _lib.stream_ptr(self.device)), "hp_ensemble_init")
        self.graph = None
        self.graph_steps = 0

    @property
    def cells(self) -> int:
        return sum(s.layout.n_valid for s in self.solvers)

    def step(self, nsteps: int = 1):
        check(lib().hp_ensemble_step(ctypes.byref(self.handle), int(nsteps), _lib.stream_ptr(self.device)),
              "hp_ensemble_step")

    def capture(self, nsteps: int):
        """Record `nsteps` ensemble steps in a HIP graph (replay() runs them)."""
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step(nsteps)
        self.graph, self.graph_steps = g, int(nsteps)
        return g

    def replay(self):
        if self.graph is None:
            raise RuntimeError("capture() first")
        self.graph.replay()

    def times(self) -> torch.Tensor:
        return torch.stack([s.ctrl[_lib.CTRL_T] for s in self.solvers]).cpu()

    def active_mask(self) -> torch.Tensor:
        """Per scenario: still below its stop time (one host synchronisation)."""
        c = torch.stack([s.ctrl for s in self.solvers]).cpu()
        t, stop = c[:, _lib.CTRL_T], c[:, _lib.CTRL_T_STOP]
        return (stop <= 0) | (t < stop)

    def active(self) -> int:
        return int(self.active_mask().sum())

    def run(self, steps_per_replay: int = 50, max_replays: int = 10_000_000, compact: float = 0.75,
            progress=None, retune_below: float = 0.25) -> int:
        """Replay the captured graph until every scenario has stopped; returns
        the number of replays.  Stopped scenarios cost a launched-and-exiting
        block per tile, so once the active count falls to `compact` of the
        scenarios in the launch maps, the maps are rebuilt over the active
        ones and the graph is captured again.  At such a rebuild, once the
        active scenarios hold less than `retune_below` of the cells the unit
        height was last chosen for, the height is timed again over them
        (autotune.tune_ensemble, heights up to the current one): the height
        that fills the GPU with many domains leaves it idle with a few small
        ones; its timed steps count as the run's own.  The states do not
        depend on the height, only the FP64 mass sums' rounding does; 0
        keeps the height.  `self.retunes` records
        (scenarios, cells, height, seconds) of each retuning.
        `progress(replays, active)` is called after each host check."""
        if self.graph is None or self.graph_steps != steps_per_replay:
            self.capture(steps_per_replay)
        tuned_cells = sum(self.solvers[m].layout.n_valid for m in self.members)
        self.retunes = []
        for k in range(max_replays):
            act = self.active_mask()
            n_act = int(act.sum())
            if progress is not None:
                progress(k, n_act)
            if n_act == 0:
                return k
            if n_act <= compact * len(self.members):
                keep = [m for m in self.members if bool(act[m])]
                cells = sum(self.solvers[m].layout.n_valid for m in keep)
                if cells < retune_below * tuned_cells:
                    t0 = time.perf_counter()
                    now = self.solvers[keep[0]].tile_rows
                    # the timed steps are the run's own (no state restored)
                    best = autotune.tune_ensemble([self.solvers[m] for m in keep], single_pass=self.single_pass,
                                                  max_rows=now if now > 0 else None, restore=False)
                    self.retunes.append((len(keep), cells, best, time.perf_counter() - t0))
                    tuned_cells = cells
                self._build(keep)
                self.capture(steps_per_replay)
            self.replay()
        raise RuntimeError("ensemble did not finish within max_replays")
