"""Ensemble of independent domains advanced together on one GPU.

The reference runs one domain per process and steps it from Python with
several host synchronisations per step (SWE_CUDA.py:780-831).  Here every
scenario keeps its own planes and device control block (its own dt, t, export
clamp and stop time), and one set of three launches - flux of every tile of
every scenario, update of every chunk of every scenario, time control of every
scenario - advances all of them by one step each.  K such steps are captured
once in a HIP graph and replayed; a scenario that has reached its stop time is
skipped by every pass, so the ensemble runs until the slowest scenario is done
with one host check per replay.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import HpEnsemble, HpGridDesc, check, lib
from .grid import GridSolver


class Ensemble:
    def __init__(self, solvers: list[GridSolver], t_stop: float | None = None):
        if not solvers:
            raise ValueError("empty ensemble")
        dev = solvers[0].device
        if any(s.device != dev for s in solvers):
            raise ValueError("all domains of an ensemble must live on one device")
        self.solvers = solvers
        self.device = dev
        if t_stop is not None:
            for s in solvers:
                s.set_t_stop(t_stop)
        n = len(solvers)
        self._descs = (HpGridDesc * n)(*[s.desc for s in solvers])
        need = int(lib().hp_ensemble_workspace_bytes(n, self._descs))
        self.work = torch.empty(max(need, 256), dtype=torch.uint8, device=dev)
        self.handle = HpEnsemble()
        check(lib().hp_ensemble_init(n, self._descs, self.work.data_ptr(), self.work.numel(),
                                     ctypes.byref(self.handle), _lib.stream_ptr(dev)), "hp_ensemble_init")
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

    def active(self) -> int:
        """Scenarios still below their stop time (one host synchronisation)."""
        c = torch.stack([s.ctrl for s in self.solvers]).cpu()
        t, stop = c[:, _lib.CTRL_T], c[:, _lib.CTRL_T_STOP]
        return int(((stop <= 0) | (t < stop)).sum())

    def run(self, steps_per_replay: int = 50, max_replays: int = 100000) -> int:
        """Replay the captured graph until every scenario has stopped; returns
        the number of replays."""
        if self.graph is None or self.graph_steps != steps_per_replay:
            self.capture(steps_per_replay)
        for k in range(max_replays):
            if self.active() == 0:
                return k
            self.replay()
        raise RuntimeError("ensemble did not finish within max_replays")
