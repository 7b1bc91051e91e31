"""Gradients of a gauge-misfit loss through the shallow-water step (no
counterpart in the reference, which has no derivatives).

Loss  L(theta) = sum_{k=1..K} sum_g (h_g^k - obs_g^k)^2  over K steps, where
theta = (Manning n of every landcover class, rain multiplier).  The time-step
schedule of the first forward run is recorded and then frozen (t_k, dt_k are
constants for the derivative, for the finite-difference checks and for every
recomputation), as the adjoint of the CFL minimum is not defined.

backward(): the state before every `checkpoint`-th step is kept from the
forward run; segments are recomputed last to first, storing the state before
each step and the step's flux, and the hand-written adjoint kernels run
the steps in reverse directly on the stored planes (order 1: one tiled
kernel per step; order 2: the update, edge and gather passes of
hp_adjoint_step).  The per-step sequences of the forward run, of a
segment's recomputation and of its reverse sweep are native loops
(hp_calib_forward, hp_calib_reverse): one call per segment.  checkpoint=1
keeps every state and flux from the forward run and recomputes nothing
(K x 6 planes of device memory).  Parameter
gradients are accumulated in FP64 (per cell for Manning, per block for the
rain multiplier) and folded at the end.  ``CalibrationLoss.apply(theta,
calib)`` exposes this as a torch.autograd.Function.

The exact adjoint of the order-1 step stays bounded over long horizons.
Order-2 runs (MUSCL-minmod with forward Euler) amplify rounding-level
differences: on the rainfall-runoff cases a 1e-15 perturbation of the depth
reaches 1e-2 within a few hundred steps, and beyond about 300 steps the
central differences of the loss depend on eps, with the slopes exact or
frozen, so the discrete order-2 model has no derivative to follow there and
both adjoints grow without bound.  Long-horizon calibration therefore uses
order 1 (the reference server's default); order-2 gradients, exact or with
frozen slopes, hold over short horizons (tested against central
differences, and for frozen slopes against one-sided differences: that
model has kinks exactly at the run's own state, where the adjoint follows
the branch the run took).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import HpAdjointDesc, HpGridDesc, check, lib
from .grid import GridSolver


class Calibration:
    def __init__(self, solver: GridSolver, gauges, steps: int, checkpoint: int = 100, obs=None,
                 slopes: str = "exact", reverse: str = "tiled", overlap: bool = True):
        """gauges: (row, col) raster cells; obs: [steps + 1, G] observed depths
        (row 0 unused) or None (set later).  slopes: "exact" differentiates
        the limited MUSCL slopes (the exact derivative of the discrete step),
        "frozen" holds them at the run's own values (the exact derivative of
        the model whose faces are the cell value plus those slopes; see the
        module notes for the horizon over which order-2 derivatives hold).
        reverse: "tiled" runs each order-1 reverse step as one tiled kernel,
        "separate" as the update, edge and gather passes (order 2 always
        does).  overlap (with
        checkpoint > 1): the recomputation of the next segment runs on a
        second stream, into a second set of segment planes, while the
        current segment is reversed."""
        if slopes not in ("exact", "frozen"):
            raise ValueError("slopes must be 'exact' or 'frozen'")
        if reverse not in ("tiled", "separate"):
            raise ValueError("reverse must be 'tiled' or 'separate'")
        self.s = s = solver
        s.ensure_out()  # the calibration runs use the single-pass step
        self.K = int(steps)
        self.ck = int(checkpoint)
        dev = s.device
        P = s.layout.pitch
        g = np.asarray(gauges, dtype=np.int64).reshape(-1, 2)
        self.G = g.shape[0]
        self.gauge_off = torch.from_numpy(g[:, 0] * P + g[:, 1]).to(dev)
        self.C = int(s.manning.numel())
        self.u0 = [t.clone() for t in (s.h, s.qx, s.qy)]
        self.hmax0 = s.h_max.clone()
        self.ctrl0 = s.ctrl.clone()
        self.sched = torch.zeros(self.K + 1, 2, dtype=torch.float64, device=dev)
        # one read-only control block per step of the frozen schedule
        self.frozen = torch.zeros(self.K + 1, _lib.CTRL_DOUBLES, dtype=torch.float64, device=dev)
        self.sched_ready = False
        self.hist = torch.zeros(self.K + 1, self.G, dtype=s.dtype, device=dev)
        self.obs = None if obs is None else torch.as_tensor(obs, dtype=torch.float64, device=dev)
        self.ckpts: dict[int, torch.Tensor] = {}
        self.tail = None  # first step of a last segment kept by run() in seg_bufs[0]
        # checkpoint states, allocated once (no allocation inside a run)
        self.ckpt_store = None if self.ck == 1 else torch.zeros(
            ((self.K + self.ck - 1) // self.ck,) + tuple(s.U.shape), dtype=s.dtype, device=dev)
        # adjoint planes
        L = s.layout
        self.cpitch = int(np.ceil((L.nx + 1) / 64) * 64)
        shp = (L.ny, P)
        z = lambda: torch.zeros(shp, dtype=s.dtype, device=dev)
        self.lam = [z(), z(), z()]
        self.cot = [z(), z(), z()]
        self.xc = torch.zeros(12, L.ny + 1, self.cpitch, dtype=s.dtype, device=dev)
        self.yc = torch.zeros(12, L.ny + 1, self.cpitch, dtype=s.dtype, device=dev)
        self.nblocks = int(lib().hp_adjoint_grad_blocks(ctypes.byref(s.desc)))
        self.grads = torch.zeros(self.nblocks, _lib.GRAD_STRIDE, dtype=torch.float64, device=dev)
        self.gn = torch.zeros(shp, dtype=torch.float64, device=dev)
        # states before each step of a segment and the steps' fluxes
        # states before each step of a segment and the steps' fluxes (with
        # checkpoint=1: of every step, filled by run())
        n_keep = self.K if self.ck == 1 else self.ck
        self.segU = torch.zeros((n_keep,) + tuple(s.U.shape), dtype=s.dtype, device=dev)
        self.segD = torch.zeros_like(self.segU)
        self.overlap = bool(overlap) and self.ck > 1 and self.K > self.ck
        self.seg_bufs = [(self.segU, self.segD)]
        if self.overlap:
            self.seg_bufs.append((torch.zeros_like(self.segU), torch.zeros_like(self.segU)))
            self.side = torch.cuda.Stream(device=dev)  # default priority
        ad = HpAdjointDesc()
        ad.lam_h, ad.lam_qx, ad.lam_qy = (t.data_ptr() for t in self.lam)
        ad.cot_h, ad.cot_qx, ad.cot_qy = (t.data_ptr() for t in self.cot)
        ad.xc, ad.yc = self.xc.data_ptr(), self.yc.data_ptr()
        ad.cpitch = self.cpitch
        ad.grads = self.grads.data_ptr()
        ad.gn = self.gn.data_ptr()
        ad.slope_mode = 1 if slopes == "frozen" else 0
        ad.flux_ready = 1
        ad.separate_passes = 1 if reverse == "separate" else 0
        # gauge index per cell: the reverse sweep adds the loss cotangent in
        # the adjoint's own update pass (one gauge per cell; otherwise a
        # separate small kernel per step does it)
        self.gauge_of = None
        if torch.unique(self.gauge_off).numel() == self.G and self.G > 0:
            self.gauge_of = torch.full(shp, -1, dtype=torch.int32, device=dev)
            self.gauge_of.view(-1)[self.gauge_off] = torch.arange(self.G, dtype=torch.int32, device=dev)
            ad.gauge_of = self.gauge_of.data_ptr()
        self.adesc = ad
        if self.C > _lib.GRAD_STRIDE - 1:
            raise ValueError("at most 15 Manning classes")

    # -- parameters ---------------------------------------------------------------
    def set_theta(self, theta):
        th = torch.as_tensor(theta, dtype=torch.float64).detach().cpu().numpy()
        self.s.set_manning(th[: self.C])
        self.s.set_rain_scale(float(th[self.C]))

    def theta0(self) -> torch.Tensor:
        return torch.cat([self.s.manning.double().cpu(), torch.tensor([self.s.rain_scale], dtype=torch.float64)])

    # -- forward ----------------------------------------------------------------------
    def _restore_initial(self):
        s = self.s
        for t, v in zip((s.h, s.qx, s.qy), self.u0):
            t.copy_(v)
        s.h_max.copy_(self.hmax0)
        s.ctrl.copy_(self.ctrl0)

    def _record_gauges(self, k):
        check(lib().hp_gather(_lib.dtype_code(self.s.dtype), self.s.h.data_ptr(), self.gauge_off.data_ptr(),
                              self.G, self.hist[k].data_ptr(), _lib.stream_ptr(self.s.device)), "hp_gather")

    def _seg_stride(self) -> int:
        return self.segU[0].numel() * self.segU.element_size()

    def _forward(self, mode, k0, n, seg=False, gauges=True, bufs=None, stream=None):
        """Steps k0 .. k0+n-1 in one native call (hp_calib_forward) of the
        single-pass step, with the states and fluxes kept in segU / segD (or
        the pair bufs) when seg is set.  mode: "record" (own CFL steps,
        schedule saved), "free" (own CFL steps, nothing saved) or "frozen"
        (the recorded schedule)."""
        s = self.s
        U, D = bufs if bufs is not None else (self.segU, self.segD)
        check(lib().hp_calib_forward(
            ctypes.byref(s.desc), self.sched.data_ptr() if mode == "record" else None,
            self.frozen.data_ptr() if mode == "frozen" else None, int(k0), int(n),
            U.data_ptr() if seg else None, D.data_ptr() if seg else None,
            self._seg_stride() if seg else 0,
            self.gauge_off.data_ptr() if gauges else None, self.G if gauges else 0,
            self.hist.data_ptr() if gauges else None,
            _lib.stream_ptr(s.device) if stream is None else stream.cuda_stream), "hp_calib_forward")

    def run(self, theta, keep_checkpoints=True) -> torch.Tensor:
        """Forward run; returns the FP64 loss (needs obs).  The first run
        records the model's own time steps; later runs follow them."""
        self.set_theta(theta)
        self._restore_initial()
        self.ckpts.clear()
        self._record_gauges(0)
        mode = "frozen" if self.sched_ready else "record"
        self.tail = None
        if keep_checkpoints and self.ck == 1:
            # every state and flux kept: nothing is recomputed in reverse
            self._forward(mode, 0, self.K, seg=True)
            self.ckpts = {0: self.segU[0]}
        else:
            last = ((self.K - 1) // self.ck) * self.ck
            for a in range(0, self.K, self.ck):
                if keep_checkpoints:
                    slot = self.ckpt_store[a // self.ck]
                    slot.copy_(self.s.U)
                    self.ckpts[a] = slot
                if keep_checkpoints and a == last:
                    # the last segment keeps its states and fluxes: the
                    # reverse sweep starts on it without recomputing it
                    self._forward(mode, a, self.K - a, seg=True, bufs=self.seg_bufs[0])
                    self.tail = a
                else:
                    self._forward(mode, a, min(self.ck, self.K - a))
        if not self.sched_ready:
            self.frozen[:, 0:2].copy_(self.sched)
            self.sched_ready = True
        return self.loss()

    def simulate(self, theta) -> torch.Tensor:
        """Forward run with the model's own CFL time steps (nothing recorded or
        frozen); returns the gauge depths [steps + 1, G] in FP64."""
        self.set_theta(theta)
        self._restore_initial()
        self._record_gauges(0)
        self._forward("free", 0, self.K)
        return self.hist.double().clone()

    def loss(self) -> torch.Tensor:
        if self.obs is None:
            raise RuntimeError("no observations set")
        d = self.hist[1:].double() - self.obs[1:]
        return (d * d).sum()

    # -- backward -----------------------------------------------------------------------
    def gradient(self) -> torch.Tensor:
        """dL/dtheta of the last run() (FP64, on the host)."""
        s = self.s
        if not self.ckpts:
            raise RuntimeError("run() with checkpoints first")
        if self.obs is None:
            raise RuntimeError("no observations set")
        for t in self.lam:
            t.zero_()
        self.grads.zero_()
        self.gn.zero_()
        h_rec = self.hist.double().contiguous()
        obs = self.obs.contiguous()
        stream = _lib.stream_ptr(s.device)
        stride = self._seg_stride()
        tail, self.tail = self.tail, None  # its planes are reused below

        def reverse(k_lo, k_hi, seg_k0, bufs):
            check(lib().hp_calib_reverse(
                ctypes.byref(s.desc), ctypes.byref(self.adesc), self.frozen.data_ptr(), int(k_lo), int(k_hi),
                int(seg_k0), bufs[0].data_ptr(), bufs[1].data_ptr(), stride, self.gauge_off.data_ptr(), self.G,
                h_rec.data_ptr(), obs.data_ptr(), stream), "hp_calib_reverse")

        if self.ck == 1:
            reverse(0, self.K, 0, self.seg_bufs[0])
        elif not self.overlap:
            for a in sorted(self.ckpts, reverse=True):
                b = min(a + self.ck, self.K)
                if a != tail:
                    # recompute the segment on the frozen schedule: the state
                    # before each step and its flux
                    s.U.copy_(self.ckpts[a])
                    self._forward("frozen", a, b - a, seg=True, gauges=False)
                reverse(a, b, a, self.seg_bufs[0])
        else:
            # segment i is recomputed on the side stream into buffer pair
            # i % 2 while segment i - 1 is reversed on the main stream
            main, side = torch.cuda.current_stream(s.device), self.side
            segs = sorted(self.ckpts, reverse=True)
            ready = [torch.cuda.Event(), torch.cuda.Event()]  # recomputed into pair j
            done = [torch.cuda.Event(), torch.cuda.Event()]   # pair j reversed

            def recompute(i):
                a = segs[i]
                j = i % 2
                if a == tail:  # kept by the forward run (in pair 0)
                    return
                if i >= 2:
                    side.wait_event(done[j])
                with torch.cuda.stream(side):
                    s.U.copy_(self.ckpts[a])
                    self._forward("frozen", a, min(a + self.ck, self.K) - a, seg=True, gauges=False,
                                  bufs=self.seg_bufs[j], stream=side)
                ready[j].record(side)

            side.wait_stream(main)
            recompute(0)
            for i, a in enumerate(segs):
                if i + 1 < len(segs):
                    recompute(i + 1)
                if a != tail:
                    main.wait_event(ready[i % 2])
                reverse(a, min(a + self.ck, self.K), a, self.seg_bufs[i % 2])
                done[i % 2].record(main)
            main.wait_stream(side)
        check(lib().hp_adjoint_finish(ctypes.byref(s.desc), ctypes.byref(self.adesc), stream), "hp_adjoint_finish")
        g = self.grads.sum(dim=0).cpu()
        return torch.cat([g[: self.C], g[_lib.GRAD_STRIDE - 1: _lib.GRAD_STRIDE]])


class CalibrationLoss(torch.autograd.Function):
    """loss = CalibrationLoss.apply(theta, calib): forward runs the model with
    theta, backward returns dL/dtheta from the adjoint."""

    @staticmethod
    def forward(ctx, theta, calib: Calibration):
        ctx.calib = calib
        return calib.run(theta).to(theta.device if theta.is_cuda else "cpu")

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.calib.gradient().to(grad_out.device)
        return g * grad_out, None
