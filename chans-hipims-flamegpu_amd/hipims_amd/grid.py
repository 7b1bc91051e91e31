"""Structured-grid state of one HiPIMS domain on the GPU.

The reference keeps the valid cells as compact 1-D arrays plus a neighbour
table (SWE_CUDA.py:115-227).  Here the same rasters are kept as 2-D planes of
``ny`` rows x ``pitch`` elements (pitch a multiple of 64), so the flux stencil
reads neighbours by address arithmetic and every row is 256-byte aligned.
``Layout`` converts between the two: ``offsets[i]`` is the plane element of
compact cell ``i`` (row-major order of ``mask > 0``), which keeps
``get_h()``-style compact views and the server's ABM cell ids working.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import HpGridDesc, check, lib

PITCH_ALIGN = 64
NODATA_CLASS = 255


@dataclass
class Layout:
    ny: int
    nx: int
    pitch: int
    valid: np.ndarray        # bool raster
    type_slot: np.ndarray    # uint8 (ny, pitch): 0 nodata, else slot
    slot_code: np.ndarray    # int32[256]: mask code of each slot
    offsets: np.ndarray      # int64[N]: plane element of compact cell i
    cell0: tuple             # (x, y) of compact cell 0, (-1, -1) if none

    @property
    def n_valid(self) -> int:
        return int(self.offsets.shape[0])

    def plane(self, raster, dtype, fill=0) -> np.ndarray:
        """Raster (ny, nx) -> plane (ny, pitch) with padding columns = fill."""
        out = np.full((self.ny, self.pitch), fill, dtype=dtype)
        out[:, : self.nx] = np.asarray(raster, dtype=dtype)
        return out


def make_layout(mask) -> Layout:
    mask = np.asarray(mask)
    if mask.ndim != 2:
        raise ValueError("mask must be a 2-D raster")
    ny, nx = mask.shape
    valid = mask > 0
    if valid[0, :].any() or valid[-1, :].any() or valid[:, 0].any() or valid[:, -1].any():
        raise ValueError(
            "valid cells on the raster's outer ring: the reference's neighbour table wraps around "
            "there (torch.roll); pad the rasters with one ring of nodata (mask <= 0)")
    codes = np.unique(mask[valid]).astype(np.int64)
    if codes.size > 255:
        raise ValueError(f"at most 255 distinct mask codes are supported, got {codes.size}")
    if codes.size and codes.max() > np.iinfo(np.int32).max:
        raise ValueError("mask codes must fit int32")
    pitch = int(math.ceil(nx / PITCH_ALIGN) * PITCH_ALIGN)
    slot_of = {int(c): k + 1 for k, c in enumerate(codes)}
    slot_code = np.zeros(256, dtype=np.int32)
    for c, k in slot_of.items():
        slot_code[k] = c
    slots = np.zeros(mask.shape, dtype=np.uint8)
    if codes.size:
        slots[valid] = np.searchsorted(codes, mask[valid]).astype(np.uint8) + 1
    type_slot = np.zeros((ny, pitch), dtype=np.uint8)
    type_slot[:, :nx] = slots
    rows, cols = np.nonzero(valid)  # row-major order == the reference's compact order
    offsets = rows.astype(np.int64) * pitch + cols.astype(np.int64)
    cell0 = (int(cols[0]), int(rows[0])) if rows.size else (-1, -1)
    return Layout(ny, nx, pitch, valid, type_slot, slot_code, offsets, cell0)


def _to_np(x):
    if x is None:
        return None
    if torch.is_tensor(x):
        return x.detach().cpu().numpy()
    return np.asarray(x)


class GridSolver:
    """One domain advanced by the HIP structured path (flux pass, update pass,
    device time control).  Inputs are the reference's rasters (mask codes,
    h, qx, qy, z) and Godunov's parameters."""

    def __init__(self, mask, h, qx, qy, z, dx, *, cfl=0.5, order=2, dtype=torch.float64, device="cuda",
                 first_dt=1.0e-4, t=0.0, export_T=math.inf, export_n=0, manning=(0.0,), landuse=None,
                 given_depth=None, given_wl=None, given_q=None, ctrl=None, tile_rows=0, t_stop=0.0,
                 mass_sums=True):
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.order = int(order)
        self.dx = float(dx)
        self.cfl = float(cfl)
        mask = _to_np(mask)
        self.layout = L = make_layout(mask)
        np_dt = np.float64 if dtype == torch.float64 else np.float32
        dev = self.device

        def plane(raster, fill=0.0):
            p = np.zeros((L.ny, L.pitch), dtype=np_dt)
            r = np.asarray(_to_np(raster), dtype=np_dt)
            p[:, : L.nx] = np.where(L.valid, r, fill)
            return torch.from_numpy(p).to(dev)

        # state and flux accumulators each in one allocation (planes are views),
        # so a whole state or flux can be saved or restored with one copy
        self.U = torch.stack([plane(h), plane(qx), plane(qy)])
        self.h, self.qx, self.qy = self.U[0], self.U[1], self.U[2]
        self.z = plane(z)
        self.h_max = self.h.clone()
        self.D = torch.zeros_like(self.U)
        self.dh, self.dqx, self.dqy = self.D[0], self.D[1], self.D[2]
        self.h_out = self.qx_out = self.qy_out = None  # allocated by step_fused()
        self.type_slot = torch.from_numpy(L.type_slot).to(dev)
        self.slot_code = torch.from_numpy(L.slot_code).to(dev)
        self.offsets = torch.from_numpy(L.offsets).to(dev)
        self.set_manning(manning)
        self.set_landuse(landuse)
        if ctrl is not None:
            # caller-owned control block (t, dt, export_T, export_n already set)
            if ctrl.dtype != torch.float64 or ctrl.numel() < _lib.CTRL_DOUBLES or ctrl.device != dev:
                raise ValueError("ctrl must be a float64 tensor of 16 entries on the solver's device")
            self.ctrl = ctrl
        else:
            self.ctrl = torch.zeros(_lib.CTRL_DOUBLES, dtype=torch.float64, device=dev)
            self.ctrl[0] = float(t)
            self.ctrl[1] = float(first_dt)
            self.ctrl[2] = float(export_T)
            self.ctrl[3] = float(export_n)
        if t_stop:
            self.ctrl[_lib.CTRL_T_STOP] = float(t_stop)
        self.tile_rows = int(tile_rows)
        self.mass_sums = bool(mass_sums)
        self.given_depth = self.given_wl = self.given_q = None
        self.set_boundary(given_depth, given_wl, given_q)
        self.station = None
        self.rain_table = None
        self.rain_field = None
        self.rain_mode = _lib.HP_RAIN_NONE
        self.rain_scale = 1.0
        self.infil = None      # (ks, phi, dtheta) class tables when infiltration is on
        self.F = None          # cumulative infiltrated depth plane
        self.dF = None         # the step's infiltrated depth plane (two-pass path)
        self._infil_pending = False
        self._desc = None
        self.work = None
        self._rebuild()

    # -- configuration -----------------------------------------------------
    def set_manning(self, manning):
        m = np.atleast_1d(np.asarray(_to_np(manning), dtype=np.float64))
        if m.size < 1 or m.size > 255:
            raise ValueError("manning must have 1..255 classes")
        if getattr(self, "_max_class", -1) >= m.size:
            raise ValueError("landuse class outside the manning table")
        self.manning = torch.tensor(m, dtype=self.dtype, device=self.device)
        self._desc = None

    def set_landuse(self, landuse):
        L = self.layout
        lu = np.full((L.ny, L.pitch), NODATA_CLASS, dtype=np.uint8)
        if landuse is None:
            cls = np.zeros((L.ny, L.nx), dtype=np.int64)
        else:
            cls = np.asarray(_to_np(landuse)).astype(np.int64)
        if (cls[L.valid] < 0).any() or (cls[L.valid] >= self.manning.numel()).any():
            raise ValueError("landuse class outside the manning table")
        self._max_class = int(cls[L.valid].max()) if L.valid.any() else -1
        lu[:, : L.nx] = np.where(L.valid, cls, NODATA_CLASS).astype(np.uint8)
        self.landuse = torch.from_numpy(lu).to(self.device)
        self._desc = None

    def set_boundary(self, given_depth=None, given_wl=None, given_q=None):
        def tab(x, default):
            a = np.asarray(_to_np(x) if x is not None else default, dtype=np.float64)
            if a.ndim == 1:
                a = a[None, :]
            return torch.tensor(a, dtype=self.dtype, device=self.device).contiguous()
        if given_depth is not None or self.given_depth is None:
            self.given_depth = tab(given_depth, [[0.0, 0.0]])
        if given_wl is not None or self.given_wl is None:
            self.given_wl = tab(given_wl, [[0.0, 0.0]])
        if given_q is not None or self.given_q is None:
            self.given_q = tab(given_q, [[0.0, 0.0, 0.0]])
        self._desc = None

    def set_station_rain(self, table, station=None):
        """Station rain: table [rows, 1+S] = [t, r_0..] in m/s; station raster
        of int16 ids (None: every cell uses station 0)."""
        tab = np.asarray(_to_np(table), dtype=np.float64)
        if tab.ndim != 2 or tab.shape[1] < 2:
            raise ValueError("rain table must be [rows, 1 + stations]")
        self.rain_table = torch.from_numpy(np.ascontiguousarray(tab)).to(self.device)
        if station is None:
            self.station = None
        else:
            L = self.layout
            st = np.zeros((L.ny, L.pitch), dtype=np.int16)
            st[:, : L.nx] = np.where(L.valid, np.asarray(_to_np(station)), 0).astype(np.int16)
            if st.max(initial=0) >= tab.shape[1] - 1 or st.min(initial=0) < 0:
                raise ValueError("station id outside the rain table")
            self.station = torch.from_numpy(st).to(self.device)
        self.rain_field = None
        self.rain_mode = _lib.HP_RAIN_STATION
        self._desc = None

    def set_rain_field(self, field):
        """Per-cell rain (m/s), the addRadarRainfall form."""
        L = self.layout
        f = np.zeros((L.ny, L.pitch), dtype=np.float32)
        f[:, : L.nx] = np.where(L.valid, np.asarray(_to_np(field), dtype=np.float32), 0.0)
        self.rain_field = torch.from_numpy(f).to(self.device)
        self.rain_table = None
        self.station = None
        self.rain_mode = _lib.HP_RAIN_FIELD
        self._desc = None

    def clear_rain(self):
        self.rain_mode = _lib.HP_RAIN_NONE
        self.rain_table = self.station = self.rain_field = None
        self._desc = None

    def set_tile_rows(self, rows: int):
        self.tile_rows = int(rows)
        self._desc = None

    def set_mass_sums(self, on: bool):
        """Keep the FP64 volume / rain / boundary-flux sums of the control
        block up to date (on by default; the reference computes none)."""
        self.mass_sums = bool(on)
        self._desc = None

    def set_t_stop(self, t_stop: float):
        """The passes of this domain do nothing once t >= t_stop (0: never)."""
        self.ctrl[_lib.CTRL_T_STOP] = float(t_stop)

    def set_rain_scale(self, s: float):
        self.rain_scale = float(s)
        self._desc = None

    def set_infiltration(self, hydraulic_conductivity, capillary_head, water_content_diff):
        """Green-Ampt infiltration with per-landuse-class tables (the
        reference's set__infiltrationField_tensor, SWE_CUDA.py:311-337); the
        cumulative depth F starts at 0.  step() and step_fused() then include
        it; with flux()/update() call infiltration() between them."""
        n = int(self.manning.numel())

        def tab(v):
            a = np.atleast_1d(np.asarray(_to_np(v), dtype=np.float64))
            if a.size == 1:
                a = np.full(n, a[0])
            if a.size < n:
                raise ValueError("infiltration tables need one entry per landuse class")
            return torch.tensor(a, dtype=self.dtype, device=self.device)
        self.infil = (tab(hydraulic_conductivity), tab(capillary_head), tab(water_content_diff))
        if self.F is None:
            self.F = torch.zeros_like(self.h)
            self.dF = torch.zeros_like(self.h)
        self._desc = None

    def clear_infiltration(self):
        self.infil = None
        self._desc = None

    # -- descriptor ----------------------------------------------------------
    def _rebuild(self):
        L = self.layout
        d = HpGridDesc()
        d.dtype = _lib.dtype_code(self.dtype)
        d.order = self.order
        d.nx, d.ny, d.pitch = L.nx, L.ny, L.pitch
        d.dx, d.cfl = self.dx, self.cfl
        d.cell0_x, d.cell0_y = L.cell0
        d.n_classes = int(self.manning.numel())
        d.rain_mode = self.rain_mode
        d.tile_rows = self.tile_rows
        d.mass_sums = int(self.mass_sums)
        for f in ("h", "qx", "qy", "z", "h_max", "dh", "dqx", "dqy", "type_slot", "landuse", "slot_code",
                  "manning"):
            setattr(d, f, getattr(self, f).data_ptr())
        for f in ("h_out", "qx_out", "qy_out"):
            setattr(d, f, _lib.ptr(getattr(self, f)))
        d.given_depth, d.gd_rows, d.gd_w = self.given_depth.data_ptr(), *self.given_depth.shape
        d.given_wl, d.gwl_rows, d.gwl_w = self.given_wl.data_ptr(), *self.given_wl.shape
        d.given_q, d.gq_rows, d.gq_w = self.given_q.data_ptr(), *self.given_q.shape
        d.station = _lib.ptr(self.station)
        if self.rain_table is not None:
            d.rain_table = self.rain_table.data_ptr()
            d.rain_rows, d.rain_cols = self.rain_table.shape
        d.rain_field = _lib.ptr(self.rain_field)
        d.rain_scale = self.rain_scale
        d.ctrl = self.ctrl.data_ptr()
        if self.infil is not None:
            d.infil_ks, d.infil_phi, d.infil_dtheta = (t.data_ptr() for t in self.infil)
            d.infil_F = self.F.data_ptr()
        need = int(lib().hp_grid_workspace_bytes(ctypes.byref(d)))
        if self.work is None or self.work.numel() < need:
            self.work = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)
        d.work = self.work.data_ptr()
        d.work_bytes = self.work.numel()
        self._desc = d

    @property
    def desc(self) -> HpGridDesc:
        if self._desc is None:
            self._rebuild()
        return self._desc

    # -- stepping --------------------------------------------------------------
    def _stream(self):
        return _lib.stream_ptr(self.device)

    def flux(self):
        check(lib().hp_grid_flux(ctypes.byref(self.desc), self._stream()), "hp_grid_flux")

    def infiltration(self):
        """Infiltration pass over the flux list of the current state; the next
        update() subtracts it (Godunov.addInfiltrationSource)."""
        if self.infil is None:
            raise RuntimeError("set_infiltration() first")
        d = self.desc
        d.infil_dF = self.dF.data_ptr()
        check(lib().hp_grid_infiltration(ctypes.byref(d), self._stream()), "hp_grid_infiltration")
        self._infil_pending = True

    def update(self, mode: int = _lib.HP_UPDATE_STEP, rain: bool = True):
        """One update pass; `rain` False leaves the configured rain source out."""
        d = self.desc
        saved = d.rain_mode
        if not rain:
            d.rain_mode = _lib.HP_RAIN_NONE
        d.infil_dF = self.dF.data_ptr() if self._infil_pending else None
        try:
            check(lib().hp_grid_update_mode(ctypes.byref(d), int(mode), self._stream()), "hp_grid_update_mode")
        finally:
            d.rain_mode = saved
            d.infil_dF = None
            self._infil_pending = False

    def step(self, nsteps: int = 1):
        d = self.desc
        d.infil_dF = self.dF.data_ptr() if self.infil is not None else None
        try:
            check(lib().hp_grid_step(ctypes.byref(d), int(nsteps), self._stream()), "hp_grid_step")
        finally:
            d.infil_dF = None

    def ensure_out(self):
        """Allocate the second state buffer of the single-pass step."""
        if self.h_out is None:
            self.h_out = torch.zeros_like(self.h)
            self.qx_out = torch.zeros_like(self.h)
            self.qy_out = torch.zeros_like(self.h)
            self._desc = None

    def step_fused(self, nsteps: int = 1):
        """nsteps single-pass steps (same results as step(); dh/dqx/dqy untouched)."""
        self.ensure_out()
        check(lib().hp_grid_step_fused(ctypes.byref(self.desc), int(nsteps), self._stream()), "hp_grid_step_fused")

    # -- views -----------------------------------------------------------------
    def compact(self, plane: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """Values of `plane` at the valid cells, in the reference's compact order."""
        n = self.layout.n_valid
        if out is None:
            out = torch.empty(n, dtype=plane.dtype, device=self.device)
        check(lib().hp_gather(_lib.dtype_code(plane.dtype), plane.data_ptr(), self.offsets.data_ptr(), n,
                              out.data_ptr(), self._stream()), "hp_gather")
        return out

    def state_written(self) -> None:
        """Record that the host wrote the state planes: a stopped domain's
        second set of planes no longer holds its state."""
        self.ctrl[_lib.CTRL_STOPPED_SYNCED] = 0.0

    def load_compact(self, plane: torch.Tensor, values: torch.Tensor) -> None:
        """Write compact-order values into `plane` at the valid cells."""
        self.state_written()
        values = values.to(device=self.device, dtype=plane.dtype).contiguous()
        check(lib().hp_scatter(_lib.dtype_code(plane.dtype), plane.data_ptr(), self.offsets.data_ptr(),
                               values.numel(), values.data_ptr(), self._stream()), "hp_scatter")

    def raster(self, plane: torch.Tensor) -> np.ndarray:
        return plane[:, : self.layout.nx].cpu().numpy()

    @property
    def wl(self) -> torch.Tensor:
        return self.z + self.h

    def ctrl_dict(self) -> dict:
        c = self.ctrl.cpu().numpy()
        return {k: float(c[i]) for i, k in enumerate(_lib.CTRL_FIELDS)}

    @property
    def t(self) -> float:
        return float(self.ctrl[0].item())

    @property
    def dt(self) -> float:
        return float(self.ctrl[1].item())
