"""CPU oracle loader: TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (the C++ restatement in ``swe_oracle.cpp``) and
restates, in numpy, the host-side parts of the reference driver that the C++
file does not cover:

* the compact valid-cell layout and the int32[5*N] neighbour table,
  ``CHANS-HiPIMS/hipims/pythonHipims/SWE_CUDA.py:115-227`` (``torch.roll`` with
  ``oppo_direction = [[-1, 1], [1, 0], [1, 1], [-1, 0]]``, restated with
  ``np.roll``, which has the same wrap-around semantics);
* the per-step wet lists (``SWE_CUDA.py:398-403`` and ``:786-788``).

Only tests, ``__graft_entry__.smoke`` and the bench's CPU-baseline leg import
this module.  Nothing here is part of the shipped library.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I = ctypes.c_int
_D = ctypes.c_double


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_num_threads.restype = _I
        L.orc_set_threads.argtypes = [_I]
        L.orc_flux_mask.argtypes = [_I, _I64, _P, _P, _P, _P]
        L.orc_flux.argtypes = [_I, _I, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                               _P, _I, _I, _P, _I, _I, _P, _I, _I, _P, _P, _P]
        L.orc_station_rain.argtypes = [_I, _I64, _P, _P, _P, _I, _P]
        L.orc_friction_update.argtypes = [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_time_control.argtypes = [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_euler_update.argtypes = [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_run_steps.restype = _I
        L.orc_infiltration.argtypes = [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_run_steps.argtypes = ([_I, _I, _I64, _I, _P] + [_P] * 10 + [_P, _P, _P, _P, _I, _I, _P]
                                    + [_P, _I, _I, _P, _I, _I, _P, _I, _I, _D, _D, _P, _P, _P, _P, _P, _P])
        L.orc_frozen_slope_loss.restype = _D
        L.orc_frozen_slope_loss.argtypes = [_I64, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D,
                                            _P, _P, _I, _P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return int(lib().orc_num_threads())


# ---------------------------------------------------------------------------
# Compact layout (SWE_CUDA.py:115-227)
# ---------------------------------------------------------------------------
OPPO_DIRECTION = [(-1, 1), (1, 0), (1, 1), (-1, 0)]  # (shift, axis) for E, N, W, S
NORMAL = np.array([[1.0, 0.0], [0.0, 1.0], [-1.0, 0.0], [0.0, -1.0]])


def build_index(mask: np.ndarray) -> np.ndarray:
    """int32[5*N] table: row 0 = mask code, rows 1..4 = compact id of the E, N,
    W, S neighbour or -1 (SWE_CUDA.py:173-222)."""
    mask = np.asarray(mask)
    valid = mask > 0
    n = int(valid.sum())
    index_mask = np.full(mask.shape, -1, dtype=np.int32)
    index_mask[valid] = np.arange(n, dtype=np.int32)
    index = np.zeros((5, n), dtype=np.int32)
    index[0] = mask[valid]
    for k, (shift, axis) in enumerate(OPPO_DIRECTION):
        index[k + 1] = np.roll(index_mask, shift, axis=axis)[valid]
    return index.reshape(-1)


def flux_list(h: np.ndarray, index: np.ndarray, t: float) -> np.ndarray:
    """fluxMask.update + nonzero (SWE_CUDA.py:398-403)."""
    n = h.shape[0]
    m = np.zeros(n, dtype=np.uint8)
    tt = np.array([t], dtype=h.dtype)
    lib().orc_flux_mask(int(h.dtype == np.float64), n, _p(m), _p(h), _p(index), _p(tt))
    return np.flatnonzero(m).astype(np.int32)


def update_list(dh: np.ndarray, h: np.ndarray) -> np.ndarray:
    """wetMask for friction/CFL: |dh| > 0 or h >= 0 (SWE_CUDA.py:786-788)."""
    return np.flatnonzero((np.abs(dh) > 0) | (h >= 0)).astype(np.int32)


def _scalar(v, dtype):
    return np.array([v], dtype=dtype)


# ---------------------------------------------------------------------------
# Single ops (same argument meaning as the reference's torch extensions)
# ---------------------------------------------------------------------------
def flux(order, wet, dh, dqx, dqy, h, wl, z, qx, qy, index, given_depth, given_wl, given_q, dx, t, dt):
    dt_ = h.dtype
    gd = np.ascontiguousarray(given_depth, dtype=dt_)
    gw = np.ascontiguousarray(given_wl, dtype=dt_)
    gq = np.ascontiguousarray(given_q, dtype=dt_)
    normal = np.ascontiguousarray(NORMAL.reshape(-1), dtype=dt_)
    lib().orc_flux(int(dt_ == np.float64), int(order), wet.shape[0], h.shape[0], _p(wet), _p(dh), _p(dqx),
                   _p(dqy), _p(h), _p(wl), _p(z), _p(qx), _p(qy), _p(index), _p(normal),
                   _p(gd), gd.shape[0], gd.shape[1], _p(gw), gw.shape[0], gw.shape[1],
                   _p(gq), gq.shape[0], gq.shape[1],
                   _p(_scalar(dx, dt_)), _p(_scalar(t, dt_)), _p(_scalar(dt, dt_)))


def station_rain(dh, station, rain, dt):
    rain = np.ascontiguousarray(rain)
    lib().orc_station_rain(int(dh.dtype == np.float64), dh.shape[0], _p(dh), _p(station), _p(rain),
                           int(rain.dtype == np.float64), _p(_scalar(dt, dh.dtype)))


def friction_update(wet, dh, dqx, dqy, dz, landuse, h, wl, qx, qy, z, manning, dt):
    manning = np.ascontiguousarray(manning, dtype=h.dtype)
    lib().orc_friction_update(int(h.dtype == np.float64), wet.shape[0], _p(wet), _p(dh), _p(dqx), _p(dqy),
                              _p(dz), _p(landuse), _p(h), _p(wl), _p(qx), _p(qy), _p(z), _p(manning),
                              _p(_scalar(dt, h.dtype)))


def time_control(wet, accel, h_max, h, qx, qy, dx, cfl):
    lib().orc_time_control(int(h.dtype == np.float64), wet.shape[0], _p(wet), _p(accel), _p(h_max), _p(h),
                           _p(qx), _p(qy), _p(_scalar(dx, h.dtype)), _p(_scalar(cfl, h.dtype)))


def euler_update(wet, dh, dqx, dqy, dz, h, wl, z, qx, qy):
    lib().orc_euler_update(int(h.dtype == np.float64), wet.shape[0], _p(wet), _p(dh), _p(dqx), _p(dqy),
                           _p(dz), _p(h), _p(wl), _p(z), _p(qx), _p(qy))


def infiltration(wet, dh, landuse, h, ks, phi, dth, F, dt):
    """addInfiltrationSource over the wet list (infiltrationCUDA_Kernel.cu:20-46)."""
    c = lambda a: np.ascontiguousarray(a, dtype=h.dtype)
    ks, phi, dth = c(ks), c(phi), c(dth)
    lib().orc_infiltration(int(h.dtype == np.float64), wet.shape[0], _p(wet), _p(dh), _p(landuse), _p(h), _p(ks),
                           _p(phi), _p(dth), _p(F), _p(_scalar(dt, h.dtype)))


# ---------------------------------------------------------------------------
# Whole-step simulation in the reference layout
# ---------------------------------------------------------------------------
@dataclass
class OracleSim:
    """Compact-layout state plus the driver scalars of ``Godunov``
    (SWE_CUDA.py:62-113).  ``step()`` runs the server sequence
    addFlux -> addStation_PrecipitationSource -> time_friction_euler_update_cuda."""
    mask: np.ndarray
    h: np.ndarray
    qx: np.ndarray
    qy: np.ndarray
    z: np.ndarray
    dx: float
    cfl: float = 0.5
    order: int = 2
    dtype: type = np.float64
    first_dt: float = 1.0e-4
    t: float = 0.0
    export_T: float = 1.0e30
    export_n: int = 0
    manning: np.ndarray = field(default_factory=lambda: np.array([0.0]))
    landuse: np.ndarray | None = None
    station: np.ndarray | None = None
    rain_table: np.ndarray | None = None
    given_depth: np.ndarray | None = None
    given_wl: np.ndarray | None = None
    given_q: np.ndarray | None = None
    # Green-Ampt per-class tables (set__infiltrationField_tensor); None: off
    hydraulic_conductivity: np.ndarray | None = None
    capillary_head: np.ndarray | None = None
    water_content_diff: np.ndarray | None = None

    def __post_init__(self):
        valid = self.mask > 0
        dt_ = self.dtype
        self.index = build_index(self.mask)
        self.N = int(valid.sum())
        self.h_c = np.ascontiguousarray(self.h[valid], dtype=dt_)
        self.qx_c = np.ascontiguousarray(self.qx[valid], dtype=dt_)
        self.qy_c = np.ascontiguousarray(self.qy[valid], dtype=dt_)
        self.z_c = np.ascontiguousarray(self.z[valid], dtype=dt_)
        self.wl_c = (self.h_c + self.z_c).astype(dt_)
        self.h_max = self.h_c.copy()
        self.dh = np.zeros(self.N, dt_)
        self.dqx = np.zeros(self.N, dt_)
        self.dqy = np.zeros(self.N, dt_)
        self.dz = np.zeros(self.N, dt_)
        self.lu = (np.zeros(self.N, np.uint8) if self.landuse is None
                   else np.ascontiguousarray(self.landuse[valid], dtype=np.uint8))
        self.man = np.ascontiguousarray(self.manning, dtype=dt_)
        self.st = None if self.station is None else np.ascontiguousarray(self.station[valid], dtype=np.int16)
        self.rt = None if self.rain_table is None else np.ascontiguousarray(self.rain_table, dtype=np.float64)
        self.gd = np.ascontiguousarray(self.given_depth if self.given_depth is not None else [[0.0, 0.0]], dtype=dt_)
        self.gw = np.ascontiguousarray(self.given_wl if self.given_wl is not None else [[0.0, 0.0]], dtype=dt_)
        self.gq = np.ascontiguousarray(self.given_q if self.given_q is not None else [[0.0, 0.0, 0.0]], dtype=dt_)
        self.ctrl = np.array([self.t, self.first_dt, float(self.export_n), self.export_T], dtype=np.float64)
        self.infil = None
        if self.hydraulic_conductivity is not None:
            self.infil = [np.ascontiguousarray(v, dtype=dt_) for v in
                          (self.hydraulic_conductivity, self.capillary_head, self.water_content_diff)]
        self.F = np.zeros(self.N, dt_)  # cumulative infiltrated depth
        self.rain_index = ctypes.c_int(0)
        self.last_seconds = 0.0
        self.exports = 0

    @property
    def time(self):
        return float(self.ctrl[0])

    @property
    def dt(self):
        return float(self.ctrl[1])

    def run(self, nsteps: int) -> int:
        secs = ctypes.c_double(0.0)
        rows, cols = (0, 0) if self.rt is None else self.rt.shape
        n_exp = lib().orc_run_steps(
            int(self.dtype == np.float64), int(self.order), self.N, int(nsteps), _p(self.index),
            _p(self.h_c), _p(self.wl_c), _p(self.qx_c), _p(self.qy_c), _p(self.z_c), _p(self.h_max),
            _p(self.dh), _p(self.dqx), _p(self.dqy), _p(self.dz), _p(self.lu), _p(self.man),
            _p(self.st) if self.rt is not None else None, _p(self.rt), rows, cols,
            ctypes.byref(self.rain_index),
            _p(self.gd), self.gd.shape[0], self.gd.shape[1], _p(self.gw), self.gw.shape[0], self.gw.shape[1],
            _p(self.gq), self.gq.shape[0], self.gq.shape[1], float(self.dx), float(self.cfl),
            *([_p(v) for v in self.infil] if self.infil else [None, None, None]), _p(self.F),
            _p(self.ctrl), ctypes.byref(secs))
        self.last_seconds = secs.value
        self.exports += n_exp
        return n_exp

    def raster(self, name: str, fill=np.nan) -> np.ndarray:
        """Scatter a compact field back onto the raster (postProcessing's mask>0 order)."""
        arr = getattr(self, name)
        out = np.full(self.mask.shape, fill, dtype=arr.dtype)
        out[self.mask > 0] = arr
        return out


def frozen_slope_loss(src: OracleSim, pert: OracleSim, sched: np.ndarray, gauges: np.ndarray,
                      obs: np.ndarray) -> float:
    """Gauge loss of `pert` run on the schedule sched[k] = (t_k, dt_k) with
    the order-2 slopes of `src` (advanced beside it on the same schedule):
    sum_k sum_g (h after step k at compact cell gauges[g] - obs[k, g])^2.
    A test construct with no counterpart in the reference: its central
    differences in pert's Manning / rain tables are the derivative that the
    frozen-slope adjoint computes.  Both sims FP64, order 2, same grid and
    stations; their states advance."""
    for s_ in (src, pert):
        if s_.dtype != np.float64 or s_.order != 2 or s_.rt is None:
            raise ValueError("frozen_slope_loss takes FP64 order-2 sims with a rain table")
    if src.rt.shape != pert.rt.shape:
        raise ValueError("both rain tables must have one shape")
    sched = np.ascontiguousarray(sched, dtype=np.float64).reshape(-1, 2)
    gauges = np.ascontiguousarray(gauges, dtype=np.int32)
    obs = np.ascontiguousarray(obs, dtype=np.float64).reshape(sched.shape[0], gauges.shape[0])
    planes = lambda s_: (ctypes.c_void_p * 5)(*[a.ctypes.data for a in (s_.h_c, s_.wl_c, s_.qx_c, s_.qy_c,
                                                                         s_.h_max)])
    a0, a1 = planes(src), planes(pert)
    rows, cols = src.rt.shape
    return float(lib().orc_frozen_slope_loss(
        src.N, sched.shape[0], _p(src.index), ctypes.cast(a0, _P), ctypes.cast(a1, _P), _p(src.z_c), _p(src.lu),
        _p(src.man), _p(pert.man), _p(src.st), _p(src.rt), _p(pert.rt), rows, cols, float(src.dx), float(src.cfl),
        _p(sched), _p(gauges), gauges.shape[0], _p(obs)))
