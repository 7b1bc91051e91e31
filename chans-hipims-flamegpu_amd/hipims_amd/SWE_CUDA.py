"""Drop-in ``Godunov`` driver (CHANS-HiPIMS/hipims/pythonHipims/SWE_CUDA.py:62-928).

Same constructor, method names, arguments and call order as the reference
class, so ``CHANS-HiPIMS_server.py`` (and the CatchFlood-style case scripts)
can use it by changing one import.  Underneath, the state lives as 2-D planes
(``hipims_amd.grid.GridSolver``) advanced by the HIP kernels; time control
(dt, t, the export clamp) runs on the device.

Behaviour kept from the reference, on purpose:
  * ``t`` advances by the dt computed for the *next* step (:818-831);
  * dt = min CFL step over the listed cells, 60 s for dry cells, capped at 10x
    the previous dt, clamped to land on export times (:799-830);
  * the flux list rule of fluxMask, including that compact cell 0 never counts
    as a wet neighbour (fluxMaskGenerator_Kernel.cu:36);
  * station rain linear in time and 0 after the table's last row (:472-495).
Differences, each documented where it matters:
  * ``wl`` is derived as z + h (the reference stores exactly that after every
    update; an inconsistent initial ``wl`` raster is not carried);
  * in float32, t and dt are held in float64 (the reference keeps them in the
    state dtype);
  * rasters must carry a ring of nodata (the reference's torch.roll neighbour
    table wraps around otherwise).
Green-Ampt infiltration (set__infiltrationField_tensor / addInfiltrationSource)
is supported; sediment transport and the explicit-friction variants are
outside this library and raise NotImplementedError.
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch

from . import _lib
from .grid import GridSolver


class Godunov:
    def __init__(self, device, dx, CFL, Export_timeStep, t=0.0, export_n=0, secondOrder=False,
                 firstTimeStep=1.0e-4, tensorType=torch.float64):
        self._device = torch.device(device) if not isinstance(device, torch.device) else device
        if self._device.type == "cuda" and self._device.index is None:
            self._device = torch.device("cuda", torch.cuda.current_device())
        self._tensorType = tensorType
        self._secondOrder = bool(secondOrder)
        self._dx = float(dx)
        self._cfl = float(CFL)
        self.export_timeStep = float(Export_timeStep)
        self._outpath = "outpath"
        self._ctrl = torch.zeros(_lib.CTRL_DOUBLES, dtype=torch.float64, device=self._device)
        self._ctrl[0] = float(t)
        self._ctrl[1] = float(firstTimeStep)
        self._ctrl[2] = self.export_timeStep
        self._ctrl[3] = float(export_n)
        self._manning = np.array([0.0])
        self._landuse_raster = None
        self._station_raster = None
        self._given = {"H_GIVEN": None, "WL_GIVEN": None, "Q_GIVEN": None}
        self._solver: GridSolver | None = None
        self._pending_rain = None  # ("station", table, has_station) | ("field",)
        self._rain_key = None
        self._rainfall_station_time_index = 0
        self._single_station = True

    # -- scalars ---------------------------------------------------------------
    @property
    def t(self) -> torch.Tensor:
        return self._ctrl[0:1]

    @t.setter
    def t(self, v):
        self._ctrl[0] = float(v.item() if torch.is_tensor(v) else v)

    @property
    def dt(self) -> torch.Tensor:
        return self._ctrl[1:2]

    @dt.setter
    def dt(self, v):
        self._ctrl[1] = float(v.item() if torch.is_tensor(v) else v)

    @property
    def dx(self) -> torch.Tensor:
        return torch.tensor([self._dx], dtype=self._tensorType, device=self._device)

    @property
    def cfl(self) -> torch.Tensor:
        return torch.tensor([self._cfl], dtype=self._tensorType, device=self._device)

    @property
    def _export_n(self) -> int:
        return int(self._ctrl[3].item())

    @property
    def solver(self) -> GridSolver:
        if self._solver is None:
            raise RuntimeError("call init__fluidField_tensor first")
        return self._solver

    # -- initialisation (SWE_CUDA.py:115-309, 448-470) ----------------------------
    def init__fluidField_tensor(self, mask, h, qx, qy, wl, z, device):
        self._mask = mask.detach().cpu().numpy() if torch.is_tensor(mask) else np.asarray(mask)
        self._solver = GridSolver(self._mask, h, qx, qy, z, self._dx, cfl=self._cfl,
                                  order=2 if self._secondOrder else 1, dtype=self._tensorType, device=self._device,
                                  manning=self._manning, ctrl=self._ctrl)
        if self._landuse_raster is not None:
            self._solver.set_landuse(self._landuse_raster)
        self._apply_boundary()

    def set__frictionField_tensor(self, manning, device):
        m = manning.detach().cpu().numpy() if torch.is_tensor(manning) else np.atleast_1d(np.asarray(manning))
        self._manning = np.asarray(m, dtype=np.float64).reshape(-1)
        if self._solver is not None:
            self._solver.set_manning(self._manning)
            if self._landuse_raster is not None:
                self._solver.set_landuse(self._landuse_raster)

    def set_landuse(self, mask, landuseMask, device):
        lu = landuseMask.detach().cpu().numpy() if torch.is_tensor(landuseMask) else np.asarray(landuseMask)
        self._landuse_raster = lu.astype(np.int64)
        if self._solver is not None:
            self._solver.set_landuse(self._landuse_raster)

    set_landuse_cpu_to_gpu = set_landuse

    def set_boundary_tensor(self, boundList, device):
        for key in ("H_GIVEN", "WL_GIVEN", "Q_GIVEN"):
            if key in boundList:
                v = boundList[key]
                self._given[key] = np.loadtxt(v) if isinstance(v, str) else np.asarray(v, dtype=np.float64)
        self._apply_boundary()

    def _apply_boundary(self):
        if self._solver is not None:
            self._solver.set_boundary(self._given["H_GIVEN"], self._given["WL_GIVEN"], self._given["Q_GIVEN"])

    def set_distributed_rainfall_station_Mask(self, mask, rainfall_station_Mask, device):
        st = rainfall_station_Mask
        self._station_raster = (st.detach().cpu().numpy() if torch.is_tensor(st) else np.asarray(st)).astype(np.int16)
        valid = (mask.detach().cpu().numpy() if torch.is_tensor(mask) else np.asarray(mask)) > 0
        # one station everywhere: the update pass then reads no per-cell station plane
        self._single_station = bool((self._station_raster[valid] == 0).all())
        self._rainfall_station_time_index = 0
        self._rain_key = None

    set_distributed_rainfall_station_Mask_cpu_to_gpu = set_distributed_rainfall_station_Mask

    def setOutPutPath(self, outpath):
        self._outpath = outpath

    # -- per-step methods --------------------------------------------------------
    def addFlux(self):
        """Flux pass over the listed cells (SWE_CUDA.py:339-424)."""
        self.solver.flux()

    def _use_station_table(self, table, station):
        tab = np.asarray(table, dtype=np.float64)
        key = (id(table), tab.shape, tab.tobytes() if tab.size <= 65536 else None, station is not None)
        if key != self._rain_key:
            self.solver.set_station_rain(tab, station)
            self._rain_key = key
        if self._pending_rain is not None:
            raise RuntimeError("one rain source per step: a rain source is already pending")
        self._pending_rain = "table"

    def addStation_PrecipitationSource(self, rainfall_ndarray_data, device):
        """Station rain, linear in time (SWE_CUDA.py:472-502); applied in the
        update pass with the dt of this step."""
        st = None if (self._station_raster is None or self._single_station) else self._station_raster
        self._use_station_table(rainfall_ndarray_data, st)

    def set_uniform_rainfall_time_index(self):
        self._rainfall_station_time_index = 0

    def add_uniform_PrecipitationSource(self, rainfall_ndarray_data, device):
        """Uniform rain from a [t, r] table (SWE_CUDA.py:517-550)."""
        tab = np.asarray(rainfall_ndarray_data, dtype=np.float64)
        self._use_station_table(tab[:, :2], None)

    def addRadarRainfall(self, rainfall_raster):
        """Per-cell rain in compact order (SWE_CUDA.py:556-557)."""
        s = self.solver
        if s.rain_field is None or s.rain_mode != _lib.HP_RAIN_FIELD:
            s.set_rain_field(np.zeros((s.layout.ny, s.layout.nx), dtype=np.float32))
        s.load_compact(s.rain_field, torch.as_tensor(rainfall_raster).reshape(-1).to(torch.float32))
        self._rain_key = None
        if self._pending_rain is not None:
            raise RuntimeError("one rain source per step: a rain source is already pending")
        self._pending_rain = "field"

    def _run_update(self, mode):
        rain = self._pending_rain is not None
        self._pending_rain = None
        self.solver.update(mode, rain=rain)

    def _export_if_due(self):
        if self._ctrl[4].item() > 0:
            self._ctrl[4] = 0.0
            # the reference exports before t += dt, naming the files t + dt,
            # which is the already advanced t here
            self._export(self._time_tag(self.t.item()))
            print("give a output")

    def time_friction_euler_update_cuda(self, device):
        """Friction + update + time control (SWE_CUDA.py:780-831)."""
        self._run_update(_lib.HP_UPDATE_STEP)
        self._export_if_due()

    def friction_euler_update_cuda(self):
        """Friction + update without time control (SWE_CUDA.py:766-778)."""
        self._run_update(_lib.HP_UPDATE_FRICTION)

    def euler_update(self):
        """Plain update without friction (SWE_CUDA.py:712-724)."""
        self._run_update(_lib.HP_UPDATE_EULER)

    def time_update_cuda(self, device):
        """Time control only, over cells with h > 1e-6 (SWE_CUDA.py:726-764)."""
        self.solver.update(_lib.HP_UPDATE_CFL, rain=False)
        self._export_if_due()

    def run_steps(self, nsteps: int, rainfall_ndarray_data=None):
        """`nsteps` server steps (flux, station rain, friction/update/time
        control) enqueued without host synchronisation; exports are checked once
        at the end, so keep export times outside the batch when they matter."""
        s = self.solver
        if rainfall_ndarray_data is not None:
            self.addStation_PrecipitationSource(rainfall_ndarray_data, self._device)
            self._pending_rain = None
        else:
            s.clear_rain()
            self._rain_key = None
        s.step(int(nsteps))
        self._export_if_due()

    def rungeKutta_update(self, rainfallMatrix, device):
        """Two Euler stages averaged (SWE_CUDA.py:870-892)."""
        self._rk2(lambda: self.add_uniform_PrecipitationSource(rainfallMatrix, device))

    def rungeKutta_update_nonUniformRain(self, rainfallMatrix, device):
        """As rungeKutta_update with station rain (SWE_CUDA.py:894-916)."""
        self._rk2(lambda: self.addStation_PrecipitationSource(rainfallMatrix, device))

    def _rk2(self, add_rain):
        s = self.solver
        h0, qx0, qy0 = s.h.clone(), s.qx.clone(), s.qy.clone()
        for _ in range(2):
            self.addFlux()
            add_rain()
            self.friction_euler_update_cuda()
        s.h.add_(h0).div_(2.0)
        s.qx.add_(qx0).div_(2.0)
        s.qy.add_(qy0).div_(2.0)

    # -- views and output (SWE_CUDA.py:290-301, 559-566, 836-868) ------------------
    def get_h(self):
        return self.solver.compact(self.solver.h)

    def get_qx(self):
        return self.solver.compact(self.solver.qx)

    def get_qy(self):
        return self.solver.compact(self.solver.qy)

    @property
    def _h_internal(self):
        return self.get_h()

    @property
    def _qx_internal(self):
        return self.get_qx()

    @property
    def _qy_internal(self):
        return self.get_qy()

    @property
    def _z_internal(self):
        return self.solver.compact(self.solver.z)

    @property
    def _wl_internal(self):
        return self.solver.compact(self.solver.wl.contiguous())

    @property
    def _h_max(self):
        return self.solver.compact(self.solver.h_max)

    def raster_update(self):
        """The compact state into raster tensors h, wl, z, qx, qy at the cells
        of `internal` (SWE_CUDA.py:918-923).  A caller that has set these
        rasters (and `internal`) gets them updated in place, as with the
        reference; missing ones are created with zeros outside the valid
        cells."""
        internal = getattr(self, "internal", None)
        if internal is None:
            internal = torch.as_tensor(self._mask > 0, device=self._device)
        for name, val in (("h", self._h_internal), ("wl", self._wl_internal), ("z", self._z_internal),
                          ("qx", self._qx_internal), ("qy", self._qy_internal)):
            r = getattr(self, name, None)
            if r is None:
                r = torch.zeros(self._mask.shape, dtype=val.dtype, device=val.device)
                setattr(self, name, r)
            r[internal] = val.to(dtype=r.dtype, device=r.device)

    def observeGauges_write(self, observe_index, dataStoreList, n):
        if n % 10 == 0:
            row = [self.t.item()]
            row += list(self.get_h()[observe_index].cpu().numpy())
            dataStoreList.append(row)

    def _time_tag(self, t: float) -> str:
        # the reference names files with str() of a state-dtype tensor
        return str(torch.tensor([t], dtype=self._tensorType, device=self._device))

    def exportField(self):
        """Snapshots named with str(t + dt), as SWE_CUDA.py:836-868."""
        self._export(self._time_tag(self.t.item() + self.dt.item()))

    def _export(self, tag: str):
        for name, val in (("h", self._h_internal), ("qx", self._qx_internal), ("qy", self._qy_internal),
                          ("wl", self._wl_internal), ("h_max", self._h_max)):
            torch.save(val, os.path.join(self._outpath, f"{name}_{tag}.pt"))

    # -- infiltration (SWE_CUDA.py:236-238, 311-337, 439-443) -----------------------
    def add_cumulativeWaterDepth_Field(self, device):
        self._cum_depth_requested = True

    def set__infiltrationField_tensor(self, hydraulic_conductivity, capillary_head, water_content_diff, device):
        """Per-landuse-class Green-Ampt parameters; the cumulative infiltrated
        depth starts at 0."""
        self.add_cumulativeWaterDepth_Field(device)
        conv = lambda v: v.detach().cpu().numpy() if torch.is_tensor(v) else v
        self.solver.set_infiltration(conv(hydraulic_conductivity), conv(capillary_head), conv(water_content_diff))

    def addInfiltrationSource(self):
        """Green-Ampt over this step's flux list, subtracted from the step's
        depth change after the rain (the reference calls it after
        addStation_PrecipitationSource, CatchFlood_main.py:164-166)."""
        self.solver.infiltration()

    @property
    def _cumulativeWaterDepth(self):
        return self.solver.compact(self.solver.F)

    # -- outside this library ------------------------------------------------------
    def _unsupported(self, *a, **k):
        raise NotImplementedError("sediment and explicit-friction paths are not part of this library")

    add_concentrationField = addFriction = _unsupported
    enableSediModule = add_sedi_source = sedi_euler_update = run_updating = _unsupported
