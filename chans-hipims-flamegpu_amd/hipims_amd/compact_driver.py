"""The reference driver's per-step sequence on the compact layout, over this
library's compact ops (SWE_CUDA.py:115-227, 339-424, 472-502, 780-831).

``Godunov`` (hipims_amd.SWE_CUDA) runs the step on structured planes with the
time control on the device.  ``CompactGodunov`` instead keeps the reference's
own data path: compact arrays plus the int32[5N] neighbour table, a wet list
built with ``nonzero`` before the flux and before the update, the station
rain interpolated on the host, ``torch.min`` of the per-cell CFL candidates
and the dt / export logic on the host with ``.item()`` reads.  Each op is the
HIP replacement of the reference's extension op (``hipims_amd.ops``), so this
class is the reference's plain per-step path on this GPU: the baseline the
batched ensemble is measured against, and a second check of the compact ops
in sequence.
"""
from __future__ import annotations

import numpy as np
import torch

from . import ops
from .preprocessing import neighbour_index


class CompactGodunov:
    def __init__(self, device, dx, CFL, Export_timeStep, t=0.0, export_n=0, secondOrder=False,
                 firstTimeStep=1.0e-4, tensorType=torch.float64):
        self._device = torch.device(device)
        self._tensorType = tensorType
        self._secondOrder = bool(secondOrder)
        T = lambda v: torch.tensor([float(v)], dtype=tensorType, device=self._device)
        self.dx = T(dx)
        self.cfl = T(CFL)
        self.t = T(t)
        self.dt = T(firstTimeStep)
        self._maxTimeStep = T(60.0)
        self.export_timeStep = float(Export_timeStep)
        self._export_n = int(export_n)
        self._normal = torch.tensor([[1.0, 0.0], [0.0, 1.0], [-1.0, 0.0], [0.0, -1.0]], dtype=tensorType,
                                    device=self._device)
        self._manning = torch.tensor([0.0], dtype=tensorType, device=self._device)
        self._given_depth = torch.tensor([[0.0, 0.0]], dtype=tensorType, device=self._device)
        self._given_wl = torch.tensor([[0.0, 0.0]], dtype=tensorType, device=self._device)
        self._given_discharge = torch.tensor([[0.0, 0.0, 0.0]], dtype=tensorType, device=self._device)
        self._rainfall_station_time_index = 0
        self._rainStationData = None

    def init__fluidField_tensor(self, mask, h, qx, qy, wl, z, device):
        mask = np.asarray(mask.cpu() if torch.is_tensor(mask) else mask)
        valid = mask > 0
        c = lambda a: torch.as_tensor(np.asarray(a.cpu() if torch.is_tensor(a) else a)[valid],
                                      dtype=self._tensorType, device=self._device).contiguous()
        self._h_internal, self._qx_internal, self._qy_internal = c(h), c(qx), c(qy)
        self._wl_internal, self._z_internal = c(wl), c(z)
        self._h_max = self._h_internal.clone()
        self._h_update = torch.zeros_like(self._h_internal)
        self._qx_update = torch.zeros_like(self._h_internal)
        self._qy_update = torch.zeros_like(self._h_internal)
        self._z_update = torch.zeros_like(self._h_internal)
        self._landuseMask = torch.zeros(self._h_internal.shape, dtype=torch.uint8, device=self._device)
        self._rainfall_station_Mask = torch.zeros(self._h_internal.shape, dtype=torch.int16, device=self._device)
        self._index = torch.from_numpy(neighbour_index(mask).reshape(-1)).to(self._device)

    def set__frictionField_tensor(self, manning, device):
        self._manning = torch.as_tensor(np.atleast_1d(np.asarray(manning, dtype=np.float64)), dtype=self._tensorType,
                                        device=self._device)

    def set_landuse(self, mask, landuseMask, device):
        mask = np.asarray(mask.cpu() if torch.is_tensor(mask) else mask)
        lu = np.asarray(landuseMask.cpu() if torch.is_tensor(landuseMask) else landuseMask)
        self._landuseMask = torch.as_tensor(lu[mask > 0], dtype=torch.uint8, device=self._device)

    def set_distributed_rainfall_station_Mask(self, mask, rainfall_station_Mask, device):
        mask = np.asarray(mask.cpu() if torch.is_tensor(mask) else mask)
        st = np.asarray(rainfall_station_Mask.cpu() if torch.is_tensor(rainfall_station_Mask)
                        else rainfall_station_Mask)
        self._rainfall_station_Mask = torch.as_tensor(st[mask > 0], dtype=torch.int16, device=self._device)
        self._rainfall_station_time_index = 0

    def set_boundary_tensor(self, boundList, device):
        if "H_GIVEN" in boundList:
            self._given_depth = torch.tensor(boundList["H_GIVEN"], dtype=self._tensorType, device=self._device)
        if "WL_GIVEN" in boundList:
            self._given_wl = torch.tensor(boundList["WL_GIVEN"], dtype=self._tensorType, device=self._device)
        if "Q_GIVEN" in boundList:
            self._given_discharge = torch.tensor(boundList["Q_GIVEN"], dtype=self._tensorType, device=self._device)

    # -- the step -------------------------------------------------------------------
    def addFlux(self):
        wet = torch.zeros_like(self._h_internal, dtype=torch.bool)
        ops.flux_mask_update(wet, self._h_internal, self._index, self.t)
        self._wetMask = torch.flatten(wet.nonzero().type(torch.int32))
        op = ops.add_flux_order2 if self._secondOrder else ops.add_flux_order1
        op(self._wetMask, self._h_update, self._qx_update, self._qy_update, self._h_internal, self._wl_internal,
           self._z_internal, self._qx_internal, self._qy_internal, self._index, self._normal, self._given_depth,
           self._given_wl, self._given_discharge, self.dx, self.t, self.dt)

    def _station_rain(self, table):
        t = self.t.item()
        i = self._rainfall_station_time_index
        if t < table[-1, 0]:
            if t < table[i + 1, 0]:
                per = (t - table[i, 0]) / (table[i + 1, 0] - table[i, 0])
                self._rainStationData = torch.from_numpy(table[i, 1:] + per * (table[i + 1, 1:] - table[i, 1:])).to(
                    device=self._device)
            else:
                self._rainfall_station_time_index += 1
                self._station_rain(table)
        else:
            self._rainStationData -= self._rainStationData

    def addStation_PrecipitationSource(self, rainfall_ndarray_data, device):
        self._station_rain(np.asarray(rainfall_ndarray_data, dtype=np.float64))
        ops.add_station_precipitation(self._h_update, self._rainfall_station_Mask, self._rainStationData, self.dt)

    def time_friction_euler_update_cuda(self, device):
        time_upper = self.dt * 10.0
        self._wetMask = torch.flatten(((self._h_update.abs() > 0.0) + (self._h_internal >= 0.0)).nonzero()).type(
            torch.int32)
        ops.add_friction_euler_update(self._wetMask, self._h_update, self._qx_update, self._qy_update,
                                      self._z_update, self._landuseMask, self._h_internal, self._wl_internal,
                                      self._qx_internal, self._qy_internal, self._z_internal, self._manning, self.dt)
        self._h_update[:] = 0.0
        self._qx_update[:] = 0.0
        self._qy_update[:] = 0.0
        self._z_update[:] = 0.0
        accel = torch.full(self._wetMask.size(), self._maxTimeStep.item(), dtype=self._tensorType,
                           device=self._device)
        ops.update_timestep(self._wetMask, accel, self._h_max, self._h_internal, self._qx_internal,
                            self._qy_internal, self.dx, self.cfl, self.t, self.dt)
        if accel.size(0) != 0:
            self.dt = torch.min(accel).reshape(1)
        self.dt = min(self.dt, time_upper)
        if (self.dt + self.t).item() >= float(self._export_n + 1) * self.export_timeStep:
            self.dt = (self._export_n + 1) * self.export_timeStep - self.t
            self._export_n += 1
        self.t = self.t + self.dt

    def get_h(self):
        return self._h_internal
