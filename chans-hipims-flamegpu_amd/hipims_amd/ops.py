"""The reference's torch-extension ops, on the reference's compact layout.

Each function keeps the module name, function name, argument order and
in-place semantics of the op it replaces, and checks its tensors the way the
reference's bindings do (TORCH_CHECK is_cuda / is_contiguous -> RuntimeError):

  fluxMask.update ........................ fluxMaskGenerator.cpp:24-35
  fluxCalculation_jh_modified_surface.addFlux   fluxCal_jh_modified_surface.cpp:28-57
  fluxCal_2ndOrder_jh_improved.addFlux ... fluxCal_2ndOrder_jh_improved.cpp:28-57
  station_PrecipitationCalculation.addStation_Precipitation
                                           stationPrecipitation_interface.cpp:26-43
  friction_implicit_andUpdate_jh.addFriction_eulerUpdate
                                           friction_implicit_andUpdate_jh_interface.cpp:29-55
  timeControl.updateTimestep ............. timeControl.cpp:27-45
  euler_update.update .................... euler_update.cpp:27-47
  infiltrationCalculation.addinfiltration  infiltration_interface.cpp:21-53

(paths under CHANS-HiPIMS/hipims/cuda/).  All launch on torch's current
stream of the tensors' device, asynchronously.
"""
from __future__ import annotations

import torch

from . import _lib
from ._lib import check, lib


def _check_input(**tensors):
    for name, t in tensors.items():
        if not torch.is_tensor(t) or not t.is_cuda:
            raise RuntimeError(f"{name} must be a CUDA tensor.")
        if not t.is_contiguous():
            raise RuntimeError(f"{name} must be contiguous.")


def _same_dtype(ref: torch.Tensor, **tensors):
    for name, t in tensors.items():
        if t.dtype != ref.dtype:
            raise RuntimeError(f"{name} has dtype {t.dtype}, expected {ref.dtype}")


def _idx(name: str, t: torch.Tensor, dtype=torch.int32):
    if t.dtype != dtype:
        raise RuntimeError(f"{name} must be {dtype}, got {t.dtype}")


def _stream(t: torch.Tensor) -> int:
    return _lib.stream_ptr(t.device)


def flux_mask_update(fluxMask, h, index, t):
    """fluxMask.update(fluxMask: bool[N], h, index: int32[5N], t)."""
    _check_input(fluxMask=fluxMask, h=h, index=index, t=t)
    if fluxMask.dtype not in (torch.bool, torch.uint8):
        raise RuntimeError("fluxMask must be torch.bool")
    _idx("index", index)
    _same_dtype(h, t=t)
    with torch.cuda.device(h.device):
        check(lib().hp_flux_mask(_lib.dtype_code(h.dtype), h.numel(), fluxMask.data_ptr(), h.data_ptr(),
                                 index.data_ptr(), t.data_ptr(), _stream(h)), "fluxMask.update")


def _add_flux(order, wetMask, h_flux, qx_flux, qy_flux, h, wl, z, qx, qy, index, normal, given_depth, given_wl,
              given_q, dx, t, dt):
    _check_input(wetMask=wetMask, h_flux=h_flux, qx_flux=qx_flux, qy_flux=qy_flux, h=h, wl=wl, z=z, qx=qx, qy=qy,
                 index=index, normal=normal, given_depth=given_depth, given_wl=given_wl, given_q=given_q, dx=dx, t=t,
                 dt=dt)
    _idx("wetMask", wetMask)
    _idx("index", index)
    _same_dtype(h, h_flux=h_flux, qx_flux=qx_flux, qy_flux=qy_flux, wl=wl, z=z, qx=qx, qy=qy, normal=normal,
                given_depth=given_depth, given_wl=given_wl, given_q=given_q, dx=dx, t=t, dt=dt)
    M, N = wetMask.numel(), h.numel()
    if M == 0:
        return
    gd = given_depth if given_depth.dim() == 2 else given_depth.view(1, -1)
    gw = given_wl if given_wl.dim() == 2 else given_wl.view(1, -1)
    gq = given_q if given_q.dim() == 2 else given_q.view(1, -1)
    with torch.cuda.device(h.device):
        check(lib().hp_flux_compact(
            _lib.dtype_code(h.dtype), order, M, N, wetMask.data_ptr(), h_flux.data_ptr(), qx_flux.data_ptr(),
            qy_flux.data_ptr(), h.data_ptr(), wl.data_ptr(), z.data_ptr(), qx.data_ptr(), qy.data_ptr(),
            index.data_ptr(), normal.data_ptr(), gd.data_ptr(), gd.shape[0], gd.shape[1], gw.data_ptr(), gw.shape[0],
            gw.shape[1], gq.data_ptr(), gq.shape[0], gq.shape[1], dx.data_ptr(), t.data_ptr(), dt.data_ptr(),
            _stream(h)), "addFlux")


def add_flux_order1(*args):
    """fluxCalculation_jh_modified_surface.addFlux(wetMask, h_flux, qx_flux, qy_flux, h, wl, z, qx, qy,
    index, normal, given_depth, given_wl, given_q, dx, t, dt)."""
    _add_flux(1, *args)


def add_flux_order2(*args):
    """fluxCal_2ndOrder_jh_improved.addFlux(same arguments as order 1)."""
    _add_flux(2, *args)


def add_station_precipitation(h_update, rainStationMask, rainStationData, dt):
    """station_PrecipitationCalculation.addStation_Precipitation(h_update, rainStationMask: int16,
    rainStationData, dt).  rainStationData may be float64 (as the reference driver builds it) or of
    h_update's dtype."""
    _check_input(h_update=h_update, rainStationMask=rainStationMask, rainStationData=rainStationData, dt=dt)
    _idx("rainStationMask", rainStationMask, torch.int16)
    _same_dtype(h_update, dt=dt)
    if rainStationData.dtype not in (torch.float64, h_update.dtype):
        raise RuntimeError("rainStationData must be float64 or h_update's dtype")
    with torch.cuda.device(h_update.device):
        check(lib().hp_station_rain(_lib.dtype_code(h_update.dtype), rainStationMask.numel(), h_update.data_ptr(),
                                    rainStationMask.data_ptr(), rainStationData.data_ptr(),
                                    int(rainStationData.dtype == torch.float64), dt.data_ptr(), _stream(h_update)),
              "addStation_Precipitation")


def add_friction_euler_update(wetMask, h_update, qx_update, qy_update, z_update, landuse, h, wl, qx, qy, z,
                              manning, dt):
    """friction_implicit_andUpdate_jh.addFriction_eulerUpdate(...)."""
    _check_input(wetMask=wetMask, landuse=landuse, h=h, qx=qx, z=z, wl=wl, qy=qy, h_update=h_update,
                 z_update=z_update, qx_update=qx_update, qy_update=qy_update, dt=dt, manning=manning)
    _idx("wetMask", wetMask)
    _idx("landuse", landuse, torch.uint8)
    _same_dtype(h, h_update=h_update, qx_update=qx_update, qy_update=qy_update, z_update=z_update, wl=wl, qx=qx,
                qy=qy, z=z, manning=manning, dt=dt)
    with torch.cuda.device(h.device):
        check(lib().hp_friction_update_compact(
            _lib.dtype_code(h.dtype), wetMask.numel(), wetMask.data_ptr(), h_update.data_ptr(), qx_update.data_ptr(),
            qy_update.data_ptr(), z_update.data_ptr(), landuse.data_ptr(), h.data_ptr(), wl.data_ptr(), qx.data_ptr(),
            qy.data_ptr(), z.data_ptr(), manning.data_ptr(), dt.data_ptr(), _stream(h)), "addFriction_eulerUpdate")


def update_timestep(wetMask, accelerator, h_max, h, qx, qy, dx, CFL, t, dt):
    """timeControl.updateTimestep(wetMask, accelerator, h_max, h, qx, qy, dx, CFL, t, dt): writes h_max and
    accelerator[i] (the CFL step of listed cell i, 60 when dry)."""
    _check_input(wetMask=wetMask, h=h, h_max=h_max, qx=qx, qy=qy, dx=dx, CFL=CFL, dt=dt, t=t)
    _idx("wetMask", wetMask)
    _same_dtype(h, accelerator=accelerator, h_max=h_max, qx=qx, qy=qy, dx=dx, CFL=CFL)
    if accelerator.numel() < wetMask.numel():
        raise RuntimeError("accelerator must have one entry per listed cell")
    with torch.cuda.device(h.device):
        check(lib().hp_time_control_compact(
            _lib.dtype_code(h.dtype), wetMask.numel(), wetMask.data_ptr(), accelerator.data_ptr(), h_max.data_ptr(),
            h.data_ptr(), qx.data_ptr(), qy.data_ptr(), dx.data_ptr(), CFL.data_ptr(), _stream(h)),
            "updateTimestep")


def euler_update(updateMask, h_update, qx_update, qy_update, z_update, h, wl, z, qx, qy):
    """euler_update.update(updateMask, h_update, qx_update, qy_update, z_update, h, wl, z, qx, qy)."""
    _check_input(updateMask=updateMask, h=h, qx=qx, qy=qy, wl=wl, z=z, qx_update=qx_update, qy_update=qy_update,
                 h_update=h_update, z_update=z_update)
    _idx("updateMask", updateMask)
    _same_dtype(h, h_update=h_update, qx_update=qx_update, qy_update=qy_update, z_update=z_update, wl=wl, z=z,
                qx=qx, qy=qy)
    with torch.cuda.device(h.device):
        check(lib().hp_euler_update_compact(
            _lib.dtype_code(h.dtype), updateMask.numel(), updateMask.data_ptr(), h_update.data_ptr(),
            qx_update.data_ptr(), qy_update.data_ptr(), z_update.data_ptr(), h.data_ptr(), wl.data_ptr(),
            z.data_ptr(), qx.data_ptr(), qy.data_ptr(), _stream(h)), "euler_update.update")


def add_infiltration(wetMask, h_update, landuse, h, hydraulic_conductivity, capillary_head, water_content_diff,
                     cumulative_depth, dt):
    """infiltrationCalculation.addinfiltration(wetMask, h_update, landuse, h, hydraulic_conductivity,
    capillary_head, water_content_diff, cumulative_depth, dt): Green-Ampt over the listed cells,
    h_update -= dF, cumulative_depth += dF."""
    _check_input(wetMask=wetMask, landuse=landuse, h=h, h_update=h_update, dt=dt,
                 hydraulic_conductivity=hydraulic_conductivity, capillary_head=capillary_head,
                 water_content_diff=water_content_diff, cumulative_depth=cumulative_depth)
    _idx("wetMask", wetMask)
    _idx("landuse", landuse, torch.uint8)
    _same_dtype(h, h_update=h_update, hydraulic_conductivity=hydraulic_conductivity, capillary_head=capillary_head,
                water_content_diff=water_content_diff, cumulative_depth=cumulative_depth, dt=dt)
    with torch.cuda.device(h.device):
        check(lib().hp_infiltration_compact(
            _lib.dtype_code(h.dtype), wetMask.numel(), wetMask.data_ptr(), h_update.data_ptr(), landuse.data_ptr(),
            h.data_ptr(), hydraulic_conductivity.data_ptr(), capillary_head.data_ptr(), water_content_diff.data_ptr(),
            cumulative_depth.data_ptr(), dt.data_ptr(), _stream(h)), "addinfiltration")
