"""The C-ABI library loads, exports every symbol include/hipims_abi.h
declares, and rejects bad arguments on the host (no GPU needed: these paths
return before any HIP call)."""
import ctypes
import os
import re

import pytest
import torch

from hipims_amd import _lib, ops
from hipims_amd._lib import HpGridDesc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hipims_abi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|void|const char\*)\s+(hp_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with the header"
    assert L.hp_abi_version() == _lib.ABI_VERSION


def test_library_links_against_torch_hip_runtime():
    # the library resolves libamdhip64.so.7 to the copy torch already loaded
    L = _lib.lib()
    maps = open("/proc/self/maps").read()
    hip = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    assert len(hip) == 1, hip
    assert _lib.LIB_PATH in maps


def test_host_side_argument_checks():
    L = _lib.lib()
    assert L.hp_grid_flux(None, None) == _lib.HP_EINVAL
    assert b"null descriptor" in L.hp_last_error()
    d = HpGridDesc()
    d.order = 3
    assert L.hp_grid_update(ctypes.byref(d), None) == _lib.HP_EINVAL
    assert b"order" in L.hp_last_error()
    d.order, d.nx, d.ny, d.pitch, d.dx = 2, 10, 10, 8, 1.0
    assert L.hp_grid_step(ctypes.byref(d), 1, None) == _lib.HP_EINVAL
    assert b"pitch" in L.hp_last_error()
    assert L.hp_grid_update_mode(None, 9, None) == _lib.HP_EINVAL
    assert L.hp_flux_mask(7, 10, None, None, None, None, None) == _lib.HP_EINVAL
    rest = [None] * 11 + [None, 0, 0, None, 0, 0, None, 0, 0] + [None] * 4
    assert L.hp_flux_compact(1, 3, 1, 1, *rest) == _lib.HP_EINVAL
    assert L.hp_flux_compact(1, 2, 0, 5, *rest) == _lib.HP_OK   # empty wet list: no-op
    assert L.hp_gather(_lib.HP_F64, None, None, 0, None, None) == _lib.HP_OK
    assert L.hp_station_rain(_lib.HP_F64, -1, None, None, None, 1, None, None) == _lib.HP_EINVAL
    assert L.hp_grid_workspace_bytes(None) == 0
    assert L.hp_probe_fma(0, None, None) == _lib.HP_EINVAL
    assert L.hp_probe_fma_per_lane() == 8 * 4096


def test_calibration_loop_argument_checks():
    L = _lib.lib()
    d = HpGridDesc()
    d.order, d.nx, d.ny, d.pitch, d.dx = 1, 64, 64, 64, 1.0
    assert L.hp_calib_forward(None, None, None, 0, 1, None, None, 0, None, 0, None, None) == _lib.HP_EINVAL
    assert L.hp_calib_forward(ctypes.byref(d), None, None, 0, -1, None, None, 0, None, 0, None, None) == _lib.HP_EINVAL
    # a segment needs both its state and its flux slots, and room for three planes
    assert L.hp_calib_forward(ctypes.byref(d), None, None, 0, 1, 16, None, 0, None, 0, None, None) == _lib.HP_EINVAL
    assert b"together" in L.hp_last_error()
    assert L.hp_calib_forward(ctypes.byref(d), None, None, 0, 1, 16, 16, 64, None, 0, None, None) == _lib.HP_EINVAL
    assert b"seg_stride" in L.hp_last_error()
    # without segments the steps need the second state buffer
    assert L.hp_calib_forward(ctypes.byref(d), None, None, 0, 1, None, None, 0, None, 0, None, None) == _lib.HP_EINVAL
    assert b"h_out" in L.hp_last_error()
    ad = _lib.HpAdjointDesc()
    assert L.hp_calib_reverse(ctypes.byref(d), ctypes.byref(ad), None, 0, 1, 0, 16, 16, 1 << 20, None, 0, None,
                              None, None) == _lib.HP_EINVAL
    ad.flux_ready = 0
    assert L.hp_calib_reverse(ctypes.byref(d), ctypes.byref(ad), 16, 0, 1, 0, 16, 16, 1 << 20, None, 0, None,
                              None, None) == _lib.HP_EINVAL
    assert b"flux_ready" in L.hp_last_error()
    ad.flux_ready = 1
    assert L.hp_calib_reverse(ctypes.byref(d), ctypes.byref(ad), 16, 2, 1, 0, 16, 16, 1 << 20, None, 0, None,
                              None, None) == _lib.HP_EINVAL


def _fake_desc(dtype=_lib.HP_F32, order=1, nx=100, ny=40, pitch=128):
    """A descriptor whose plane pointers are aligned placeholders: it passes the
    host-side checks, and nothing here reaches a launch."""
    d = HpGridDesc()
    d.dtype, d.order, d.nx, d.ny, d.pitch, d.dx, d.n_classes = dtype, order, nx, ny, pitch, 1.0, 6
    for k, name in enumerate(("h", "qx", "qy", "z", "h_max", "dh", "dqx", "dqy", "type_slot", "landuse",
                              "slot_code", "manning", "ctrl")):
        setattr(d, name, (k + 1) << 12)
    return d


def test_adjoint_and_ensemble_argument_checks():
    L = _lib.lib()
    d = _fake_desc()
    # grad rows cover the update chunks and the tiles of the single-pass
    # reverse step (64 x 8 in FP32, 64 x 4 in FP64)
    assert L.hp_adjoint_grad_blocks(ctypes.byref(d)) >= max((128 * 40 + 1023) // 1024, 2 * 5)
    d.dtype = _lib.HP_F64
    assert L.hp_adjoint_grad_blocks(ctypes.byref(d)) >= 2 * 10
    d.dtype = _lib.HP_F32
    ad = _lib.HpAdjointDesc()
    for k, name in enumerate(("lam_h", "lam_qx", "lam_qy", "cot_h", "cot_qx", "cot_qy", "xc", "yc", "grads", "gn")):
        setattr(ad, name, (k + 64) << 12)
    ad.cpitch, ad.flux_ready = 128, 1
    ad.separate_passes = 2
    assert L.hp_adjoint_step(ctypes.byref(d), ctypes.byref(ad), None) == _lib.HP_EINVAL
    assert b"separate_passes" in L.hp_last_error()
    ad.separate_passes = 0
    ad.gauge_of = 1 << 20  # a gauge plane without its loss rows
    assert L.hp_adjoint_step(ctypes.byref(d), ctypes.byref(ad), None) == _lib.HP_EINVAL
    assert b"loss_h" in L.hp_last_error()
    # the single-pass ensemble needs every domain's second set of planes
    descs = (HpGridDesc * 2)(_fake_desc(), _fake_desc(nx=60, pitch=64))
    e = _lib.HpEnsemble()
    need = L.hp_ensemble_workspace_bytes(2, descs)
    assert need > 0
    assert L.hp_ensemble_init(2, descs, 2, 1 << 24, need, ctypes.byref(e), None) == _lib.HP_EINVAL
    assert b"single_pass" in L.hp_last_error()
    assert L.hp_ensemble_init(2, descs, 1, 1 << 24, need, ctypes.byref(e), None) == _lib.HP_EINVAL
    assert b"h_out" in L.hp_last_error()
    assert L.hp_ensemble_init(2, descs, 0, 1 << 24, need - 1, ctypes.byref(e), None) == _lib.HP_ESPACE
    # an even step count is asked of a single-pass handle before anything runs
    e.n_domains, e.doms, e.single_pass = 2, 1 << 24, 1
    assert L.hp_ensemble_step(ctypes.byref(e), 3, None) == _lib.HP_EINVAL
    assert b"even" in L.hp_last_error()


def test_ops_reject_host_tensors_like_the_reference():
    h = torch.zeros(4, dtype=torch.float64)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        ops.flux_mask_update(torch.zeros(4, dtype=torch.bool), h, torch.zeros(20, dtype=torch.int32),
                             torch.zeros(1, dtype=torch.float64))
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        ops.update_timestep(*([torch.zeros(1)] * 10))


def test_top_level_shim_modules_resolve_to_the_hip_ops():
    import euler_update
    import fluxCal_2ndOrder_jh_improved
    import fluxCalculation_jh_modified_surface
    import fluxMask
    import friction_implicit_andUpdate_jh
    import station_PrecipitationCalculation
    import timeControl
    assert fluxMask.update is ops.flux_mask_update
    assert fluxCalculation_jh_modified_surface.addFlux is ops.add_flux_order1
    assert fluxCal_2ndOrder_jh_improved.addFlux is ops.add_flux_order2
    assert station_PrecipitationCalculation.addStation_Precipitation is ops.add_station_precipitation
    assert friction_implicit_andUpdate_jh.addFriction_eulerUpdate is ops.add_friction_euler_update
    assert timeControl.updateTimestep is ops.update_timestep
    assert euler_update.update is ops.euler_update
