"""ctypes binding of libhipims_amd.so (declared in include/hipims_abi.h).

torch is imported before the library is opened: torch ships its own HIP
runtime under the same soname (libamdhip64.so.7), so loading the library
afterwards binds it to the runtime torch already uses and device pointers and
streams are shared.  There is no CPU fallback: if the library is missing the
ops raise.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import torch  # noqa: F401  (must be loaded before the library, see above)

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libhipims_amd.so")
CSRC_DIR = os.path.join(PKG_DIR, "csrc")

HP_OK, HP_EINVAL, HP_ELAUNCH, HP_ESPACE = 0, -1, -2, -3
HP_F32, HP_F64 = 0, 1
HP_RAIN_NONE, HP_RAIN_STATION, HP_RAIN_FIELD = 0, 1, 2
HP_UPDATE_STEP, HP_UPDATE_FRICTION, HP_UPDATE_EULER, HP_UPDATE_CFL = 0, 1, 2, 3
ABI_VERSION = 10

_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_I = ctypes.c_int


class HpCtrl(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("export_T", ctypes.c_double),
                ("export_n", ctypes.c_double), ("export_pending", ctypes.c_double), ("volume", ctypes.c_double),
                ("rain_volume", ctypes.c_double), ("flux_volume", ctypes.c_double), ("steps", ctypes.c_double),
                ("t_stop", ctypes.c_double), ("stopped_synced", ctypes.c_double),
                ("reserved", ctypes.c_double * 5)]


CTRL_FIELDS = [f for f, _ in HpCtrl._fields_[:10]]
CTRL_T, CTRL_DT, CTRL_EXPORT_T, CTRL_EXPORT_N, CTRL_PENDING, CTRL_VOLUME = 0, 1, 2, 3, 4, 5
CTRL_RAIN, CTRL_FLUX, CTRL_STEPS, CTRL_T_STOP, CTRL_STOPPED_SYNCED = 6, 7, 8, 9, 10
CTRL_DOUBLES = 16


class HpGridDesc(ctypes.Structure):
    _fields_ = [
        ("dtype", _I32), ("order", _I32), ("nx", _I32), ("ny", _I32), ("pitch", _I64),
        ("dx", ctypes.c_double), ("cfl", ctypes.c_double),
        ("cell0_x", _I32), ("cell0_y", _I32), ("n_classes", _I32), ("rain_mode", _I32),
        ("tile_rows", _I32), ("mass_sums", _I32),
        ("h", _P), ("qx", _P), ("qy", _P), ("z", _P), ("h_max", _P),
        ("dh", _P), ("dqx", _P), ("dqy", _P),
        ("h_out", _P), ("qx_out", _P), ("qy_out", _P),
        ("type_slot", _P), ("landuse", _P), ("slot_code", _P), ("manning", _P),
        ("given_depth", _P), ("given_wl", _P), ("given_q", _P),
        ("gd_rows", _I32), ("gd_w", _I32), ("gwl_rows", _I32), ("gwl_w", _I32), ("gq_rows", _I32), ("gq_w", _I32),
        ("station", _P), ("rain_table", _P), ("rain_rows", _I32), ("rain_cols", _I32),
        ("rain_field", _P), ("rain_scale", ctypes.c_double),
        ("ctrl", _P), ("work", _P), ("work_bytes", _I64),
        ("infil_ks", _P), ("infil_phi", _P), ("infil_dtheta", _P), ("infil_F", _P), ("infil_dF", _P),
    ]


class HpEnsemble(ctypes.Structure):
    _fields_ = [("n_domains", _I32), ("dtype", _I32), ("order", _I32), ("tile_rows", _I32),
                ("n_flux_tiles", _I32), ("n_update_blocks", _I32),
                ("doms", _P), ("flux_map", _P), ("update_map", _P), ("finalize_map", _P), ("partials", _P),
                ("single_pass", _I32), ("fused_finalize_map", _P)]


class HpAdjointDesc(ctypes.Structure):
    _fields_ = [("lam_h", _P), ("lam_qx", _P), ("lam_qy", _P), ("cot_h", _P), ("cot_qx", _P), ("cot_qy", _P),
                ("xc", _P), ("yc", _P), ("cpitch", _I64), ("grads", _P), ("gn", _P), ("slope_mode", _I32),
                ("flux_ready", _I32), ("gauge_of", _P), ("loss_h", _P), ("loss_obs", _P),
                ("separate_passes", _I32)]


GRAD_STRIDE = 16
_lib = None

# name -> (restype, argtypes); every symbol include/hipims_abi.h declares
SIGNATURES = {
    "hp_abi_version": (_I, []),
    "hp_last_error": (ctypes.c_char_p, []),
    "hp_struct_sizes": (None, [ctypes.POINTER(ctypes.c_int64)]),
    "hp_flux_mask": (_I, [_I, _I64, _P, _P, _P, _P, _P]),
    "hp_flux_compact": (_I, [_I, _I, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                             _P, _I32, _I32, _P, _I32, _I32, _P, _I32, _I32, _P, _P, _P, _P]),
    "hp_station_rain": (_I, [_I, _I64, _P, _P, _P, _I, _P, _P]),
    "hp_friction_update_compact": (_I, [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "hp_time_control_compact": (_I, [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "hp_euler_update_compact": (_I, [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "hp_infiltration_compact": (_I, [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "hp_grid_workspace_bytes": (_I64, [ctypes.POINTER(HpGridDesc)]),
    "hp_grid_infiltration": (_I, [ctypes.POINTER(HpGridDesc), _P]),
    "hp_grid_flux": (_I, [ctypes.POINTER(HpGridDesc), _P]),
    "hp_grid_update": (_I, [ctypes.POINTER(HpGridDesc), _P]),
    "hp_grid_update_mode": (_I, [ctypes.POINTER(HpGridDesc), _I, _P]),
    "hp_grid_step": (_I, [ctypes.POINTER(HpGridDesc), _I32, _P]),
    "hp_grid_step_fused": (_I, [ctypes.POINTER(HpGridDesc), _I32, _P]),
    "hp_grid_step_fused_noalloc": (_I, [ctypes.POINTER(HpGridDesc), _I32, _P]),
    "hp_grid_noalloc_max_bytes": (_I64, []),
    "hp_grid_step_fused_timed": (_I, [ctypes.POINTER(HpGridDesc), _I32, _I32, ctypes.POINTER(ctypes.c_float),
                                      ctypes.POINTER(ctypes.c_float), _P]),
    "hp_ensemble_workspace_bytes": (_I64, [_I32, ctypes.POINTER(HpGridDesc)]),
    "hp_ensemble_init": (_I, [_I32, ctypes.POINTER(HpGridDesc), _I32, _P, _I64, ctypes.POINTER(HpEnsemble), _P]),
    "hp_ensemble_step": (_I, [ctypes.POINTER(HpEnsemble), _I32, _P]),
    "hp_adjoint_grad_blocks": (_I64, [ctypes.POINTER(HpGridDesc)]),
    "hp_adjoint_step": (_I, [ctypes.POINTER(HpGridDesc), ctypes.POINTER(HpAdjointDesc), _P]),
    "hp_adjoint_finish": (_I, [ctypes.POINTER(HpGridDesc), ctypes.POINTER(HpAdjointDesc), _P]),
    "hp_adjoint_add_loss": (_I, [_I, _P, _P, _P, _P, _I32, ctypes.c_double, _P]),
    "hp_calib_forward": (_I, [ctypes.POINTER(HpGridDesc), _P, _P, _I32, _I32, _P, _P, _I64, _P, _I32, _P, _P]),
    "hp_calib_reverse": (_I, [ctypes.POINTER(HpGridDesc), ctypes.POINTER(HpAdjointDesc), _P, _I32, _I32, _I32, _P,
                              _P, _I64, _P, _I32, _P, _P, _P]),
    "hp_probe_fma_per_lane": (_I64, []),
    "hp_probe_fma": (_I, [_I32, _P, _P]),
    "hp_gather": (_I, [_I, _P, _P, _I64, _P, _P]),
    "hp_scatter": (_I, [_I, _P, _P, _I64, _P, _P]),
}


def build(force: bool = True) -> str:
    """Compile libhipims_amd.so for gfx950 with the committed Makefile."""
    args = ["make", "-s", "-C", CSRC_DIR]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    return LIB_PATH


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.hp_abi_version() != ABI_VERSION:
            raise RuntimeError("libhipims_amd.so ABI version mismatch")
        sizes = (ctypes.c_int64 * 3)()
        L.hp_struct_sizes(sizes)
        mine = (ctypes.sizeof(HpCtrl), ctypes.sizeof(HpGridDesc), ctypes.sizeof(HpEnsemble))
        if tuple(sizes) != mine:
            raise RuntimeError(f"struct layout mismatch: library {tuple(sizes)} vs bindings {mine}")
        _lib = L
    return _lib


class HipimsError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    if rc != HP_OK:
        msg = lib().hp_last_error()
        raise HipimsError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def dtype_code(dt: torch.dtype) -> int:
    if dt == torch.float32:
        return HP_F32
    if dt == torch.float64:
        return HP_F64
    raise TypeError(f"unsupported dtype {dt}; expected torch.float32 or torch.float64")


def stream_ptr(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t) -> int | None:
    return None if t is None else t.data_ptr()
