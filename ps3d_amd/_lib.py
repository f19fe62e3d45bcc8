"""ctypes binding of the C-ABI in include/ps3d.h (libps3d.so, built for gfx950).

There is no fallback: if the HIP library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

from .params import CParams

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libps3d.so")

# include/ps3d.h
PS3D_OK = 0
PS3D_ERR_ARG, PS3D_ERR_HIP, PS3D_ERR_STATE, PS3D_ERR_INTERNAL = 1, 2, 3, 4
PS3D_U, PS3D_V, PS3D_W, PS3D_B, PS3D_P = 0, 1, 2, 3, 4
PS3D_DU, PS3D_DV, PS3D_DW, PS3D_DB = 5, 6, 7, 8
PS3D_FS, PS3D_KSQR = 9, 10
PS3D_HARMONIC, PS3D_BIHARMONIC_H, PS3D_BIHARMONIC_V = 0, 1, 2

_c_int, _c_i64, _c_dbl, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)

# name -> (restype, argtypes); every function declared in include/ps3d.h
class OptBalanceParams(ctypes.Structure):
    """ps3d_opt_balance_params"""
    _fields_ = [("period", ctypes.c_double), ("average", ctypes.c_double), ("tol", ctypes.c_double),
                ("max_iterations", ctypes.c_int32), ("average_times", ctypes.c_int32),
                ("time_average", ctypes.c_int32), ("reserved", ctypes.c_int32)]


# int (*)(ps3d_model *m, void *user, int itt, int tau, void *stream)
FORCING_FN = ctypes.CFUNCTYPE(_c_int, _vp, _vp, _c_int, _c_int, _vp)

PROTOTYPES = {
    "ps3d_last_error": (_c_int, [ctypes.c_char_p, _c_int]),
    "ps3d_version": (ctypes.c_char_p, []),
    "ps3d_kernel_names": (ctypes.c_char_p, []),
    "ps3d_create": (_c_int, [ctypes.POINTER(CParams), _dp, _vp, ctypes.POINTER(_vp)]),
    "ps3d_destroy": (_c_int, [_vp]),
    "ps3d_set_params": (_c_int, [_vp, ctypes.POINTER(CParams)]),
    "ps3d_set_field": (_c_int, [_vp, _c_int, _dp]),
    "ps3d_get_field": (_c_int, [_vp, _c_int, _dp]),
    "ps3d_field_size": (_c_int, [_vp, _c_int, ctypes.POINTER(_c_i64)]),
    "ps3d_integrate": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int]),
    "ps3d_run": (_c_int, [_vp, _c_int, _i32p, _c_int]),
    "ps3d_synchronize": (_c_int, [_vp]),
    "ps3d_set_fused_tendency": (_c_int, [_vp, _c_int]),
    "ps3d_set_fft_plan": (_c_int, [_vp, _i32p]),
    "ps3d_get_fft_plan": (_c_int, [_vp, _i32p]),
    "ps3d_tendency": (_c_int, [_vp, _c_int]),
    "ps3d_buoyancy_advection": (_c_int, [_vp, _c_int]),
    "ps3d_momentum_advection": (_c_int, [_vp, _c_int]),
    "ps3d_friction": (_c_int, [_vp]),
    "ps3d_mixing": (_c_int, [_vp, _c_int, _c_int, _c_dbl, _c_dbl]),
    "ps3d_ab_update": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int]),
    "ps3d_border_exchg": (_c_int, [_vp, _c_int]),
    "ps3d_pressure_solver": (_c_int, [_vp]),
    "ps3d_my_fft": (_c_int, [_vp, _c_int, _dp, _dp]),
    "ps3d_my_ifft": (_c_int, [_vp, _c_int, _dp, _dp]),
    "ps3d_spectral_solve": (_c_int, [_vp, _dp]),
    "ps3d_profile_kernel": (_c_int, [_vp, ctypes.c_char_p]),
    "ps3d_profile_read": (_c_int, [_vp, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_dbl)]),
    "ps3d_tune": (_c_int, [_vp, _c_int, ctypes.c_char_p, _c_int]),
    "ps3d_fft_create": (_c_int, [_c_int, _c_int, _c_int, _vp, ctypes.POINTER(_vp)]),
    "ps3d_fft_destroy": (_c_int, [_vp]),
    "ps3d_fft_forward": (_c_int, [_vp, ctypes.c_char_p, _dp, _dp]),
    "ps3d_fft_backward": (_c_int, [_vp, ctypes.c_char_p, _dp, _dp]),
    "ps3d_tuning": (_c_int, [_vp, ctypes.c_char_p, _c_int]),
    "ps3d_diagnose": (_c_int, [_vp, _dp, _i32p]),
    "ps3d_hbm_probe": (_c_int, [_c_i64, _c_int, _c_int, _dp]),
    "ps3d_fma_probe": (_c_int, [_c_int, _dp]),
    "ps3d_diagnose_noalloc": (_c_int, [ctypes.POINTER(CParams), _vp, _vp, _vp, _vp, _vp, _dp, _i32p]),
    "ps3d_set_forcing": (_c_int, [_vp, _vp, _vp]),
    "ps3d_device_ptr": (_c_int, [_vp, _c_int, _c_int, ctypes.POINTER(_vp)]),
    "ps3d_set_wavemaker": (_c_int, [_vp, _c_dbl, ctypes.POINTER(_dp), ctypes.POINTER(_dp)]),
    "ps3d_clear_wavemaker": (_c_int, [_vp]),
    "ps3d_set_background_flow": (_c_int, [_vp, _dp, _dp]),
    "ps3d_set_body_force": (_c_int, [_vp, ctypes.POINTER(_dp)]),
    "ps3d_write_restart": (_c_int, [_vp, ctypes.c_char_p, _c_int, _c_dbl, _c_int, _c_int, _c_int]),
    "ps3d_read_restart": (_c_int, [_vp, ctypes.c_char_p, _i32p, _dp, _i32p, _i32p, _i32p]),
    "ps3d_restart_write_host": (_c_int, [ctypes.c_char_p, _c_int, _c_int, _c_int, _c_int, _c_dbl, _c_int, _c_int,
                                         _c_int, _dp]),
    "ps3d_restart_read_host": (_c_int, [ctypes.c_char_p, _c_int, _c_int, _c_int, _i32p, _dp, _i32p, _i32p, _i32p,
                                        _dp]),
    "ps3d_diag_balance": (_c_int, [_vp, _c_int, _dp, _dp, _dp, ctypes.POINTER(_dp)]),
    "ps3d_opt_balance": (_c_int, [_vp, ctypes.POINTER(OptBalanceParams), _dp, _dp, _dp, ctypes.POINTER(_dp), _dp,
                                  _i32p]),
    "ps3d_ensemble_create": (_c_int, [_c_int, ctypes.POINTER(CParams), _vp, ctypes.POINTER(_vp)]),
    "ps3d_ensemble_destroy": (_c_int, [_vp]),
    "ps3d_ensemble_set_field": (_c_int, [_vp, _c_int, _c_int, _dp]),
    "ps3d_ensemble_get_field": (_c_int, [_vp, _c_int, _c_int, _dp]),
    "ps3d_ensemble_set_background_flow": (_c_int, [_vp, _c_int, _dp, _dp]),
    "ps3d_ensemble_set_body_force": (_c_int, [_vp, _c_int, ctypes.POINTER(_dp)]),
    "ps3d_ensemble_set_fft_plan": (_c_int, [_vp, _c_int, _i32p]),
    "ps3d_ensemble_get_fft_plan": (_c_int, [_vp, _c_int, _i32p]),
    "ps3d_ensemble_run": (_c_int, [_vp, _c_int, _i32p, _c_int]),
    "ps3d_ensemble_synchronize": (_c_int, [_vp]),
    "ps3d_ensemble_tune": (_c_int, [_vp, ctypes.c_char_p, _c_int]),
    "ps3d_ensemble_describe": (_c_int, [_vp, ctypes.c_char_p, _c_int]),
}

_lib = None


class Ps3dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ps3d error {code}: {msg}")
        self.code = code


def lib_path() -> str:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load libps3d.so; raises if it is missing (no CPU fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: the HIP library has not been built "
                "(run python -c 'import __graft_entry__ as g; g.build()')")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(2048)
    load().ps3d_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(code: int) -> None:
    if code != PS3D_OK:
        raise Ps3dError(code, last_error())
