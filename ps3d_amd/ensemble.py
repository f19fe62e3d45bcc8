"""Batched ensemble of ps3D models on one GPU (ps3d_ensemble_* of the C-ABI).

Members may differ in size and parameters.  Each stage of the step is one
launch over all members and a step can be replayed as one captured
hipGraph; each member's fields are those a single Model with the same
parameters computes, bit for bit.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check
from .params import CParams, ModelParams

FIELDS = {"u": _lib.PS3D_U, "v": _lib.PS3D_V, "w": _lib.PS3D_W, "b": _lib.PS3D_B, "p": _lib.PS3D_P,
          "du": _lib.PS3D_DU, "dv": _lib.PS3D_DV, "dw": _lib.PS3D_DW, "db": _lib.PS3D_DB}


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.f_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Ensemble:
    def __init__(self, prms: list[ModelParams], stream: int | None = None):
        self.lib = _lib.load()
        self.prms = list(prms)
        arr = (CParams * len(prms))(*[CParams.from_params(p) for p in prms])
        self._carr = arr
        h = ctypes.c_void_p()
        check(self.lib.ps3d_ensemble_create(len(prms), arr, ctypes.c_void_p(stream) if stream else None,
                                            ctypes.byref(h)))
        self.h = h
        self.itt, self.tau, self.taum1, self.taum2 = 0, 3, 2, 1

    def __len__(self):
        return len(self.prms)

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.ps3d_ensemble_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, member: int, name: str, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.float64)
        check(self.lib.ps3d_ensemble_set_field(self.h, member, FIELDS[name], _dp(a)))

    def get(self, member: int, name: str) -> np.ndarray:
        shp = self.prms[member].shape
        if name in ("du", "dv", "dw", "db"):
            shp = shp + (3,)
        a = np.zeros(shp, order="F")
        check(self.lib.ps3d_ensemble_get_field(self.h, member, FIELDS[name], _dp(a)))
        return a

    def set_background_flow(self, member: int, u_back=None, b_back=None):
        """Advection by the background flow (u_back, 0, 0) and of b_back for
        one member (ps3d_ensemble_set_background_flow); u_back None: off."""
        ub = None if u_back is None else np.asfortranarray(u_back, dtype=np.float64)
        bb = None if b_back is None else np.asfortranarray(b_back, dtype=np.float64)
        self._keep = (ub, bb)
        check(self.lib.ps3d_ensemble_set_background_flow(self.h, member, None if ub is None else _dp(ub),
                                                         None if bb is None else _dp(bb)))

    def set_body_force(self, member: int, force: dict | None):
        """Steady body force of one member ({'du': .., 'dw': ..};
        ps3d_ensemble_set_body_force); None: off."""
        force = force or {}
        arrs = [None if force.get(n) is None else np.asfortranarray(force[n], dtype=np.float64)
                for n in ("du", "dv", "dw", "db")]
        tab = (ctypes.POINTER(ctypes.c_double) * 4)(*[None if a is None else _dp(a) for a in arrs])
        check(self.lib.ps3d_ensemble_set_body_force(self.h, member, tab))

    def fft_plan(self, member: int) -> tuple:
        """Radix plans (x, y, z) of one member's FP32 transforms
        (ps3d_ensemble_get_fft_plan)."""
        a = (ctypes.c_int32 * 3)()
        check(self.lib.ps3d_ensemble_get_fft_plan(self.h, member, a))
        return tuple(a)

    def set_fft_plan(self, member: int, plans):
        check(self.lib.ps3d_ensemble_set_fft_plan(self.h, member, (ctypes.c_int32 * 3)(*plans)))

    def run(self, nsteps: int, use_graph: bool = True):
        st = (ctypes.c_int32 * 4)(self.itt, self.tau, self.taum1, self.taum2)
        check(self.lib.ps3d_ensemble_run(self.h, nsteps, st, int(use_graph)))
        self.itt, self.tau, self.taum1, self.taum2 = list(st)

    def synchronize(self):
        check(self.lib.ps3d_ensemble_synchronize(self.h))

    def tune(self, cache_path: str | None = None, reps: int = 5):
        check(self.lib.ps3d_ensemble_tune(self.h, cache_path.encode() if cache_path else None, reps))

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(8192)
        n = self.lib.ps3d_ensemble_describe(self.h, buf, len(buf))
        if n < 0:
            check(-n)
        return buf.value.decode()
