"""Model parameters: the part of ``main_module`` (for_src/main_module.f90:16-59)
that a ``set_parameter`` hook sets and the time step reads.

Defaults are main_module's initialisers.  ``eps_ab`` has no initialiser in the
reference (main_module.f90:56) and is therefore zero unless a config sets it.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, fields, replace

PS3D_FP64 = 0
PS3D_MIXED = 1


@dataclass
class ModelParams:
    nx: int
    ny: int
    nz: int
    Lx: float
    Ly: float
    Lz: float
    dt: float
    f0: float = 0.0
    N0: float = 0.0
    dx: float = 0.0     # 0 -> Lx/nx (what every config computes, e.g. eddy1.f90:28)
    dy: float = 0.0
    dz: float = 0.0
    Ro: float = 1.0
    dsqr: float = 1.0
    eps_ab: float = 0.0
    Ah: float = 0.0
    Av: float = 0.0
    Kh: float = 0.0
    Kv: float = 0.0
    Ahbi: float = 0.0
    Avbi: float = 0.0
    Khbi: float = 0.0
    Kvbi: float = 0.0
    enable_nonlinear: bool = True
    enable_vertical_boundaries: bool = False
    enable_AB_3_order: bool = False
    precision: int = PS3D_FP64

    def __post_init__(self):
        if not self.dx:
            self.dx = self.Lx / self.nx
        if not self.dy:
            self.dy = self.Ly / self.ny
        if not self.dz:
            self.dz = self.Lz / self.nz

    def replace(self, **kw) -> "ModelParams":
        return replace(self, **kw)

    @property
    def shape(self):
        """host array shape of a 3-D field including the onx=2 halo"""
        return (self.nx + 4, self.ny + 4, self.nz + 4)

    @property
    def npoints(self) -> int:
        return self.nx * self.ny * self.nz


class CParams(ctypes.Structure):
    """``ps3d_params`` of include/ps3d.h"""
    _fields_ = ([("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                 ("enable_nonlinear", ctypes.c_int32),
                 ("enable_vertical_boundaries", ctypes.c_int32),
                 ("enable_AB_3_order", ctypes.c_int32)]
                + [(n, ctypes.c_double) for n in
                   ("Lx", "Ly", "Lz", "dx", "dy", "dz", "dt", "f0", "N0", "Ro", "dsqr", "eps_ab",
                    "Ah", "Av", "Kh", "Kv", "Ahbi", "Avbi", "Khbi", "Kvbi")]
                + [("precision", ctypes.c_int32), ("reserved", ctypes.c_int32 * 7)])

    @classmethod
    def from_params(cls, p: ModelParams) -> "CParams":
        c = cls()
        for name, ctype in cls._fields_:
            if name == "reserved":
                continue
            val = getattr(p, name)
            setattr(c, name, float(val) if ctype is ctypes.c_double else int(val))
        return c
