"""ps3d_amd -- the ps3D model time step on AMD Instinct MI355X (gfx950).

The product is the C-ABI library ``libps3d.so`` (include/ps3d.h) built from
the HIP sources in ``csrc/``; this package is its host-side mirror:

* ``params.ModelParams``  -- main_module parameters (set_parameter surface)
* ``model.Model``         -- one model on the GPU, reference-layout host arrays
* ``configs``             -- synthetic scenarios (eddy, jets) with seeded noise
* ``build``               -- compiles the library (hipcc, gfx950)
* ``fortran/``            -- ISO_C_BINDING interfaces and the drop-in
                             integrate / pressure_solver for the Fortran driver
"""
from .params import PS3D_FP64, PS3D_MIXED, ModelParams  # noqa: F401

__version__ = "0.1"


def __getattr__(name):
    if name in ("Model", "kernel_names"):
        from . import model
        return getattr(model, name)
    raise AttributeError(name)
