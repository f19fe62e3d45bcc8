"""The Fortran drop-in, run as the reference's driver runs it.

oracle/_ref/libps3d_dropin.so holds the reference's own main_module,
timing_module and parallel_mpi (compiled unchanged), the drop-in
integrate / pressure_solver of ps3d_amd/fortran/integrate_gpu.f90 with its
ISO_C_BINDING interfaces, and a driver (oracle/dropin_glue.f90) that sets up
main_module and steps it like the main loop of main.f90:138-162.  The state
goes in through main_module's arrays and comes back through them, so this
checks the whole boundary: module layout, parameters, ksqr, tau rotation,
the device model behind the Fortran calls and the host refresh.  FP64 must
equal the CPU oracle bit for bit; the mixed form must equal the mixed model
driven from Python bit for bit.
"""
import ctypes
import dataclasses
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DROPIN = os.path.join(ROOT, "oracle", "_ref", "libps3d_dropin.so")
NAMES = ("u", "v", "w", "b", "p")
REAL = ("Lx", "Ly", "Lz", "dx", "dy", "dz", "dt", "f0", "N0", "Ro", "dsqr", "eps_ab",
        "Ah", "Av", "Kh", "Kv", "Ahbi", "Avbi", "Khbi", "Kvbi")


def _load():
    if not os.path.exists(DROPIN):
        pytest.skip("the drop-in driver is built with the reference sources present (oracle/Makefile: dropin)")
    lib = ctypes.CDLL(DROPIN)
    lib.dropin_setup.argtypes = [ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]
    lib.dropin_ptr.argtypes = [ctypes.c_int32]
    lib.dropin_ptr.restype = ctypes.c_void_p
    lib.dropin_run.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    return lib


def _view(lib, fid, shape):
    n = int(np.prod(shape))
    buf = (ctypes.c_double * n).from_address(lib.dropin_ptr(fid))
    return np.ndarray(shape, dtype=np.float64, buffer=buf, order="F")


def _fortran_run(lib, prm, fields, ksqr, steps, mixed, force=None, autotune=False):
    iprm = (ctypes.c_int32 * 7)(prm.nx, prm.ny, prm.nz, int(prm.enable_nonlinear),
                                int(prm.enable_vertical_boundaries), int(prm.enable_AB_3_order), int(mixed))
    rprm = (ctypes.c_double * 20)(*[float(getattr(prm, n)) for n in REAL])
    lib.dropin_setup(iprm, rprm)
    shape = (prm.nx + 4, prm.ny + 4, prm.nz + 4)
    for fid, n in enumerate(NAMES[:4]):
        _view(lib, fid, shape)[...] = fields[n]
    _view(lib, 5, ksqr.shape)[...] = ksqr
    if autotune:
        lib.dropin_set_autotune()
    if force is not None:
        lib.dropin_set_forcing()
        _view(lib, 6, shape)[...] = force
    st = (ctypes.c_int32 * 4)()
    lib.dropin_run(steps, st)
    out = {n: _view(lib, fid, shape).copy() for fid, n in enumerate(NAMES)}
    lib.dropin_finish()
    return out, tuple(st)


CASES = {
    # fused friction + AB, rigid lid
    "lid_fused": dict(nx=32, ny=24, nz=16, lid=True, Avbi=0.0, Kvbi=0.0),
    # vertical biharmonic friction: the unfused friction chain, periodic z
    "periodic_unfused": dict(nx=24, ny=20, nz=10, lid=False),
}


@pytest.mark.parametrize("autotune", [False, True])
@pytest.mark.parametrize("case", list(CASES))
def test_fortran_dropin_matches_the_oracle(case, autotune):
    import oracle
    from ps3d_amd import configs
    lib = _load()
    c = dict(CASES[case])
    nx, ny, nz, lid = c.pop("nx"), c.pop("ny"), c.pop("nz"), c.pop("lid")
    prm, fields = configs.random_state(nx, ny, nz, seed=31, lid=lid, **c)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        orc.f[n][...] = a
    steps = 12
    print(f"fortran drop-in case={case} steps={steps} autotune={autotune}", flush=True)
    got, st = _fortran_run(lib, prm, fields, orc.f["ksqr"].copy(), steps, mixed=False, autotune=autotune)
    orc.run(steps)
    assert st == (orc.itt, orc.tau, orc.taum1, orc.taum2)
    for n in NAMES:
        assert np.array_equal(got[n], orc.f[n]), (n, float(np.abs(got[n] - orc.f[n]).max()))
    orc.close()


@pytest.mark.parametrize("autotune", [False, True])
def test_fortran_dropin_mixed_matches_the_python_mixed_model(autotune):
    """The Fortran drop-in in the mixed form, bit for bit against the Python
    mixed model.  With the autotuner (run by the drop-in before the mixed
    cold start) the launch shapes do not change results, but the radix
    plans of the FP32 transforms do: the drop-in must then equal the Python
    model with one of the plan choices open to the tuner (x, y: lengths 32
    and 64 both have a factor 8)."""
    import itertools
    import oracle
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    lib = _load()
    prm, fields = configs.eddy(64, 64, 32, seed=33)
    orc = oracle.OracleModel(prm)   # ksqr as the reference's driver would set it
    ksqr = orc.f["ksqr"].copy()
    orc.close()
    steps = 8
    print(f"fortran drop-in mixed steps={steps} autotune={autotune}", flush=True)
    got, _ = _fortran_run(lib, prm, fields, ksqr, steps, mixed=True, autotune=autotune)
    plans = [(px, py, 0) for px, py in itertools.product((0, 1), repeat=2)] if autotune else [(0, 0, 0)]
    matched = []
    for pl in plans:
        m = Model(dataclasses.replace(prm, precision=PS3D_MIXED))
        m.fft_plan = pl
        for n, a in fields.items():
            m.f[n][...] = a
        m.push()
        m.run(steps)
        m.pull(list(NAMES))
        if all(np.array_equal(got[n], m.f[n]) for n in NAMES):
            matched.append(pl)
        elif not autotune:
            for n in NAMES:
                assert np.array_equal(got[n], m.f[n]), (n, float(np.abs(got[n] - m.f[n]).max()))
        m.close()
    print(f"  equal to the Python mixed model with plans {matched}", flush=True)
    assert matched


def test_fortran_dropin_host_forcing_path():
    """A config whose set_forcing works on the host mirrors (here db += a
    fixed array, whole arrays): the drop-in runs tendency and friction on
    the device, brings fields and tendencies to the host, calls set_forcing
    there (integrate.f90:82) and finishes the step on the device.  The same
    forcing as a host round-trip hook of the Python model must give the same
    bits."""
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    import oracle
    lib = _load()
    prm, fields = configs.random_state(32, 24, 16, seed=35, lid=True, Avbi=0.0, Kvbi=0.0)
    rng = np.random.default_rng(35)
    force = 1e-7 * rng.standard_normal((prm.nx + 4, prm.ny + 4, prm.nz + 4))
    orc = oracle.OracleModel(prm)
    ksqr = orc.f["ksqr"].copy()
    orc.close()
    steps = 6
    print(f"fortran drop-in host forcing steps={steps}", flush=True)
    got, _ = _fortran_run(lib, prm, fields, ksqr, steps, mixed=False, force=force)
    m = Model(prm)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()

    def hook(mm, itt, tau, stream):
        mm.synchronize()
        mm.pull(["db"])
        mm.f["db"][..., tau - 1] = mm.f["db"][..., tau - 1] + force
        mm.push(["db"])

    m.set_forcing(hook)
    m.run(steps)
    m.set_forcing(None)
    m.pull(list(NAMES))
    for n in NAMES:
        assert np.array_equal(got[n], m.f[n]), (n, float(np.abs(got[n] - m.f[n]).max()))
    # and the forcing did act
    ref = _fortran_run(lib, prm, fields, ksqr, steps, mixed=False)[0]
    assert not np.array_equal(ref["b"], got["b"])
    m.close()


@pytest.mark.parametrize("lid", [True, False])
def test_fortran_dropin_pressure_solver(lid):
    """pressure_solver called on its own through the drop-in (host arrays
    pushed, projected on the device, pulled back) equals the oracle's
    pressure_solver.f90:2-120 bit for bit."""
    import oracle
    from ps3d_amd import configs
    lib = _load()
    prm, fields = configs.random_state(32, 16, 16, seed=37, lid=lid, Avbi=0.0, Kvbi=0.0)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        orc.f[n][...] = a
    iprm = (ctypes.c_int32 * 7)(prm.nx, prm.ny, prm.nz, int(prm.enable_nonlinear),
                                int(prm.enable_vertical_boundaries), int(prm.enable_AB_3_order), 0)
    rprm = (ctypes.c_double * 20)(*[float(getattr(prm, n)) for n in REAL])
    lib.dropin_setup(iprm, rprm)
    shape = (prm.nx + 4, prm.ny + 4, prm.nz + 4)
    for fid, n in enumerate(NAMES[:4]):
        _view(lib, fid, shape)[...] = fields[n]
    _view(lib, 5, orc.f["ksqr"].shape)[...] = orc.f["ksqr"]
    print(f"fortran drop-in pressure_solver lid={lid}", flush=True)
    lib.dropin_pressure()
    orc.pressure_solver()
    for fid, n in enumerate(("u", "v", "w", "b", "p")):
        got = _view(lib, fid, shape)
        assert np.array_equal(got, orc.f[n]), (n, float(np.abs(got - orc.f[n]).max()))
    lib.dropin_finish()
    orc.close()


JETS = os.path.join(ROOT, "oracle", "_ref", "libps3d_jets.so")


def test_reference_config_linked_unchanged():
    """The reference's own for_config/jets_unscaled.f90 (set_parameter,
    set_initial_conditions, set_forcing), compiled unchanged and linked with
    the drop-in, run in main.f90's order (oracle/config_glue.f90): its
    parameters and initial state come from its own code, the steps run on
    the GPU through the drop-in integrate, and the result equals the oracle
    stepping the same state, bit for bit.  (100 x 100 x 10, rigid lid: FFT
    lengths 50 and 100 take the generic transform kernels.)"""
    import oracle
    from ps3d_amd.params import ModelParams
    if not os.path.exists(JETS):
        pytest.skip("the config driver is built with the reference sources present (oracle/Makefile: config)")
    lib = ctypes.CDLL(JETS)
    lib.config_ptr.argtypes = [ctypes.c_int32]
    lib.config_ptr.restype = ctypes.c_void_p
    lib.config_run.argtypes = [ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    lib.config_setup()
    iprm = (ctypes.c_int32 * 6)()
    rprm = (ctypes.c_double * 20)()
    lib.config_params(iprm, rprm)
    nx, ny, nz, nonlin, lid, ab3 = list(iprm)
    prm = ModelParams(nx=nx, ny=ny, nz=nz, enable_nonlinear=bool(nonlin), enable_vertical_boundaries=bool(lid),
                      enable_AB_3_order=bool(ab3), **{n: rprm[q] for q, n in enumerate(REAL)})
    orc = oracle.OracleModel(prm)

    def view(fid, shape):
        n = int(np.prod(shape))
        buf = (ctypes.c_double * n).from_address(lib.config_ptr(fid))
        return np.ndarray(shape, dtype=np.float64, buffer=buf, order="F")

    view(5, orc.f["ksqr"].shape)[...] = orc.f["ksqr"]
    lib.config_initial_conditions()
    shape = (nx + 4, ny + 4, nz + 4)
    for fid, n in enumerate(NAMES[:4]):
        orc.f[n][...] = view(fid, shape)
    assert np.abs(orc.f["u"]).max() > 0.05 and np.abs(orc.f["b"]).max() > 0.0
    steps = 20
    print(f"jets_unscaled.f90 through the drop-in: {nx}x{ny}x{nz}, lid={bool(lid)}, {steps} steps", flush=True)
    st = (ctypes.c_int32 * 4)()
    lib.config_run(steps, st)
    orc.run(steps)
    assert tuple(st) == (orc.itt, orc.tau, orc.taum1, orc.taum2)
    for fid, n in enumerate(NAMES):
        got = view(fid, shape)
        assert np.array_equal(got, orc.f[n]), (n, float(np.abs(got - orc.f[n]).max()))
    lib.config_finish()
    orc.close()
