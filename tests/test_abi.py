"""The C-ABI library loads and exports what include/ps3d.h declares; argument
errors come back as codes (no GPU needed: validation precedes device work)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ps3d.h")


def header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ps3d_\w+)\s*\(", txt)))


def test_header_declares_the_api():
    names = header_functions()
    for must in ("ps3d_create", "ps3d_integrate", "ps3d_run", "ps3d_pressure_solver", "ps3d_my_fft",
                 "ps3d_set_field", "ps3d_get_field", "ps3d_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from ps3d_amd import _lib
    lib = _lib.load()
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.lib_path()], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ps3d_\w+)", out))
    assert set(header_functions()) <= exported


def test_python_binding_covers_the_header():
    from ps3d_amd import _lib
    assert set(header_functions()) == set(_lib.PROTOTYPES)


def test_version_and_kernel_names():
    from ps3d_amd import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.ps3d_version()
    names = lib.ps3d_kernel_names().decode().split(",")
    assert "tendency" in names and "thomas" in names


@pytest.mark.parametrize("bad,msg", [
    (dict(nx=33), "nx must be even"),
    (dict(ny=14), "products of 2, 3 and 5"),
    (dict(nz=1), "nz must be"),
    (dict(precision=7), "precision"),
    (dict(dt=0.0), "dt"),
    # the mixed form: fused friction only, compiled-in FFT lengths only
    (dict(precision=1, Avbi=1e-2), "vertical biharmonic"),
    (dict(precision=1, nx=36), "compiled-in FFT lengths"),
    (dict(precision=1, nz=18, enable_vertical_boundaries=False), "compiled-in FFT lengths"),
])
def test_create_rejects_bad_parameters(bad, msg):
    from ps3d_amd import _lib
    from ps3d_amd.params import CParams, ModelParams
    lib = _lib.load()
    base = dict(nx=32, ny=32, nz=16, Lx=1.0, Ly=1.0, Lz=1.0, dt=0.1, enable_vertical_boundaries=True)
    base.update(bad)
    prm = ModelParams(**base)
    h = ctypes.c_void_p()
    code = lib.ps3d_create(ctypes.byref(CParams.from_params(prm)), None, None, ctypes.byref(h))
    assert code == 1                      # PS3D_ERR_ARG
    assert not h.value
    assert msg in _lib.last_error()


def test_null_model_is_an_error():
    from ps3d_amd import _lib
    lib = _lib.load()
    assert lib.ps3d_integrate(None, 0, 3, 2, 1) == 1
    assert "null model" in _lib.last_error()


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_cparams_layout_matches_header(tmp_path):
    """the ctypes mirror of ps3d_params has the C compiler's layout"""
    from ps3d_amd.params import CParams
    prog = tmp_path / "lay.c"
    fields = [n for n, _ in CParams._fields_]
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ps3d.h"\nint main(void){\n'
                    'printf("%zu\\n", sizeof(ps3d_params));\n'
                    + "".join(f'printf("%zu\\n", offsetof(ps3d_params, {n}));\n' for n in fields)
                    + "return 0;}\n")
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    vals = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(CParams)
    assert vals[1:] == [getattr(CParams, n).offset for n in fields]


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not available")
def test_fortran_interface_module_compiles(tmp_path):
    src = os.path.join(ROOT, "ps3d_amd", "fortran", "ps3d_capi.f90")
    r = subprocess.run(["amdflang", "-c", src, "-module-dir", str(tmp_path), "-o", str(tmp_path / "c.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", str(tmp_path / "c.o")], capture_output=True, text=True).stdout
    assert "ps3d_error_message" in nm


@pytest.mark.skipif(shutil.which("amdflang") is None, reason="amdflang not available")
def test_fortran_dropin_links_against_reference_module(tmp_path):
    """integrate_gpu.f90 compiles against the reference's own main_module and
    exports the two symbols the reference driver links (integrate_,
    pressure_solver_)."""
    import oracle
    refobj = os.path.join(os.path.dirname(oracle.REF_SO), "obj")
    if not os.path.exists(os.path.join(refobj, "main_module.mod")):
        pytest.skip("reference build (oracle/_ref) not present")
    d = str(tmp_path)
    capi = os.path.join(ROOT, "ps3d_amd", "fortran", "ps3d_capi.f90")
    shim = os.path.join(ROOT, "ps3d_amd", "fortran", "integrate_gpu.f90")
    for src, obj in ((capi, "capi.o"), (shim, "shim.o")):
        r = subprocess.run(["amdflang", "-c", src, "-I", refobj, "-I", d, "-module-dir", d, "-o",
                            os.path.join(d, obj)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", os.path.join(d, "shim.o")], capture_output=True, text=True).stdout
    assert re.search(r"\bT integrate_\b", nm) and re.search(r"\bT pressure_solver_\b", nm)
    for sym in ("ps3d_integrate", "ps3d_create", "ps3d_pressure_solver"):
        assert re.search(rf"\bU {sym}\b", nm), sym


@pytest.mark.parametrize("bad,msg", [
    (dict(Avbi=1e-2), "vertical biharmonic"),
    (dict(nx=36), "compiled-in FFT lengths"),
    (dict(precision=1), "share one precision form"),
    (dict(precision=7), "unknown precision form"),
])
def test_ensemble_create_rejects_bad_members(bad, msg):
    from ps3d_amd import _lib
    from ps3d_amd.params import CParams, ModelParams
    lib = _lib.load()
    good = ModelParams(nx=32, ny=32, nz=16, Lx=1.0, Ly=1.0, Lz=1.0, dt=0.1, enable_vertical_boundaries=True)
    base = dict(nx=32, ny=32, nz=16, Lx=1.0, Ly=1.0, Lz=1.0, dt=0.1, enable_vertical_boundaries=True)
    base.update(bad)
    arr = (CParams * 2)(CParams.from_params(good), CParams.from_params(ModelParams(**base)))
    h = ctypes.c_void_p()
    code = lib.ps3d_ensemble_create(2, arr, None, ctypes.byref(h))
    assert code == 1
    assert not h.value
    assert "member 1" in _lib.last_error() and msg in _lib.last_error()


def test_mixed_ensemble_rejects_tall_lid_members():
    """A mixed ensemble's Thomas stage keeps 32 columns of FP64 factors in
    LDS (at most 160 KiB per work-group on gfx950): rigid-lid members of more
    than 213 levels are refused at creation, before any device work."""
    from ps3d_amd import _lib
    from ps3d_amd.params import CParams, ModelParams, PS3D_MIXED
    lib = _lib.load()
    base = dict(nx=32, ny=32, Lx=1.0, Ly=1.0, Lz=1.0, dt=0.1, enable_vertical_boundaries=True,
                precision=PS3D_MIXED)
    arr = (CParams * 2)(CParams.from_params(ModelParams(nz=16, **base)),
                        CParams.from_params(ModelParams(nz=256, **base)))
    h = ctypes.c_void_p()
    code = lib.ps3d_ensemble_create(2, arr, None, ctypes.byref(h))
    assert code == 1
    assert not h.value
    assert "member 1" in _lib.last_error() and "at most 213 levels" in _lib.last_error()
