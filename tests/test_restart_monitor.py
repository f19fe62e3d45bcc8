"""Restart records (restart.f90:2-82) and the device monitors
(diagnose.F90:11-53).

The record layout is pinned against the reference's own write_restart /
read_restart compiled in oracle/_ref: a file written by the library is
byte-identical to the reference's for the same state, and each side reads the
other's file back bit for bit.  The monitors are checked against a numpy
restatement of diagnose.F90's loop (sums to 1e-13: only the summation order
differs).
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

needs_ref = pytest.mark.skipif(not oracle.reference_available(), reason="reference build not present")


def _state(prm, seed=0):
    rng = np.random.default_rng(seed)
    shp = prm.shape
    f = {n: rng.standard_normal(shp) for n in ("u", "v", "w", "b", "p")}
    for n in ("du", "dv", "dw", "db"):
        f[n] = rng.standard_normal(shp + (3,))
    return f


def _flat(f):
    return np.concatenate([np.asfortranarray(f[n]).ravel(order="F")
                           for n in ("u", "v", "w", "b", "p", "du", "dv", "dw", "db")])


@needs_ref
def test_restart_records_match_the_reference(tmp_path, monkeypatch):
    from ps3d_amd import _lib, configs
    lib = _lib.load()
    prm, _ = configs.random_state(12, 10, 6, seed=1)
    f = _state(prm)
    ref = oracle.ReferenceModel(prm)
    for n, a in f.items():
        ref.f[n][...] = a
    ref.lib.ref_set_time(37, 2, 1, 3)
    ref.lib.ref_set_clock(12345.625)
    monkeypatch.chdir(tmp_path)
    ref.lib.write_restart_()
    ref_file = tmp_path / "restart_pe=000000000.dta"
    assert ref_file.exists()
    # ours, from host arrays, byte for byte
    ours = tmp_path / "ours.dta"
    state = np.ascontiguousarray(_flat(f))
    rc = lib.ps3d_restart_write_host(str(ours).encode(), prm.nx, prm.ny, prm.nz, 37, 12345.625, 2, 1, 3,
                                     state.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, _lib.last_error()
    assert ours.read_bytes() == ref_file.read_bytes()
    # ours reads the reference's file
    back = np.zeros_like(state)
    itt, tau, t1, t2 = (ctypes.c_int32() for _ in range(4))
    tm = ctypes.c_double()
    rc = lib.ps3d_restart_read_host(str(ref_file).encode(), prm.nx, prm.ny, prm.nz, ctypes.byref(itt),
                                    ctypes.byref(tm), ctypes.byref(tau), ctypes.byref(t1), ctypes.byref(t2),
                                    back.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0, _lib.last_error()
    assert (itt.value, tau.value, t1.value, t2.value, tm.value) == (37, 2, 1, 3, 12345.625)
    assert np.array_equal(back, state)
    # the reference reads ours
    for n in f:
        ref.f[n][...] = 0.0
    os.replace(ours, ref_file)
    ref.lib.read_restart_()
    for n, a in f.items():
        assert np.array_equal(ref.f[n], a), n


def test_restart_rejects_a_different_grid(tmp_path):
    from ps3d_amd import _lib, configs
    lib = _lib.load()
    prm, _ = configs.random_state(12, 10, 6, seed=1)
    state = np.ascontiguousarray(_flat(_state(prm)))
    path = str(tmp_path / "r.dta").encode()
    dp = state.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.ps3d_restart_write_host(path, prm.nx, prm.ny, prm.nz, 1, 0.0, 3, 2, 1, dp) == 0
    big = np.zeros(state.size * 2)
    i = [ctypes.c_int32() for _ in range(4)]
    tm = ctypes.c_double()
    rc = lib.ps3d_restart_read_host(path, prm.nx + 2, prm.ny, prm.nz, ctypes.byref(i[0]), ctypes.byref(tm),
                                    ctypes.byref(i[1]), ctypes.byref(i[2]), ctypes.byref(i[3]),
                                    big.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 1 and "grid size" in _lib.last_error()


def diagnose_numpy(f, prm):
    """diagnose.F90:11-53, loop order i fastest, as the reference sums"""
    I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
    u, v, w, b = (np.asarray(f[n])[I] for n in ("u", "v", "w", "b"))
    sp = np.sqrt(u ** 2 + v ** 2)
    ke = float(np.sum((0.5 * (u ** 2 + v ** 2 + prm.dsqr * w ** 2)).ravel(order="F")))
    if prm.N0 > 0:
        pe = float(np.sum((0.5 * b ** 2 / prm.N0 ** 2).ravel(order="F")))
    else:
        zk = (np.arange(1, prm.nz + 1) - 0.5) * prm.dz
        pe = float(np.sum((b * zk[None, None, :]).ravel(order="F")))
    n = prm.nx * prm.ny * prm.nz
    umax, wmax = float(sp.max()), float(np.abs(w).max())
    um = max(umax, wmax)
    pecl_h = min(um * prm.dx / (1e-32 + prm.Ah), um * prm.dx ** 3 / (1e-32 + prm.Ahbi))
    pecl_v = min(um * prm.dz / (1e-32 + prm.Av), um * prm.dz ** 3 / (1e-32 + prm.Avbi))
    return [ke / n, pe / n, umax, wmax, umax * prm.dt / prm.dx, wmax * prm.dt / prm.dz, pecl_h, pecl_v]


@pytest.mark.gpu
@pytest.mark.parametrize("mixed", [False, True])
def test_device_monitors_and_restart_roundtrip(tmp_path, mixed):
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields = configs.jets(64, 64, 16, seed=3)
    if mixed:
        prm = prm.replace(precision=PS3D_MIXED)
    m = Model(prm)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()
    m.run(7)
    m.pull(["u", "v", "w", "b"])
    out = m.diagnose()
    ref = diagnose_numpy(m.f, prm)
    print("monitors", out, flush=True)
    for a, b in zip(out["values"], ref):
        assert a == pytest.approx(b, rel=1e-13, abs=1e-300)
    assert not out["nan"]
    if mixed:
        m.close()
        return
    # restart: write, run on, read back, run again: the same states
    path = str(tmp_path / "r.dta")
    m.write_restart(path, time=700.0)
    m.run(3)
    m.pull(["u", "w", "p"])
    after = {n: m.f[n].copy() for n in ("u", "w", "p")}
    st = m.read_restart(path)
    assert st["itt"] == 7 and st["time"] == 700.0
    m.run(3)
    m.pull(["u", "w", "p"])
    for n, a in after.items():
        assert np.array_equal(m.f[n], a), n
    m.close()
