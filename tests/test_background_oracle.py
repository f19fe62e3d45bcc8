"""Pinning of the oracle's background-flow advection (orc_background_flow,
lee_wave1.f90:239-314) and body force on the CPU.

With a constant u_back = U and fields that vary in x only (periodic), the
flux divergences reduce to centred differences: du = -Ro U (u(i+1) -
u(i-1))/dx (the u flux carries the factor 2 of the reference's "2*0.25"),
dv = -Ro (U/2)(v(i+1) - v(i-1))/dx, dw = -dsqr Ro (U/2)(w(i+1) - w(i-1))/dx,
db = -Ro (U/2)(b(i+1) - b(i-1))/dx, and with b_back(x) and u = c the extra
db = -Ro (c/2)(b_back(i+1) - b_back(i-1))/dx.  For sin(k x) the centred
difference is 2 sin(k dx) cos(k x)/dx exactly, so the expected tendencies
are closed forms independent of the code.  A wrong factor, sign, operand or
index fails them.
"""
import math

import numpy as np
import pytest

from ps3d_amd import configs


def _x_only(prm, f):
    """a full array (halo included) holding f(x_i) at every (j, k)"""
    i = np.arange(-1, prm.nx + 3)
    return np.asfortranarray(np.broadcast_to(f(i)[:, None, None], prm.shape).copy())


@pytest.mark.parametrize("Ro,dsqr", [(1.0, 1.0), (0.7, 0.04)])
def test_background_flow_closed_form(oracle_mod, Ro, dsqr):
    prm, _ = configs.random_state(32, 8, 6, seed=1, lid=False, Ro=Ro, dsqr=dsqr)
    orc = oracle_mod.OracleModel(prm)
    U, dx, L = 0.3, prm.dx, prm.Lx
    k = 2 * math.pi / L * 3
    amp = {"u": 1.0, "v": 0.5, "w": 0.25, "b": 2.0}
    for n, a in amp.items():
        orc.f[n][...] = _x_only(prm, lambda i, a=a: a * np.sin(k * (i - 0.5) * dx))
    tau = 2
    for n in ("du", "dv", "dw", "db"):
        orc.f[n][...] = 0.0
    orc.background_flow(np.full(prm.shape, U, order="F"), None, tau=tau)
    I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
    xi = (np.arange(1, prm.nx + 1) - 0.5) * dx
    cd = (2 * math.sin(k * dx) * np.cos(k * xi) / dx)[:, None, None]     # (f(i+1) - f(i-1))/dx of sin
    want = {"du": -Ro * U * amp["u"] * cd, "dv": -Ro * (U / 2) * amp["v"] * cd,
            "dw": -dsqr * Ro * (U / 2) * amp["w"] * cd, "db": -Ro * (U / 2) * amp["b"] * cd}
    for n, w in want.items():
        got = orc.f[n][..., tau - 1][I]
        assert np.allclose(got, np.broadcast_to(w, got.shape), rtol=1e-12, atol=1e-12 * np.abs(w).max()), n
    # the background buoyancy term: u = c, b = 0, b_back = sin
    c = 0.2
    orc.f["u"][...] = c
    orc.f["b"][...] = 0.0
    orc.f["db"][...] = 0.0
    bb = _x_only(prm, lambda i: np.sin(k * (i - 0.5) * dx))
    orc.background_flow(np.full(prm.shape, U, order="F"), bb, tau=tau)
    got = orc.f["db"][..., tau - 1][I]
    w = np.broadcast_to(-Ro * (c / 2) * cd, got.shape)
    assert np.allclose(got, w, rtol=1e-12, atol=1e-12 * np.abs(w).max())
    orc.close()


def test_body_force_adds_to_the_slot(oracle_mod):
    prm, fields = configs.random_state(12, 8, 6, seed=2, lid=True)
    orc = oracle_mod.OracleModel(prm)
    rng = np.random.default_rng(0)
    f = {"dw": rng.standard_normal(prm.shape), "db": rng.standard_normal(prm.shape)}
    before = {n: orc.f[n][..., 0].copy() for n in ("du", "dw", "db")}
    fw, fb = np.asfortranarray(f["dw"]), np.asfortranarray(f["db"])   # kept alive across the call
    tab = (oracle_mod._dp * 4)(None, None, oracle_mod._as_dp(fw), oracle_mod._as_dp(fb))
    orc.lib.orc_body_force(orc.h, 1, tab)
    assert np.array_equal(orc.f["du"][..., 0], before["du"])
    assert np.array_equal(orc.f["dw"][..., 0], before["dw"] + f["dw"])
    assert np.array_equal(orc.f["db"][..., 0], before["db"] + f["db"])
    orc.close()
