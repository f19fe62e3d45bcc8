"""Pinning of the balancing-diagnostic oracle (oracle/balance.py).

* The discrete normal modes (omega2, coeff_p, vec_q, vec_p; rigid lid and
  periodic) against the reference's own eigenvectors.f90, compiled unchanged
  in oracle/_ref, at spectral points of real grids and at the special points
  (kx = ky = 0, kz = 0, Nyquist) where the reference's case rules apply.
* The vertical-mode tables (my_fft.f90:112-127) by orthonormality and by the
  difference / average relations that define phi and psib from psi.
* The projections (diag_balance.f90) by their algebra: the zero-order
  projection is idempotent; a fast-mode eigenvector has no slow part.
"""
import ctypes

import numpy as np
import pytest

import oracle
from oracle import balance

needs_ref = pytest.mark.skipif(not oracle.reference_available(), reason="reference build not present")


def _prm(lid, nx=16, ny=12, nz=8):
    from ps3d_amd import configs
    prm, fields = configs.random_state(nx, ny, nz, seed=3, lid=lid, Avbi=0.0, Kvbi=0.0)
    return prm, fields


def _ref_call(lib, name, kx, ky, kz, prm, s=None, nout=4):
    d = ctypes.c_double
    args = [d(kx), d(ky), d(kz), d(prm.dx), d(prm.dy), d(prm.dz)]
    if s is not None:
        args.append(d(s))
    args += [d(prm.f0), d(prm.N0), d(prm.dsqr)]
    if nout == 4:
        outs = [(ctypes.c_double * 2)() for _ in range(4)]
        getattr(lib, name)(*[ctypes.byref(a) for a in args], *outs)
        return [complex(o[0], o[1]) for o in outs]
    out = ctypes.c_double()
    getattr(lib, name)(*[ctypes.byref(a) for a in args], ctypes.byref(out))
    return out.value


@needs_ref
@pytest.mark.parametrize("lid", [True, False])
def test_modes_match_reference_eigenvectors(lid):
    lib = oracle.load_reference()
    prm, _ = _prm(lid)
    orc = oracle.OracleModel(prm)
    kx, ky, kz = orc.f["kx"].copy(), orc.f["ky"].copy(), orc.f["kz"].copy()
    M = balance.Modes(prm, kx, ky, kz)
    suf = "vbc" if lid else "discrete"
    om2 = M.omega2()
    vals = {s: (M.vec_q(s), M.vec_p(s), M.coeff_p(s)) for s in (0.0, 1.0, -1.0)}
    rng = np.random.default_rng(0)
    pts = [(0, 0, 0), (0, 0, 1), (1, 0, 0), (0, 1, 0), (len(kx) - 1, 0, 0), (0, len(ky) // 2, 0),
           (len(kx) - 1, len(ky) // 2, len(kz) - 1)]
    pts += [tuple(int(rng.integers(n)) for n in (len(kx), len(ky), len(kz))) for _ in range(40)]
    for (i, j, k) in pts:
        ref = _ref_call(lib, f"omega2_{suf}_", kx[i], ky[j], kz[k], prm, nout=1)
        assert om2[i, j, k] == pytest.approx(ref, rel=1e-14, abs=1e-300)
        for s in (0.0, 1.0, -1.0):
            q, p, c = vals[s]
            rq = _ref_call(lib, f"vec_q_{suf}_", kx[i], ky[j], kz[k], prm, s)
            rp = _ref_call(lib, f"vec_p_{suf}_", kx[i], ky[j], kz[k], prm, s)
            rc = _ref_call(lib, f"coeff_p_{suf}_", kx[i], ky[j], kz[k], prm, s, nout=1)
            assert c[i, j, k] == pytest.approx(rc, rel=1e-13, abs=1e-300), (i, j, k, s)
            for n in range(4):
                for a, b in ((q[n][i, j, k], rq[n]), (p[n][i, j, k], rp[n])):
                    if np.isfinite(b):
                        assert abs(a - b) <= 1e-12 * max(abs(b), 1e-300), (i, j, k, s, n, a, b)
    orc.close()


@pytest.mark.parametrize("nz", [4, 7, 16])
def test_vertical_modes_orthonormal(nz):
    Lz = 500.0
    dz = Lz / nz
    phi, psi, psib = balance.vertical_modes(nz, Lz)
    G = phi.T @ phi * dz
    assert np.allclose(G, np.eye(nz), atol=1e-13)
    Gs = psi[:, 1:].T @ psi[:, 1:] * dz
    assert np.allclose(Gs, np.eye(nz - 1), atol=1e-13)
    assert np.all(psi[:, 0] == 0) and np.all(psib[:, 0] == 0)
    # phi(k) = (psi(k) - psi(k-1)) / (2 sin), psib(k) = (psi(k) + psi(k-1)) / (2 cos)
    for k2 in range(2, nz + 1):
        sn = 2 * np.sin((k2 - 1) * np.pi / (2 * nz))
        cs = 2 * np.cos((k2 - 1) * np.pi / (2 * nz))
        prev = np.concatenate([[0.0], psi[:-1, k2 - 1]])
        assert np.allclose(phi[:, k2 - 1] * sn, psi[:, k2 - 1] - prev, atol=1e-14)
        assert np.allclose(psib[:, k2 - 1] * cs, psi[:, k2 - 1] + prev, atol=1e-14)


@pytest.mark.parametrize("lid", [True, False])
def test_zero_order_projection_is_idempotent(lid):
    prm, fields = _prm(lid)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        orc.f[n][...] = a
    bal = balance.Balance(orc)
    bal.zero_order()
    first = {n: bal.r[n + "0"].copy() for n in ("u", "v", "w", "b")}
    for n in ("u", "v", "w", "b"):
        orc.f[n][...] = first[n]
    bal.zero_order()
    for n in ("u", "v", "w", "b"):
        scale = np.abs(first[n]).max() + 1e-300
        assert np.abs(bal.r[n + "0"] - first[n]).max() <= 1e-11 * scale, n
    orc.close()


@pytest.mark.parametrize("lid", [True, False])
def test_fast_mode_has_no_slow_part(lid):
    """a state made of one +omega eigenvector q+ (one wavenumber) projects to
    zero under the slow projection p0 . (q+) = 0 (biorthogonality)."""
    prm, _ = _prm(lid)
    orc = oracle.OracleModel(prm)
    M = balance.Modes(prm, orc.f["kx"], orc.f["ky"], orc.f["kz"])
    qp = M.vec_q(1.0)
    p0 = M.vec_p(0.0)
    dot = sum(p0[c] * qp[c] for c in range(4))
    nrm = sum(M.vec_p(1.0)[c] * qp[c] for c in range(4))
    # the modes with a vertical structure (kz > 0); the barotropic rigid-lid
    # mode (kz = 0) has no gravity-wave partner
    sel = np.broadcast_to((np.asarray(orc.f["kz"]) ** 2 > 0)[None, None, :], dot.shape)
    big = max(float(np.abs(qp[c]).max()) for c in range(4)) * max(float(np.abs(p0[c]).max()) for c in range(4))
    assert np.abs(dot[sel]).max() <= 1e-12 * big
    # and p+ . q+ = 1 where the fast mode exists
    live = sel & (np.abs(nrm) > 0)
    assert np.allclose(nrm[live], 1.0, rtol=1e-9)
    orc.close()
