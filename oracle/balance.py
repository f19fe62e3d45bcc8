"""Normal-mode balancing diagnostic restated in numpy -- TEST INFRASTRUCTURE ONLY.

Follows for_src/diag_balance.f90 (Warn et al. slaving to third order),
for_src/eigenvectors.f90 (discrete normal modes: omega2, coeff_p, vec_q,
vec_p for the rigid-lid (_vbc) and periodic (_discrete) cases) and
for_src/my_fft.f90:112-127, :211-252 (vertical-mode tables and transforms),
on top of the C oracle (advection, integrate, 2-D/3-D FFTs).  Only tests/
may import it, as the checker of the GPU diagnostic.

Pinning: the eigenvector functions against the reference's own
eigenvectors.f90 compiled in oracle/_ref (tests/test_balance_oracle.py); the
vertical modes by orthonormality (phi, psi are orthonormal under the dz
inner product, so the vertical-mode transform pair is the identity) and the
discrete relations psi(k)-psi(k-1) = 2 sin() phi(k); the projections by
their algebra (the zero-order projection is idempotent, a slow-mode
eigenvector is reproduced, a fast mode is removed).  The orchestration
(diag_balance.f90:89-362) calls P3DFFT and cannot be built here.  Its
first, second and third orders are pinned by the slaving relations they
satisfy, with every piece taken from the pinned oracle step run as a black
box (tests/test_balance_pinning.py): the rate of each fast correction under
the linear step plus the nonlinear forcing of that order (the step with and
without its nonlinear terms) equals the rate of change of the lower-order
slaved amplitude; and a state independent of x with v = w = 0 has no
corrections.  Where the periodic form's kz = 0 fast modes
(eigenvectors.f90:87-92) turn with the opposite sign, the reference's
correction has the opposite sign as well, and so does this one.
"""
from __future__ import annotations

import numpy as np

ONX = 2


# ---- eigenvectors.f90:224-270 ------------------------------------------------
def hat_ksqr(k, d):
    return 2.0 * (1.0 - np.cos(k * d)) / d ** 2


def hat_epm(k, d):
    return 0.5 * (1 + np.cos(k * d))


def hat_kp(k, d):
    return -1j * (np.exp(1j * k * d) - 1.0) / d


def hat_ep(k, d):
    return (np.exp(1j * k * d) + 1.0) * 0.5


def hat_em(k, d):
    return (np.exp(-1j * k * d) + 1.0) * 0.5


# ---- eigenvectors.f90:124-218 (rigid lid) and :6-118 (periodic) ----------------
class Modes:
    """omega2, coeff_p, vec_q, vec_p on the spectral grid (nx/2+1, ny, nz) of
    one model (arrays broadcast over kx, ky, kz)."""

    def __init__(self, prm, kx, ky, kz):
        self.prm = prm
        self.vbc = bool(prm.enable_vertical_boundaries)
        self.kx = np.asarray(kx, dtype=np.float64)[:, None, None]
        self.ky = np.asarray(ky, dtype=np.float64)[None, :, None]
        self.kz = np.asarray(kz, dtype=np.float64)[None, None, :]

    # eigenvectors.f90:124-141 / :6-21
    def omega2(self):
        p = self.prm
        kx, ky, kz = self.kx, self.ky, self.kz
        k2 = hat_ksqr(kx, p.dx) + hat_ksqr(ky, p.dy)
        with np.errstate(divide="ignore", invalid="ignore"):
            if self.vbc:
                arg = kz / 2 * p.dz
                m2 = 4 * np.sin(arg) ** 2 / p.dz ** 2
                N2 = p.N0 ** 2 * np.cos(arg) ** 2
                f2 = p.f0 ** 2 * hat_epm(kx, p.dx) * hat_epm(ky, p.dy)
                K = p.dsqr * k2 + m2
                om2 = (N2 * k2 + f2 * m2) / K
            else:
                K = p.dsqr * k2 + hat_ksqr(kz, p.dz)
                om2 = (hat_epm(kz, p.dz) * p.N0 ** 2 * k2
                       + p.f0 ** 2 * hat_ksqr(kz, p.dz) * hat_epm(kx, p.dx) * hat_epm(ky, p.dy)) / K
        return np.where(K > 0.0, om2, 0.0)

    def _common(self):
        p = self.prm
        kx, ky, kz = self.kx, self.ky, self.kz
        k2 = hat_ksqr(kx, p.dx) + hat_ksqr(ky, p.dy)
        f2 = p.f0 ** 2 * hat_epm(kx, p.dx) * hat_epm(ky, p.dy)
        if self.vbc:
            arg = kz / 2 * p.dz
            m2 = 4 * np.sin(arg) ** 2 / p.dz ** 2
            N2 = p.N0 ** 2 * np.cos(arg) ** 2
        else:
            arg = None
            m2 = hat_ksqr(kz, p.dz)
            N2 = p.N0 ** 2 * hat_epm(kz, p.dz)
        return k2, f2, m2, N2, arg

    # eigenvectors.f90:144-165 / :24-43
    def coeff_p(self, s):
        p = self.prm
        om2 = self.omega2()
        k2, f2, m2, N2, _ = self._common()
        d = k2 * N2 + f2 * m2
        with np.errstate(divide="ignore", invalid="ignore"):
            if s == 0.0:
                val = f2 * N2 / d
            else:
                val = (N2 - p.dsqr * om2) * (om2 - f2) / (2 * d)
        return np.where(d > 0.0, val, 0.0)

    # eigenvectors.f90:168-206 / :46-93
    def vec_q(self, s):
        p = self.prm
        kx, ky, kz = self.kx, self.ky, self.kz
        om2 = self.omega2()
        om = np.sqrt(om2)
        k2, f2, m2, N2, arg = self._common()
        shp = np.broadcast_shapes(kx.shape, ky.shape, kz.shape)
        hor = np.broadcast_to(kx ** 2 + ky ** 2 > 0.0, shp)
        ver = np.broadcast_to(kz ** 2 > 0.0, shp)
        with np.errstate(divide="ignore", invalid="ignore"):
            fxa = f2 - s ** 2 * om2
            qx = (-1j * p.f0 * hat_ep(kx, p.dx) * hat_em(ky, p.dy) * hat_kp(ky, p.dy)
                  + s * om * hat_kp(kx, p.dx)) / fxa
            qy = (+1j * p.f0 * hat_em(kx, p.dx) * hat_ep(ky, p.dy) * hat_kp(kx, p.dx)
                  + s * om * hat_kp(ky, p.dy)) / fxa
            fxb = N2 - p.dsqr * s ** 2 * om2
            if self.vbc:
                lam = 2 * np.sin(arg) / p.dz
                qz = 1j * lam * s * om / fxb
                qb = -lam * np.cos(arg) * p.N0 ** 2 / fxb
            else:
                qz = hat_kp(kz, p.dz) * s * om / fxb
                qb = 1j * hat_em(kz, p.dz) * hat_kp(kz, p.dz) * p.N0 ** 2 / fxb
        qx = np.broadcast_to(qx, shp).astype(np.complex128)
        qy = np.broadcast_to(qy, shp).astype(np.complex128)
        qz = np.where(ver, np.broadcast_to(qz, shp), 0.0).astype(np.complex128)
        qb = np.where(ver, np.broadcast_to(qb, shp), 0.0).astype(np.complex128)
        qx = np.where(hor, qx, -1j * s)
        qy = np.where(hor, qy, s ** 2 + 0j)
        qz = np.where(hor, qz, 0.0)
        qb = np.where(hor, qb, 0.0)
        f2b = np.broadcast_to(f2, shp)
        N2b = np.broadcast_to(N2, shp)
        om2b = np.broadcast_to(om2, shp)
        if s == 0.0:
            z = (f2b == 0.0) | (N2b == 0.0)
        else:
            z = om2b == 0.0
        qx, qy, qz, qb = (np.where(z, 0.0, a) for a in (qx, qy, qz, qb))
        if not self.vbc and s != 0.0:
            # eigenvectors.f90:87-92
            h = hor & ~ver
            qx = np.where(h, 0.0, qx)
            qy = np.where(h, 0.0, qy)
            qz = np.where(h, 1j * s, qz)
            qb = np.where(h, s ** 2 * p.N0 * np.sqrt(p.dsqr), qb)
        return qx, qy, qz, qb

    # eigenvectors.f90:209-218 / :96-118
    def vec_p(self, s):
        p = self.prm
        c = self.coeff_p(s)
        qx, qy, qz, qb = self.vec_q(s)
        with np.errstate(divide="ignore", invalid="ignore"):
            px = c * np.conj(qx)
            py = c * np.conj(qy)
            pz = c * np.conj(qz) * p.dsqr
            pb = c * np.conj(qb) / p.N0 ** 2
        if not self.vbc:
            shp = px.shape
            hor = np.broadcast_to(self.kx ** 2 + self.ky ** 2 > 0.0, shp)
            ver = np.broadcast_to(self.kz ** 2 > 0.0, shp)
            a = ~hor & ver
            px = np.where(a, 1j * s / 2.0, px)
            py = np.where(a, s ** 2 / 2.0 + 0j, py)
            if s != 0.0:
                h = hor & ~ver
                px = np.where(h, 0.0, px)
                py = np.where(h, 0.0, py)
                pz = np.where(h, -1j * s / 2.0, pz)
                pb = np.where(h, s ** 2 / (2.0 * p.N0 * np.sqrt(p.dsqr)), pb)
        return px, py, pz, pb


# ---- my_fft.f90:112-127 ------------------------------------------------------
def vertical_modes(nz, Lz):
    """(phi, psi, psib), each (nz, nz): [k-1, mode-1]."""
    pi = np.pi
    phi = np.zeros((nz, nz))
    psi = np.zeros((nz, nz))
    psib = np.zeros((nz, nz))
    phi[:, 0] = np.sqrt(1.0 / Lz)
    for k2 in range(2, nz + 1):
        for k in range(1, nz + 1):
            psi[k - 1, k2 - 1] = np.sqrt(2.0 / Lz) * np.sin((k2 - 1) * k * pi / nz)
        sn = 2 * np.sin((k2 - 1) * pi / (2 * nz))
        cs = 2 * np.cos((k2 - 1) * pi / (2 * nz))
        phi[0, k2 - 1] = psi[0, k2 - 1] / sn
        psib[0, k2 - 1] = psi[0, k2 - 1] / cs
        for k in range(2, nz + 1):
            phi[k - 1, k2 - 1] = (psi[k - 1, k2 - 1] - psi[k - 2, k2 - 1]) / sn
            psib[k - 1, k2 - 1] = (psi[k - 1, k2 - 1] + psi[k - 2, k2 - 1]) / cs
    return phi, psi, psib


def border_exchg(a: np.ndarray) -> None:
    """parallel_mpi.f90:70-211 on one rank, in place (z, then y, then x)."""
    nx, ny, nz = a.shape[0] - 4, a.shape[1] - 4, a.shape[2] - 4
    for k in (1, 2):
        a[:, :, 2 - k] = a[:, :, 2 + nz - k]
        a[:, :, 1 + nz + k] = a[:, :, 1 + k]
    for j in (1, 2):
        a[:, 2 - j, :] = a[:, 2 + ny - j, :]
        a[:, 1 + ny + j, :] = a[:, 1 + j, :]
    for i in (1, 2):
        a[2 - i, :, :] = a[2 + nx - i, :, :]
        a[1 + nx + i, :, :] = a[1 + i, :, :]


class Balance:
    """diag_balance.f90 on one oracle model (its current state u, v, w, b).

    After run(): u0..b3 (full arrays with halo) as in module_diag_balance."""

    def __init__(self, orc, filter_width=None):
        self.orc = orc
        self.prm = orc.prm
        self.modes = Modes(orc.prm, orc.f["kx"], orc.f["ky"], orc.f["kz"])
        self.filter_width = filter_width
        if self.modes.vbc:
            self.phi, self.psi, self.psib = vertical_modes(self.prm.nz, self.prm.Lz)
        self.s = {}        # stored spectral arrays (I_00, f1p, ...)
        self.r = {}        # real outputs u0..b3, pTu0..

    # diag_balance.f90:421-438 / my_fft.f90:235-252
    def forward(self, fields):
        p = self.prm
        I = (slice(2, 2 + p.nx), slice(2, 2 + p.ny), slice(2, 2 + p.nz))
        out = []
        for n, a in enumerate(fields):
            qp = np.asfortranarray(a[I])
            if self.modes.vbc:
                Qloc = self.orc.my_fft(qp, three_d=False)
                mode = (self.phi, self.phi, self.psi, self.psib)[n]
                QS = np.zeros_like(Qloc)
                for k1 in range(p.nz):
                    for k2 in range(p.nz):
                        QS[:, :, k1] = QS[:, :, k1] + Qloc[:, :, k2] * mode[k2, k1] * p.dz
            else:
                QS = self.orc.my_fft(qp, three_d=True)
            out.append(QS)
        return out

    # diag_balance.f90:441-462 / my_fft.f90:211-230
    def backward(self, spec):
        p = self.prm
        out = []
        for n, QS in enumerate(spec):
            if self.modes.vbc:
                mode = (self.phi, self.phi, self.psi, self.psib)[n]
                Qloc = np.zeros_like(QS)
                for k1 in range(p.nz):
                    for k2 in range(p.nz):
                        Qloc[:, :, k2] = Qloc[:, :, k2] + QS[:, :, k1] * mode[k2, k1]
                qp = self.orc.my_ifft(Qloc, three_d=False)
            else:
                qp = self.orc.my_ifft(QS, three_d=True)
            a = np.zeros(self.orc.shape, order="F")
            a[2:2 + p.nx, 2:2 + p.ny, 2:2 + p.nz] = qp
            border_exchg(a)
            out.append(a)
        return out

    # diag_balance.f90:500-535
    def nonlin(self, u, v, w, b):
        o, p = self.orc, self.prm
        save = {n: o.f[n].copy() for n in ("u", "v", "w", "b")}
        slot = {n: o.f[n][..., 0].copy() for n in ("du", "dv", "dw", "db")}
        for n, a in zip(("u", "v", "w", "b"), (u, v, w, b)):
            o.f[n][...] = a
        o.buoyancy_advection(tau=1)
        o.momentum_advection(tau=1)
        db = o.f["db"][..., 0] / p.Ro
        du = o.f["du"][..., 0] / p.Ro
        dv = o.f["dv"][..., 0] / p.Ro
        dw = o.f["dw"][..., 0] / p.Ro / p.dsqr
        for n in save:
            o.f[n][...] = save[n]
        for n in slot:
            o.f[n][..., 0] = slot[n]
        S = self.forward((du, dv, dw, db))
        if self.filter_width is not None:
            kz = np.asarray(o.f["kz"])
            fxa = abs(kz[p.nz // 2 + self.filter_width])    # kz(nz/2+1+width), 1-based
            cut = (kz ** 2 >= fxa ** 2)[None, None, :]
            S = [np.where(cut, 0.0, a) for a in S]
        return S

    # diag_balance.f90:465-497 (note: itt stays 0, so every step is Euler)
    def model_run(self, u, v, w, b, nsteps):
        o = self.orc
        names = ("u", "v", "w", "b", "p", "du", "dv", "dw", "db")
        save = {n: o.f[n].copy() for n in names}
        st = (o.itt, o.tau, o.taum1, o.taum2)
        for n, a in zip(("u", "v", "w", "b"), (u, v, w, b)):
            o.f[n][...] = a
        o.itt, o.taum2, o.taum1, o.tau = 0, 1, 2, 3
        for _ in range(nsteps):
            o.integrate()
            o.tau = o.tau % 3 + 1
            o.taum1 = o.taum1 % 3 + 1
            o.taum2 = o.taum2 % 3 + 1
        out = [o.f[n].copy() for n in ("u", "v", "w", "b")]
        for n in names:
            o.f[n][...] = save[n]
        o.itt, o.tau, o.taum1, o.taum2 = st
        return out

    def _dot(self, pv, S):
        return pv[0] * S[0] + pv[1] * S[1] + pv[2] * S[2] + pv[3] * S[3]

    # diag_balance.f90:109-131
    def zero_order(self):
        o = self.orc
        S = self.forward((o.f["u"], o.f["v"], o.f["w"], o.f["b"]))
        q = self.modes.vec_q(0.0)
        pv = self.modes.vec_p(0.0)
        g0 = self._dot(pv, S)
        self.r["u0"], self.r["v0"], self.r["w0"], self.r["b0"] = self.backward([g0 * qq for qq in q])

    # diag_balance.f90:134-176
    def first_order(self):
        p, M, s, r = self.prm, self.modes, self.s, self.r
        S = self.nonlin(r["u0"], r["v0"], r["w0"], r["b0"])
        p0, pp, pm = M.vec_p(0.0), M.vec_p(1.0), M.vec_p(-1.0)
        qp, qm = M.vec_q(1.0), M.vec_q(-1.0)
        om = np.sqrt(M.omega2())
        s["I_00"] = 1j * self._dot(p0, S)
        s["I_0p"] = 1j * self._dot(pp, S)
        s["I_0m"] = 1j * self._dot(pm, S)
        with np.errstate(divide="ignore", invalid="ignore"):
            s["f1p"] = np.where(om > 0.0, s["I_0p"] / om, 0.0)
            s["f1m"] = np.where(om > 0.0, -s["I_0m"] / om, 0.0)
        out = [p.Ro * (s["f1p"] * qp[c] + s["f1m"] * qm[c]) for c in range(4)]
        r["u1"], r["v1"], r["w1"], r["b1"] = self.backward(out)

    # diag_balance.f90:181-268
    def second_order(self):
        p, M, s, r = self.prm, self.modes, self.s, self.r
        Ro = p.Ro
        pp, pm = M.vec_p(1.0), M.vec_p(-1.0)
        qp, qm, q0 = M.vec_q(1.0), M.vec_q(-1.0), M.vec_q(0.0)
        om = np.sqrt(M.omega2())
        S = self.nonlin(r["u1"] / Ro, r["v1"] / Ro, r["w1"] / Ro, r["b1"] / Ro)
        s["I_f1p"] = 1j * self._dot(pp, S)
        s["I_f1m"] = 1j * self._dot(pm, S)
        S = self.nonlin(r["u0"] + r["u1"] / Ro, r["v0"] + r["v1"] / Ro, r["w0"] + r["w1"] / Ro,
                        r["b0"] + r["b1"] / Ro)
        s["I_0f1p"] = 1j * self._dot(pp, S)
        s["I_0f1m"] = 1j * self._dot(pm, S)
        pTg0 = -1j * s["I_00"]
        r["pTu0"], r["pTv0"], r["pTw0"], r["pTb0"] = self.backward([pTg0 * q0[c] for c in range(4)])
        S = self.nonlin(r["pTu0"], r["pTv0"], r["pTw0"], r["pTb0"])
        s["I_pT0p"] = 1j * self._dot(pp, S)
        s["I_pT0m"] = 1j * self._dot(pm, S)
        S = self.nonlin(r["u0"] + r["pTu0"], r["v0"] + r["pTv0"], r["w0"] + r["pTw0"], r["b0"] + r["pTb0"])
        I_0pT0p = 1j * self._dot(pp, S)
        I_0pT0m = 1j * self._dot(pm, S)
        with np.errstate(divide="ignore", invalid="ignore"):
            pTf1p = (I_0pT0p - s["I_0p"] - s["I_pT0p"]) / om
            pTf1m = -(I_0pT0m - s["I_0m"] - s["I_pT0m"]) / om
            f2p = (s["I_0f1p"] - s["I_0p"] - s["I_f1p"] - 1j * pTf1p) / om
            f2m = -(s["I_0f1m"] - s["I_0m"] - s["I_f1m"] - 1j * pTf1m) / om
        s["f2p"] = np.where(om > 0, f2p, 0.0)
        s["f2m"] = np.where(om > 0, f2m, 0.0)
        out = [Ro ** 2 * (s["f2p"] * qp[c] + s["f2m"] * qm[c]) for c in range(4)]
        r["u2"], r["v2"], r["w2"], r["b2"] = self.backward(out)

    # diag_balance.f90:273-362
    def third_order(self):
        p, M, s, r = self.prm, self.modes, self.s, self.r
        Ro, dt = p.Ro, p.dt
        pp, pm = M.vec_p(1.0), M.vec_p(-1.0)
        S = self.nonlin(r["u2"] / Ro ** 2, r["v2"] / Ro ** 2, r["w2"] / Ro ** 2, r["b2"] / Ro ** 2)
        s["I_f2p"] = 1j * self._dot(pp, S)
        s["I_f2m"] = 1j * self._dot(pm, S)
        S = self.nonlin(r["u0"] + r["u2"] / Ro ** 2, r["v0"] + r["v2"] / Ro ** 2, r["w0"] + r["w2"] / Ro ** 2,
                        r["b0"] + r["b2"] / Ro ** 2)
        s["I_0f2p"] = 1j * self._dot(pp, S)
        s["I_0f2m"] = 1j * self._dot(pm, S)
        u3 = self.model_run(r["u0"] + r["u1"] + r["u2"], r["v0"] + r["v1"] + r["v2"],
                            r["w0"] + r["w1"] + r["w2"], r["b0"] + r["b1"] + r["b2"], 10)
        S = self.forward(u3)
        q0, p0 = M.vec_q(0.0), M.vec_p(0.0)
        g0 = self._dot(p0, S)
        r["u0"], r["v0"], r["w0"], r["b0"] = self.backward([g0 * q0[c] for c in range(4)])
        pTf2p, pTf2m = s["f2p"], s["f2m"]
        self.first_order()
        self.second_order()
        s["pTf2p"] = (s["f2p"] - pTf2p) / (10.0 * dt) / Ro
        s["pTf2m"] = (s["f2m"] - pTf2m) / (10.0 * dt) / Ro
        self.zero_order()
        self.first_order()
        self.second_order()
        qp, qm = M.vec_q(1.0), M.vec_q(-1.0)
        om = np.sqrt(M.omega2())
        with np.errstate(divide="ignore", invalid="ignore"):
            f3p = (s["I_0f2p"] + s["I_f1p"] - s["I_f2p"] - s["I_0p"] - 1j * s["pTf2p"]) / om
            f3m = -(s["I_0f2m"] + s["I_f1m"] - s["I_f2m"] - s["I_0m"] - 1j * s["pTf2m"]) / om
        f3p = np.where(om > 0, f3p, 0.0)
        f3m = np.where(om > 0, f3m, 0.0)
        out = [Ro ** 3 * (f3p * qp[c] + f3m * qm[c]) for c in range(4)]
        r["u3"], r["v3"], r["w3"], r["b3"] = self.backward(out)

    # diag_balance.f90:89-106
    def run(self):
        self.zero_order()
        self.first_order()
        self.second_order()
        self.third_order()
        return self.r


# ---- diag_opt_balance.f90 -----------------------------------------------------
def ramp(theta):
    """diag_opt_balance.f90:367-378"""
    import math
    t1 = 1.0 / max(1e-32, theta)
    t2 = 1.0 / max(1e-32, 1.0 - theta)
    return math.exp(-t1) / (math.exp(-t1) + math.exp(-t2))


class OptBalance:
    """diag_opt_balance.f90:55-179 on one oracle model: backward integration
    with the nonlinear terms ramped off (Ro -> 0, dt < 0, no biharmonic
    friction), projection on the geostrophic mode, forward integration with
    them ramped on, the geostrophic base point put back (or the time-average
    variant, enable_diag_opt_time_ave), until the norm of the change is
    below tol.  The model's state is restored at the end; run() returns
    (u_bal, v_bal, w_bal, b_bal) and the norm per iteration."""

    def __init__(self, orc, period, max_iterations, tol, time_average=False, average=1.0, average_times=1):
        self.orc = orc
        self.prm = orc.prm
        self.bal = Balance(orc)
        self.period, self.max_its, self.tol = period, max_iterations, tol
        self.time_average, self.average, self.average_times = time_average, average, average_times

    def _set(self, dt, Ro, bih=True):
        p = self.prm
        if bih:
            self.orc.set_dynamic(dt, Ro, p.Ahbi, p.Avbi, p.Khbi, p.Kvbi)
        else:
            self.orc.set_dynamic(dt, Ro, 0.0, 0.0, 0.0, 0.0)

    def _steps(self, n, dt, ro_of, bih=True, accumulate=None):
        o = self.orc
        o.itt, o.taum2, o.taum1, o.tau = 0, 1, 2, 3
        for s in range(n):
            self._set(dt, ro_of(s), bih)
            o.integrate()
            o.tau = o.tau % 3 + 1
            o.taum1 = o.taum1 % 3 + 1
            o.taum2 = o.taum2 % 3 + 1
            if accumulate is not None:
                for a, q in zip(accumulate, ("u", "v", "w", "b")):
                    a += o.f[q]

    def _time_ave(self):
        # :212-235
        acc = [np.zeros(self.orc.shape, order="F") for _ in range(4)]
        self._steps(self.n_end_ave, abs(self.dt_loc), lambda s: 0.0, accumulate=acc)
        return [a / self.n_end_ave for a in acc]

    def _set_state(self, fields):
        for q, a in zip(("u", "v", "w", "b"), fields):
            self.orc.f[q][...] = a

    def _state(self):
        return [self.orc.f[q].copy() for q in ("u", "v", "w", "b")]

    def _geo(self, fields, g0_base=None, keep=False):
        M = self.bal.modes
        S = self.bal.forward(fields)
        q, pv = M.vec_q(0.0), M.vec_p(0.0)
        g0 = self.bal._dot(pv, S)
        if g0_base is None:
            out = self.bal.backward([g0 * qq for qq in q])
        else:
            out = self.bal.backward([S[c] + (g0_base - g0) * q[c] for c in range(4)])
        return (out, g0) if keep else out

    def _norm(self, cur, prev):
        # :342-363
        p = self.prm
        I = (slice(2, 2 + p.nx), slice(2, 2 + p.ny), slice(2, 2 + p.nz))
        u, v, w, b = (a[I] for a in cur)
        ub, vb, wb, bb = (a[I] for a in prev)
        N02 = p.N0 * p.N0
        e1 = np.sum(u ** 2 + v ** 2 + p.dsqr * w ** 2 + b ** 2 / N02)
        e2 = np.sum(ub ** 2 + vb ** 2 + p.dsqr * wb ** 2 + bb ** 2 / N02)
        d = np.sum((u - ub) ** 2 + (v - vb) ** 2 + p.dsqr * (w - wb) ** 2 + (b - bb) ** 2 / N02)
        return 2.0 * np.sqrt(d) / (np.sqrt(e1) + np.sqrt(e2))

    def run(self):
        o, p = self.orc, self.prm
        names = ("u", "v", "w", "b", "p", "du", "dv", "dw", "db")
        save = {n: o.f[n].copy() for n in names}
        st = (o.itt, o.tau, o.taum1, o.taum2)
        self.dt_loc, Ro_loc = p.dt, p.Ro
        u_bal = self._state()
        self.n_end = int(self.period / self.dt_loc) + 1
        self.n_end_ave = int(self.average / self.dt_loc) + 1
        if self.time_average:
            for _ in range(self.average_times):
                base = self._time_ave()
                self._set_state(base)
            self._set_state(u_bal)
        else:
            _, g0_base = self._geo(self._state(), keep=True)
        norms = []
        for n in range(1, self.max_its + 1):
            # :190-198 backward, friction off, Ro ramped to 0
            self._steps(self.n_end, -abs(self.dt_loc),
                        lambda s: (1.0 - ramp(s * self.dt_loc / self.period)) * Ro_loc, bih=False)
            if self.time_average:
                for _ in range(self.average_times):
                    self._set_state(self._time_ave())
            else:
                self._set_state(self._geo(self._state()))
            self._steps(self.n_end, abs(self.dt_loc), lambda s: ramp(s * self.dt_loc / self.period) * Ro_loc)
            if self.time_average:
                keep = self._state()
                for _ in range(self.average_times):
                    ave = self._time_ave()
                    self._set_state(ave)
                self._set_state([keep[c] - ave[c] + base[c] for c in range(4)])
            else:
                self._set_state(self._geo(self._state(), g0_base=g0_base))
            cur = self._state()
            nd = self._norm(cur, u_bal)
            norms.append(nd)
            u_bal = cur
            if n > 1 and nd < self.tol:
                break
        for n_ in names:
            o.f[n_][...] = save[n_]
        self._set(self.dt_loc, Ro_loc)
        o.itt, o.tau, o.taum1, o.taum2 = st
        return u_bal, norms
