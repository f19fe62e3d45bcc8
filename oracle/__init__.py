"""CPU oracle for the ps3D time step -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
may import this package, and only as the checker (or, for cpu_baseline, as the
timed CPU port).  The product path (``ps3d_amd``) never imports it.

Two libraries are wrapped here:

* ``libps3d_oracle.so`` -- the plain-C restatement of the reference step
  (``ps3d_oracle.c``), always built by ``make -C oracle``.
* ``_ref/libps3d_ref.so`` -- the reference's own Fortran (main_module.f90,
  advection.f90, mixing.f90, integrate.f90, parallel_mpi.f90,
  eigenvectors.f90) compiled unchanged by ``make -C oracle ref`` where the
  reference sources exist, used only to pin the restatement.

Parity pinning of each oracle function (reference file:line it follows, and
what checks it):

======================  ==============================  =====================================
oracle function         follows                         pinned by
======================  ==============================  =====================================
buoyancy_advection_2nd  advection.f90:2-48              reference build, bit for bit
momentum_advection_2nd  advection.f90:53-184            reference build, bit for bit
harmonic                mixing.f90:225-274              reference build, bit for bit
biharmonic_h            mixing.f90:96-158               reference build, bit for bit
biharmonic_v            mixing.f90:164-217              reference build, bit for bit
border_exchg_3d         parallel_mpi.f90:70-211         reference build, bit for bit
integrate_pre           integrate.f90:2-111             reference build (its own integrate),
                                                        bit for bit, several steps, all AB orders
hat_ksqr / wavenumbers  eigenvectors.f90:224-229,       reference build's hat_ksqr; closed form
                        my_fft.f90:87-156
my_fft_2d/3d, ifft      my_fft.f90:159-205              numpy.fft (an independent FFT) and the
                                                        DFT definition; round trip; Parseval
solve_tridiag           pressure_solver.f90:126-162     dense solve (numpy.linalg) on random
                                                        systems, incl. the pressure operator
spectral_solve          pressure_solver.f90:30-71       applying the discrete vertical operator
                                                        to the solution returns the RHS; zero
                                                        mode = double cumulative sum
pressure_solver         pressure_solver.f90:2-120       discrete divergence after projection is
                                                        round-off; p solves the discrete Poisson
                                                        problem; projection is idempotent
integrate / run         integrate.f90, main.f90:139-161 integrate_pre (above) + pressure_solver
restart records         restart.f90:2-82                the reference's own write_restart /
                                                        read_restart from the same build: byte-
                                                        identical files, both directions
======================  ==============================  =====================================
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libps3d_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libps3d_ref.so")
ONX = 2

# field ids (ps3d_oracle.h)
U, V, W, B, P, DU, DV, DW, DB, FLUX_EAST, FLUX_NORTH, FLUX_TOP, FS, KSQR, KX, KY, KZ = range(17)
FIELD_NAMES = ["u", "v", "w", "b", "p", "du", "dv", "dw", "db",
               "flux_east", "flux_north", "flux_top", "FS", "ksqr", "kx", "ky", "kz"]


class _OrcParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("enable_nonlinear", ctypes.c_int32),
                ("enable_vertical_boundaries", ctypes.c_int32),
                ("enable_AB_3_order", ctypes.c_int32)] + [
        (n, ctypes.c_double) for n in
        ("Lx", "Ly", "Lz", "dx", "dy", "dz", "dt", "f0", "N0", "Ro", "dsqr", "eps_ab",
         "Ah", "Av", "Kh", "Kv", "Ahbi", "Avbi", "Khbi", "Kvbi")]


REAL_PARAMS = ("Lx", "Ly", "Lz", "dx", "dy", "dz", "dt", "f0", "N0", "Ro", "dsqr", "eps_ab",
               "Ah", "Av", "Kh", "Kv", "Ahbi", "Avbi", "Khbi", "Kvbi")
INT_PARAMS = ("nx", "ny", "nz", "enable_nonlinear", "enable_vertical_boundaries",
              "enable_AB_3_order")


def _cparams(prm) -> _OrcParams:
    c = _OrcParams()
    for n in INT_PARAMS:
        setattr(c, n, int(getattr(prm, n)))
    for n in REAL_PARAMS:
        setattr(c, n, float(getattr(prm, n)))
    return c


_orc_lib = None
_ref_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def _as_dp(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_dp)


def load_oracle():
    """Load the C restatement; raises if it has not been built."""
    global _orc_lib
    if _orc_lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        lib = ctypes.CDLL(ORACLE_SO)
        vp = ctypes.c_void_p
        lib.orc_create.restype = vp
        lib.orc_create.argtypes = [ctypes.POINTER(_OrcParams)]
        lib.orc_destroy.argtypes = [vp]
        lib.orc_ptr.restype = _dp
        lib.orc_ptr.argtypes = [vp, ctypes.c_int]
        lib.orc_field_len.restype = ctypes.c_int64
        lib.orc_field_len.argtypes = [vp, ctypes.c_int]
        for n in ("orc_buoyancy_advection_2nd",):
            getattr(lib, n).argtypes = [vp, _dp]
        lib.orc_momentum_advection_2nd.argtypes = [vp, _dp, _dp, _dp]
        lib.orc_harmonic.argtypes = [vp, _dp, ctypes.c_double, ctypes.c_double]
        lib.orc_biharmonic_h.argtypes = [vp, _dp, ctypes.c_double]
        lib.orc_biharmonic_v.argtypes = [vp, _dp, ctypes.c_double]
        lib.orc_border_exchg_3d.argtypes = [vp, _dp]
        for n in ("orc_integrate_pre", "orc_integrate"):
            getattr(lib, n).argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.orc_pressure_solver.argtypes = [vp]
        lib.orc_wavemaker.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_dp), ctypes.POINTER(_dp)]
        lib.orc_set_wavemaker.argtypes = [vp, ctypes.c_double, ctypes.POINTER(_dp), ctypes.POINTER(_dp)]
        lib.orc_set_forcing.argtypes = [vp, vp, vp]
        lib.orc_set_dynamic.argtypes = [vp] + [ctypes.c_double] * 6
        lib.orc_background_flow.argtypes = [vp, ctypes.c_int, _dp, _dp]
        lib.orc_body_force.argtypes = [vp, ctypes.c_int, ctypes.POINTER(_dp)]
        lib.orc_set_background.argtypes = [vp, _dp, _dp, ctypes.POINTER(_dp)]
        lib.orc_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
        for n in ("orc_my_fft_2d", "orc_my_ifft_2d", "orc_my_fft_3d", "orc_my_ifft_3d"):
            getattr(lib, n).argtypes = [vp, _dp, _dp]
        lib.orc_spectral_solve.argtypes = [vp, _dp]
        lib.orc_solve_tridiag.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int]
        lib.orc_cfft.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        lib.orc_hat_ksqr.restype = ctypes.c_double
        lib.orc_hat_ksqr.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.orc_wavenumbers.argtypes = [vp]
        _orc_lib = lib
    return _orc_lib


def reference_available() -> bool:
    return os.path.exists(REF_SO)


def load_reference():
    """Load the reference-Fortran build; returns None if it was not built."""
    global _ref_lib
    if _ref_lib is None:
        if not os.path.exists(REF_SO):
            return None
        lib = ctypes.CDLL(REF_SO)
        vp = ctypes.c_void_p
        lib.ref_setup.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
        lib.ref_ptr.restype = vp
        lib.ref_ptr.argtypes = [ctypes.c_int]
        lib.ref_set_time.argtypes = [ctypes.c_int] * 4
        lib.ref_mixing.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]
        lib.ref_border_exchg.argtypes = [ctypes.c_int]
        lib.ref_hat_ksqr.restype = ctypes.c_double
        lib.ref_hat_ksqr.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.ref_ab3.restype = ctypes.c_double
        lib.ref_ab3.argtypes = [ctypes.c_int]
        if hasattr(lib, "ref_set_clock"):
            lib.ref_set_clock.argtypes = [ctypes.c_double]
        _ref_lib = lib
    return _ref_lib


def _view(ptr, shape):
    n = int(np.prod(shape))
    buf = (ctypes.c_double * n).from_address(ctypes.cast(ptr, ctypes.c_void_p).value)
    return np.ndarray(shape, dtype=np.float64, buffer=buf, order="F")


class OracleModel:
    """One instance of the C restatement; fields are numpy views (Fortran
    order, halo included: numpy index = Fortran index + 1)."""

    def __init__(self, prm):
        self.lib = load_oracle()
        self.prm = prm
        self._c = _cparams(prm)
        self.h = self.lib.orc_create(ctypes.byref(self._c))
        nx, ny, nz = prm.nx, prm.ny, prm.nz
        self.shape = (nx + 2 * ONX, ny + 2 * ONX, nz + 2 * ONX)
        self.nkx = nx // 2 + 1
        shapes = {U: self.shape, V: self.shape, W: self.shape, B: self.shape, P: self.shape,
                  DU: self.shape + (3,), DV: self.shape + (3,), DW: self.shape + (3,),
                  DB: self.shape + (3,), FLUX_EAST: self.shape, FLUX_NORTH: self.shape,
                  FLUX_TOP: self.shape, FS: (2, self.nkx, ny, nz),
                  KSQR: (self.nkx, ny, 1 if prm.enable_vertical_boundaries else nz),
                  KX: (self.nkx,), KY: (ny,), KZ: (nz,)}
        self.f = {}
        for fid, shp in shapes.items():
            self.f[FIELD_NAMES[fid]] = _view(self.lib.orc_ptr(self.h, fid), shp)
        self.itt, self.tau, self.taum1, self.taum2 = 0, 3, 2, 1

    def __getattr__(self, name):
        f = self.__dict__.get("f")
        if f is not None and name in f:
            return f[name]
        raise AttributeError(name)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def FS_complex(self):
        a = self.f["FS"]
        return a[0] + 1j * a[1]

    def slot(self, name, t):
        return self.f[name][..., t - 1]

    # --- routines -------------------------------------------------------
    def buoyancy_advection(self, tau=None):
        tau = tau or self.tau
        self.lib.orc_buoyancy_advection_2nd(self.h, _as_dp(self.f["db"][..., tau - 1]))

    def momentum_advection(self, tau=None):
        tau = tau or self.tau
        du, dv, dw = (self.f[n][..., tau - 1] for n in ("du", "dv", "dw"))
        # slot views of an F-ordered 4-D array are F-contiguous
        self.lib.orc_momentum_advection_2nd(self.h, _as_dp(du), _as_dp(dv), _as_dp(dw))

    def harmonic(self, name, H, Vv):
        self.lib.orc_harmonic(self.h, _as_dp(self.f[name]), H, Vv)

    def biharmonic_h(self, name, H):
        self.lib.orc_biharmonic_h(self.h, _as_dp(self.f[name]), H)

    def biharmonic_v(self, name, Vv):
        self.lib.orc_biharmonic_v(self.h, _as_dp(self.f[name]), Vv)

    def border_exchg(self, name):
        self.lib.orc_border_exchg_3d(self.h, _as_dp(self.f[name]))

    def integrate_pre(self):
        self.lib.orc_integrate_pre(self.h, self.itt, self.tau, self.taum1, self.taum2)

    def pressure_solver(self):
        self.lib.orc_pressure_solver(self.h)

    def background_flow(self, u_back, b_back=None, tau=None):
        """lee_wave1.f90:239-314 into the tau slot"""
        ub = np.asfortranarray(u_back, dtype=np.float64)
        bb = None if b_back is None else np.asfortranarray(b_back, dtype=np.float64)
        self.lib.orc_background_flow(self.h, tau or self.tau, _as_dp(ub), None if bb is None else _as_dp(bb))

    def set_background(self, u_back=None, b_back=None, force=None):
        """background flow and body force ({'du': .., 'dw': ..}) as the set_forcing hook"""
        keep = []
        def arr(a):
            if a is None:
                return None
            a = np.asfortranarray(a, dtype=np.float64)
            keep.append(a)
            return _as_dp(a)
        force = force or {}
        tab = (_dp * 4)(*[arr(force.get(n)) for n in ("du", "dv", "dw", "db")])
        self._bg_keep = (keep, tab)
        self.lib.orc_set_background(self.h, arr(u_back), arr(b_back), tab)

    def set_dynamic(self, dt, Ro, Ahbi, Avbi, Khbi, Kvbi):
        """dt, Ro and the biharmonic coefficients of the next steps"""
        self.lib.orc_set_dynamic(self.h, dt, Ro, Ahbi, Avbi, Khbi, Kvbi)

    def _amp_tables(self, re, im):
        self._wm_keep = []
        rp, ip = (_dp * 4)(), (_dp * 4)()
        for q, n in enumerate(("u", "v", "w", "b")):
            if n in re:
                a = np.asfortranarray(re[n], dtype=np.float64)
                b = np.asfortranarray(im[n], dtype=np.float64)
                self._wm_keep += [a, b]
                rp[q], ip[q] = _as_dp(a), _as_dp(b)
        return rp, ip

    def wavemaker(self, omega, re, im, itt=None):
        """for_src/config.f90:224-237 on the oracle's fields (whole arrays)"""
        rp, ip = self._amp_tables(re, im)
        self.lib.orc_wavemaker(self.h, self.itt if itt is None else itt, omega, rp, ip)

    def set_wavemaker(self, omega, re, im):
        """the wavemaker as the set_forcing hook of integrate"""
        rp, ip = self._amp_tables(re, im)
        self._wm_tables = (rp, ip)
        self.lib.orc_set_wavemaker(self.h, omega, rp, ip)

    def integrate(self):
        self.lib.orc_integrate(self.h, self.itt, self.tau, self.taum1, self.taum2)

    def rotate(self):
        """main.f90:156-160"""
        self.tau = self.tau % 3 + 1
        self.taum1 = self.taum1 % 3 + 1
        self.taum2 = self.taum2 % 3 + 1
        self.itt += 1

    def run(self, nsteps):
        st = (ctypes.c_int32 * 4)(self.itt, self.tau, self.taum1, self.taum2)
        self.lib.orc_run(self.h, nsteps, st)
        self.itt, self.tau, self.taum1, self.taum2 = list(st)

    def my_fft(self, qp: np.ndarray, three_d: bool) -> np.ndarray:
        """qp real (nx,ny,nz) -> complex (nx/2+1,ny,nz)"""
        prm = self.prm
        qp = np.asfortranarray(qp, dtype=np.float64)
        out = np.zeros((2, self.nkx, prm.ny, prm.nz), order="F")
        fn = self.lib.orc_my_fft_3d if three_d else self.lib.orc_my_fft_2d
        fn(self.h, _as_dp(qp), _as_dp(out))
        return out[0] + 1j * out[1]

    def my_ifft(self, qs: np.ndarray, three_d: bool) -> np.ndarray:
        prm = self.prm
        buf = np.zeros((2, self.nkx, prm.ny, prm.nz), order="F")
        buf[0] = qs.real
        buf[1] = qs.imag
        out = np.zeros((prm.nx, prm.ny, prm.nz), order="F")
        fn = self.lib.orc_my_ifft_3d if three_d else self.lib.orc_my_ifft_2d
        fn(self.h, _as_dp(buf), _as_dp(out))
        return out

    def spectral_solve(self, qs: np.ndarray) -> np.ndarray:
        prm = self.prm
        buf = np.zeros((2, self.nkx, prm.ny, prm.nz), order="F")
        buf[0] = qs.real
        buf[1] = qs.imag
        self.lib.orc_spectral_solve(self.h, _as_dp(buf))
        return buf[0] + 1j * buf[1]


def solve_tridiag(a, b, c, d):
    """pressure_solver.f90:126-162 via the oracle (complex d)."""
    lib = load_oracle()
    n = len(b)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    dd = np.ascontiguousarray(np.stack([np.real(d), np.imag(d)], axis=-1), dtype=np.float64)
    x = np.zeros((n, 2))
    lib.orc_solve_tridiag(_as_dp(a), _as_dp(b), _as_dp(c), _as_dp(dd), _as_dp(x), n)
    return x[:, 0] + 1j * x[:, 1]


def cfft(x: np.ndarray, sign: int) -> np.ndarray:
    lib = load_oracle()
    buf = np.ascontiguousarray(np.stack([x.real, x.imag], axis=-1), dtype=np.float64)
    lib.orc_cfft(len(x), _as_dp(buf), sign)
    return buf[:, 0] + 1j * buf[:, 1]


class ReferenceModel:
    """The reference's own Fortran (module globals: one instance at a time)."""

    def __init__(self, prm):
        self.lib = load_reference()
        if self.lib is None:
            raise RuntimeError("reference build missing (make -C oracle ref)")
        self.prm = prm
        iprm = (ctypes.c_int * 6)(*[int(getattr(prm, n)) for n in INT_PARAMS])
        rprm = (ctypes.c_double * 20)(*[float(getattr(prm, n)) for n in REAL_PARAMS])
        self.lib.ref_setup(iprm, rprm)
        nx, ny, nz = prm.nx, prm.ny, prm.nz
        self.shape = (nx + 2 * ONX, ny + 2 * ONX, nz + 2 * ONX)
        self.f = {}
        for fid in range(12):
            shp = self.shape + ((3,) if DU <= fid <= DB else ())
            self.f[FIELD_NAMES[fid]] = _view(self.lib.ref_ptr(fid), shp)
        self.itt, self.tau, self.taum1, self.taum2 = 0, 3, 2, 1

    def __getattr__(self, name):
        f = self.__dict__.get("f")
        if f is not None and name in f:
            return f[name]
        raise AttributeError(name)

    def _sync_time(self):
        self.lib.ref_set_time(self.itt, self.tau, self.taum1, self.taum2)

    def buoyancy_advection(self):
        self._sync_time()
        self.lib.ref_buoyancy_advection()

    def momentum_advection(self):
        self._sync_time()
        self.lib.ref_momentum_advection()

    _ids = {"u": 0, "v": 1, "w": 2, "b": 3, "p": 4}

    def harmonic(self, name, H, Vv):
        self.lib.ref_mixing(0, self._ids[name], H, Vv)

    def biharmonic_h(self, name, H):
        self.lib.ref_mixing(1, self._ids[name], H, 0.0)

    def biharmonic_v(self, name, Vv):
        self.lib.ref_mixing(2, self._ids[name], 0.0, Vv)

    def border_exchg(self, name):
        self.lib.ref_border_exchg(self._ids[name])

    def integrate_pre(self):
        self._sync_time()
        self.lib.ref_integrate()

    def rotate(self):
        self.tau = self.tau % 3 + 1
        self.taum1 = self.taum1 % 3 + 1
        self.taum2 = self.taum2 % 3 + 1
        self.itt += 1

    def hat_ksqr(self, k, d):
        return self.lib.ref_hat_ksqr(k, d)
