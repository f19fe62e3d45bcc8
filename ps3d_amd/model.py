"""Host-side mirror of the reference's step interface, over the C-ABI.

``Model`` holds host copies of the ``main_module`` fields in their reference
layout (Fortran order, halo onx=2 included: numpy index = Fortran index + 1)
and drives the device model through libps3d.so:

=====================  ===============================================
Model method           reference routine it stands for
=====================  ===============================================
``integrate()``        integrate.f90:2-115 (empty set_forcing)
``run(n)``             main.f90:139-161 (loop without diagnostics)
``tendency()``         integrate.f90:12-54
``friction()``         integrate.f90:59-78
``ab_update()``        integrate.f90:87-111
``pressure_solver()``  pressure_solver.f90:2-120
``buoyancy_advection`` advection.f90:2-48
``momentum_advection`` advection.f90:53-184
``harmonic`` ...       mixing.f90:225-274, :96-158, :164-217
``border_exchg``       parallel_mpi.f90:70-81
``my_fft/my_ifft``     my_fft.f90:159-205
``spectral_solve``     pressure_solver.f90:30-71
``set_params``         assignments to main_module's dt, Ro, friction
                       between steps (diag_opt_balance.f90:190-205)
=====================  ===============================================

``FftPlan`` stands for P3DFFT's p3dfft_ftran_r2c / p3dfft_btran_c2r as
my_fft.f90:159-252 calls them (unnormalised, one GPU).

Device state is authoritative between ``push()`` and ``pull()``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check
from .params import PS3D_MIXED, CParams, ModelParams

FIELDS = {"u": _lib.PS3D_U, "v": _lib.PS3D_V, "w": _lib.PS3D_W, "b": _lib.PS3D_B,
          "p": _lib.PS3D_P, "du": _lib.PS3D_DU, "dv": _lib.PS3D_DV, "dw": _lib.PS3D_DW,
          "db": _lib.PS3D_DB}
PROGNOSTIC = ("u", "v", "w", "b")


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Model:
    """One ps3D model instance on the GPU (``cuda:0`` / HIP device 0)."""

    def __init__(self, prm: ModelParams, ksqr: np.ndarray | None = None, stream: int | None = None):
        self.lib = _lib.load()
        self.prm = prm
        self._cprm = CParams.from_params(prm)
        h = ctypes.c_void_p()
        kp = None
        if ksqr is not None:
            self._ksqr = np.asfortranarray(ksqr, dtype=np.float64)
            kp = _dp(self._ksqr)
        check(self.lib.ps3d_create(ctypes.byref(self._cprm), kp,
                                   ctypes.c_void_p(stream) if stream else None, ctypes.byref(h)))
        self.h = h
        shp = prm.shape
        self.f = {n: np.zeros(shp, order="F") for n in ("u", "v", "w", "b", "p")}
        for n in ("du", "dv", "dw", "db"):
            self.f[n] = np.zeros(shp + (3,), order="F")
        # driver time-step state, main.f90:112-113
        self.itt, self.tau, self.taum1, self.taum2 = 0, 3, 2, 1

    # ---- lifetime ------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            check(self.lib.ps3d_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __getattr__(self, name):
        f = self.__dict__.get("f")
        if f is not None and name in f:
            return f[name]
        raise AttributeError(name)

    # ---- host <-> device -----------------------------------------------
    def _default_names(self):
        # a mixed-precision model exchanges only the prognostic fields and p
        if self.prm.precision == PS3D_MIXED:
            return ("u", "v", "w", "b", "p")
        return tuple(FIELDS)

    def push(self, names=None):
        for n in names or self._default_names():
            check(self.lib.ps3d_set_field(self.h, FIELDS[n], _dp(self.f[n])))

    def pull(self, names=None):
        for n in names or self._default_names():
            check(self.lib.ps3d_get_field(self.h, FIELDS[n], _dp(self.f[n])))

    def get_spectral(self) -> np.ndarray:
        prm = self.prm
        buf = np.zeros((2, prm.nx // 2 + 1, prm.ny, prm.nz), order="F")
        check(self.lib.ps3d_get_field(self.h, _lib.PS3D_FS, _dp(buf)))
        return buf[0] + 1j * buf[1]

    def get_ksqr(self) -> np.ndarray:
        n = ctypes.c_int64()
        check(self.lib.ps3d_field_size(self.h, _lib.PS3D_KSQR, ctypes.byref(n)))
        buf = np.zeros(n.value)
        check(self.lib.ps3d_get_field(self.h, _lib.PS3D_KSQR, _dp(buf)))
        return buf

    def synchronize(self):
        check(self.lib.ps3d_synchronize(self.h))

    # ---- the step --------------------------------------------------------
    def rotate(self):
        """main.f90:156-160"""
        self.tau = self.tau % 3 + 1
        self.taum1 = self.taum1 % 3 + 1
        self.taum2 = self.taum2 % 3 + 1
        self.itt += 1

    def integrate(self):
        check(self.lib.ps3d_integrate(self.h, self.itt, self.tau, self.taum1, self.taum2))

    def set_params(self, **changes):
        """dt, Ro, eps_ab, f0, N0, friction coefficients or the nonlinear / AB3
        switches of the next steps"""
        prm = self.prm.replace(**changes)
        c = CParams.from_params(prm)
        check(self.lib.ps3d_set_params(self.h, ctypes.byref(c)))
        self.prm, self._cprm = prm, c

    def run(self, nsteps: int, use_graph: bool = False):
        st = (ctypes.c_int32 * 4)(self.itt, self.tau, self.taum1, self.taum2)
        check(self.lib.ps3d_run(self.h, nsteps, st, int(use_graph)))
        self.itt, self.tau, self.taum1, self.taum2 = list(st)

    def tendency(self):
        check(self.lib.ps3d_tendency(self.h, self.tau))

    def buoyancy_advection(self):
        check(self.lib.ps3d_buoyancy_advection(self.h, self.tau))

    def momentum_advection(self):
        check(self.lib.ps3d_momentum_advection(self.h, self.tau))

    def friction(self):
        check(self.lib.ps3d_friction(self.h))

    def harmonic(self, name, hmix, vmix):
        check(self.lib.ps3d_mixing(self.h, _lib.PS3D_HARMONIC, FIELDS[name], hmix, vmix))

    def biharmonic_h(self, name, hmix):
        check(self.lib.ps3d_mixing(self.h, _lib.PS3D_BIHARMONIC_H, FIELDS[name], hmix, 0.0))

    def biharmonic_v(self, name, vmix):
        check(self.lib.ps3d_mixing(self.h, _lib.PS3D_BIHARMONIC_V, FIELDS[name], 0.0, vmix))

    def ab_update(self):
        check(self.lib.ps3d_ab_update(self.h, self.itt, self.tau, self.taum1, self.taum2))

    def border_exchg(self, name):
        check(self.lib.ps3d_border_exchg(self.h, FIELDS[name]))

    def pressure_solver(self):
        check(self.lib.ps3d_pressure_solver(self.h))

    # ---- building blocks ---------------------------------------------------
    def my_fft(self, qp: np.ndarray, three_d: bool) -> np.ndarray:
        prm = self.prm
        qp = np.asfortranarray(qp, dtype=np.float64)
        out = np.zeros((2, prm.nx // 2 + 1, prm.ny, prm.nz), order="F")
        check(self.lib.ps3d_my_fft(self.h, int(three_d), _dp(qp), _dp(out)))
        return out[0] + 1j * out[1]

    def my_ifft(self, qs: np.ndarray, three_d: bool) -> np.ndarray:
        prm = self.prm
        buf = np.zeros((2, prm.nx // 2 + 1, prm.ny, prm.nz), order="F")
        buf[0], buf[1] = qs.real, qs.imag
        out = np.zeros((prm.nx, prm.ny, prm.nz), order="F")
        check(self.lib.ps3d_my_ifft(self.h, int(three_d), _dp(buf), _dp(out)))
        return out

    def spectral_solve(self, qs: np.ndarray) -> np.ndarray:
        prm = self.prm
        buf = np.zeros((2, prm.nx // 2 + 1, prm.ny, prm.nz), order="F")
        buf[0], buf[1] = qs.real, qs.imag
        check(self.lib.ps3d_spectral_solve(self.h, _dp(buf)))
        return buf[0] + 1j * buf[1]

    # ---- measurement ---------------------------------------------------------
    def diag_balance(self, filter_width: int | None = None, kx=None, ky=None, kz=None, outputs=None) -> dict:
        """Normal-mode balancing of the model's current state
        (diag_balance.f90:89-362): {'u0': .., 'v0': .., ..., 'b3': ..}, full
        arrays with halo (outputs: the names wanted, default all 16).  The
        model's own fields are left unchanged."""
        shp = self.prm.shape
        names = [f"{c}{o}" for o in range(4) for c in ("u", "v", "w", "b")]
        want = set(names if outputs is None else outputs)
        out = {n: np.zeros(shp, order="F") for n in names if n in want}
        tab = (ctypes.POINTER(ctypes.c_double) * 16)(*[_dp(out[n]) if n in out else None for n in names])
        ks = []
        for k in (kx, ky, kz):
            if k is None:
                ks.append(None)
            else:
                a = np.ascontiguousarray(k, dtype=np.float64)
                ks.append(a)
        self._ks = ks
        ptr = [None if k is None else _dp(k) for k in ks]
        check(self.lib.ps3d_diag_balance(self.h, -1 if filter_width is None else int(filter_width),
                                         ptr[0], ptr[1], ptr[2], tab))
        return out

    def opt_balance(self, period: float, max_iterations: int, tol: float, time_average: bool = False,
                    average: float = 1.0, average_times: int = 1, kx=None, ky=None, kz=None) -> dict:
        """Optimal balance of the model's current state
        (diag_opt_balance.f90:55-179): {'u': u_bal, 'v': .., 'w': .., 'b': ..,
        'norms': [norm_diff per iteration], 'iterations': n}.  The model's own
        state is left unchanged."""
        shp = self.prm.shape
        out = {n: np.zeros(shp, order="F") for n in ("u", "v", "w", "b")}
        tab = (ctypes.POINTER(ctypes.c_double) * 4)(*[_dp(out[n]) for n in ("u", "v", "w", "b")])
        norms = np.zeros(max(1, int(max_iterations)))
        its = ctypes.c_int32()
        ob = _lib.OptBalanceParams(float(period), float(average), float(tol), int(max_iterations),
                                   int(average_times), int(bool(time_average)), 0)
        ks = [None if k is None else np.ascontiguousarray(k, dtype=np.float64) for k in (kx, ky, kz)]
        ptr = [None if k is None else _dp(k) for k in ks]
        check(self.lib.ps3d_opt_balance(self.h, ctypes.byref(ob), ptr[0], ptr[1], ptr[2], tab, _dp(norms),
                                        ctypes.byref(its)))
        out["iterations"] = its.value
        out["norms"] = list(norms[:its.value])
        return out

    def tune(self, cache_path: str | None = None, reps: int = 5):
        """Autotune the launch shapes (ps3d_tune); the next step's slot is scratch."""
        check(self.lib.ps3d_tune(self.h, self.tau, cache_path.encode() if cache_path else None, reps))

    def tuning(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        n = self.lib.ps3d_tuning(self.h, buf, len(buf))
        if n < 0:
            check(-n)
        return buf.value.decode()

    # --- forcing (integrate.f90:82) -----------------------------------------
    def set_forcing(self, fn):
        """Install fn(model, itt, tau, stream) -> None as the set_forcing hook
        (ps3d_set_forcing); None removes it.  fn runs between the friction
        block and the Adams-Bashforth update of every step; it may launch work
        on `stream` on device_ptr() arrays, or synchronize and use pull/push
        on u, v, w, b and the tau slot of du..db (the host round trip)."""
        if fn is None:
            self._forcing_cb = None
            check(self.lib.ps3d_set_forcing(self.h, None, None))
            return

        def tramp(_m, _user, itt, tau, stream):
            try:
                fn(self, itt, tau, stream)
                return 0
            except Exception as e:   # reported as the step's error
                self._forcing_error = e
                return 1

        self._forcing_cb = _lib.FORCING_FN(tramp)
        check(self.lib.ps3d_set_forcing(self.h, ctypes.cast(self._forcing_cb, ctypes.c_void_p), None))

    def set_fused_tendency(self, on: bool):
        """Form the tendencies inside the friction + AB kernel (True) or in a
        kernel of their own (False, the default): ps3d_set_fused_tendency."""
        check(self.lib.ps3d_set_fused_tendency(self.h, 1 if on else 0))

    @property
    def fft_plan(self) -> tuple:
        """Radix plans (x, y, z) of the FP32 transforms (ps3d_get_fft_plan):
        0 radix 4 first, 1 radix 8 first."""
        a = (ctypes.c_int32 * 3)()
        check(self.lib.ps3d_get_fft_plan(self.h, a))
        return tuple(a)

    @fft_plan.setter
    def fft_plan(self, plans):
        check(self.lib.ps3d_set_fft_plan(self.h, (ctypes.c_int32 * 3)(*plans)))

    def device_ptr(self, name: str, slot: int | None = None) -> int:
        """Device address of a field array (ps3d_device_ptr)."""
        out = ctypes.c_void_p()
        check(self.lib.ps3d_device_ptr(self.h, FIELDS[name], 0 if slot is None else slot, ctypes.byref(out)))
        return out.value

    def set_wavemaker(self, omega: float, re: dict, im: dict):
        """Device wavemaker forcing (ps3d_set_wavemaker); re/im map u, v, w, b
        to arrays of the field layout (missing names are not forced)."""
        keep = []
        rp = (_lib._dp * 4)()
        ip = (_lib._dp * 4)()
        for q, n in enumerate(("u", "v", "w", "b")):
            if n in re:
                a = np.asfortranarray(re[n], dtype=np.float64)
                b = np.asfortranarray(im[n], dtype=np.float64)
                assert a.shape == self.f[n].shape and b.shape == self.f[n].shape, n
                keep += [a, b]
                rp[q] = _dp(a)
                ip[q] = _dp(b)
        check(self.lib.ps3d_set_wavemaker(self.h, float(omega), rp, ip))

    def clear_wavemaker(self):
        check(self.lib.ps3d_clear_wavemaker(self.h))

    def set_background_flow(self, u_back=None, b_back=None):
        """Advection by the background flow (u_back, 0, 0) and of b_back,
        lee_wave1.f90:239-314 (ps3d_set_background_flow); u_back None: off."""
        ub = None if u_back is None else np.asfortranarray(u_back, dtype=np.float64)
        bb = None if b_back is None else np.asfortranarray(b_back, dtype=np.float64)
        check(self.lib.ps3d_set_background_flow(self.h, None if ub is None else _dp(ub),
                                                None if bb is None else _dp(bb)))

    def set_body_force(self, force: dict | None):
        """Steady body force added to the tendencies ({'du': .., 'dw': ..};
        ps3d_set_body_force); None: off."""
        force = force or {}
        arrs = [None if force.get(n) is None else np.asfortranarray(force[n], dtype=np.float64)
                for n in ("du", "dv", "dw", "db")]
        tab = (ctypes.POINTER(ctypes.c_double) * 4)(*[None if a is None else _dp(a) for a in arrs])
        check(self.lib.ps3d_set_body_force(self.h, tab))

    def diagnose(self) -> dict:
        """diagnose.F90:11-53 on the device: KE, PE (per point), max |u_h|,
        max |w|, CFL and Peclet numbers, and whether a NaN was seen."""
        out = np.zeros(8)
        nan = ctypes.c_int32()
        check(self.lib.ps3d_diagnose(self.h, _dp(out), ctypes.byref(nan)))
        names = ("ke", "pe", "umax", "wmax", "cfl_h", "cfl_v", "pecl_h", "pecl_v")
        d = dict(zip(names, out.tolist()))
        d["values"] = out.tolist()
        d["nan"] = bool(nan.value)
        return d

    def write_restart(self, path: str, time: float = 0.0):
        """restart.f90:2-22 of the model's state and the driver's counters"""
        check(self.lib.ps3d_write_restart(self.h, path.encode(), self.itt, float(time), self.tau, self.taum1,
                                          self.taum2))

    def read_restart(self, path: str) -> dict:
        """restart.f90:27-82; sets the driver's counters from the file"""
        v = [ctypes.c_int32() for _ in range(4)]
        tm = ctypes.c_double()
        check(self.lib.ps3d_read_restart(self.h, path.encode(), ctypes.byref(v[0]), ctypes.byref(tm),
                                         ctypes.byref(v[1]), ctypes.byref(v[2]), ctypes.byref(v[3])))
        self.itt, self.tau, self.taum1, self.taum2 = (x.value for x in v)
        return {"itt": self.itt, "time": tm.value, "tau": self.tau, "taum1": self.taum1, "taum2": self.taum2}

    def profile_kernel(self, name: str | None):
        check(self.lib.ps3d_profile_kernel(self.h, name.encode() if name else None))

    def profile_read(self):
        n = ctypes.c_int64()
        ms = ctypes.c_double()
        check(self.lib.ps3d_profile_read(self.h, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value


def kernel_names() -> list[str]:
    return _lib.load().ps3d_kernel_names().decode().split(",")


class FftPlan:
    """The device transforms behind my_fft.f90 (P3DFFT's r2c / c2r on one GPU):
    forward(qp, op) with op "ffn" (x, y) or "fft" (x, y, z), backward(qs, op)
    with op "nff" or "fft"; unnormalised."""

    def __init__(self, nx: int, ny: int, nz: int):
        self.lib = _lib.load()
        self.shape = (nx, ny, nz)
        h = ctypes.c_void_p()
        check(self.lib.ps3d_fft_create(nx, ny, nz, None, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            check(self.lib.ps3d_fft_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, qp: np.ndarray, op: str = "fft") -> np.ndarray:
        nx, ny, nz = self.shape
        qp = np.asfortranarray(qp, dtype=np.float64)
        out = np.zeros((2, nx // 2 + 1, ny, nz), order="F")
        check(self.lib.ps3d_fft_forward(self.h, op.encode(), _dp(qp), _dp(out)))
        return out[0] + 1j * out[1]

    def backward(self, qs: np.ndarray, op: str = "fft") -> np.ndarray:
        nx, ny, nz = self.shape
        buf = np.zeros((2, nx // 2 + 1, ny, nz), order="F")
        buf[0], buf[1] = qs.real, qs.imag
        out = np.zeros((nx, ny, nz), order="F")
        check(self.lib.ps3d_fft_backward(self.h, op.encode(), _dp(buf), _dp(out)))
        return out
