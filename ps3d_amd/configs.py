"""Synthetic scenario configs written against the reference's plug-in surface
(``set_parameter`` + ``set_initial_conditions``, for_config/*.f90).

Each builder returns ``(ModelParams, {"u","v","w","b": host arrays})`` with the
main_module layout (halo included, already exchanged, lid applied as
main.f90:119-122 does after set_initial_conditions).

Noise is drawn from SplitMix64 keyed on (seed, field, i, j, k): integer-only
and independent of evaluation order, so any implementation reproduces the
same initial state bit for bit.
"""
from __future__ import annotations

import math

import numpy as np

from .params import PS3D_FP64, ModelParams

FIELD_ID = {"u": 1, "v": 2, "w": 3, "b": 4, "R": 5}   # R: per-member scalars
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser over uint64 counters (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def _counters(seed: int, field: str, nx: int, ny: int, nz: int, stream: int = 0) -> np.ndarray:
    i = np.arange(nx, dtype=np.uint64)[:, None, None]
    j = np.arange(ny, dtype=np.uint64)[None, :, None]
    k = np.arange(nz, dtype=np.uint64)[None, None, :]
    lin = i + np.uint64(nx) * (j + np.uint64(ny) * k)
    key = (np.uint64(seed) * np.uint64(64) + np.uint64(FIELD_ID[field] * 4 + stream)) << np.uint64(40)
    return lin + key


def uniform(seed: int, field: str, nx: int, ny: int, nz: int, stream: int = 0) -> np.ndarray:
    """U(-1, 1) on the interior, shape (nx, ny, nz)."""
    z = splitmix64(_counters(seed, field, nx, ny, nz, stream))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -52) - 1.0


def normal(seed: int, field: str, nx: int, ny: int, nz: int) -> np.ndarray:
    """N(0, 1) by Box-Muller from two uniform streams."""
    u1 = (splitmix64(_counters(seed, field, nx, ny, nz, 1)) >> np.uint64(11)).astype(np.float64)
    u2 = (splitmix64(_counters(seed, field, nx, ny, nz, 2)) >> np.uint64(11)).astype(np.float64)
    u1 = (u1 + 1.0) * 2.0 ** -53          # (0, 1]
    u2 = u2 * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def border_exchg(a: np.ndarray) -> None:
    """parallel_mpi.f90:70-211 on one rank (z, then y, then x), in place."""
    nx, ny, nz = a.shape[0] - 4, a.shape[1] - 4, a.shape[2] - 4
    for k in (1, 2):
        a[:, :, 1 + nz + k] = a[:, :, 1 + k]
        a[:, :, 1 + 1 - k] = a[:, :, 1 + nz - k + 1]
    for j in (1, 2):
        a[:, 1 + ny + j, :] = a[:, 1 + j, :]
        a[:, 1 + 1 - j, :] = a[:, 1 + ny - j + 1, :]
    for i in (1, 2):
        a[1 + nx + i, :, :] = a[1 + i, :, :]
        a[1 + 1 - i, :, :] = a[1 + nx - i + 1, :, :]


def _finish(prm: ModelParams, interior: dict) -> dict:
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    out = {}
    for n in ("u", "v", "w", "b"):
        a = np.zeros(prm.shape, order="F")
        a[2:2 + nx, 2:2 + ny, 2:2 + nz] = interior[n]
        border_exchg(a)
        out[n] = a
    if prm.enable_vertical_boundaries:       # main.f90:119-122
        out["w"][:, :, :2] = 0.0
        out["w"][:, :, 1 + nz:] = 0.0
    return out


def eddy(nx=64, ny=64, nz=32, seed=1, radius_frac=1.0 / 6.0, noise=1e-3, cfl=0.3,
         precision=PS3D_FP64, biharmonic=True):
    """Baroclinic eddy on an f-plane with constant N^2 and a rigid lid.

    Scales follow for_config/eddy1.f90:15-40 (f0 = 1e-4, N0 = 30 f0, H = 2000 m,
    Lx = Ly = 20 N0 H / (pi f0), A_bi = dx^4 / 30 days).  Buoyancy anomaly
    0.1 N0^2 H exp(-r^2/R^2) exp(-depth/H) with R = radius_frac*Lx, velocities
    in thermal-wind balance (zero at the bottom), plus noise*U(-1,1) on u, v, b.
    dt = cfl*dx/max|u|, capped at 0.5/N0 for the explicit AB3 gravity waves.
    """
    f0 = 1e-4
    N0 = 30 * f0
    H = 2000.0
    Lz = H
    L = 20 * N0 * H / (math.pi * f0)
    dx, dy, dz = L / nx, L / ny, Lz / nz
    A = 0.1 * N0 ** 2 * H
    R = radius_frac * L
    xc = yc = 0.5 * L

    def G(x, y):
        return np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / R ** 2)

    xi = (np.arange(1, nx + 1) - 0.5) * dx      # cell centres
    yj = (np.arange(1, ny + 1) - 0.5) * dy
    zk = (np.arange(1, nz + 1) - 0.5) * dz      # height above the bottom
    E = np.exp(-(Lz - zk) / H)
    E0 = math.exp(-Lz / H)
    X, Y = np.meshgrid(xi, yj, indexing="ij")
    b = A * G(X, Y)[:, :, None] * E[None, None, :]
    # u at (i+1/2, j), v at (i, j+1/2): f u_z = -b_y, f v_z = b_x, zero at z = 0
    Xu, Yu = np.meshgrid(xi + 0.5 * dx, yj, indexing="ij")
    Xv, Yv = np.meshgrid(xi, yj + 0.5 * dy, indexing="ij")
    gy = -2 * (Yu - yc) / R ** 2 * G(Xu, Yu)
    gx = -2 * (Xv - xc) / R ** 2 * G(Xv, Yv)
    u = -(A / f0) * gy[:, :, None] * (H * (E - E0))[None, None, :]
    v = (A / f0) * gx[:, :, None] * (H * (E - E0))[None, None, :]
    u = u + noise * uniform(seed, "u", nx, ny, nz)
    v = v + noise * uniform(seed, "v", nx, ny, nz)
    b = b + noise * uniform(seed, "b", nx, ny, nz)
    umax = float(np.sqrt(u ** 2 + v ** 2).max())
    dt = min(cfl * dx / umax, 0.5 / N0)
    bih = dx ** 4 / (30 * 86400.0) if biharmonic else 0.0
    prm = ModelParams(nx=nx, ny=ny, nz=nz, Lx=L, Ly=L, Lz=Lz, dx=dx, dy=dy, dz=dz, dt=dt,
                      f0=f0, N0=N0, Ahbi=bih, Khbi=bih, enable_vertical_boundaries=True,
                      enable_AB_3_order=True, eps_ab=0.0, precision=precision)
    return prm, _finish(prm, {"u": u, "v": v, "w": np.zeros_like(u), "b": b})


def jets(nx=256, ny=256, nz=64, seed=2, umax=0.1, noise=1e-3, friction=True, precision=PS3D_FP64,
         width_frac=0.02, noise_fields=("u", "v"), stable_N0=False):
    """Two zonal jets of opposite sign, geostrophically balanced, rigid lid.

    The set-up of for_config/jets_unscaled.f90:5-95 at any resolution:
    Lx = Ly = 500 km, Lz = 500 m with the lid, f0 = 1e-4, N0 = 50 f0,
    dt = 0.5/N0; u = umax (jets at y = Ly/4, 3Ly/4 of Gaussian width
    width_frac Lx + 0.02 sin(10 pi x/Lx) sin(2 pi y/Ly)) cos(pi z/Lz); b from
    the thermal-wind cumsum in y (:80-95).  Added: N(0, noise) on the
    noise_fields, and harmonic (Ah, Av, Kh, Kv) and horizontal biharmonic
    (Ahbi, Khbi) friction sized to the grid.  (Stronger, wider jets -- 0.5
    m/s over Ly/10 -- or noise on b of 1e-3 make the thermal-wind
    stratification statically unstable at N0 = 50 f0, d(b)/dz ~ 1.7e-4 >
    N0^2 = 2.5e-5, and the run convects and blows up within ~50 steps;
    stable_N0 raises N0 until N0^2 is four times the largest -d(b)/dz of the
    jets' buoyancy, with dt = 0.5/N0 following it.)
    """
    f0 = 1e-4
    N0 = 50 * f0
    Lx = Ly = 500e3
    Lz = 500.0
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    width = Lx * width_frac
    xi = (np.arange(1, nx + 1) - 1) * dx
    yj = (np.arange(1, ny + 1) - 1) * dy
    zk = (np.arange(1, nz + 1) - 1) * dz
    prof = (np.exp(-(yj - 3 * Ly / 4) ** 2 / width ** 2) - np.exp(-(yj - Ly / 4) ** 2 / width ** 2))
    wave = 0.02 * np.sin(xi / Lx * 10 * math.pi)[:, None] * np.sin(yj / Ly * 2 * math.pi)[None, :]
    u = umax * (prof[None, :] + wave)[:, :, None] * np.cos(zk / Lz * math.pi)[None, None, :]
    # b_y = -f u_z by one-sided/centred differences and a cumulative sum in y
    up = np.concatenate([u[:, :, :1], u, u[:, :, -1:]], axis=2)
    b = np.empty_like(u)
    for k in range(nz):
        if k == 0:
            b[:, :, k] = -dy * (up[:, :, k + 2] - up[:, :, k + 1]) / dz * f0
        elif k == nz - 1:
            b[:, :, k] = -dy * (up[:, :, k + 1] - up[:, :, k]) / dz * f0
        else:
            b[:, :, k] = -dy * (up[:, :, k + 2] - up[:, :, k]) / (2 * dz) * f0
    b = np.cumsum(b, axis=1)
    if stable_N0:
        bz_min = float(np.min(np.diff(b, axis=2))) / dz
        N0 = max(N0, math.sqrt(max(0.0, -4.0 * bz_min)))
    dt = 0.5 / N0
    fields = {"u": u, "v": np.zeros_like(u), "w": np.zeros_like(u), "b": b}
    for n in noise_fields:
        fields[n] = fields[n] + noise * normal(seed, n, nx, ny, nz)
    kw = {}
    if friction:
        kw = dict(Ah=1e-3 * dx ** 2 / dt, Av=1e-2 * dz ** 2 / dt, Kh=1e-3 * dx ** 2 / dt,
                  Kv=1e-2 * dz ** 2 / dt, Ahbi=dx ** 4 / (30 * 86400.0), Khbi=dx ** 4 / (30 * 86400.0))
    prm = ModelParams(nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Lz=Lz, dx=dx, dy=dy, dz=dz, dt=dt,
                      f0=f0, N0=N0, enable_vertical_boundaries=True, enable_AB_3_order=True,
                      eps_ab=0.0, precision=precision, **kw)
    return prm, _finish(prm, fields)


def jets_c2(nx=256, ny=256, nz=64, seed=2, precision=PS3D_FP64):
    """The 256x256x64 jets state as the benchmark configuration describes
    it: jets of 0.5 m/s and width Ly/10, N(0, 1e-3) on u, v, w and b; N0
    raised for static stability (see jets: 269 f0, dt = 18.6 s at
    256x256x64)."""
    return jets(nx, ny, nz, seed=seed, umax=0.5, noise=1e-3, precision=precision, width_frac=0.1,
                noise_fields=("u", "v", "w", "b"), stable_N0=True)


def jets_eddies(nx=1024, ny=1024, nz=128, seed=5, n_eddies=8, eddy_amp=0.02, radius_frac=1.0 / 20.0,
                precision=PS3D_FP64):
    """Jets and eddy turbulence at scale: the two balanced jets of jets()
    plus n_eddies Gaussian buoyancy eddies at positions drawn from the seed,
    each b' = eddy_amp N0^2 Lz exp(-r^2/R^2) exp(-depth/Lz) with R =
    radius_frac Lx, the velocities in discrete thermal-wind balance (zero at
    the bottom) as in eddy().  The eddies keep d(b')/dz well below N0^2 and
    their speed below the jets'."""
    prm, f = jets(nx, ny, nz, seed=seed, precision=precision)
    nx_, ny_, nz_ = prm.nx, prm.ny, prm.nz
    L, Lz, f0, N0 = prm.Lx, prm.Lz, prm.f0, prm.N0
    dx, dy, dz = prm.dx, prm.dy, prm.dz
    A = eddy_amp * N0 ** 2 * Lz
    R = radius_frac * L
    pos = 0.5 * (uniform(seed, "R", n_eddies, 2, 1, stream=3)[:, :, 0] + 1.0) * L   # (n_eddies, 2) in [0, L)
    xi = (np.arange(1, nx_ + 1) - 0.5) * dx
    yj = (np.arange(1, ny_ + 1) - 0.5) * dy
    zk = (np.arange(1, nz_ + 1) - 0.5) * dz
    E = np.exp(-(Lz - zk) / Lz)
    E0 = math.exp(-1.0)

    def per(d, P):   # periodic distance
        return (d + 0.5 * P) % P - 0.5 * P

    b = np.zeros((nx_, ny_))
    gy = np.zeros((nx_, ny_))
    gx = np.zeros((nx_, ny_))
    for xc, yc in pos:
        X, Y = np.meshgrid(per(xi - xc, L), per(yj - yc, L), indexing="ij")
        b += np.exp(-(X ** 2 + Y ** 2) / R ** 2)
        Xu, Yu = np.meshgrid(per(xi + 0.5 * dx - xc, L), per(yj - yc, L), indexing="ij")
        gy += -2 * Yu / R ** 2 * np.exp(-(Xu ** 2 + Yu ** 2) / R ** 2)
        Xv, Yv = np.meshgrid(per(xi - xc, L), per(yj + 0.5 * dy - yc, L), indexing="ij")
        gx += -2 * Xv / R ** 2 * np.exp(-(Xv ** 2 + Yv ** 2) / R ** 2)
    I = (slice(2, 2 + nx_), slice(2, 2 + ny_), slice(2, 2 + nz_))
    inner = {n: np.array(f[n][I]) for n in ("u", "v", "w", "b")}
    inner["b"] += A * b[:, :, None] * E[None, None, :]
    inner["u"] += -(A / f0) * gy[:, :, None] * (Lz * (E - E0))[None, None, :]
    inner["v"] += (A / f0) * gx[:, :, None] * (Lz * (E - E0))[None, None, :]
    return prm, _finish(prm, inner)


def lee_wave(nx=512, ny=128, nz=128, seed=3, U=0.1, N0=1e-3, noise=1e-4, precision=PS3D_FP64):
    """Stratified flow U past a Gaussian bottom obstacle, the obstacle as a
    body force: a new configuration written against the plug-in surface.
    The background flow is advected as in for_config/lee_wave1.f90:239-314
    (constant u_back = U, no background buoyancy).  The obstacle
    h(x) = 0.05 H exp(-((x - Lx/4)/(Lx/20))^2) lifts the stratification by
    w = U h'(x); its effect on the buoyancy, -N0^2 U h'(x) exp(-z/(0.05 H))
    (z the height above the bottom), is the body force on db.  Rigid lid,
    H = 2 km, Lx = 10 km (obstacle width 500 m against the lee wavelength
    2 pi U/N0 = 630 m), square cells in x-y, f0 = 1e-4, dt at an advective
    CFL of 0.2, harmonic and horizontal biharmonic friction sized to the
    grid, noise * U(-1, 1) on u, v, w (seed).

    Returns (prm, fields, forcing) with forcing = {'u_back': .., 'force':
    {'db': ..}} for Model.set_background_flow / set_body_force."""
    H = 2000.0
    Lx = 10e3
    dx = Lx / nx
    Ly = dx * ny
    dz = H / nz
    f0 = 1e-4
    dt = 0.2 * dx / U
    kw = dict(Ah=1e-3 * dx ** 2 / dt, Av=1e-3 * dz ** 2 / dt, Kh=1e-3 * dx ** 2 / dt, Kv=1e-3 * dz ** 2 / dt,
              Ahbi=dx ** 4 / (30 * 86400.0), Khbi=dx ** 4 / (30 * 86400.0))
    prm = ModelParams(nx=nx, ny=ny, nz=nz, Lx=Lx, Ly=Ly, Lz=H, dx=dx, dy=dx, dz=dz, dt=dt, f0=f0, N0=N0,
                      enable_vertical_boundaries=True, enable_AB_3_order=True, eps_ab=0.0, precision=precision,
                      **kw)
    inner = {n: noise * uniform(seed, n, nx, ny, nz) for n in ("u", "v", "w")}
    inner["b"] = np.zeros((nx, ny, nz))
    fields = _finish(prm, inner)
    xi = (np.arange(-1, nx + 3) - 0.5) * dx                 # cell centres, halo included
    zk = (np.arange(-1, nz + 3) - 0.5) * dz
    h0, x0, sig = 0.05 * H, Lx / 4, Lx / 20
    dh = -2 * (xi - x0) / sig ** 2 * h0 * np.exp(-((xi - x0) / sig) ** 2)
    decay = np.exp(-np.clip(zk, 0.0, None) / (0.05 * H))
    fb = (-N0 ** 2 * U) * dh[:, None, None] * np.ones(ny + 4)[None, :, None] * decay[None, None, :]
    forcing = {"u_back": np.full(prm.shape, U, order="F"), "force": {"db": np.asfortranarray(fb)}}
    return prm, fields, forcing


def random_state(nx, ny, nz, seed=7, lid=True, amp=0.1, **kw):
    """A generic test state: smooth large-scale fields plus noise, with every
    friction operator switched on (kw override any parameter)."""
    L = 1.0e5
    Lz = 1.0e3
    base = dict(nx=nx, ny=ny, nz=nz, Lx=L, Ly=L * ny / nx, Lz=Lz, dt=60.0, f0=1e-4, N0=2e-3,
                Ro=1.0, dsqr=1.0, eps_ab=0.01, Ah=20.0, Av=1e-3, Kh=10.0, Kv=5e-4,
                Ahbi=1e7, Avbi=1e-2, Khbi=5e6, Kvbi=5e-3, enable_vertical_boundaries=lid,
                enable_AB_3_order=True)
    base.update(kw)
    prm = ModelParams(**base)
    xi = np.arange(nx)[:, None, None] / nx
    yj = np.arange(ny)[None, :, None] / ny
    zk = (np.arange(nz)[None, None, :] + 0.5) / nz
    fields = {}
    for n, ph in zip(("u", "v", "w", "b"), (0.1, 0.7, 1.3, 2.1)):
        smooth = np.sin(2 * math.pi * (xi + ph)) * np.cos(2 * math.pi * (yj - ph)) * np.cos(math.pi * zk)
        fields[n] = amp * (smooth + 0.1 * uniform(seed, n, nx, ny, nz))
    fields["w"] *= 0.01
    fields["b"] *= 1e-3
    return prm, _finish(prm, fields)


def wavemaker_amplitudes(prm: ModelParams, seed: int = 11, amp: float = 1e-5):
    """Amplitude fields (q_r, q_i) of the wavemaker forcing
    (for_src/config.f90:224-237) for u, v, w, b: uniform noise of size amp on
    the interior, halos exchanged as set_initial_conditions leaves them."""
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    re, im = {}, {}
    for n in ("u", "v", "w", "b"):
        for d, stream in ((re, 1), (im, 2)):
            a = np.zeros(prm.shape, order="F")
            a[2:2 + nx, 2:2 + ny, 2:2 + nz] = amp * uniform(seed, n, nx, ny, nz, stream=stream)
            border_exchg(a)
            d[n] = a
    return re, im


def ensemble_c4(per_size: int = 16, sizes=((64, 64, 32), (96, 96, 48), (128, 128, 64)), precision=PS3D_FP64):
    """Ensemble of baroclinic eddies (the 64x64x32 eddy physics) in three
    sizes, per_size members each; member m has seed 1000+m and eddy radius
    R_m = U(1/8, 1/4) Lx drawn from that seed."""
    out = []
    m = 0
    for (nx, ny, nz) in sizes:
        for _ in range(per_size):
            seed = 1000 + m
            u01 = 0.5 * (float(uniform(seed, "R", 1, 1, 1)[0, 0, 0]) + 1.0)
            out.append(eddy(nx, ny, nz, seed=seed, radius_frac=0.125 + 0.125 * u01, precision=precision))
            m += 1
    return out

