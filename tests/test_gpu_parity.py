"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Every kernel evaluates the same operations in the same order as the oracle
(stencils: the reference's own expressions; FFT: the same Stockham
factorisation, twiddles and butterflies), with FMA contraction off and
divisions correctly rounded, so the GPU must agree with the oracle BIT FOR
BIT -- per kernel, per transform and over 100 whole steps.  The drop-in
criterion (<= 1e-12 relative L2 per prognostic field after 100 steps) is
asserted as well, and the measured error is printed.
"""
import numpy as np
import pytest

from conftest import interior, rel_l2

pytestmark = pytest.mark.gpu

CASES = {
    "lid": dict(nx=32, ny=24, nz=16, lid=True),
    "periodic": dict(nx=24, ny=20, nz=10, lid=False),
    # no vertical biharmonic friction: integrate takes the fused
    # friction + Adams-Bashforth kernel
    "lid_fused": dict(nx=96, ny=40, nz=12, lid=True, Avbi=0.0, Kvbi=0.0),
    "periodic_fused": dict(nx=30, ny=18, nz=20, lid=False, Avbi=0.0, Kvbi=0.0),
    # every FFT length compiled in (x: 16, y: 32, z: 16); the cases above mix
    # compiled-in and generic lengths
    "periodic_pow2": dict(nx=32, ny=32, nz=16, lid=False, Avbi=0.0, Kvbi=0.0),
    # three levels: the kernels leave the halo to a separate exchange pass
    # (they write halo images themselves only when every extent is >= 4)
    "thin_lid_fused": dict(nx=12, ny=8, nz=3, lid=True, Avbi=0.0, Kvbi=0.0),
}


def _pair(case, seed=5, **kw):
    import oracle
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    c = dict(CASES[case])
    nx, ny, nz, lid = c.pop("nx"), c.pop("ny"), c.pop("nz"), c.pop("lid")
    c.update(kw)
    prm, fields = configs.random_state(nx, ny, nz, seed=seed, lid=lid, **c)
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    rng = np.random.default_rng(seed)
    for n, a in fields.items():
        gpu.f[n][...] = a
        orc.f[n][...] = a
    for n in ("du", "dv", "dw", "db"):
        t = 1e-6 * rng.standard_normal(gpu.f[n].shape)
        gpu.f[n][...] = t
        orc.f[n][...] = t
    gpu.push()
    return prm, gpu, orc


@pytest.mark.parametrize("case", list(CASES))
def test_advection_bitwise(case):
    prm, gpu, orc = _pair(case)
    print(f"advection case={case}", flush=True)
    gpu.buoyancy_advection()
    gpu.momentum_advection()
    orc.buoyancy_advection()
    orc.momentum_advection()
    gpu.pull(["du", "dv", "dw", "db"])
    t = gpu.tau - 1
    for n in ("du", "dv", "dw", "db"):
        assert np.array_equal(interior(gpu.f[n][..., t], prm), interior(orc.f[n][..., t], prm)), n
    gpu.close()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("op", ["harmonic", "biharmonic_h", "biharmonic_v"])
def test_mixing_bitwise(case, op):
    prm, gpu, orc = _pair(case)
    print(f"mixing case={case} op={op}", flush=True)
    for n in ("u", "w", "b"):
        if op == "harmonic":
            gpu.harmonic(n, 17.0, 2e-3)
            orc.harmonic(n, 17.0, 2e-3)
        elif op == "biharmonic_h":
            gpu.biharmonic_h(n, 3e7)
            orc.biharmonic_h(n, 3e7)
        else:
            gpu.biharmonic_v(n, 4e-2)
            orc.biharmonic_v(n, 4e-2)
    gpu.pull(["u", "w", "b"])
    for n in ("u", "w", "b"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n     # halo included: stale halos kept
    gpu.close()


@pytest.mark.parametrize("case", list(CASES))
def test_border_exchg_bitwise(case):
    prm, gpu, orc = _pair(case)
    rng = np.random.default_rng(11)
    a = rng.standard_normal(prm.shape)
    gpu.f["b"][...] = a
    orc.f["b"][...] = a
    gpu.push(["b"])
    gpu.border_exchg("b")
    orc.border_exchg("b")
    gpu.pull(["b"])
    assert np.array_equal(gpu.f["b"], orc.f["b"])
    gpu.close()


@pytest.mark.parametrize("case", list(CASES))
def test_integrate_pre_pressure_bitwise(case):
    """integrate.f90:12-111 (tendency, friction with stale halos, AB, halo)
    over 4 steps (Euler, AB2, AB3, AB3), bit for bit; the pressure solve is
    left out so FFT rounding cannot enter."""
    prm, gpu, orc = _pair(case)
    for step in range(4):
        print(f"integrate_pre case={case} step={step}", flush=True)
        gpu.tendency()
        gpu.friction()
        gpu.ab_update()
        orc.integrate_pre()
        gpu.pull()
        for n in ("u", "v", "w", "b"):
            assert np.array_equal(gpu.f[n], orc.f[n]), (step, n)
        t = gpu.tau - 1
        for n in ("du", "dv", "dw", "db"):
            assert np.array_equal(interior(gpu.f[n][..., t], prm), interior(orc.f[n][..., t], prm)), (step, n)
        gpu.rotate()
        orc.rotate()
    gpu.close()


@pytest.mark.parametrize("case", list(CASES))
def test_my_fft_matches_oracle(case):
    prm, gpu, orc = _pair(case)
    rng = np.random.default_rng(3)
    x = rng.standard_normal((prm.nx, prm.ny, prm.nz))
    modes = [False] if prm.enable_vertical_boundaries else [False, True]
    for td in modes:
        qs = gpu.my_fft(x, td)
        ref = orc.my_fft(x, td)
        assert np.array_equal(qs, ref), td
        back = gpu.my_ifft(ref, td)
        assert np.array_equal(back, orc.my_ifft(ref, td)), td
        assert np.abs(back - x).max() < 1e-12, td
    gpu.close()


@pytest.mark.parametrize("case", list(CASES))
def test_spectral_solve_matches_oracle(case):
    prm, gpu, orc = _pair(case)
    rng = np.random.default_rng(4)
    qs = rng.standard_normal((prm.nx // 2 + 1, prm.ny, prm.nz)) + 1j * rng.standard_normal(
        (prm.nx // 2 + 1, prm.ny, prm.nz))
    out = gpu.spectral_solve(qs)
    ref = orc.spectral_solve(qs)
    # Thomas and the zero-mode recurrence use the reference's operations: exact
    assert np.array_equal(out, ref)
    gpu.close()


@pytest.mark.parametrize("case", list(CASES))
def test_pressure_solver_matches_oracle(case):
    prm, gpu, orc = _pair(case)
    gpu.pressure_solver()
    orc.pressure_solver()
    gpu.pull(["u", "v", "w", "p"])
    for n in ("u", "v", "w", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    gpu.close()


@pytest.mark.parametrize("case", list(CASES))
def test_integrate_100_steps(case):
    """The drop-in criterion: <= 1e-12 relative L2 per field after 100 steps."""
    prm, gpu, orc = _pair(case)
    print(f"100 steps case={case}", flush=True)
    gpu.run(100)
    orc.run(100)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        e = rel_l2(interior(gpu.f[n], prm), interior(orc.f[n], prm))
        print(f"  {n}: rel L2 = {e:.2e}", flush=True)
        assert e <= 1e-12, n
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    gpu.close()


def test_eddy_200_steps_c1_shape():
    """64x64x32 baroclinic eddy (rigid lid, AB3, biharmonic friction), the
    200 steps of that configuration: bit for bit, and within the drop-in
    bound."""
    import oracle
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.eddy(64, 64, 32, seed=1)
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
        orc.f[n][...] = a
    gpu.push()
    gpu.run(200)
    orc.run(200)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        e = rel_l2(interior(gpu.f[n], prm), interior(orc.f[n], prm))
        print(f"eddy64 {n}: rel L2 = {e:.2e}", flush=True)
        assert e <= 1e-12, n
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    gpu.close()


def test_divergence_roundoff_c2_size():
    """256x256x64 jets: after the projection the discrete divergence
    (pressure_solver.f90:16-23) is at round-off relative to the terms that
    cancel in it -- the largest being the pressure-gradient update of w,
    dt/dsqr * delta_z p / dz, differenced once more in z."""
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.jets(256, 256, 64, seed=2)
    gpu = Model(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
    gpu.push()
    gpu.run(5)
    gpu.pull(["u", "v", "w", "p"])
    u, v, w, p = gpu.f["u"], gpu.f["v"], gpu.f["w"], gpu.f["p"]
    I, J, K = slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz)
    du = (u[I, J, K] - u[1:1 + prm.nx, J, K]) / prm.dx
    dv = (v[I, J, K] - v[I, 1:1 + prm.ny, K]) / prm.dy
    dw = (w[I, J, K] - w[I, J, 1:1 + prm.nz]) / prm.dz
    div = du + dv + dw
    dpz = np.abs(p[I, J, 3:3 + prm.nz] - p[I, J, K]).max()
    scale = max(np.abs(du).max(), np.abs(dv).max(), np.abs(dw).max(),
                prm.dt / prm.dsqr * dpz / prm.dz ** 2)
    print(f"max|div| = {np.abs(div).max():.3e}, term scale = {scale:.3e}", flush=True)
    assert np.abs(div).max() <= 1e-12 * scale
    assert np.isfinite(u).all() and np.isfinite(w).all()
    gpu.close()


@pytest.mark.parametrize("case", ["lid_fused", "periodic_pow2"])
def test_tuned_shapes_give_the_same_bits(case, tmp_path):
    """ps3d_tune changes launch shapes only: a tuned model steps bit for bit
    like the oracle, and the cached choice is reused."""
    prm, gpu, orc = _pair(case)
    cache = str(tmp_path / "tune.txt")
    gpu.tune(cache, reps=2)
    print(gpu.tuning(), flush=True)
    gpu.run(6)
    orc.run(6)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    text = open(cache).read()
    assert "tend_friction_ab" in text or "tendency" in text
    gpu.close()


@pytest.mark.parametrize("nz", [48, 128])
def test_every_thomas_form_gives_the_same_bits(nz, tmp_path):
    """Each form of the FP64 Thomas solve, forced through a tuning cache --
    whole columns in registers (nz of 16, 32, 48, 64), the real and imaginary
    parts of a column on two lanes (nz of 16, 32, 48, 64, 96, 128; 208
    columns, so the last work-group is partly empty), 16/32/64 columns per
    wave on the fly, the factor tables with 16/32/64 columns (d' in LDS or,
    where that exceeds 64 KiB, in FS) -- steps bit for bit like the oracle."""
    import oracle
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.random_state(24, 16, nz, seed=41, lid=True, Avbi=0.0, Kvbi=0.0)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        orc.f[n][...] = a
    orc.run(3)
    shape = f"24x16x{nz}:lid:fp64"
    forms = ([1] if nz in (16, 32, 48, 64) else []) + [2, 16, 32, 64, -16, -32, -64]
    for form in forms:
        print(f"nz {nz}: Thomas form {form}", flush=True)
        cache = tmp_path / f"thomas_{nz}_{form}.txt"
        cache.write_text(f"friction_ab {shape} 8\ntendency {shape} 8\nthomas {shape} {form}\n")
        m = Model(prm)
        for n, a in fields.items():
            m.f[n][...] = a
        m.push()
        m.tune(str(cache), reps=1)
        assert f"thomas {shape} {form}" in m.tuning()
        m.run(3)
        m.pull()
        for n in ("u", "v", "w", "b", "p"):
            assert np.array_equal(m.f[n], orc.f[n]), (form, n)
        m.close()
    orc.close()


FUSED = ["lid_fused", "periodic_fused", "periodic_pow2", "thin_lid_fused"]


@pytest.mark.parametrize("case", FUSED)
def test_fused_tendency_kernel_100_steps(case):
    """The step with the tendencies formed inside the friction + AB kernel
    (k_tend_fric_ab) instead of written to slot tau by a kernel of their own
    and read back: bit for bit with the oracle as well."""
    prm, gpu, orc = _pair(case)
    gpu.set_fused_tendency(True)
    gpu.run(100)
    orc.run(100)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    gpu.close()


@pytest.mark.parametrize("case", FUSED)
@pytest.mark.parametrize("steps", [1, 2, 5])
def test_fused_tendency_writes_slot_tau(case, steps):
    """Euler, AB2 and AB3 steps of the fused kernel: the tendency slots it
    leaves (interior: the contract of the tendency arrays) and every field
    equal the oracle's bit for bit."""
    prm, gpu, orc = _pair(case, seed=9)
    gpu.set_fused_tendency(True)
    gpu.run(steps)
    orc.run(steps)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    for n in ("du", "dv", "dw", "db"):
        g, o = gpu.f[n], orc.f[n]
        for slot in range(3):
            assert np.array_equal(interior(g[..., slot], prm), interior(o[..., slot], prm)), (n, slot)
    gpu.close()
