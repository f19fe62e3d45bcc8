"""GPU against the oracle at the sizes the benchmarks and configurations use.

* Every FFT length compiled into fft.hip (8 .. 1024; eddy1.f90's 120 x 120
  and jets_scaled.f90's 100 x 100 among them) and a set of lengths that take
  the generic kernels (18, 36, 45, 90, 150), in x (r2c of
  twice the length), in y and, periodic, in z: my_fft / my_ifft
  (my_fft.f90:159-205) through the model, and the P3DFFT-style plan
  (ps3d_fft_*) scaled as my_fft.f90 scales it -- bit for bit.
* The bench workload, 256 x 256 x 64 jets: ten steps, the pressure solver
  and the spectral solve on its state -- bit for bit.
* Batched ensemble members of each size of the 48-member ensemble
  (64 x 64 x 32, 96 x 96 x 48, 128 x 128 x 64) against the oracle over 20
  steps: bit for bit in FP64 (graph replay), within 1e-5 in the mixed form.
* ps3d_set_params between steps (dt, Ro, friction changed as
  diag_opt_balance.f90 changes them) against the oracle's set_dynamic.
* The restart file the device writes, byte for byte against the
  reference's own write_restart (restart.f90:2-22, oracle/_ref).
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

COMPILED = [8, 10, 12, 15, 16, 20, 24, 30, 32, 40, 48, 50, 60, 64, 80, 96, 100, 120, 128, 160, 240, 256, 512, 1024]
GENERIC = [18, 36, 45, 90, 150]


def _qp(shape, seed):
    rng = np.random.default_rng(seed)
    return np.asfortranarray(rng.standard_normal(shape))


def _fft_case(prm, three_d, seed):
    from ps3d_amd.model import FftPlan, Model
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    qp = _qp((prm.nx, prm.ny, prm.nz), seed)
    a = gpu.my_fft(qp, three_d)
    b = orc.my_fft(qp, three_d)
    assert np.array_equal(a, b), "my_fft"
    back_g = gpu.my_ifft(b, three_d)
    back_o = orc.my_ifft(b, three_d)
    assert np.array_equal(back_g, back_o), "my_ifft"
    # the P3DFFT-style plan, unnormalised; my_fft.f90 scales after it
    plan = FftPlan(prm.nx, prm.ny, prm.nz)
    if three_d:
        q = plan.forward(qp, "fft")
        scaled = (q.real * prm.dx * prm.dy * prm.dz) + 1j * (q.imag * prm.dx * prm.dy * prm.dz)
        r = plan.backward(b, "fft") / (prm.Lx * prm.Ly * prm.Lz)
    else:
        q = plan.forward(qp, "ffn")
        scaled = (q.real * prm.dx * prm.dy) + 1j * (q.imag * prm.dx * prm.dy)
        r = plan.backward(b, "nff") / (prm.Lx * prm.Ly)
    assert np.array_equal(scaled, b), "plan forward"
    assert np.array_equal(r, back_o), "plan backward"
    plan.close()
    gpu.close()
    orc.close()


@pytest.mark.parametrize("n", COMPILED + GENERIC)
def test_fft_lengths_x_and_y(n):
    from ps3d_amd import configs
    print(f"fft length {n}: x (nx = {2 * n})" + (" and y" if n % 2 == 0 else ""), flush=True)
    prm, _ = configs.random_state(2 * n, 6, 3, seed=1, lid=True)
    _fft_case(prm, False, seed=n)
    if n % 2 == 0:   # ny is even (main.f90:45-55); odd lengths occur as nx/2 only
        prm, _ = configs.random_state(8, n, 3, seed=1, lid=True)
        _fft_case(prm, False, seed=n + 1)


@pytest.mark.parametrize("n", [n for n in COMPILED + GENERIC if n % 2 == 0])
def test_fft_lengths_z(n):
    from ps3d_amd import configs
    print(f"fft length {n}: z (periodic)", flush=True)
    prm, _ = configs.random_state(8, 6, n, seed=1, lid=False)
    _fft_case(prm, True, seed=n + 2)


def test_fft_plan_rejects_bad_op():
    from ps3d_amd import _lib
    from ps3d_amd.model import FftPlan
    plan = FftPlan(16, 8, 6)
    with pytest.raises(_lib.Ps3dError, match="op string"):
        plan.forward(np.zeros((16, 8, 6)), "nff")
    with pytest.raises(_lib.Ps3dError, match="op string"):
        plan.backward(np.zeros((9, 8, 6), dtype=complex), "ffn")
    plan.close()


@pytest.fixture(scope="module")
def jets256():
    from ps3d_amd import configs
    return configs.jets(256, 256, 64)


def test_bench_workload_ten_steps_bitwise(jets256):
    from ps3d_amd.model import Model
    prm, fields = jets256
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
        orc.f[n][...] = a
    gpu.push()
    print("256x256x64 jets: 10 steps", flush=True)
    gpu.run(10)
    orc.run(10)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    for n in ("du", "dv", "dw", "db"):
        I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
        assert np.array_equal(gpu.f[n][I], orc.f[n][I]), n
    # the building blocks on this state
    qp = np.asfortranarray(orc.f["u"][2:2 + prm.nx, 2:2 + prm.ny, 2:2 + prm.nz])
    qs = orc.my_fft(qp, False)
    assert np.array_equal(gpu.my_fft(qp, False), qs)
    assert np.array_equal(gpu.my_ifft(qs, False), orc.my_ifft(qs, False))
    assert np.array_equal(gpu.spectral_solve(qs), orc.spectral_solve(qs))
    gpu.pressure_solver()
    orc.pressure_solver()
    gpu.pull(["u", "v", "w", "p"])
    for n in ("u", "v", "w", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    gpu.close()
    orc.close()


def test_c2_state_steps_bitwise():
    """The benchmark configuration's own jets state (0.5 m/s over Ly/10,
    N(0, 1e-3) on u, v, w, b; N0 raised for static stability,
    configs.jets_c2) at 256x256x64: 5 steps bit for bit against the oracle;
    then 300 more on the GPU stay bounded (the state is stable)."""
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.jets_c2(256, 256, 64)
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
        orc.f[n][...] = a
    gpu.push()
    print("256x256x64 C2 jets: 5 steps against the oracle", flush=True)
    gpu.run(5)
    orc.run(5)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    orc.close()
    gpu.run(300)
    gpu.pull(["u", "v", "w"])
    umax = max(float(np.abs(gpu.f[n]).max()) for n in ("u", "v", "w"))
    print(f"  after 305 steps: max |u, v, w| = {umax:.3f} m/s", flush=True)
    assert np.isfinite(umax) and umax < 1.0
    gpu.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_ensemble_members_match_oracle(mixed):
    from ps3d_amd import configs
    from ps3d_amd.ensemble import Ensemble
    from ps3d_amd.params import PS3D_FP64, PS3D_MIXED
    prec = PS3D_MIXED if mixed else PS3D_FP64
    members = [configs.eddy(nx, ny, nz, seed=1000 + q, radius_frac=0.125 + 0.03 * q, precision=prec)
               for q, (nx, ny, nz) in enumerate(((64, 64, 32), (96, 96, 48), (128, 128, 64)))]
    ens = Ensemble([p for p, _ in members])
    for q, (_, f) in enumerate(members):
        for n, a in f.items():
            ens.set(q, n, a)
    print(f"ensemble of 3 sizes, mixed={mixed}: 20 steps", flush=True)
    ens.run(20, use_graph=True)
    ens.synchronize()
    for q, (prm, f) in enumerate(members):
        orc = oracle.OracleModel(prm.replace(precision=PS3D_FP64))
        for n, a in f.items():
            orc.f[n][...] = a
        orc.run(20)
        for n in ("u", "v", "w", "b", "p"):
            got = ens.get(q, n)
            if mixed:
                I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
                ref = orc.f[n][I]
                err = float(np.linalg.norm(got[I] - ref) / np.linalg.norm(ref))
                print(f"  member {q} {prm.nx}x{prm.ny}x{prm.nz} {n}: rel L2 {err:.2e}", flush=True)
                assert err <= 1e-5, (q, n)
            else:
                assert np.array_equal(got, orc.f[n]), (q, n)
        orc.close()
    ens.close()


def test_set_params_between_steps_matches_oracle():
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.random_state(32, 24, 12, seed=4, lid=True)
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
        orc.f[n][...] = a
    gpu.push()
    gpu.run(3)
    orc.run(3)
    # backward, Ro ramped down, biharmonic friction off (diag_opt_balance.f90:190-205)
    for ro in (0.75, 0.5, 0.25):
        gpu.set_params(dt=-prm.dt, Ro=ro * prm.Ro, Ahbi=0.0, Avbi=0.0, Khbi=0.0, Kvbi=0.0)
        orc.set_dynamic(-prm.dt, ro * prm.Ro, 0.0, 0.0, 0.0, 0.0)
        gpu.run(1)
        orc.run(1)
    gpu.set_params(dt=prm.dt, Ro=prm.Ro, Ahbi=prm.Ahbi, Avbi=prm.Avbi, Khbi=prm.Khbi, Kvbi=prm.Kvbi)
    orc.set_dynamic(prm.dt, prm.Ro, prm.Ahbi, prm.Avbi, prm.Khbi, prm.Kvbi)
    gpu.run(2)
    orc.run(2)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    # the grid cannot change
    from ps3d_amd import _lib
    with pytest.raises(_lib.Ps3dError, match="can change"):
        gpu.set_params(dx=2 * prm.dx)
    gpu.close()
    orc.close()


@pytest.mark.skipif(not oracle.reference_available(), reason="reference build not present")
def test_device_restart_matches_reference_write_restart(tmp_path, monkeypatch):
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.jets(48, 40, 12, seed=3)
    gpu = Model(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
    gpu.push()
    gpu.run(5)
    ours = str(tmp_path / "ours.dta")
    gpu.write_restart(ours, time=5 * prm.dt)
    gpu.pull()
    # the reference's own write_restart on the same main_module state
    ref = oracle.ReferenceModel(prm)
    for n in ("u", "v", "w", "b", "p", "du", "dv", "dw", "db"):
        ref.f[n][...] = gpu.f[n]
    ref.lib.ref_set_time(gpu.itt, gpu.tau, gpu.taum1, gpu.taum2)
    ref.lib.ref_set_clock(5 * prm.dt)
    monkeypatch.chdir(tmp_path)
    ref.lib.write_restart_()
    theirs = tmp_path / "restart_pe=000000000.dta"
    assert open(ours, "rb").read() == theirs.read_bytes()
    gpu.close()
