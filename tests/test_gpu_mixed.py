"""The mixed-precision form (PS3D_MIXED) against the FP64 drop-in.

FP32 stencils and transforms, FP64 spectral division / Thomas elimination
and FP64-accumulated (hi, lo) prognostic state.  Stated accuracy: relative
L2 per field (u, v, w, b, p) and of the total energy (diagnose.F90:14-24)
within 1e-5 of the FP64 path after 1000 steps; the FP64 GPU path is itself
bit-identical to the CPU oracle (test_gpu_parity.py), and the oracle is
the reference for the small cases here.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _models(make, **kw):
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED, PS3D_FP64
    prm, fields = make(**kw)
    prm_m = prm.replace(precision=PS3D_MIXED)
    g64 = Model(prm.replace(precision=PS3D_FP64))
    gmx = Model(prm_m)
    for n, a in fields.items():
        g64.f[n][...] = a
        gmx.f[n][...] = a
    g64.push()
    gmx.push()
    return prm, g64, gmx


def _jets(nx, ny, nz):
    from ps3d_amd import configs
    return lambda **kw: configs.jets(nx, ny, nz, seed=2, **kw)


def test_mixed_roundtrip_and_state_errors():
    from ps3d_amd import _lib, configs
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields = configs.jets(32, 32, 16, seed=3, precision=PS3D_MIXED)
    m = Model(prm)
    rng = np.random.default_rng(0)
    ref = {}
    for n in ("u", "v", "w", "b", "p"):
        ref[n] = rng.standard_normal(prm.shape)
        m.f[n][...] = ref[n]
    m.push(["u", "v", "w", "b", "p"])
    m.pull(["u", "v", "w", "b", "p"])
    for n in ("u", "v", "w", "b"):
        # (hi, lo) FP32 pair: 48 significant bits
        assert np.all(np.abs(m.f[n] - ref[n]) <= 2.0 ** -46 * np.abs(ref[n])), n
    assert np.all(np.abs(m.f["p"] - ref["p"]) <= 2.0 ** -23 * np.abs(ref["p"]))
    # building blocks are FP64-only
    rc = m.lib.ps3d_tendency(m.h, 1)
    assert rc == _lib.PS3D_ERR_STATE
    buf = ctypes.create_string_buffer(256)
    m.lib.ps3d_last_error(buf, 256)
    assert b"PS3D_MIXED" in buf.value
    m.close()


@pytest.mark.parametrize("size", [(32, 32, 16), (64, 64, 32)])
def test_mixed_vs_oracle_100_steps(size):
    """Mixed GPU vs the CPU oracle (FP64) after 1 and 100 steps."""
    import oracle
    from ps3d_amd import configs, diagnose
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields = configs.jets(*size, seed=4)
    gmx = Model(prm.replace(precision=PS3D_MIXED))
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gmx.f[n][...] = a
        orc.f[n][...] = a
    gmx.push()
    done = 0
    for upto in (1, 100):
        gmx.run(upto - done)
        orc.run(upto - done)
        done = upto
        gmx.pull()
        ref = {n: orc.f[n] for n in ("u", "v", "w", "b", "p")}
        err = diagnose.compare(gmx.f, ref, prm)
        print(f"mixed {size} step {upto}: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
        for k, v in err.items():
            assert v <= TOL, (upto, k, v)
    gmx.close()
    orc.close()


@pytest.mark.parametrize("case", ["periodic_pow2", "eddy"])
def test_mixed_other_configs(case):
    """Periodic z (3-D spectral division) and the eddy config (no harmonic
    friction) against the FP64 GPU path after 100 steps."""
    from ps3d_amd import configs, diagnose
    if case == "eddy":
        make = lambda **kw: configs.eddy(64, 64, 32, seed=1, **kw)
    else:
        make = lambda **kw: configs.random_state(32, 32, 16, seed=7, lid=False, Avbi=0.0, Kvbi=0.0, **kw)
    prm, g64, gmx = _models(make)
    g64.run(100)
    gmx.run(100)
    g64.pull()
    gmx.pull()
    err = diagnose.compare(gmx.f, g64.f, prm)
    print(f"mixed {case} step 100: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
    for k, v in err.items():
        assert v <= TOL, (k, v)
    g64.close()
    gmx.close()


def test_mixed_1000_steps_c2_size():
    """256x256x64 jets, 1000 steps: the stated accuracy of the mixed form."""
    from ps3d_amd import diagnose
    prm, g64, gmx = _models(_jets(256, 256, 64))
    done = 0
    for upto in (1, 100, 1000):
        g64.run(upto - done)
        gmx.run(upto - done)
        done = upto
        g64.pull()
        gmx.pull()
        err = diagnose.compare(gmx.f, g64.f, prm)
        print(f"mixed C2 step {upto}: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
        for k, v in err.items():
            assert v <= TOL, (upto, k, v)
    g64.close()
    gmx.close()


@pytest.mark.parametrize("case", ["eddy1_120x120x10", "jets_100x100x20"])
def test_mixed_at_reference_config_sizes(case):
    """The mixed form at the grid sizes of the reference's own configs --
    eddy1.f90:15 (120 x 120 x 10: transforms of 60 and 120) and
    jets_scaled.f90:9 / jets_unscaled.f90:17 (100 x 100 x 20: 50 and 100),
    lengths with factors 3 and 5 -- against the CPU oracle (FP64) after 1
    and 100 steps, within the mixed form's bound."""
    import oracle
    from ps3d_amd import configs, diagnose
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    if case.startswith("eddy1"):
        prm, fields = configs.eddy(120, 120, 10, seed=1)
    else:
        prm, fields = configs.jets(100, 100, 20, seed=2)
    gmx = Model(prm.replace(precision=PS3D_MIXED))
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gmx.f[n][...] = a
        orc.f[n][...] = a
    gmx.push()
    done = 0
    for upto in (1, 100):
        gmx.run(upto - done)
        orc.run(upto - done)
        done = upto
        gmx.pull()
        ref = {n: orc.f[n] for n in ("u", "v", "w", "b", "p")}
        err = diagnose.compare(gmx.f, ref, prm)
        print(f"mixed {case} step {upto}: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
        for k, v in err.items():
            assert v <= TOL, (upto, k, v)
    gmx.close()
    orc.close()


@pytest.mark.parametrize("case", ["lid_128x128x32", "periodic_64x64x32"])
def test_radix8_plans_vs_fp64(case):
    """The FP32 transforms with radix-8 passes first (plan 1) on every axis
    that has a factor 8 -- x (n/2 = 64 or 32), y and, periodic, z: within
    the stated accuracy of the FP64 path after 100 steps, and not the same
    bits as plan 0 (so the other factorisation ran)."""
    from ps3d_amd import configs, diagnose
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    if case.startswith("lid"):
        make = lambda **kw: configs.jets(128, 128, 32, seed=5, **kw)
        plans = (1, 1, 0)
    else:
        make = lambda **kw: configs.random_state(64, 64, 32, seed=8, lid=False, Avbi=0.0, Kvbi=0.0, **kw)
        plans = (1, 1, 1)
    prm, g64, gmx = _models(make)
    gmx.fft_plan = plans
    assert gmx.fft_plan == plans
    _, fields = make()
    g0 = Model(prm.replace(precision=PS3D_MIXED))
    for n, a in fields.items():
        g0.f[n][...] = a
    g0.push()
    assert g0.fft_plan == (0, 0, 0)
    for m in (g64, gmx, g0):
        m.run(100)
        m.pull()
    err = diagnose.compare(gmx.f, g64.f, prm)
    print(f"plans {plans} {case} step 100: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
    for k, v in err.items():
        assert v <= TOL, (k, v)
    assert any(not np.array_equal(gmx.f[n], g0.f[n]) for n in ("u", "v", "w", "b", "p"))
    for m in (g64, gmx, g0):
        m.close()


def test_fft_plan_rules_and_tuning(tmp_path):
    """FP64 models keep plan 0 (the oracle's factorisation); plans are 0 or
    1; the tuner of a mixed model picks a plan for each transform axis with
    a choice (lengths with a factor 8) and reports it."""
    from ps3d_amd import _lib, configs
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields = configs.jets(64, 64, 32, seed=6)
    m64 = Model(prm)
    with pytest.raises(_lib.Ps3dError, match="plan 0"):
        m64.fft_plan = (1, 0, 0)
    assert m64.fft_plan == (0, 0, 0)
    m64.close()
    m = Model(prm.replace(precision=PS3D_MIXED))
    with pytest.raises(_lib.Ps3dError, match="0 or 1"):
        m.fft_plan = (2, 0, 0)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()
    m.run(3)
    m.tune(str(tmp_path / "tune.txt"), reps=3)
    desc = m.tuning()
    print(desc, flush=True)
    lines = {ln.split()[0]: int(ln.split()[2]) for ln in desc.strip().splitlines()}
    assert lines["fft_plan_x"] in (0, 1) and lines["fft_plan_y"] in (0, 1)
    assert (lines["fft_plan_x"], lines["fft_plan_y"]) == m.fft_plan[:2]
    cached = open(tmp_path / "tune.txt").read()
    assert "fft_plan_x 64x64x32:lid:mixed" in cached and "fft_plan_y 64x64x32:lid:mixed" in cached
    m.close()


def test_mixed_c2_state_100_steps():
    """The mixed form on the benchmark configuration's own jets state
    (configs.jets_c2, 256x256x64) against the FP64 path after 100 steps."""
    from ps3d_amd import configs, diagnose
    prm, g64, gmx = _models(lambda **kw: configs.jets_c2(256, 256, 64, **kw))
    g64.run(100)
    gmx.run(100)
    g64.pull()
    gmx.pull()
    err = diagnose.compare(gmx.f, g64.f, prm)
    print("mixed C2 state step 100: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
    for k, v in err.items():
        assert v <= TOL, (k, v)
    g64.close()
    gmx.close()


@pytest.mark.parametrize("nz", [48, 128])
def test_every_mixed_thomas_form_gives_the_same_bits(nz, tmp_path):
    """Each form of the mixed form's Thomas solve (FP32 right-hand side, FP64
    elimination), forced through a tuning cache -- whole columns in
    registers (nz of 16-64), the real and imaginary parts on two lanes, 32
    columns per wave on the fly, the factor tables with FP64 d' in a scratch
    array (the model has one where the on-the-fly form's LDS exceeds 64 KiB)
    -- steps bit for bit like the others, and within the stated bound of the
    FP64 path."""
    from ps3d_amd import configs, diagnose
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields = configs.random_state(32, 16, nz, seed=43, lid=True, Avbi=0.0, Kvbi=0.0)
    g64 = Model(prm)
    for n, a in fields.items():
        g64.f[n][...] = a
    g64.push()
    g64.run(3)
    g64.pull()
    prm_m = prm.replace(precision=PS3D_MIXED)
    shape = f"32x16x{nz}:lid:mixed"
    forms = ([1] if nz <= 64 else []) + [2, 32] + ([-64] if 24 * 32 * nz > 64 * 1024 else [])
    ref = None
    for form in forms:
        print(f"mixed nz {nz}: Thomas form {form}", flush=True)
        cache = tmp_path / f"thomas_{nz}_{form}.txt"
        cache.write_text(f"friction_ab {shape} 8\ntendency {shape} 8\nthomas {shape} {form}\n"
                         f"fft_plan_x {shape} 0\nfft_plan_y {shape} 0\n")
        m = Model(prm_m)
        for n, a in fields.items():
            m.f[n][...] = a
        m.push()
        m.tune(str(cache), reps=1)
        assert f"thomas {shape} {form}" in m.tuning()
        m.run(3)
        m.pull()
        state = {n: m.f[n].copy() for n in ("u", "v", "w", "b", "p")}
        m.close()
        if ref is None:
            ref = state
            err = diagnose.compare(state, g64.f, prm)
            print(f"mixed nz {nz} step 3: " + " ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)
            for k, v in err.items():
                assert v <= TOL, (k, v)
        else:
            for n in ("u", "v", "w", "b", "p"):
                assert np.array_equal(state[n], ref[n]), (form, n)
    g64.close()
