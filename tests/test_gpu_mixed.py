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
