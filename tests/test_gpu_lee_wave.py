"""Background-flow advection (lee_wave1.f90:239-314) and a steady body force
on the GPU: the FP64 form against the oracle with the same hook, bit for bit,
and the lee-wave configuration at 512x128x128 in the mixed form against the
FP64 form over its 2000 steps, reported at steps 1, 100, 1000 and 2000
(<= 1e-5 relative L2 per field and in total energy).  Both forms form the forcing terms from the post-friction
state, as the reference does (integrate.f90:82): the FP64 form runs the
fused friction chain alone, the forcing, then the Adams-Bashforth pass; the
mixed form the same three passes on its (hi, lo) state.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _forcing(prm, seed=4):
    from ps3d_amd import configs
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    ub = np.zeros(prm.shape, order="F")
    ub[2:2 + nx, 2:2 + ny, 2:2 + nz] = 0.05 + 0.02 * configs.uniform(seed, "u", nx, ny, nz, stream=2)
    configs.border_exchg(ub)
    bb = np.zeros(prm.shape, order="F")
    bb[2:2 + nx, 2:2 + ny, 2:2 + nz] = 1e-4 * configs.uniform(seed, "b", nx, ny, nz, stream=2)
    configs.border_exchg(bb)
    fw = np.zeros(prm.shape, order="F")
    fw[2:2 + nx, 2:2 + ny, 2:2 + nz] = 1e-7 * configs.uniform(seed, "w", nx, ny, nz, stream=2)
    return ub, bb, {"dw": fw}


@pytest.mark.parametrize("lid", [True, False])
@pytest.mark.parametrize("with_bb", [False, True])
def test_background_flow_matches_oracle(lid, with_bb):
    import oracle
    from ps3d_amd import configs
    from ps3d_amd.model import Model
    prm, fields = configs.random_state(32, 24, 12, seed=6, lid=lid, Avbi=0.0, Kvbi=0.0)
    print(f"background flow lid={lid} b_back={with_bb}", flush=True)
    ub, bb, force = _forcing(prm)
    bb = bb if with_bb else None
    gpu = Model(prm)
    orc = oracle.OracleModel(prm)
    for n, a in fields.items():
        gpu.f[n][...] = a
        orc.f[n][...] = a
    gpu.push()
    gpu.set_background_flow(ub, bb)
    gpu.set_body_force(force)
    orc.set_background(ub, bb, force)
    gpu.run(20)
    orc.run(20)
    gpu.pull()
    for n in ("u", "v", "w", "b", "p"):
        assert np.array_equal(gpu.f[n], orc.f[n]), n
    gpu.close()
    orc.close()


def test_lee_wave_mixed_vs_fp64():
    from ps3d_amd import configs, diagnose
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields, forcing = configs.lee_wave(512, 128, 128, seed=3)
    print(f"lee wave {prm.nx}x{prm.ny}x{prm.nz} dt={prm.dt}", flush=True)
    models = []
    for p in (prm, prm.replace(precision=PS3D_MIXED)):
        m = Model(p)
        for n, a in fields.items():
            m.f[n][...] = a
        m.push()
        m.set_background_flow(forcing["u_back"])
        m.set_body_force(forcing["force"])
        models.append(m)
    g64, gmx = models
    done = 0
    worst = 0.0
    for upto in (1, 100, 1000, 2000):
        g64.run(upto - done)
        gmx.run(upto - done)
        done = upto
        g64.pull()
        gmx.pull()
        errs = diagnose.compare(gmx.f, g64.f, prm)
        print(f"step {upto}: {errs}", flush=True)
        worst = max(worst, max(errs.values()))
    # the flow is forced: w has grown well above the initial noise (1e-4)
    I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
    print("max |w|", float(np.abs(g64.f["w"][I]).max()), flush=True)
    assert np.abs(g64.f["w"][I]).max() > 1e-3
    assert worst < 1e-5
    g64.close()
    gmx.close()
