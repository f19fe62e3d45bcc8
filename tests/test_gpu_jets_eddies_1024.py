"""Jets and eddy turbulence at 1024x1024x128 (the large single instance):
500 steps in the mixed form against the FP64 form, then the FP64 balancing
diagnostic (diag_balance.f90) on both final states.

The FP64 GPU step is the oracle's bit for bit (test_gpu_parity.py) at sizes
the CPU oracle finishes; at this size the FP64 GPU run stands in for it.
Stated bars: the mixed form within 1e-5 relative L2 per field (u, v, w, b,
p) and in total energy after 500 steps; the zero-order balanced state of the
mixed run within 1e-5 of the FP64 run's (the projection is linear, so the
state gap carries over); the first order is reported (it divides nonlinear
terms by the wave frequency and amplifies the state gap, measured ~100x at
round-off level in test_gpu_balance.py) and bounded at 1e-3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STEPS = 500


def _rel(a, b, prm):
    I = (slice(2, 2 + prm.nx), slice(2, 2 + prm.ny), slice(2, 2 + prm.nz))
    return float(np.linalg.norm(a[I] - b[I]) / np.linalg.norm(b[I]))


def test_jets_eddies_1024_mixed_and_balance():
    from ps3d_amd import configs, diagnose
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_MIXED
    prm, fields = configs.jets_eddies(1024, 1024, 128, seed=5)
    print(f"jets+eddies {prm.nx}x{prm.ny}x{prm.nz}, {STEPS} steps", flush=True)
    names = ("u", "v", "w", "b", "p")
    sel = ("u0", "v0", "w0", "b0", "u1", "b1")

    ref = Model(prm)
    for n, a in fields.items():
        ref.f[n][...] = a
    ref.push()
    ref.run(STEPS)
    ref.pull(list(names))
    final64 = {n: ref.f[n].copy() for n in names}
    bal64 = ref.diag_balance(outputs=sel)
    print("fp64 done", flush=True)

    mix = Model(prm.replace(precision=PS3D_MIXED))
    for n, a in fields.items():
        mix.f[n][...] = a
    mix.push()
    mix.run(STEPS)
    mix.pull(list(names))
    finalmx = {n: mix.f[n].copy() for n in names}
    mix.close()
    errs = diagnose.compare(finalmx, final64, prm)
    print("mixed vs fp64 after", STEPS, "steps:", errs, flush=True)
    assert max(errs.values()) < 1e-5

    for n in ("u", "v", "w", "b"):
        ref.f[n][...] = finalmx[n]
    ref.push(["u", "v", "w", "b"])
    balmx = ref.diag_balance(outputs=sel)
    ref.close()
    gaps = {n: _rel(balmx[n], bal64[n], prm) for n in sel if np.linalg.norm(bal64[n]) > 0}
    print("diag_balance gap (mixed final state vs fp64 final state):", gaps, flush=True)
    assert max(gaps[n] for n in ("u0", "v0", "b0")) < 1e-5
    assert max(gaps[n] for n in ("u1", "b1")) < 1e-3
