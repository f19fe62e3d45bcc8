"""The batched ensemble against single models.

Members of different sizes, boundaries and parameters advance together, one
launch per stage (and as a replayed hipGraph); each member must come out
BIT FOR BIT equal to a single Model with the same parameters and initial
state -- which is itself bit-identical to the CPU oracle
(test_gpu_parity.py).  The autotuner changes tile shapes only, so results
stay identical after tuning.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "w", "b", "p")


def _members():
    from ps3d_amd import configs
    out = []
    out.append(configs.random_state(32, 32, 16, seed=11, lid=True, Avbi=0.0, Kvbi=0.0))
    out.append(configs.random_state(32, 16, 16, seed=12, lid=False, Avbi=0.0, Kvbi=0.0))
    out.append(configs.eddy(64, 64, 32, seed=13))
    out.append(configs.random_state(96, 48, 12, seed=14, lid=True, Avbi=0.0, Kvbi=0.0, eps_ab=0.0))
    out.append(configs.jets(128, 64, 20, seed=15))
    return out


def _single(prm, fields, steps):
    from ps3d_amd.model import Model
    m = Model(prm)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()
    m.run(steps)
    m.pull()
    out = {n: m.f[n].copy() for n in NAMES}
    m.close()
    return out


def _ensemble(members):
    from ps3d_amd.ensemble import Ensemble
    ens = Ensemble([p for p, _ in members])
    for i, (_, fields) in enumerate(members):
        for n, a in fields.items():
            ens.set(i, n, a)
    return ens


@pytest.mark.parametrize("use_graph", [False, True])
def test_ensemble_matches_single_models(use_graph):
    members = _members()
    ens = _ensemble(members)
    print(f"ensemble of {len(members)} members, graph={use_graph}", flush=True)
    ens.run(3, use_graph=use_graph)      # Euler, AB2, AB3
    ens.run(9, use_graph=use_graph)
    for i, (prm, fields) in enumerate(members):
        ref = _single(prm, fields, 12)
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, ref[n]), (i, n, float(np.abs(got - ref[n]).max()))
    ens.close()


def test_ensemble_tuning_keeps_results(tmp_path):
    members = _members()
    ens = _ensemble(members)
    ens.run(4)
    cache = str(tmp_path / "tune.txt")
    ens.tune(cache, reps=3)
    desc = ens.describe()
    print(desc, flush=True)
    assert "tendency" in desc and "friction_ab" in desc
    ens.run(6)
    for i, (prm, fields) in enumerate(members):
        ref = _single(prm, fields, 10)
        for n in NAMES:
            assert np.array_equal(ens.get(i, n), ref[n]), (i, n)
    ens.close()
    # a second ensemble takes the cached choice without timing
    ens2 = _ensemble(members[:2])
    ens2.tune(cache, reps=3)
    assert ens2.describe().splitlines()[0] in desc
    ens2.close()
