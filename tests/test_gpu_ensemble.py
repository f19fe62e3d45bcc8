"""The batched ensemble against single models.

Members of different sizes, boundaries and parameters advance together, one
launch per stage (and as a replayed hipGraph); each member must come out
BIT FOR BIT equal to a single Model with the same parameters and initial
state -- which is itself bit-identical to the CPU oracle
(test_gpu_parity.py).  The autotuner changes tile shapes only, so results
stay identical after tuning.  An ensemble of mixed-precision members must
likewise equal single mixed models bit for bit, through the cold start and
the two correction steps that follow it.
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
    assert "ens_tendency" in desc and "ens_friction_ab" in desc
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


def _mixed_members():
    import dataclasses
    from ps3d_amd import configs
    from ps3d_amd.params import PS3D_MIXED
    out = []
    for prm, fields in (configs.random_state(32, 32, 16, seed=21, lid=True, Avbi=0.0, Kvbi=0.0),
                        configs.random_state(32, 16, 16, seed=22, lid=False, Avbi=0.0, Kvbi=0.0),
                        configs.eddy(64, 64, 32, seed=23),
                        configs.jets(128, 64, 20, seed=25)):
        out.append((dataclasses.replace(prm, precision=PS3D_MIXED), fields))
    return out


@pytest.mark.parametrize("use_graph", [False, True])
def test_mixed_ensemble_matches_single_mixed_models(use_graph):
    members = _mixed_members()
    ens = _ensemble(members)
    print(f"mixed ensemble of {len(members)} members, graph={use_graph}", flush=True)
    ens.run(2, use_graph=use_graph)      # cold start + the two correction steps
    ens.run(6, use_graph=use_graph)
    for i, (prm, fields) in enumerate(members):
        ref = _single(prm, fields, 8)
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, ref[n]), (i, n, float(np.abs(got - ref[n]).max()))
    ens.close()


def test_mixed_ensemble_tuning_keeps_results(tmp_path):
    members = _mixed_members()
    ens = _ensemble(members)
    ens.run(3)
    ens.tune(str(tmp_path / "tune.txt"), reps=2)
    desc = ens.describe()
    print(desc, flush=True)
    assert ":mixed" in desc
    ens.run(3)
    for i, (prm, fields) in enumerate(members):
        ref = _single(prm, fields, 6)
        for n in NAMES:
            assert np.array_equal(ens.get(i, n), ref[n]), (i, n)
    ens.close()


def test_ensemble_rejects_mixed_forms():
    import dataclasses
    from ps3d_amd.params import PS3D_MIXED
    from ps3d_amd.ensemble import Ensemble
    members = _members()[:2]
    prms = [members[0][0], dataclasses.replace(members[1][0], precision=PS3D_MIXED)]
    with pytest.raises(Exception, match="share one precision form"):
        Ensemble(prms)


@pytest.mark.parametrize("mixed", [False, True])
def test_c4_ensemble_500_steps(mixed):
    """The 48-member ensemble of three sizes over its 500 steps (one launch
    per stage, graph replay, tuned): one member of each size equals a
    single model of its form run alone, bit for bit."""
    from ps3d_amd import configs
    from ps3d_amd.params import PS3D_FP64, PS3D_MIXED
    members = configs.ensemble_c4(precision=PS3D_MIXED if mixed else PS3D_FP64)
    ens = _ensemble(members)
    print(f"C4 ensemble, {len(members)} members, mixed={mixed}, 500 steps", flush=True)
    ens.run(3)
    ens.tune(None, reps=2)
    ens.run(497, use_graph=True)
    for i in (0, 16, 47):
        prm, fields = members[i]
        ref = _single(prm, fields, 500)
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, ref[n]), (i, n, float(np.abs(got - ref[n]).max()))
    ens.close()
