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


def _single(prm, fields, steps, plans=None, plans_from=0):
    """A single model run for steps steps; plans: the FP32 radix plans it
    switches to after plans_from steps (as an ensemble tuned then)."""
    from ps3d_amd.model import Model
    m = Model(prm)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()
    if plans is not None:
        m.run(plans_from)
        m.fft_plan = plans
        steps -= plans_from
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


def test_mixed_ensemble_radix8_plans_match_single_models():
    """Members set to the radix-8 plan on every axis (plan 1; the lengths
    without a factor 8 run plan 0 either way) equal single mixed models with
    the same plans bit for bit; a FP64 ensemble refuses plan 1."""
    members = _mixed_members()
    ens = _ensemble(members)
    for i in range(len(members)):
        ens.set_fft_plan(i, (1, 1, 1))
        assert ens.fft_plan(i) == (1, 1, 1)
    ens.run(8, use_graph=True)
    for i, (prm, fields) in enumerate(members):
        ref = _single(prm, fields, 8, plans=(1, 1, 1), plans_from=0)
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, ref[n]), (i, n, float(np.abs(got - ref[n]).max()))
    ens.close()
    from ps3d_amd import _lib
    e64 = _ensemble(_members()[:1])
    with pytest.raises(_lib.Ps3dError, match="plan 0"):
        e64.set_fft_plan(0, (1, 0, 0))
    e64.close()


def test_mixed_ensemble_tuning_keeps_results(tmp_path):
    members = _mixed_members()
    ens = _ensemble(members)
    ens.run(3)
    ens.tune(str(tmp_path / "tune.txt"), reps=2)
    desc = ens.describe()
    print(desc, flush=True)
    assert ":mixed" in desc
    # the radix plans picked per transform length (lengths with a factor 8)
    assert "ens_fft_plan_x 64:mixed" in desc and "ens_fft_plan_y 64:mixed" in desc
    ens.run(3)
    for i, (prm, fields) in enumerate(members):
        ref = _single(prm, fields, 6, plans=ens.fft_plan(i), plans_from=3)
        for n in NAMES:
            assert np.array_equal(ens.get(i, n), ref[n]), (i, n)
    ens.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_tall_members_take_the_register_thomas_forms(mixed):
    """Rigid-lid members of 96 and 128 levels beside one of 32: every member
    has a register Thomas length, so the stage runs k_ens_thomas_reg with
    whole columns for 32 levels and the real and imaginary parts on two
    lanes for 96 and 128 (32 columns per work-group; 16x16 and 24x32
    grids leave the last work-group of a member partly empty); members
    bit-identical to single models of their form."""
    import dataclasses
    from ps3d_amd import configs
    from ps3d_amd.params import PS3D_MIXED
    members = [configs.random_state(16, 16, 128, seed=31, lid=True, Avbi=0.0, Kvbi=0.0),
               configs.random_state(32, 24, 96, seed=32, lid=True, Avbi=0.0, Kvbi=0.0),
               configs.eddy(64, 64, 32, seed=33)]
    if mixed:
        members = [(dataclasses.replace(p, precision=PS3D_MIXED), f) for p, f in members]
    ens = _ensemble(members)
    print(f"tall members, mixed={mixed}", flush=True)
    ens.run(4)
    ens.run(4, use_graph=True)
    for i, (prm, fields) in enumerate(members):
        ref = _single(prm, fields, 8)
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, ref[n]), (i, n, float(np.abs(got - ref[n]).max()))
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
        ref = _single(prm, fields, 500, plans=ens.fft_plan(i) if mixed else None, plans_from=3)
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, ref[n]), (i, n, float(np.abs(got - ref[n]).max()))
    ens.close()


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("use_graph", [False, True])
def test_forced_members_match_single_forced_models(use_graph, mixed):
    """Forcing in the batched ensemble: members with background-flow
    advection (lee_wave1.f90:239-314) and a body force, next to unforced
    ones, each bit-identical to a single model of its form with the same
    forcing -- the friction chain into the third buffers ((hi, lo) pairs in
    the mixed form), the pre-friction halos copied, then forcing +
    Adams-Bashforth, as a single forced model steps; the mixed form through
    its cold start (forcing on the pre-friction state) and correction
    steps."""
    import dataclasses
    from ps3d_amd import configs
    from ps3d_amd.ensemble import Ensemble
    from ps3d_amd.model import Model
    from ps3d_amd.params import PS3D_FP64, PS3D_MIXED
    from test_gpu_lee_wave import _forcing
    members = [configs.random_state(32, 24, 12, seed=6, lid=True, Avbi=0.0, Kvbi=0.0),
               configs.random_state(24, 16, 8, seed=7, lid=False, Avbi=0.0, Kvbi=0.0),
               configs.random_state(16, 16, 16, seed=8, lid=True, Avbi=0.0, Kvbi=0.0)]
    members = [(dataclasses.replace(p, precision=PS3D_MIXED if mixed else PS3D_FP64), f) for p, f in members]
    forcing = {0: _forcing(members[0][0], seed=4), 1: _forcing(members[1][0], seed=5)}
    forcing[1] = (forcing[1][0], None, forcing[1][2])   # no background buoyancy
    ens = Ensemble([p for p, _ in members])
    for i, (_, fields) in enumerate(members):
        for k, a in fields.items():
            ens.set(i, k, a)
    for i, (ub, bb, force) in forcing.items():
        ens.set_background_flow(i, ub, bb)
        ens.set_body_force(i, force)
    ens.run(12, use_graph=use_graph)
    singles = []
    for i, (prm, fields) in enumerate(members):
        m = Model(prm)
        for k, a in fields.items():
            m.f[k][...] = a
        m.push()
        if i in forcing:
            ub, bb, force = forcing[i]
            m.set_background_flow(ub, bb)
            m.set_body_force(force)
        m.run(12)
        m.pull()
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, m.f[n]), (i, n, float(np.abs(got - m.f[n]).max()))
        singles.append(m)
    # forcing switched off for member 0 (its background flow and force) and
    # member 1's body force: 4 more steps
    ens.set_background_flow(0, None)
    ens.set_body_force(0, None)
    ens.set_body_force(1, None)
    singles[0].set_background_flow(None)
    singles[0].set_body_force(None)
    singles[1].set_body_force(None)
    ens.run(4, use_graph=use_graph)
    for i, m in enumerate(singles):
        m.run(4)
        m.pull()
        for n in NAMES:
            got = ens.get(i, n)
            assert np.array_equal(got, m.f[n]), ("after switching off", i, n, float(np.abs(got - m.f[n]).max()))
        m.close()
    ens.close()
