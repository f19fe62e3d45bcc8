"""ps3d_run with use_graph: the steps replayed from a captured hipGraph of six
steps give the same bits as the same steps queued one by one -- FP64 and mixed,
rigid lid and periodic, with background forcing, across calls whose step counts
are not multiples of six, and after the launch shapes or radix plans change
(which takes a new capture)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("u", "v", "w", "b", "p")


def _model(prm, fields, forcing=None):
    from ps3d_amd.model import Model
    m = Model(prm)
    for n, a in fields.items():
        m.f[n][...] = a
    m.push()
    if forcing is not None:
        ub, bb, force = forcing
        m.set_background_flow(ub, bb)
        m.set_body_force(force)
    return m


def _case(name):
    import dataclasses
    from ps3d_amd import configs
    from ps3d_amd.params import PS3D_MIXED
    lid = not name.startswith("periodic")
    prm, fields = configs.random_state(32, 24, 16, seed=51, lid=lid, Avbi=0.0, Kvbi=0.0)
    if "mixed" in name:
        prm = dataclasses.replace(prm, precision=PS3D_MIXED)
    forcing = None
    if "forced" in name:
        from test_gpu_lee_wave import _forcing
        forcing = _forcing(prm, seed=9)
    return prm, fields, forcing


@pytest.mark.parametrize("name", ["lid", "periodic", "lid_mixed", "periodic_mixed", "lid_forced",
                                  "lid_forced_mixed"])
def test_graph_replay_gives_the_same_bits(name):
    prm, fields, forcing = _case(name)
    print(f"graph replay: {name}", flush=True)
    eager = _model(prm, fields, forcing)
    graph = _model(prm, fields, forcing)
    for n in (5, 13, 1, 12, 7):
        eager.run(n)
        graph.run(n, use_graph=True)
    assert (graph.itt, graph.tau, graph.taum1, graph.taum2) == (eager.itt, eager.tau, eager.taum1, eager.taum2)
    eager.pull()
    graph.pull()
    for n in NAMES:
        assert np.array_equal(graph.f[n], eager.f[n]), (name, n)
    eager.close()
    graph.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_recaptured_when_the_step_changes(mixed, tmp_path):
    """A captured graph is not replayed once the launch shapes (and, mixed,
    the radix plans) change: the next call captures again."""
    prm, fields, _ = _case("lid_mixed" if mixed else "lid")
    eager = _model(prm, fields)
    graph = _model(prm, fields)
    eager.run(14)
    graph.run(14, use_graph=True)
    shape = f"{prm.nx}x{prm.ny}x{prm.nz}:lid:{'mixed' if mixed else 'fp64'}"
    cache = tmp_path / "shapes.txt"
    lines = [f"tendency {shape} 4", f"friction_ab {shape} 2"]
    if mixed:
        lines += [f"fft_plan_x {shape} 1", f"fft_plan_y {shape} 1"]
    else:
        lines += [f"thomas {shape} 16"]
    cache.write_text("\n".join(lines) + "\n")
    for m in (eager, graph):
        m.tune(str(cache), reps=1)
    assert "friction_ab " + shape + " 2" in graph.tuning()
    eager.run(13)
    graph.run(13, use_graph=True)
    eager.pull()
    graph.pull()
    for n in NAMES:
        assert np.array_equal(graph.f[n], eager.f[n]), n
    eager.close()
    graph.close()
