"""The CPU side of bench.py: the baseline leg (the oracle stepped in several
threads at once) and the algorithmic byte counts the roofline uses."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_cpu_baseline_threads_step_independent_instances():
    import bench
    from ps3d_amd import configs
    prm, fields = configs.random_state(32, 16, 8, seed=3, lid=True)
    one = bench.cpu_baseline(prm, fields, budget_s=0.2, threads=1)
    two = bench.cpu_baseline(prm, fields, budget_s=0.2, threads=2)
    assert one["cores"] == 1 and two["cores"] == 2
    assert one["kind"] == "port" and one["unit"] == "grid-pt steps/s"
    assert one["value"] > 0 and two["value"] > 0
    assert abs(two["per_thread_value"] * 2 - two["value"]) < 1e-6 * two["value"]


def test_algorithmic_bytes_of_the_friction_kernel():
    import bench
    prm, _ = bench.build_workload()
    n = prm.nx * prm.ny * prm.nz
    # friction + AB3 of four fields: read f and three tendency slots, write f
    assert bench.algorithmic_bytes("friction_ab", prm) == 4 * (2 + 3) * 8 * n


def test_algorithmic_bytes_of_the_fused_tendency_kernel():
    import bench
    prm, _ = bench.build_workload()
    n = prm.nx * prm.ny * prm.nz
    # per field: read f and the two older slots, write slot tau and f
    assert bench.algorithmic_bytes("tend_friction_ab", prm) == 4 * (1 + 2 + 1 + 1) * 8 * n
    # the separate tendency kernel it replaces: read u, v, w, b; write four slots
    assert bench.algorithmic_bytes("tendency", prm) == 8 * 8 * n


def test_reference_cpu_leg_times_the_reference_beside_the_port():
    import pytest
    import bench
    import oracle
    if not oracle.reference_available():
        pytest.skip("reference Fortran build (oracle/_ref) not present")
    from ps3d_amd import configs
    prm, _ = configs.jets(32, 32, 16, seed=2)
    r = bench.reference_cpu_leg(prm, 1, budget_s=0.2, small=True)
    assert r["kind"] == "reference" and r["cores"] == 1
    part = r["integrate_before_pressure"]
    assert part["reference"] > 0 and part["port"] > 0 and r["pressure_solve_port"] > 0
    # a step is the two parts one after the other
    expect = 1.0 / (1.0 / part["reference"] + 1.0 / r["pressure_solve_port"])
    assert abs(r["value"] - expect) <= 1e-9 * expect
