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
