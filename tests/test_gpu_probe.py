"""The measurement probes the benchmark reports beside the step: HBM rates
of the library's streaming kernels (ps3d_hbm_probe) and the vector FMA
issue rates (ps3d_fma_probe).  They must be positive and below MI355X's
published peaks (8 TB/s HBM; 78.6 TFLOP/s FP64 vector, 157.3 FP32)."""
import pytest

pytestmark = pytest.mark.gpu


def test_probes_report_rates_below_the_peaks():
    import bench
    print("hbm and fma probes", flush=True)
    hbm = bench.hbm_probe(64, reps=2)
    for k in ("copy", "copy_nt", "read", "read_nt", "write", "write_nt"):
        assert 0 < hbm[k] < 8000 * 1.05, (k, hbm[k])
    fma = bench.fma_probe(reps=2)
    assert 0 < fma["fp64_fma"] < 78.6 * 1.05, fma
    assert 0 < fma["fp32_fma"] < 157.3 * 1.05, fma
