"""bench.py keeps its command-line contract (checked on the CPU: argument parsing and the helper
entry points exist; the measurements themselves need the GPU)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_helpers_exist():
    sys.path.insert(0, ROOT)
    import bench
    for name in ("copy_calibration", "cpu_baseline", "error_reports", "batched_sweep", "stress_slab",
                 "large_batch", "measured_traffic", "timed_steps", "main"):
        assert callable(getattr(bench, name)), name


def test_bench_rejects_more_than_one_gpu():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 2 and "one GPU" in res.stderr


def test_measured_traffic_reads_committed_summary():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic("k_g2p_advance")
    assert t is None or t > 100_000_000
    assert bench.measured_traffic("no_such_kernel") is None


def test_valu_floor_reads_the_instruction_profile():
    """roofline.valu_floor prices the committed SQ_INSTS_VALU of a kernel at 4 cycles per wave64
    instruction on 1024 SIMDs at 2.4 GHz, against the measured launch time."""
    sys.path.insert(0, ROOT)
    import bench
    for kernel in ("k_g2p_advance", "k_p2g_chunk", "k_grid"):
        vf = bench.valu_floor(kernel, 100.0)
        assert vf is not None, kernel
        assert vf["floor_us"] == pytest.approx(vf["instructions"] * 4 / (1024 * 2.4e9) * 1e6)
        assert vf["frac"] == pytest.approx(vf["floor_us"] / 100.0)
        assert 0.0 < vf["floor_us"] < 100.0
    assert bench.valu_floor("k_no_such_kernel", 1.0) is None
