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


def test_stage_figures_cover_every_launch_of_the_stage():
    """The P2G stage timer spans the P2G and the fold of continuation partials: its traffic and VALU
    floor are the sums over both launches."""
    sys.path.insert(0, ROOT)
    import bench
    assert bench.STAGE_LAUNCHES["k_p2g_chunk"] == ("k_p2g_chunk", "k_fold_partials")
    p2g, fold = bench.measured_traffic("k_p2g_chunk"), bench.measured_traffic("k_fold_partials")
    if p2g is not None and fold is not None:
        assert bench.stage_traffic("k_p2g_chunk") == p2g + fold
    assert bench.stage_traffic("k_g2p_advance") == bench.measured_traffic("k_g2p_advance")
    rows, col = {}, None  # the INSTS_VALU column of the instruction-counter table
    with open(bench.INSTR_SUMMARY) as f:
        for line in f:
            if not line.startswith("|"):
                col = None
                continue
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if "INSTS_VALU" in cells:
                col = cells.index("INSTS_VALU")
            elif col is not None and cells[0].startswith("snow::"):
                rows[cells[0]] = int(cells[col])
    vf = bench.valu_floor("k_p2g_chunk", 100.0)
    assert vf["instructions"] == rows["snow::k_p2g_chunk"] + rows["snow::k_fold_partials"]
    assert bench.valu_floor("k_grid", 100.0)["instructions"] == rows["snow::k_grid<false, true>"]


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
