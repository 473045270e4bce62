"""Step plan and argument handling of bench.py (CPU only, no library call).

The driver runs `bench.py --gpus 1 --steps K --warmup W` with any K and W; the
plan must never refuse them: the warm-up pass is clamped to the load path and
the timed pass times load steps 1..min(K, n_steps) from the unloaded state.
"""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bench_module():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.parametrize("n_steps,steps,warmup,expect", [
    (20, 20, 5, (5, 20)),    # the driver's command on the 20-step path
    (20, 19, 1, (1, 19)),
    (20, 50, 30, (20, 20)),  # more than the path holds: clamped, not refused
    (20, 3, 0, (0, 3)),
    (10, 0, 0, (0, 1)),      # at least one timed step
    (40, 5, 2, (2, 5)),
])
def test_step_plan(n_steps, steps, warmup, expect):
    assert bench_module().step_plan(n_steps, steps, warmup) == expect


def test_step_plan_rejects_an_empty_path():
    with pytest.raises(ValueError):
        bench_module().step_plan(0, 1, 0)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_help_lists_the_driver_arguments():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0
    for a in ("--gpus", "--steps", "--warmup"):
        assert a in r.stdout


def test_algorithmic_bytes_of_the_node_spmv():
    # nnz values, one column-node index per 3x3 block, 4 words per node, x and y once
    m = bench_module()
    info = {"n": 6_440_067, "nnz": 286_222_473, "ne": 12_582_912, "nn": 2_146_689, "ncolors": 24, "spmv": "node"}
    b = m.kernel_bytes(info, "spmv", 4)
    assert b == 286_222_473 * 4 + (286_222_473 // 9) * 4 + 2_146_689 * 16 + 6_440_067 * 8


def test_committed_traffic_summary_feeds_every_roofline_class():
    """bench.py's roofline.traffic comes from the machine-readable lines of the
    committed PMC summary: every class in both precisions must be there, and
    each figure must be at least 0.85 of the class's algorithmic bytes at
    128^3 hexes (the matrix values alone, which no cache holds, are 0.84 of
    the SpMV's; its vectors may partly come from the 256 MB last-level cache)."""
    import bench

    info = {"n": 6440067, "nnz": 286222473, "ne": 12582912, "nn": 2146689, "spmv": "node", "sweep": "chain"}
    for prec, v in (("fp32", 4), ("fp64", 8)):
        for kind in ("spmv", "ilu_fwd", "ilu_bwd", "ilu_factor"):
            t = bench.measured_traffic(kind, prec, (128, 128, 128))
            assert t is not None, (kind, prec)
            assert t >= 0.85 * bench.kernel_bytes(info, kind, v), (kind, prec, t)
    assert bench.measured_traffic("spmv", "fp32", (64, 64, 64)) is None
