"""The benchmark's contract and its roofline arithmetic, on the CPU: more
than one GPU is refused with a message, and the dominant kernel's prices
follow from a synthetic timing record (figures worked by hand below)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 1" in r.stderr
    assert r.stdout == ""


def test_fused_schedule_roofline():
    """512^3, 20 steps of one fused stage-1+2 launch (1.5 ms) and one
    stage-3 launch (1.7 ms): stage 3 is the dominant kernel."""
    n, steps = 512 ** 3, 20
    tm = {"stage_launches": 2 * steps, "fused_launches": steps,
          "stage_cells": 3 * n * steps, "stage_ms": steps * (1.5 + 1.7),
          "fused_ms": steps * 1.5}
    r = bench.kernel_roofline(tm, "fp32c16", True, 512)
    assert r["kernel"].startswith("shm::stage_kernel")
    assert r["avg_launch_ms"] == pytest.approx(1.7)
    # 4 components x (12 words of 4 B + a 2 B remainder read and written)
    # per step, over 3 stages: 69.33 B per cell-stage.
    alg = 4 * (12 * 4 + 2 * 2) / 3 * n
    assert r["algorithmic_bytes_per_launch"] == pytest.approx(alg)
    assert r["achieved"] == pytest.approx(alg / 1.7e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["peak"] == 8000.0 and r["unit"] == "GB/s" and r["bound"] == "hbm"
    # Stage 3 reads u0 and k2 (4 B each) and u0's remainder (2 B), writes u
    # (4 B) and its remainder (2 B): 16 B per component.
    assert r["moved_bytes_per_launch"] == pytest.approx(4 * 16 * n)
    assert r["traffic"] == bench.TRAFFIC_BYTES_512["stage3"]
    o = r["other_kernel"]
    assert o["avg_launch_ms"] == pytest.approx(1.5)
    # Two cell-stages per cell; moves u0 in and k2 out (4 B each).
    assert o["algorithmic_bytes_per_launch"] == pytest.approx(2 * alg)
    assert o["moved_bytes_per_launch"] == pytest.approx(4 * 8 * n)


def test_unfused_schedule_roofline():
    n = 64 ** 3
    tm = {"stage_launches": 30, "fused_launches": 0, "stage_cells": 30 * n,
          "stage_ms": 30 * 0.01, "fused_ms": 0.0}
    r = bench.kernel_roofline(tm, "fp64", False, 64)
    assert r["avg_launch_ms"] == pytest.approx(0.01)
    # 4 components x 12 words of 8 B per step over 3 stages: 128 B.
    assert r["algorithmic_bytes_per_launch"] == pytest.approx(128 * n)
    assert r["traffic"] is None


def test_ensemble_bytes():
    """RKN4 on one field and its momentum: stage 1 reads 2 words and writes
    1, stage 2 reads 3 and writes 1, stage 3 reads 4 and writes 2; with FP32
    + bf16 both remainders are read and written once (4 x 2 B)."""
    words = (2 + 1) + (3 + 1) + (4 + 2)
    assert bench.ENSEMBLE_BYTES_PER_CELL_STEP["fp32c16"] == words * 4 + 4 * 2
    assert bench.ENSEMBLE_BYTES_PER_CELL_STEP["fp64"] == words * 8
