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
    stage-3 launch (1.7 ms): stage 3 is the dominant kernel, priced at what
    it must move."""
    n, steps = 512 ** 3, 20
    tm = {"stage_launches": 2 * steps, "fused_launches": steps,
          "stage_cells": 3 * n * steps, "stage_ms": steps * (1.5 + 1.7),
          "fused_ms": steps * 1.5}
    launch = dict(bench.PROFILED_SHAPE)
    r = bench.headline_roofline(tm, "fp32c16", launch, n)
    assert r["kernel"].startswith("shm::stage_kernel")
    assert r["avg_launch_ms"] == pytest.approx(1.7)
    # Stage 3 reads u0 and k2 (4 B each) and u0's remainder (2 B), writes u
    # (4 B) and its remainder (2 B): 16 B per component, 4 components.
    assert r["bytes_per_launch"] == pytest.approx(4 * 16 * n)
    assert r["achieved"] == pytest.approx(4 * 16 * n / 1.7e-3 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0) and r["frac"] < 1
    assert r["peak"] == 8000.0 and r["unit"] == "GB/s" and r["bound"] == "hbm"
    assert r["traffic"] == bench.TRAFFIC_BYTES_512["stage3"]
    assert r["traffic_source"] == bench.TRAFFIC_PROFILE
    o = r["other_kernel"]
    assert o["avg_launch_ms"] == pytest.approx(1.5)
    # Moves u0 in and k2 out (4 B each).
    assert o["bytes_per_launch"] == pytest.approx(4 * 8 * n)
    assert o["frac"] < 1
    # Another launch shape than the profiled one (any field): no traffic.
    for field, v in (("fused_variant", 0), ("variant", 4), ("xcd_split", 0),
                     ("tile_z", 0)):
        r2 = bench.headline_roofline(tm, "fp32c16", dict(launch, **{field: v}), n)
        assert r2["traffic"] is None, field
    # The headline's default launch string is the profiled shape.
    keys = ("variant", "tile_z", "xcd_split", "persistent", "zc_minor", "fused",
            "fused_variant")
    assert dict(zip(keys, map(int, bench.HEADLINE_LAUNCH.split(",")))) == \
        bench.PROFILED_SHAPE


def test_unfused_schedule_roofline():
    n = 64 ** 3
    tm = {"stage_launches": 30, "fused_launches": 0, "stage_cells": 30 * n,
          "stage_ms": 30 * 0.01, "fused_ms": 0.0}
    r = bench.headline_roofline(tm, "fp64", {"fused": 0}, n)
    assert r["avg_launch_ms"] == pytest.approx(0.01)
    # Per component: stage 1 u0 in, k1 out; stage 2 u0, k1 in, k2 out;
    # stage 3 u0, k2 in, u out: 8 words of 8 B over 3 launches.
    assert r["bytes_per_launch"] == pytest.approx(4 * 8 * 8 * n / 3)
    assert r["traffic"] is None


def test_compensated_stencil_bytes():
    """With FP64 tiles the stages read u0's remainder wherever they read u0:
    SSPRK3 stage one 4 + 2 + 4 B per component, stage three as FP32C16."""
    assert bench.ssprk_bytes("fp32cs", "stage1", 4) == 4 * 10
    assert bench.ssprk_bytes("fp32cs", "fused12", 4) == 4 * 10
    assert bench.ssprk_bytes("fp32cs", "stage3", 4) == bench.ssprk_bytes("fp32c16", "stage3", 4)


def test_rkn_bytes():
    """RKN4 on one field and its momentum: stage 1 reads 2 words and writes
    1, stage 2 reads 3 and writes 1, stage 3 reads 4 and writes 2; with FP32
    + bf16 both remainders are read and written once (4 x 2 B); with FP64
    tiles of the pair the first two stages also read both remainders."""
    words = (2 + 1) + (3 + 1) + (4 + 2)
    assert bench.rkn_bytes_per_step("fp32c16", 1) == words * 4 + 4 * 2
    assert bench.rkn_bytes_per_step("fp64", 1) == words * 8
    assert bench.rkn_bytes_per_step("fp32cs", 1) == words * 4 + 4 * 2 + 4 * 2
    assert bench.rkn_bytes_per_step("fp32c16", 7) == 7 * (words * 4 + 4 * 2)
