"""C-ABI boundary checks that need no GPU: every function include/shm.h
declares is exported by every project library, the product refuses to run
without a device instead of falling back to host code, and no product
source reaches the oracle."""
from __future__ import annotations

import os
import re
import subprocess
import sys

import pytest

import sledgehamr_amd as shm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shm.h")


def header_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shm_[a-z0-9_]+)\s*\(", text)))


def _library(project):
    if not os.path.exists(shm.library_path(project)):
        from sledgehamr_amd import build
        build.build([project])
    return shm.load_library(project)


def test_header_matches_binding_table():
    assert header_functions() == sorted(shm.EXPORTS)


@pytest.mark.parametrize("project", shm.PROJECTS)
def test_library_exports_every_declared_function(project):
    lib = _library(project)
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def _has_gpu() -> bool:
    try:
        out = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60)
        return "gfx9" in out.stdout
    except (OSError, subprocess.TimeoutExpired):
        return False


def test_create_without_device_fails_loudly():
    """No host fallback: without a GPU the context cannot be created."""
    if _has_gpu():
        pytest.skip("a GPU is present")
    _library("RealScalar")
    with pytest.raises(shm.ShmError) as e:
        shm.Simulation("RealScalar", {"project.name": "RealScalar",
                                      "amr.coarse_level_grid_size": 8,
                                      "sim.L": 1.0, "integrator.type": 21})
    assert e.value.code == 2


def test_product_sources_do_not_reach_the_oracle():
    pat = re.compile(r"(import\s+oracle|from\s+oracle|oracle_py|liboracle|"
                     r"#include\s*[<\"][^>\"]*oracle|[/\\]oracle[/\\])")
    offenders = []
    for base in ("sledgehamr_amd", "include"):
        for dirpath, _, files in os.walk(os.path.join(ROOT, base)):
            for f in files:
                if f.endswith((".py", ".h", ".hpp", ".cpp", ".hip")):
                    p = os.path.join(dirpath, f)
                    if pat.search(open(p, errors="replace").read()):
                        offenders.append(os.path.relpath(p, ROOT))
    assert not offenders, offenders


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
