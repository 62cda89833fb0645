"""Builds the gfx950 libraries: one ``libshm_<Project>.so`` per project.

The project registry below plays the role of the reference's Make.projects
(Make.projects:1-18): each entry names the project header (compiled
unchanged), its class, and which kernel variants to instantiate.

Usage: ``python -m sledgehamr_amd.build [-j N] [--projects A,B]``.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from dataclasses import dataclass

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
BUILD = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"


@dataclass(frozen=True)
class Project:
    name: str          # library name and project.name in the inputs file
    header: str        # header path relative to the repository root
    cls: str           # fully qualified class
    r_mask: int        # bit r: stencil radius r instantiated
    gw_modes: int = 1  # bit 0: without the GW tensor, bit 1: with it
    ko: bool = False   # Kreiss-Oliger variants


PROJECTS = [
    Project("RealScalar", "sledgehamr_amd/projects/RealScalar/RealScalar.h",
            "RealScalar::RealScalar", 0b1110, ko=True),
    Project("ComplexScalar", "sledgehamr_amd/projects/ComplexScalar/ComplexScalar.h",
            "ComplexScalar::ComplexScalar", 0b0100),
    Project("PhaseTransitionGW",
            "sledgehamr_amd/projects/PhaseTransitionGW/PhaseTransitionGW.h",
            "PhaseTransitionGW::PhaseTransitionGW", 0b1000, gw_modes=2),
]

# Multiply-adds are fused per source expression (-ffp-contract=on), not
# wherever the optimised code happens to place a multiply next to an add:
# kernels that evaluate the same right-hand side in different code (the
# fused stages 1 + 2 and the separate stage launches, the cells of a tile
# and the ring cells of its neighbour) then round identically.
COMMON_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=on", "-munsafe-fp-atomics",
    "-Wno-unused-result", "-Wno-unknown-pragmas",
    "-I", INCLUDE, "-I", os.path.join(INCLUDE, "sledgehamr"), "-I", CSRC,
]

RUNTIME_SOURCES = ["shm_runtime.cpp", "shm_capi.cpp", "shm_io.cpp"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {cmd[-1]}")


def _compile_runtime(src: str) -> str:
    os.makedirs(BUILD, exist_ok=True)
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    _run([HIPCC, *COMMON_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src),
          "-o", obj])
    return obj


def _project_units(p: Project) -> list[tuple[str, list[str]]]:
    """(source, extra flags) of the project's translation units: one per
    (precision, stencil radius) holding the fused stage kernels, and a main
    unit with the tagging/probe kernels and the project registration."""
    os.makedirs(BUILD, exist_ok=True)
    units = [("main", [])]
    for prec in (0, 1, 2, 3):
        for r in (1, 2, 3):
            if (p.r_mask >> r) & 1:
                units.append((f"p{prec}_r{r}",
                              [f"-DSHM_TU_PREC={prec}", f"-DSHM_TU_R={r}"]))
    out = []
    for tag, flags in units:
        tu = os.path.join(BUILD, f"proj_{p.name}_{tag}.hip")
        text = (
            f'#include "{os.path.join(ROOT, p.header)}"\n'
            f"#define SHM_PROJECT_CLASS {p.cls}\n"
            '#include "shm_project_impl.hpp"\n'
        )
        if not os.path.exists(tu) or open(tu).read() != text:
            with open(tu, "w") as f:
                f.write(text)
        out.append((tu, flags))
    return out


def _compile_unit(p: Project, tu: str, flags: list[str]) -> str:
    obj = os.path.splitext(tu)[0] + ".o"
    _run([HIPCC, *COMMON_FLAGS, f"-DSHM_R_MASK={p.r_mask}",
          f"-DSHM_GW_MODES={p.gw_modes}", f"-DSHM_WITH_KO={int(p.ko)}", *flags,
          "-c", tu, "-o", obj])
    return obj


def _link(p: Project, pobjs: list[str], robjs: list[str]) -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, f"libshm_{p.name}.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *pobjs, *robjs,
          "-L/opt/rocm/lib", "-lhipfft", "-ldl", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    return out


def build(projects: list[str] | None = None, jobs: int = 8) -> list[str]:
    """Compiles the runtime and every selected project; returns .so paths."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found")
    sel = [p for p in PROJECTS if projects is None or p.name in projects]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        rfut = [ex.submit(_compile_runtime, s) for s in RUNTIME_SOURCES]
        pfut = {p.name: [ex.submit(_compile_unit, p, tu, fl)
                         for tu, fl in _project_units(p)] for p in sel}
        robjs = [f.result() for f in rfut]
        pobjs = {k: [f.result() for f in v] for k, v in pfut.items()}
    return [_link(p, pobjs[p.name], robjs) for p in sel]


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--projects", default="")
    a = ap.parse_args()
    sel = [s for s in a.projects.split(",") if s] or None
    for path in build(sel, a.j):
        print(path)


if __name__ == "__main__":
    main()
