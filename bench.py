#!/usr/bin/env python3
"""Benchmark of the per-level time step on one MI355X.

Workload (default): complex scalar on a 512^3 periodic level, 4th-order
Laplacian, V = (lambda/4)(|psi|^2-1)^2 with lambda = 1 and radiation-era
damping 2/eta, low-storage SSPRK3 (integrator.type = 10), L = 128
(dx = 0.25), cfl = 0.2 (dt = 0.05), eta starting at 0.1; initial phases
seeded uniform in [0, 2 pi), moduli 1 + N(0, 0.05), momenta 0.

A "step" is one SSPRK3 step of the level (3 RK stages).  `value` is
cell-updates per second per RK stage of the FP32-compensated path (FP32
state + bf16 remainder; valid cells x stages / wall time), inputs resident in
HBM.  The same seeded run on the FP64 path gives `fp64` and the field error;
`fp32c` is the FP32 + FP32-remainder variant.

roofline: the dominant stage kernel (longest mean launch; the other one in
`other_kernel`), its algorithmic bytes per launch over its mean launch time,
measured with HIP events on the library's stream inside the timed region.
Algorithmic bytes per cell-stage are those of the minimal low-storage
schedule that processes one stage per pass (every stored array read once and
written at most once per stage: 12 words per component and step, plus one
read and one write of the remainder per step): 69.3 B for FP32 + bf16
remainder, 74.7 B with an FP32 remainder, 128 B for FP64 (C = 4).  This
library moves fewer: stage states are rebuilt from u0 and k (8 words per
component and step: 48 / 53.3 / 85.3 B per cell-stage), and on a level of
periodic boxes stages one and two run in one pass that keeps k1 on chip (5
words: 32 / 37.3 / 53.3 B); those bytes give `moved_frac`.  `traffic` is the
HBM bytes per launch measured by rocprofv3 counters (DRAM read and write
requests of the same stage launches).

cpu_baseline: the FP64 oracle in the reference's unfused structure (Rhs pass,
LinComb passes, ghost fill per stage; OpenMP), whole SSPRK3 steps of the same
512^3 workload on the host cores after one warm-up step.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = ("cell-updates/s per RK stage + % HBM roofline; FP32-comp vs FP64 "
          "at ≤1e-6 rel err")
# Algorithmic bytes per cell-stage, C = 4 components, SSPRK3: 12 words per
# component and step (each stored array read once, written at most once per
# stage), plus one read and one write of the remainder per step.
C = 4
BYTES_PER_CELL_STAGE = {"fp32c16": (12 * 4 + 2 * 2) * C / 3.0,
                        "fp32c": (12 * 4 + 2 * 4) * C / 3.0,
                        "fp64": 12 * 8 * C / 3.0}
# Bytes this schedule moves: stage states rebuilt from u0 and k, so per
# component and step stage 1 reads u0, writes k1; stage 2 reads u0, k1, writes
# k2; stage 3 reads u0, k2, writes u (8 words), plus the remainder.
MOVED_BYTES_PER_CELL_STAGE = {"fp32c16": (8 * 4 + 2 * 2) * C / 3.0,
                              "fp32c": (8 * 4 + 2 * 4) * C / 3.0,
                              "fp64": 8 * 8 * C / 3.0}
# With stages one and two fused (k1 kept on chip): stage 1+2 reads u0 and
# writes k2, stage 3 reads u0, k2 and writes u (5 words), plus the remainder.
MOVED_BYTES_FUSED = {"fp32c16": (5 * 4 + 2 * 2) * C / 3.0,
                     "fp32c": (5 * 4 + 2 * 4) * C / 3.0,
                     "fp64": 5 * 8 * C / 3.0}


def moved_bytes(prec_key, fused):
    return (MOVED_BYTES_FUSED if fused else MOVED_BYTES_PER_CELL_STAGE)[prec_key]


# Bytes per valid cell each kernel of the fused schedule moves: the fused
# stage-one-and-two launch reads u0 (leading words only) and writes k2; the
# stage-three launch reads u0 with its remainder and k2 and writes u with its
# remainder.
WORD = {"fp32c16": (4, 2), "fp32c": (4, 4), "fp64": (8, 0)}


def moved_per_cell(prec_key, kernel):
    w, r = WORD[prec_key]
    return C * (2 * w if kernel == "fused12" else 3 * w + 2 * r)


# HBM bytes per launch at 512^3, FP32 + bf16 remainder, fused schedule: DRAM
# read + write requests counted by rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B
# TCC_EA0_WRREQ_WRITE_DRAM_32B (x 32 B), profiles/stage_kernel_traffic_complex512.md.
TRAFFIC_BYTES_512 = {"fused12": (74270667 + 67108864) * 32.0,
                     "stage3": (174360952 + 100663296) * 32.0}


def kernel_record(name, avg_ms, alg_bytes, moved, traffic):
    rec = {"kernel": name, "avg_launch_ms": avg_ms,
           "algorithmic_bytes_per_launch": alg_bytes,
           "achieved": alg_bytes / (avg_ms * 1e-3) / 1e9,
           "moved_bytes_per_launch": moved,
           "moved_achieved": moved / (avg_ms * 1e-3) / 1e9, "traffic": traffic}
    rec["frac"] = rec["achieved"] / HBM_PEAK_GBS
    rec["moved_frac"] = rec["moved_achieved"] / HBM_PEAK_GBS
    if traffic:
        rec["traffic_achieved"] = traffic / (avg_ms * 1e-3) / 1e9
    return rec


def kernel_roofline(tm, prec_key, fused, N):
    """Roofline of the dominant stage kernel (longest mean launch) from the
    HIP-event timings.  `achieved` prices a cell-stage at the algorithmic
    bytes; `moved_*` at what this schedule must read and write."""
    launches = max(tm["stage_launches"], 1)
    contract = BYTES_PER_CELL_STAGE[prec_key]
    if fused and tm["fused_launches"]:
        fl = tm["fused_launches"]
        # Each fused launch updates ncell cells twice, each stage-3 launch once.
        ncell = tm["stage_cells"] / (2 * fl + (launches - fl))
        t = TRAFFIC_BYTES_512 if N == 512 and prec_key == "fp32c16" else {}
        recs = [kernel_record("shm::fused12_kernel (SSPRK3 stages 1+2)",
                              tm["fused_ms"] / fl, 2 * contract * ncell,
                              moved_per_cell(prec_key, "fused12") * ncell,
                              t.get("fused12")),
                kernel_record("shm::stage_kernel (SSPRK3 stage 3)",
                              (tm["stage_ms"] - tm["fused_ms"]) / max(launches - fl, 1),
                              contract * ncell,
                              moved_per_cell(prec_key, "stage3") * ncell,
                              t.get("stage3"))]
        recs.sort(key=lambda r: -r["avg_launch_ms"])
        main, other = recs
        res = {"bound": "hbm", "achieved": main["achieved"], "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "frac": main["frac"], "traffic": main["traffic"]}
        res.update(main)
        res["other_kernel"] = other
        res["note"] = ("per step: one fused stage-1+2 launch (two cell-stages per "
                       "cell, k1 and u1 kept on chip) and one stage-3 launch")
        return res
    cells_pl = tm["stage_cells"] / launches
    main = kernel_record("shm::stage_kernel (SSPRK3 stages 1..3, mean)",
                         tm["stage_ms"] / launches, contract * cells_pl,
                         moved_bytes(prec_key, False) * cells_pl, None)
    res = {"bound": "hbm", "achieved": main["achieved"], "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": main["frac"], "traffic": None}
    res.update(main)
    return res


def inputs(N, L, cfl, t0):
    return {
        "project.name": "ComplexScalar",
        "project.lambda": 1.0,
        "sim.t_start": t0,
        "sim.t_end": 2.0,
        "sim.L": L,
        "sim.cfl": cfl,
        "amr.coarse_level_grid_size": N,
        "amr.nghost": 2,
        "integrator.type": 10,
    }


def run_gpu(shm, inp, prec, steps, warmup, seed, launch=None):
    sim = shm.Simulation("ComplexScalar", inp, precision=prec)
    sim.init_state(0, 1, seed, [0.05])
    if launch:
        sim.set_launch(*launch)
    else:
        sim.autotune(0)
    sim.advance(0, warmup)
    sim.sync()
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.advance(0, steps)
    sim.sync()
    t1 = time.perf_counter()
    tm = sim.timing()
    sim.set_timing(False)
    return sim, t1 - t0, tm


def graph_ms_per_step(sim, steps):
    """Wall time per step with each level step replayed from a captured
    graph (the first step after enabling captures it)."""
    sim.set_graphs(True)
    sim.advance(0, 2)
    sim.sync()
    t0 = time.perf_counter()
    sim.advance(0, steps)
    sim.sync()
    ms = (time.perf_counter() - t0) / steps * 1e3
    sim.set_graphs(False)
    return ms


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inp, seed):
    """FP64 oracle, whole SSPRK3 steps after one warm-up step, until about
    10 s of work (at most 8 steps)."""
    from oracle_py import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    orc = Oracle("ComplexScalar", inp)
    orc.init_state(0, 1, seed, [0.05])
    n = orc.ncells(0)
    orc.advance(0, 1)
    steps = 0
    t0 = time.perf_counter()
    while steps < 8 and (steps == 0 or time.perf_counter() - t0 < 10.0):
        orc.advance(0, 1)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": n * 3 * steps / dt, "unit": "cell-updates/s",
            "cores": threads, "kind": "port",
            "sample": f"{steps} SSPRK3 step(s) (3 stages each) of the same "
                      f"{round(n ** (1/3))}^3 workload after 1 warm-up step, FP64 "
                      "oracle in the reference's unfused structure (Rhs pass, "
                      "LinComb passes, ghost fill per stage), OpenMP "
                      f"{threads} threads on {cpu_model()}, {dt:.1f} s"}


# Bytes per cell and RKN4 step of the ensemble (real scalar: one field and
# its momentum; three stages, each one pass): stage 1 reads phi0, pi0 and
# writes F0; stage 2 reads phi0, pi0, F0 and writes F1; stage 3 reads phi0,
# pi0, F0, F1 and the two remainders and writes phi, pi and their remainders.
# No pass can do with less, so these are both the algorithmic and the moved
# bytes.
ENSEMBLE_BYTES_PER_CELL_STEP = {"fp32c16": 3 * 4 + 4 * 4 + 6 * 4 + 4 * 2,
                                "fp64": 3 * 8 + 4 * 8 + 6 * 8}


def ensemble(shm, seed, steps=200, warmup=10, timing_steps=5):
    """Ensemble of 64 independent real-scalar realizations (edges 32..160,
    seeded), RKN4, 4th-order Laplacian, one launch per stage over the box
    table, steps replayed from captured graphs: p50 / p99 wall time per step
    (each step synchronised), aggregate cell-updates/s, the same for FP64,
    the stage-kernel time per step (HIP events, eager steps) and the
    FP32-compensated field error against FP64 after the same steps."""
    from cases import ensemble_boxes, ensemble_fields, ensemble_inputs
    boxes = ensemble_boxes(64, 32, 160, seed)
    inp = ensemble_inputs(boxes)
    fields = ensemble_fields(boxes, seed)
    out, final = {}, {}
    for prec in (shm.PREC_FP32C16, shm.PREC_FP64):
        key = "fp32c16" if prec == shm.PREC_FP32C16 else "fp64"
        sim = shm.Simulation("RealScalar", inp, precision=prec)
        sim.set_boxes(0, boxes)
        for c, f in enumerate(fields):
            sim.upload(0, c, f)
        sim.autotune(0)
        progress(f"ensemble {key}: autotuned")
        sim.set_graphs(True)
        sim.advance(0, warmup)
        sim.sync()
        lat = []
        for _ in range(steps):
            t0 = time.perf_counter()
            sim.advance(0, 1)
            sim.sync()
            lat.append(time.perf_counter() - t0)
        progress(f"ensemble {key}: timed steps")
        sim.set_graphs(False)
        sim.set_timing(True)
        sim.advance(0, timing_steps)
        sim.sync()
        tm = sim.timing()
        sim.set_timing(False)
        lat = np.array(lat) * 1e3
        cells = sim.ncells(0)
        bpc = ENSEMBLE_BYTES_PER_CELL_STEP[key]
        kern_ms = tm["stage_ms"] / timing_steps
        rec = {"p50_ms": float(np.percentile(lat, 50)),
               "p99_ms": float(np.percentile(lat, 99)),
               "value": cells * 3 * steps / (lat.sum() * 1e-3),
               "unit": "cell-updates/s",
               "bytes_per_cell_stage": bpc / 3.0,
               "stage_kernel_ms_per_step": kern_ms,
               "kernel_frac": bpc * cells / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
               "launch": sim.get_launch(0)}
        rec["step_frac"] = rec["value"] * bpc / 3.0 / 1e9 / HBM_PEAK_GBS
        final[key] = [sim.download(0, c) for c in range(2)]
        if prec == shm.PREC_FP32C16:
            out = {"workload": "64 independent real-scalar boxes, edges 32..160 "
                               "(seeded), dx 1/64, RKN4, 4th-order Laplacian, "
                               "FP32 + bf16 remainder, steps replayed from "
                               "captured graphs",
                   "boxes": len(boxes), "cells": cells, "steps": steps, **rec}
        else:
            out["fp64"] = rec
            out["speedup_fp32c_vs_fp64"] = out["value"] / rec["value"]
        sim.close()
    err = 0.0
    for c in range(2):
        x32, x64 = final["fp32c16"][c], final["fp64"][c]
        err = max(err, float(np.max(np.abs(x32 - x64)) /
                             max(np.max(np.abs(x64)), 1e-300)))
    out["rel_err_vs_fp64"] = err
    return out


def amr_batched(shm, seed, steps=6, autotune=True):
    """Batched AMR levels: complex scalar on a 256^3 base with two levels
    of 2000 boxes each (edges 16..32 on level 1, 16..52 on level 2, seeded),
    3 ghost cells, 40 seeded string loops, SSPRK3, truncation-error tagging
    at te_crit 1e-3 on its schedule, autotuned launches, sub-cycled coarse
    steps replayed from captured graphs.  Rate = valid cells x stages x level
    steps per coarse step / wall time; the FP32-compensated fields are
    compared with an FP64 run of the same steps."""
    from cases import amr_batched_boxes
    lev1, lev2 = amr_batched_boxes(seed % 100000)
    inp = {"project.name": "ComplexScalar", "project.lambda": 1.0, "sim.t_start": 0.1,
           "sim.t_end": 2.0, "sim.L": 64.0, "sim.cfl": 0.2,
           "amr.coarse_level_grid_size": 256, "amr.nghost": 3, "integrator.type": 10,
           "amr.max_refinement_levels": 2, "amr.te_crit": 1e-3, "amr.regrid_dt": 0.2}
    out, fields = {}, {}
    for prec in (shm.PREC_FP32C16, shm.PREC_FP64):
        sim = shm.Simulation("ComplexScalar", inp, precision=prec)
        sim.set_boxes(1, lev1)
        sim.set_boxes(2, lev2)
        for lev in range(3):
            sim.init_state(lev, 5, seed, [40, 8.0, 40.0, 1.5])
        sim.sync()
        progress("batched AMR: layout and initial state")
        if autotune:
            for lev in range(3):
                sim.autotune(lev)
        progress("batched AMR: autotuned" if autotune else "batched AMR: default launches")
        sim.set_graphs(True)
        sim.evolve(2)
        sim.sync()
        progress("batched AMR: warm-up steps")
        cells = [sim.ncells(lev) for lev in range(3)]
        work = 3 * (cells[0] + 2 * cells[1] + 4 * cells[2])
        t0 = time.perf_counter()
        sim.evolve(steps)
        sim.sync()
        ms = (time.perf_counter() - t0) / steps * 1e3
        sim.set_graphs(False)
        sim.set_timing(True)
        sim.evolve(2)
        sim.sync()
        tm = sim.timing()
        sim.set_timing(False)
        rate = work / (ms * 1e-3)
        key = "fp32c16" if prec == shm.PREC_FP32C16 else "fp64"
        rec = {"ms_per_coarse_step": ms, "value": rate, "unit": "cell-updates/s",
               "step_frac": BYTES_PER_CELL_STAGE[key] * rate / 1e9 / HBM_PEAK_GBS,
               "stage_kernel_ms_per_coarse_step": tm["stage_ms"] / 2,
               "ghost_fill_ms_per_coarse_step": tm["halo_ms"] / 2,
               "launch": [sim.get_launch(lev) for lev in range(3)]}
        if prec == shm.PREC_FP32C16:
            out = {"workload": "complex scalar, 256^3 base + 2 levels (ratio 2, "
                               "sub-cycled), 2000 boxes per fine level (seeded "
                               "edges 16..32 / 16..52), 3 ghost cells, 40 string "
                               "loops, SSPRK3, TE tagging at 1e-3, graphs",
                   "boxes": [1, len(lev1), len(lev2)], "cells": cells,
                   "cell_stages_per_coarse_step": work, "steps": steps, **rec}
        else:
            out["fp64"] = rec
            out["speedup_fp32c_vs_fp64"] = out["value"] / rate
        fields[key] = [[sim.download(lev, c) for c in range(4)] for lev in range(3)]
        sim.close()
        progress(f"batched AMR {key}: {ms:.2f} ms per coarse step")
    err = 0.0
    for lev in range(3):
        for c in range(4):
            a32, a64 = fields["fp32c16"][lev][c], fields["fp64"][lev][c]
            err = max(err, float(np.max(np.abs(a32 - a64)) /
                                 max(np.max(np.abs(a64)), 1e-300)))
    out["rel_err_vs_fp64"] = err
    return out


_T0 = time.perf_counter()


def progress(what):
    """One line per finished phase on stderr (stdout carries the JSON)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {what}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--seed", type=int, default=20240401)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-amr", action="store_true")
    ap.add_argument("--only-amr", action="store_true",
                    help="run only the batched AMR block (for profiling) and print it")
    ap.add_argument("--only-ensemble", action="store_true",
                    help="run only the ensemble block (for profiling) and print it")
    ap.add_argument("--amr-no-autotune", action="store_true",
                    help="default launch shapes in the batched AMR block (a profile "
                         "then holds no autotuner launches)")
    ap.add_argument("--launch", default="",
                    help="fixed stage launch shape 'variant,tile_z,xcd_split,"
                         "persistent,zc_minor[,fused,fused_variant]' instead "
                         "of the autotuner")
    a = ap.parse_args()
    launch = [int(v) for v in a.launch.split(",")] if a.launch else None
    if launch is not None and len(launch) not in (5, 7):
        ap.error("--launch takes five or seven integers")
    if a.gpus != 1:
        print("bench.py measures one GPU: each level runs on a single MI355X "
              "(sharding a level across GPUs is not implemented); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import sledgehamr_amd as shm

    if a.only_ensemble:
        print(json.dumps({"ensemble": ensemble(shm, a.seed)}))
        return
    if a.only_amr:
        print(json.dumps({"amr": amr_batched(shm, a.seed,
                                             autotune=not a.amr_no_autotune)}))
        return

    N = a.n
    inp = inputs(N, L=N * 0.25, cfl=0.2, t0=0.1)
    stages = 3
    sim32, sec32, tm32 = run_gpu(shm, inp, shm.PREC_FP32C16, a.steps, a.warmup,
                                 a.seed, launch)
    cells = sim32.ncells(0)
    rate32 = cells * stages * a.steps / sec32
    progress(f"fp32c16 {N}^3: {sec32 / a.steps * 1e3:.3f} ms/step")

    result = {
        "metric": METRIC,
        "value": rate32,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": sec32 / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 + bf16 remainder storage (compensated); stencil sums f32, "
                 "stage sums f64",
        "data": "synthetic (seeded random-phase complex field)",
        "config": {"workload": f"complex scalar {N}^3 periodic, 4th-order "
                               "Laplacian, SSPRK3, lambda=1, radiation damping",
                   "N": N, "components": 4, "integrator": "SSPRK3 (type 10)",
                   "L": N * 0.25, "cfl": 0.2, "seed": a.seed},
    }

    # Roofline of the dominant stage kernel.
    fused32 = bool(sim32.get_launch(0)["fused"])
    result["roofline"] = kernel_roofline(tm32, "fp32c16", fused32, N)
    result["roofline"]["stages_one_two_fused"] = fused32
    # Whole-step rate priced at the algorithmic bytes.
    result["step_frac"] = (BYTES_PER_CELL_STAGE["fp32c16"] * rate32 / 1e9
                           / HBM_PEAK_GBS)
    result["launch"] = sim32.get_launch(0)
    result["halo_ms_per_step"] = tm32["halo_ms"] / a.steps
    # Achievable bandwidth on this GPU (plain float4 kernels over 4 GiB).
    result["measured_bandwidth"] = sim32.stream_probe(4 << 30, 10)

    if not a.no_fp64:
        sim64, sec64, tm64 = run_gpu(shm, inp, shm.PREC_FP64, a.steps, a.warmup,
                                     a.seed, launch)
        rate64 = cells * stages * a.steps / sec64
        err = 0.0
        for c in range(4):
            x32 = sim32.download(0, c)
            x64 = sim64.download(0, c)
            err = max(err, float(np.max(np.abs(x32 - x64)) /
                                 max(np.max(np.abs(x64)), 1e-300)))
            del x32, x64
        progress(f"fp64 {N}^3: {sec64 / a.steps * 1e3:.3f} ms/step")
        fused64 = bool(sim64.get_launch(0)["fused"])
        r64 = kernel_roofline(tm64, "fp64", fused64, 0)
        result["fp64"] = {"value": rate64, "ms_per_step": sec64 / a.steps * 1e3,
                          "step_frac": BYTES_PER_CELL_STAGE["fp64"] * rate64
                          / 1e9 / HBM_PEAK_GBS,
                          "kernel_frac": r64["frac"],
                          "kernel_moved_frac": r64["moved_frac"],
                          "launch": sim64.get_launch(0)}
        result["speedup_fp32c_vs_fp64"] = rate32 / rate64
        result["fp32c_rel_err_vs_fp64"] = err
        simc, secc, tmc = run_gpu(shm, inp, shm.PREC_FP32C, a.steps, a.warmup,
                                  a.seed, launch)
        ratec = cells * stages * a.steps / secc
        errc = 0.0
        for c in range(4):
            xc = simc.download(0, c)
            x64 = sim64.download(0, c)
            errc = max(errc, float(np.max(np.abs(xc - x64)) /
                                   max(np.max(np.abs(x64)), 1e-300)))
            del xc, x64
        progress(f"fp32c {N}^3: {secc / a.steps * 1e3:.3f} ms/step")
        result["fp32c"] = {"value": ratec,
                           "step_frac": BYTES_PER_CELL_STAGE["fp32c"] * ratec
                           / 1e9 / HBM_PEAK_GBS,
                           "rel_err_vs_fp64": errc,
                           "remainder": "fp32"}
        simc.close()
        sim64.close()
    # After the comparisons above (this advances the FP32 run further).
    result["graphs"] = {"ms_per_step": graph_ms_per_step(sim32, a.steps),
                        "note": "same workload, level steps replayed from "
                                "captured HIP graphs, no kernel timing"}
    sim32.close()
    progress("graph replays")

    if not a.no_ensemble:
        result["ensemble"] = ensemble(shm, a.seed)
        progress("ensemble")
    if not a.no_amr:
        result["amr"] = amr_batched(shm, a.seed, autotune=not a.amr_no_autotune)
        progress("batched AMR levels")

    if not a.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(inp, a.seed)
        progress("cpu baseline")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()
