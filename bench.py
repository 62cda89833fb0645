#!/usr/bin/env python3
"""Benchmark of the per-level time step on one MI355X.

Headline workload (default): complex scalar on a 512^3 periodic level,
4th-order Laplacian, V = (lambda/4)(|psi|^2-1)^2 with lambda = 1 and
radiation-era damping 2/eta, low-storage SSPRK3 (integrator.type = 10),
L = 128 (dx = 0.25), cfl = 0.2 (dt = 0.05), eta starting at 0.1; initial
phases seeded uniform in [0, 2 pi), moduli 1 + N(0, 0.05), momenta 0.

A "step" is one SSPRK3 step of the level (3 RK stages).  `value` is
cell-updates per second per RK stage of the FP32-compensated path (FP32 state
+ bf16 remainder; valid cells x stages / wall time), inputs resident in HBM.
The same seeded run in FP64 gives `fp64` and the field error; `fp32cs` is the
same storage with stencils formed in FP64 from value + remainder.

roofline: the dominant stage kernel (longest mean launch; the other one in
`other_kernel`), priced at the bytes that kernel must move per launch: every
array it reads once and every array it writes once, over the level's valid
cells.  On the headline level (one periodic box) a step is two launches: the
fused stages one and two read u0 (leading words) and write k2 (32 B per cell
with C = 4 in FP32 + bf16), stage three reads u0 with its remainder and k2
and writes u with its remainder (64 B per cell).  Launch times are HIP
events on the library's stream inside the timed region.  `traffic` is the
DRAM bytes of one launch of that kernel measured with rocprofv3 counters in
the profile named by `traffic_source` (same workload, kernel and launch
shape), not in this run; null when this run's shape differs.  The FP32 +
bf16 headline runs in that profiled shape by default (the fastest of a sweep
of the autotuner's candidates, within 1 % of its pick; `--launch auto` runs
the autotuner); FP32CS and FP64 take the autotuner's shapes.

Further blocks, each the one-GPU workload of a configuration with its own
FP64 comparison, rate, stage-kernel HBM fraction and CPU baseline:
  cfg0      real scalar 64^3, 2nd order, RKN4, 100 steps from a seeded
            Gaussian random field;
  gw        first-order phase transition with gravitational waves, 384^3,
            14 components, 6th order, RKN4, 64 seeded bubbles, spectrum in
            256 log bins every 20 steps;
  ensemble  64 independent real-scalar boxes 32^3..160^3, RKN4, latency
            p50/p99;
  amr       256^3 base + 2 levels of about 2000 boxes each (edges 16..64),
            string loops, sub-cycled SSPRK3, truncation-error tagging.

cpu_baseline: the FP64 oracle in the reference's unfused structure (Rhs pass,
LinComb passes, ghost fill per stage; OpenMP with OMP_PROC_BIND=spread,
OMP_PLACES=cores), whole steps of the same workload on the host cores after
one warm-up step, for about 10 s.
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

# Word sizes per precision: state word, remainder, stage-array word, and
# whether the stencils read the remainder too.
WORDS = {"fp64": (8, 0, 8, False), "fp32c16": (4, 2, 4, False),
         "fp32cs": (4, 2, 4, True)}


def ssprk_bytes(key, kernel, C):
    """Bytes per valid cell one SSPRK3 launch must move (each array it reads
    read once, each it writes written once)."""
    w, r, k, cs = WORDS[key]
    u0 = w + (r if cs else 0)  # u0 as the stage states read it
    return C * {"fused12": u0 + k, "stage1": u0 + k, "stage2": u0 + 2 * k,
                "stage3": (w + r) + k + (w + r)}[kernel]


def rkn_bytes_per_step(key, npair, nstages=3):
    """Bytes per valid cell and RKN step (stages one pass each): stage i
    reads phi0, pi0 and F_0..F_{i-1}; non-final stages write F_i, the final
    one reads the remainders and writes phi, pi with theirs."""
    w, r, k, cs = WORDS[key]
    u0 = w + (r if cs else 0)
    b = 0
    for i in range(nstages):
        if i < nstages - 1:
            b += 2 * u0 + i * k + k
        else:
            b += 2 * (w + r) + i * k + 2 * (w + r)
    return npair * b


# DRAM bytes per launch (rocprofv3 --pmc TCC_EA0_RDREQ_DRAM_32B and
# TCC_EA0_WRREQ_WRITE_DRAM_32B, x 32 B) of the headline kernels at 512^3,
# FP32 + bf16 remainder, in the launch shapes the profile used.
TRAFFIC_PROFILE = "profiles/stage_kernel_traffic_complex512.md"
TRAFFIC_BYTES_512 = {"fused12": 2.3767e9 + 2.1475e9,
                     "stage3": 5.6065e9 + 3.2212e9}
# The launch shape those counters were measured in (every field must match),
# which is also the headline's default launch shape (HEADLINE_LAUNCH).
PROFILED_SHAPE = {"variant": 2, "tile_z": 512, "xcd_split": 1, "persistent": 0,
                  "zc_minor": 0, "fused": 1, "fused_variant": 2}
TRAFFIC_SHAPE = {"fused12": PROFILED_SHAPE, "stage3": PROFILED_SHAPE}
HEADLINE_LAUNCH = "2,512,1,0,0,1,2"


def kernel_record(name, avg_ms, nbytes, traffic=None, source=None):
    rec = {"kernel": name, "avg_launch_ms": avg_ms, "bytes_per_launch": nbytes,
           "achieved": nbytes / (avg_ms * 1e-3) / 1e9}
    rec["frac"] = rec["achieved"] / HBM_PEAK_GBS
    rec["traffic"] = traffic
    if traffic:
        rec["traffic_source"] = source
        rec["traffic_over_bytes"] = traffic / nbytes
    return rec


def headline_roofline(tm, key, launch, cells, C=4):
    """Roofline of the dominant stage kernel of the SSPRK3 level from the
    HIP-event timings of the timed steps."""
    launches = max(tm["stage_launches"], 1)
    if launch["fused"] and tm["fused_launches"]:
        fl = tm["fused_launches"]
        fused_ms = tm["fused_ms"] / fl
        s3_ms = (tm["stage_ms"] - tm["fused_ms"]) / max(launches - fl, 1)
        traf = {}
        if cells == 512 ** 3 and key == "fp32c16":
            for kern, shape in TRAFFIC_SHAPE.items():
                if all(launch.get(f) == v for f, v in shape.items()):
                    traf[kern] = TRAFFIC_BYTES_512[kern]
        recs = [kernel_record("shm::fused12_kernel (SSPRK3 stages 1+2)", fused_ms,
                              ssprk_bytes(key, "fused12", C) * cells,
                              traf.get("fused12"), TRAFFIC_PROFILE),
                kernel_record("shm::stage_kernel (SSPRK3 stage 3)", s3_ms,
                              ssprk_bytes(key, "stage3", C) * cells,
                              traf.get("stage3"), TRAFFIC_PROFILE)]
        recs.sort(key=lambda r: -r["avg_launch_ms"])
        main, other = recs
        note = ("per step: one fused stage-1+2 launch (k1 and u1 kept on chip) "
                "and one stage-3 launch")
    else:
        ms = tm["stage_ms"] / launches
        nbytes = (ssprk_bytes(key, "stage1", C) + ssprk_bytes(key, "stage2", C) +
                  ssprk_bytes(key, "stage3", C)) * cells / 3.0
        main, other = kernel_record("shm::stage_kernel (SSPRK3 stages 1..3, mean)",
                                    ms, nbytes), None
        note = "per step: three stage launches"
    res = {"bound": "hbm", "achieved": main["achieved"], "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": main["frac"], "traffic": main["traffic"]}
    res.update(main)
    res["bytes_basis"] = ("bytes the kernel must move per launch: each array it "
                          "reads read once, each it writes written once, over the "
                          "valid cells")
    if other:
        res["other_kernel"] = other
    res["note"] = note
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def omp_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_run(make, step, nsteps_max, cells_per_step, what, budget=10.0):
    """Oracle timing: one warm-up step, then whole steps until about
    `budget` seconds (at most nsteps_max)."""
    orc = make()
    step(orc)
    n = 0
    t0 = time.perf_counter()
    while n < nsteps_max and (n == 0 or time.perf_counter() - t0 < budget):
        step(orc)
        n += 1
    dt = time.perf_counter() - t0
    return {"value": cells_per_step * n / dt, "unit": "cell-updates/s",
            "cores": omp_threads(), "kind": "port",
            "sample": f"{n} step(s) of {what} after 1 warm-up step, FP64 oracle in the "
                      "reference's unfused structure (Rhs pass, LinComb passes, ghost "
                      f"fill per stage), OpenMP {omp_threads()} threads "
                      f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                      f"OMP_PLACES={os.environ.get('OMP_PLACES')}) on {cpu_model()} "
                      f"({os.cpu_count()} logical CPUs visible), {dt:.1f} s"}


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


_T0 = time.perf_counter()


def progress(what):
    """One line per finished phase on stderr (stdout carries the JSON)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {what}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# Headline: complex scalar 512^3, SSPRK3.
# ---------------------------------------------------------------------------

def complex_inputs(N, L, cfl, t0):
    return {
        "project.name": "ComplexScalar", "project.lambda": 1.0, "sim.t_start": t0,
        "sim.t_end": 2.0, "sim.L": L, "sim.cfl": cfl, "amr.coarse_level_grid_size": N,
        "amr.nghost": 2, "integrator.type": 10,
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


def headline(shm, a, launch):
    N = a.n
    inp = complex_inputs(N, L=N * 0.25, cfl=0.2, t0=0.1)
    sims = {}
    out = {}
    for key, prec in (("fp32c16", shm.PREC_FP32C16), ("fp32cs", shm.PREC_FP32CS),
                      ("fp64", shm.PREC_FP64)):
        if key != "fp32c16" and a.no_fp64:
            continue
        # The headline precision runs in the given shape (by default the
        # profiled one); the others take the autotuner's pick.
        sim, sec, tm = run_gpu(shm, inp, prec, a.steps, a.warmup, a.seed,
                               launch if key == "fp32c16" else None)
        cells = sim.ncells(0)
        ms = sec / a.steps * 1e3
        lc = sim.get_launch(0)
        roof = headline_roofline(tm, key, lc, cells)
        step_bytes = (ssprk_bytes(key, "fused12", 4) + ssprk_bytes(key, "stage3", 4)
                      if lc["fused"] and tm["fused_launches"] else
                      ssprk_bytes(key, "stage1", 4) + ssprk_bytes(key, "stage2", 4) +
                      ssprk_bytes(key, "stage3", 4)) * cells
        out[key] = {"value": cells * 3 * a.steps / sec, "ms_per_step": ms, "roofline": roof,
                    "step_frac": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "halo_ms_per_step": tm["halo_ms"] / a.steps, "launch": lc}
        sims[key] = sim
        progress(f"{key} {N}^3: {ms:.3f} ms/step")
    if "fp64" in sims:
        for key in ("fp32c16", "fp32cs"):
            out[key]["rel_err_vs_fp64"] = max(
                rel_err(sims[key].download(0, c), sims["fp64"].download(0, c))
                for c in range(4))
            out[key]["speedup_vs_fp64"] = out[key]["value"] / out["fp64"]["value"]
    bw = sims["fp32c16"].stream_probe(4 << 30, 10)
    graphs = graph_ms_per_step(sims["fp32c16"], a.steps)
    for s in sims.values():
        s.close()
    return inp, out, bw, graphs


# ---------------------------------------------------------------------------
# cfg 0: real scalar 64^3, RKN4, 100 steps.
# ---------------------------------------------------------------------------

def cfg0(shm, seed, cpu):
    from cases import gaussian_random_field, real_scalar
    inp = real_scalar(N=64, order=1, integrator=21, L=1.0, cfl=0.2, lam=1.0, nghost=1)
    phi = gaussian_random_field(64, seed=seed, kcut=16, std=0.1)
    out, fields = {}, {}
    for key, prec in (("fp32c16", shm.PREC_FP32C16), ("fp32cs", shm.PREC_FP32CS),
                      ("fp64", shm.PREC_FP64)):
        sim = shm.Simulation("RealScalar", inp, precision=prec)
        sim.upload(0, 0, phi)
        sim.upload(0, 1, np.zeros_like(phi))
        sim.autotune(0)
        sim.set_graphs(True)
        sim.sync()
        t0 = time.perf_counter()
        sim.advance(0, 100)
        sim.sync()
        sec = time.perf_counter() - t0
        cells = sim.ncells(0)
        out[key] = {"value": cells * 3 * 100 / sec, "ms_per_step": sec * 10,
                    "launch": sim.get_launch(0)}
        fields[key] = [sim.download(0, c) for c in range(2)]
        sim.close()
    for key in ("fp32c16", "fp32cs"):
        out[key]["rel_err_vs_fp64"] = max(rel_err(fields[key][c], fields["fp64"][c])
                                          for c in range(2))
        out[key]["speedup_vs_fp64"] = out[key]["value"] / out["fp64"]["value"]
    res = {"workload": "real scalar 64^3 periodic, L = 1, 2nd-order Laplacian, RKN4, "
                       "100 steps from a seeded Gaussian random field (|phi_k| ~ 1/k, "
                       "k <= 16, std 0.1), steps replayed from captured graphs",
           "note": "262 144 cells: launch-bound on one MI355X (a step is three "
                   "launches of about 1 K workgroups)", **out}
    if cpu:
        from oracle_py import Oracle

        def make():
            o = Oracle("RealScalar", inp)
            o.upload(0, 0, phi)
            o.upload(0, 1, np.zeros_like(phi))
            return o
        res["cpu_baseline"] = cpu_run(make, lambda o: o.advance(0, 1), 100, 64 ** 3 * 3,
                                      "the 64^3 RKN4 workload", budget=10.0)
    return res


# ---------------------------------------------------------------------------
# cfg 3: first-order phase transition with GW, 384^3, RKN4.
# ---------------------------------------------------------------------------

def gw_block(shm, seed, steps, cpu, N=384):
    from cases import bubble_params, phase_transition
    inp = phase_transition(N=N, integrator=21, L=N * 0.5, cfl=0.2,
                           **{"shm.allow_any_grid_size": 1})
    out, fields, spectra = {}, {}, {}
    npair = 7
    for key, prec in (("fp32cs", shm.PREC_FP32CS), ("fp32c16", shm.PREC_FP32C16),
                      ("fp64", shm.PREC_FP64)):
        sim = shm.Simulation("PhaseTransitionGW", inp, precision=prec)
        sim.init_state(0, 3, seed, bubble_params(64, -1e30, 0.0))
        sim.autotune(0)
        # The first spectrum creates the FFT plan and k tables of the
        # context (set-up, kept for later calls): outside the timed steps.
        sim.gw_spectrum(mode=2, nbins=256, projection_type=3)
        sim.sync()
        progress(f"gw {key}: autotuned")
        spec_ms = []
        t = sim.level_info(0)[1]
        sim.set_timing(True)
        t0 = time.perf_counter()
        for s in range(1, steps + 1):
            sim.advance(0, 1)
            t_new = sim.level_info(0)[1]
            sim.init_state(0, 4, seed, bubble_params(64, t, t_new))
            t = t_new
            if s % 20 == 0:
                sim.sync()
                ts = time.perf_counter()
                b, kv = sim.gw_spectrum(mode=2, nbins=256, projection_type=3)
                spec_ms.append((time.perf_counter() - ts) * 1e3)
                spectra.setdefault(key, []).append(b)
        sim.sync()
        sec = time.perf_counter() - t0
        tm = sim.timing()
        sim.set_timing(False)
        cells = sim.ncells(0)
        kern_ms = tm["stage_ms"] / steps
        bps = rkn_bytes_per_step(key, npair)
        out[key] = {"value": cells * 3 * steps / sec, "ms_per_step": sec / steps * 1e3,
                    "stage_kernel_ms_per_step": kern_ms,
                    "stage_kernel_frac": bps * cells / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "bytes_per_cell_step": bps,
                    "spectrum_ms": spec_ms, "launch": sim.get_launch(0)}
        fields[key] = [sim.download(0, c) for c in range(14)]
        sim.close()
        progress(f"gw {key}: {sec / steps * 1e3:.2f} ms/step")
    for key in ("fp32cs", "fp32c16"):
        out[key]["rel_err_vs_fp64"] = {
            "phi_pi": max(rel_err(fields[key][c], fields["fp64"][c]) for c in range(2)),
            "u_ij": max(rel_err(fields[key][c], fields["fp64"][c]) for c in range(2, 14)),
            "spectrum": max(rel_err(x, y) for x, y in zip(spectra[key], spectra["fp64"]))}
        out[key]["speedup_vs_fp64"] = out[key]["value"] / out["fp64"]["value"]
    res = {"workload": f"first-order phase transition with GW, {N}^3 periodic, "
                       "L = N/2, cfl 0.2, 14 components (phi, pi, u_ij, du_ij), "
                       "6th-order Laplacian and gradient, RKN4, 64 seeded bubbles "
                       "(wall width 2 cells, phi_in = 1) injected at their nucleation "
                       f"times, spectrum in 256 log |k| bins every 20 steps, {steps} steps",
           "bytes_basis": "stage kernels priced at the bytes each stage must move "
                          "(phi0, pi0 and earlier forces in, the force or the new state "
                          "out), summed over a step, over the HIP-event stage time",
           **out}
    if cpu:
        from oracle_py import Oracle

        def make():
            o = Oracle("PhaseTransitionGW", inp)
            o.init_state(0, 3, seed, bubble_params(64, -1e30, 0.0))
            return o
        res["cpu_baseline"] = cpu_run(make, lambda o: o.advance(0, 1), 3, N ** 3 * 3,
                                      f"the {N}^3 GW workload (RKN4)", budget=20.0)
    return res


# ---------------------------------------------------------------------------
# cfg 4: ensemble of 64 independent real-scalar boxes.
# ---------------------------------------------------------------------------

def ensemble(shm, seed, cpu, steps=200, warmup=10, timing_steps=5):
    """Ensemble of 64 independent real-scalar realizations (edges 32..160,
    seeded), RKN4, 4th-order Laplacian, one launch per stage over the box
    table, steps replayed from captured graphs: p50 / p99 wall time per step
    (each step synchronised), aggregate cell-updates/s, stage-kernel time
    per step (HIP events, eager steps) and the field error against FP64."""
    from cases import ensemble_boxes, ensemble_fields, ensemble_inputs
    boxes = ensemble_boxes(64, 32, 160, seed)
    inp = ensemble_inputs(boxes)
    fields = ensemble_fields(boxes, seed)
    out, final = {}, {}
    for key, prec in (("fp32c16", shm.PREC_FP32C16), ("fp32cs", shm.PREC_FP32CS),
                      ("fp64", shm.PREC_FP64)):
        sim = shm.Simulation("RealScalar", inp, precision=prec)
        sim.set_boxes(0, boxes)
        for c, f in enumerate(fields):
            sim.upload(0, c, f)
        sim.autotune(0)
        sim.set_graphs(True)
        sim.advance(0, warmup)
        sim.sync()
        lat = []
        for _ in range(steps):
            t0 = time.perf_counter()
            sim.advance(0, 1)
            sim.sync()
            lat.append(time.perf_counter() - t0)
        sim.set_graphs(False)
        sim.set_timing(True)
        sim.advance(0, timing_steps)
        sim.sync()
        tm = sim.timing()
        sim.set_timing(False)
        lat = np.array(lat) * 1e3
        cells = sim.ncells(0)
        bps = rkn_bytes_per_step(key, 1)
        kern_ms = tm["stage_ms"] / timing_steps
        out[key] = {"p50_ms": float(np.percentile(lat, 50)),
                    "p99_ms": float(np.percentile(lat, 99)),
                    "value": cells * 3 * steps / (lat.sum() * 1e-3),
                    "bytes_per_cell_step": bps, "stage_kernel_ms_per_step": kern_ms,
                    "stage_kernel_frac": bps * cells / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "launch": sim.get_launch(0)}
        out[key]["step_frac"] = out[key]["value"] * bps / 3.0 / 1e9 / HBM_PEAK_GBS
        final[key] = [sim.download(0, c) for c in range(2)]
        sim.close()
        progress(f"ensemble {key}: p50 {out[key]['p50_ms']:.3f} ms")
    for key in ("fp32c16", "fp32cs"):
        out[key]["rel_err_vs_fp64"] = max(rel_err(final[key][c], final["fp64"][c])
                                          for c in range(2))
        out[key]["speedup_vs_fp64"] = out[key]["value"] / out["fp64"]["value"]
    res = {"workload": "64 independent real-scalar boxes, edges 32..160 (seeded), dx "
                       "1/64, RKN4, 4th-order Laplacian, steps replayed from captured "
                       "graphs", "boxes": len(boxes), "cells": int(sum((b[3] + 1) ** 3
                                                                        for b in boxes)),
           "steps": steps, **out}
    if cpu:
        from oracle_py import Oracle

        def make():
            o = Oracle("RealScalar", inp)
            o.set_boxes(0, boxes)
            for c, f in enumerate(fields):
                o.upload(0, c, f)
            return o
        res["cpu_baseline"] = cpu_run(make, lambda o: o.advance(0, 1), 20,
                                      res["cells"] * 3, "the 64-box ensemble (RKN4)")
    return res


# ---------------------------------------------------------------------------
# cfg 2: batched AMR levels.
# ---------------------------------------------------------------------------

def amr_inputs():
    return {"project.name": "ComplexScalar", "project.lambda": 1.0, "sim.t_start": 0.1,
            "sim.t_end": 2.0, "sim.L": 64.0, "sim.cfl": 0.2,
            "amr.coarse_level_grid_size": 256, "amr.nghost": 3, "integrator.type": 10,
            "amr.max_refinement_levels": 2, "amr.te_crit": 1e-3, "amr.regrid_dt": 0.2}


def amr_batched(shm, seed, cpu, steps=6, autotune=True, precs=None):
    """Batched AMR levels: complex scalar on a 256^3 base with two levels of
    about 2000 boxes each (edges 16..64, seeded), 3 ghost cells, 40 seeded
    string loops, SSPRK3, truncation-error tagging at te_crit 1e-3 on its
    schedule, autotuned launches, sub-cycled coarse steps replayed from
    captured graphs.  Rate = valid cells x stages x level steps per coarse
    step / wall time; the FP32-compensated fields are compared with an FP64
    run of the same steps."""
    from cases import amr_batched_boxes_wide
    lev1, lev2 = amr_batched_boxes_wide(seed % 100000)
    inp = amr_inputs()
    out, fields = {}, {}
    precs = precs or (("fp32cs", shm.PREC_FP32CS), ("fp32c16", shm.PREC_FP32C16),
                      ("fp64", shm.PREC_FP64))
    for key, prec in precs:
        sim = shm.Simulation("ComplexScalar", inp, precision=prec)
        sim.set_boxes(1, lev1)
        sim.set_boxes(2, lev2)
        for lev in range(3):
            sim.init_state(lev, 5, seed, [40, 8.0, 40.0, 1.5])
        sim.sync()
        if autotune:
            for lev in range(3):
                sim.autotune(lev)
        progress(f"batched AMR {key}: set up")
        sim.set_graphs(True)
        sim.evolve(2)
        sim.sync()
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
        # Bytes the stage kernels must move per coarse step: each level step
        # is the three unfused SSPRK3 stages (the fused pass on level 0).
        lc = [sim.get_launch(lev) for lev in range(3)]
        per_cell = (ssprk_bytes(key, "stage1", 4) + ssprk_bytes(key, "stage2", 4) +
                    ssprk_bytes(key, "stage3", 4))
        step_bytes = per_cell * (cells[0] + 2 * cells[1] + 4 * cells[2])
        kern_ms = tm["stage_ms"] / 2
        out[key] = {"ms_per_coarse_step": ms, "value": work / (ms * 1e-3),
                    "stage_kernel_ms_per_coarse_step": kern_ms,
                    "stage_kernel_frac": step_bytes / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "ghost_fill_ms_per_coarse_step": tm["halo_ms"] / 2,
                    "step_frac": step_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "launch": lc}
        fields[key] = [[sim.download(lev, c) for c in range(4)] for lev in range(3)]
        sim.close()
        progress(f"batched AMR {key}: {ms:.2f} ms per coarse step")
    if "fp64" in fields:
        for key in fields:
            if key == "fp64":
                continue
            out[key]["rel_err_vs_fp64"] = max(
                rel_err(fields[key][lev][c], fields["fp64"][lev][c])
                for lev in range(3) for c in range(4))
            out[key]["speedup_vs_fp64"] = out[key]["value"] / out["fp64"]["value"]
    res = {"workload": "complex scalar, 256^3 base + 2 levels (ratio 2, sub-cycled), "
                       "about 2000 boxes per fine level (seeded edges 16..64 on both), "
                       "3 ghost cells, 40 string loops, SSPRK3, TE tagging at 1e-3, graphs",
           "boxes": [1, len(lev1), len(lev2)], "cells": cells,
           "cell_stages_per_coarse_step": work, "steps": steps, **out}
    # No CPU figure here: the oracle's box-to-box ghost fill scans every box
    # for every ghost cell, which at 2 x 2000 boxes is not a stand-in for
    # the reference's CPU path.
    res["cpu_baseline"] = None
    return res


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
    ap.add_argument("--no-gw", action="store_true")
    ap.add_argument("--no-cfg0", action="store_true")
    ap.add_argument("--gw-steps", type=int, default=40)
    ap.add_argument("--only", default="",
                    help="run one block alone and print it: amr, ensemble, gw, cfg0 "
                         "(for profiling)")
    ap.add_argument("--amr-no-autotune", action="store_true",
                    help="default launch shapes in the batched AMR block")
    ap.add_argument("--launch", default=HEADLINE_LAUNCH,
                    help="stage launch shape of the headline workload "
                         "'variant,tile_z,xcd_split,persistent,zc_minor[,fused,"
                         "fused_variant]' (default: the shape its DRAM traffic "
                         "was profiled in, the fastest of a sweep of the "
                         "autotuner's candidates); 'auto': the autotuner")
    a = ap.parse_args()
    launch = ([int(v) for v in a.launch.split(",")]
              if a.launch and a.launch != "auto" else None)
    if launch is not None and len(launch) not in (5, 7):
        ap.error("--launch takes five or seven integers")
    if a.gpus != 1:
        print("bench.py measures one GPU: each level runs on a single MI355X "
              "(sharding a level across GPUs is not implemented); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    # The CPU baseline's OpenMP runtime reads these when the oracle loads.
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    cpu = not a.no_cpu_baseline

    import sledgehamr_amd as shm

    if a.only:
        blocks = {"amr": lambda: amr_batched(shm, a.seed, cpu, autotune=not a.amr_no_autotune),
                  "ensemble": lambda: ensemble(shm, a.seed, cpu),
                  "gw": lambda: gw_block(shm, a.seed, a.gw_steps, cpu),
                  "cfg0": lambda: cfg0(shm, a.seed, cpu)}
        print(json.dumps({a.only: blocks[a.only]()}))
        return

    inp, h, bw, graphs = headline(shm, a, launch)
    N = a.n
    main32 = h["fp32c16"]
    result = {
        "metric": METRIC,
        "value": main32["value"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": main32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state + bf16 remainder (compensated); FP32 stencil loads, FP64 "
                 "ring and stage sums",
        "data": "synthetic (seeded random-phase complex field)",
        "config": {"workload": f"complex scalar {N}^3 periodic, 4th-order Laplacian, "
                               "SSPRK3, lambda=1, radiation damping",
                   "N": N, "components": 4, "integrator": "SSPRK3 (type 10)",
                   "L": N * 0.25, "cfl": 0.2, "seed": a.seed},
        "roofline": main32["roofline"],
        "step_frac": main32["step_frac"],
        "launch": main32["launch"],
        "halo_ms_per_step": main32["halo_ms_per_step"],
        "measured_bandwidth": bw,
        "graphs": {"ms_per_step": graphs, "note": "same workload, level steps replayed "
                                                  "from captured HIP graphs, no kernel timing"},
    }
    for key in ("fp32c16", "fp32cs"):
        if "rel_err_vs_fp64" in h.get(key, {}):
            result.setdefault("rel_err_vs_fp64", {})[key] = h[key]["rel_err_vs_fp64"]
    if "fp64" in h:
        result["speedup_fp32c_vs_fp64"] = main32["speedup_vs_fp64"]
        result["fp64"] = {k: v for k, v in h["fp64"].items()}
        result["fp32cs"] = {k: v for k, v in h["fp32cs"].items()}
    progress("headline")

    if not a.no_cfg0:
        result["cfg0"] = cfg0(shm, a.seed, cpu)
        progress("cfg0")
    if not a.no_gw:
        result["gw"] = gw_block(shm, a.seed, a.gw_steps, cpu)
        progress("gw")
    if not a.no_ensemble:
        result["ensemble"] = ensemble(shm, a.seed, cpu)
        progress("ensemble")
    if not a.no_amr:
        result["amr"] = amr_batched(shm, a.seed, cpu, autotune=not a.amr_no_autotune)
        progress("batched AMR levels")

    if cpu:
        from oracle_py import Oracle

        def make():
            o = Oracle("ComplexScalar", inp)
            o.init_state(0, 1, a.seed, [0.05])
            return o
        result["cpu_baseline"] = cpu_run(make, lambda o: o.advance(0, 1), 8, N ** 3 * 3,
                                         f"the same {N}^3 workload (SSPRK3, 3 stages)")
        progress("cpu baseline")
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result))


if __name__ == "__main__":
    main()
