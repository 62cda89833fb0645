"""Spectra and projections of project output functions on the GPU against the
oracle (Spectrum::Compute, output_types/spectrum.cpp:12-113;
Projection::Compute, output_types/projection.cpp:11-118).  The oracle's
versions are pinned in test_oracle_pins.py (plane waves, Parseval, column
sums of fields linear in z).

Tolerances: relative to the largest bin / pixel, 1e-12 for FP64.  On the
uploaded state, 1e-8 for the FP32-compensated precisions, whose output kernels
read value + remainder (32 significant bits with a bf16 remainder); after
steps, the 1e-6 field bar of the FP32-compensated path.  Counts are exact.
"""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import real_scalar
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

TOL = {shm.PREC_FP64: 1e-12, shm.PREC_FP32C: 1e-8, shm.PREC_FP32C16: 1e-8,
       shm.PREC_FP32CS: 1e-8}
TOL_STEPPED = {shm.PREC_FP64: 1e-12, shm.PREC_FP32C: 1e-6, shm.PREC_FP32C16: 1e-6,
               shm.PREC_FP32CS: 1e-6}
LEV1 = [[8, 8, 8, 15, 15, 15], [16, 8, 8, 23, 15, 15], [4, 20, 0, 11, 27, 7]]
LEV2 = [[20, 20, 20, 27, 27, 27], [30, 20, 20, 43, 27, 27]]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _fields(n, seed):
    rng = np.random.default_rng(seed)
    return [0.5 + 0.3 * rng.standard_normal(n), 0.2 * rng.standard_normal(n)]


def _pair(prec, levels=0, **extra):
    inp = real_scalar(N=16, order=2, nghost=2, lam=1.0, L=8.0,
                      **dict({"sim.t_start": 0.75, "project.output_scale": 1.5,
                              "amr.max_refinement_levels": levels}, **extra))
    return shm.Simulation("RealScalar", inp, precision=prec), Oracle("RealScalar", inp)


@pytest.mark.parametrize("prec", list(TOL))
@pytest.mark.parametrize("mgs", [16, 8])
def test_spectrum_matches_oracle(prec, mgs):
    """Every output function (GradientEnergy reads neighbours through the
    ghost cells), on one box and on 8 boxes, before and after two steps."""
    sim, orc = _pair(prec, **{"amr.max_grid_size": mgs})
    assert sim.output_functions() == ["ScaledPhi", "dPhi2", "GradientEnergy"]
    n = sim.ncells(0)
    for c, f in enumerate(_fields(n, 8)):
        sim.upload(0, c, f)
        orc.upload(0, c, f)
    for step in range(2):
        for fn in range(3):
            gb, gk = sim.spectrum(fn)
            ob, ok = orc.spectrum(fn)
            assert np.array_equal(gk, ok)
            tol = TOL[prec] if step == 0 else TOL_STEPPED[prec]
            assert _rel(gb, ob) < tol, (step, fn, _rel(gb, ob))
        sim.advance(0, 2)
        orc.advance(0, 2)


@pytest.mark.parametrize("prec", list(TOL))
@pytest.mark.parametrize("mode", [0, 1])
def test_projection_matches_oracle(prec, mode):
    """Three levels, before and after one coarse step (fine levels hold
    their own evolved data), every projected depth."""
    sim, orc = _pair(prec, levels=2)
    for lev, bx in ((1, LEV1), (2, LEV2)):
        sim.set_boxes(lev, bx)
        orc.set_boxes(lev, bx)
    for lev in range(3):
        for c, f in enumerate(_fields(sim.ncells(lev), 20 + lev)):
            sim.upload(lev, c, f)
            orc.upload(lev, c, f)
    for stepped in (False, True):
        if stepped:
            sim.evolve(1)
            orc.evolve(1)
        tol = TOL_STEPPED[prec] if stepped else TOL[prec]
        for ml in (-1, 0, 1):
            for fn in range(3):
                gd, gc = sim.projection(fn, mode, ml)
                od, oc = orc.projection(fn, mode, ml)
                assert gd.shape == od.shape == (16 * 2 ** (2 if ml < 0 else ml),) * 2
                assert np.array_equal(gc, oc), (ml, fn)
                assert _rel(gd, od) < tol, (stepped, ml, fn, _rel(gd, od))


def test_outputs_refused_without_functions_or_domain():
    cs = shm.Simulation("ComplexScalar", {
        "project.name": "ComplexScalar", "project.lambda": 1.0, "sim.t_start": 0.1,
        "sim.t_end": 1.0, "sim.L": 8.0, "sim.cfl": 0.2,
        "amr.coarse_level_grid_size": 16, "amr.nghost": 2, "integrator.type": 10})
    assert cs.output_functions() == []
    with pytest.raises(shm.ShmError, match="no output functions"):
        cs.spectrum(0)
    sim, _ = _pair(shm.PREC_FP64)
    with pytest.raises(shm.ShmError):
        sim.spectrum(3)
    with pytest.raises(shm.ShmError):
        sim.projection(0, mode=2)
