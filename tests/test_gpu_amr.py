"""Coarse/fine ghost fill on the GPU against the oracle (FillPatchTwoLevels
and FillIntermediatePatch, level_synchronizer.cpp:92-209, GetLevelData at
:613-632): the Rhs of a fine level whose stencils reach into interpolated
ghost cells, and sub-cycled multi-level steps of every integrator, where
stage states at coarse/fine ghost cells are the coarser level interpolated
at the stage times.  The oracle's fill is pinned in test_oracle_pins.py
(exact quadratics, parent values, time windows).

Tolerances: relative L-infinity against max |oracle|; 1e-13 (FP64 Rhs),
1e-5 (Rhs from FP32 stencil inputs), 1e-7 (Rhs from FP64 stencil tiles of the
stored pair), 1e-12 (FP64) / 1e-6 (FP32-compensated) for fields after steps.

Sub-cycled fine levels run the FP32-compensated storage with FP64 stencil
tiles (FP32CS): with FP32 stencil inputs (FP32C, FP32C16) the rounding of the
leading words acts as a forcing of about dt/dx^2 times 3e-8 per stage, and
at coarse/fine ghost cells the stage states are re-rounded at every stage,
which leaves 1.65e-6 on fine-level momenta next to box faces after two
coarse steps; the stencils of FP32CS read hi + lo.
"""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import AMR_LEV1 as LEV1
from cases import AMR_LEV2 as LEV2
from cases import AMR_LEV2_Q as LEV2_Q
from cases import real_scalar, rel_linf
from cases import smooth_boxes as _smooth
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

ALLP = [shm.PREC_FP64, shm.PREC_FP32C, shm.PREC_FP32C16, shm.PREC_FP32CS]
PRECS = [(shm.PREC_FP64, 1e-12), (shm.PREC_FP32CS, 1e-6)]
RHS_TOL = {shm.PREC_FP64: 1e-13, shm.PREC_FP32C: 1e-5, shm.PREC_FP32C16: 1e-5,
           shm.PREC_FP32CS: 1e-7}

def _setup(inp, prec, levels):
    sim = shm.Simulation("RealScalar", inp, precision=prec)
    orc = Oracle("RealScalar", inp)
    for lev, bx in levels.items():
        sim.set_boxes(lev, bx)
        orc.set_boxes(lev, bx)
    return sim, orc


@pytest.mark.parametrize("interp", [0, 1, 2, 4])
@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("prec", ALLP)
def test_fine_rhs_reads_interpolated_ghosts(order, interp, prec):
    """Random coarse and fine data (coarse/fine ghosts differ strongly from
    the fine data), so every wrong ghost value changes the Rhs."""
    inp = real_scalar(N=16, order=order, nghost=order, lam=1.0, L=16.0,
                      **{"amr.max_refinement_levels": 1,
                         "amr.interpolation_type": interp})
    sim, orc = _setup(inp, prec, {1: LEV1})
    rng = np.random.default_rng(40 + order)
    for lev in (0, 1):
        for c in range(2):
            d = rng.standard_normal(sim.ncells(lev)) * 0.3
            sim.upload(lev, c, d)
            orc.upload(lev, c, d)
    sim.fill_rhs(1, 0.0)
    orc.fill_rhs(1, 0.0)
    for c in range(2):
        err = rel_linf(sim.rhs(1, c), orc.rhs(1, c))
        assert err < RHS_TOL[prec], (c, err)


@pytest.mark.parametrize("integrator", [10, 21, 22, 11])
@pytest.mark.parametrize("prec,tol", PRECS)
def test_subcycled_levels_match_oracle(integrator, prec, tol):
    """Three levels, two coarse steps: level 1 sub-cycles twice and level 2
    four times per coarse step, with coarse states interpolated in time at
    every stage of the finer levels."""
    inp = real_scalar(N=16, order=2, nghost=2, lam=1.0, L=16.0, cfl=0.25,
                      integrator=integrator,
                      **{"amr.max_refinement_levels": 2})
    sim, orc = _setup(inp, prec, {1: LEV1, 2: LEV2})
    for lev, bx in ((0, [[0, 0, 0, 15, 15, 15]]), (1, LEV1), (2, LEV2)):
        h = 16.0 / (16 * 2 ** lev)
        for c, ph in ((0, 0.3), (1, 1.1)):
            d = _smooth(bx, h, 16.0, ph) * (1.0 if c == 0 else 0.2)
            sim.upload(lev, c, d)
            orc.upload(lev, c, d)
    sim.evolve(2)
    orc.evolve(2)
    for lev in range(3):
        assert sim.level_info(lev)[1] == orc.time(lev)
        for c in range(2):
            err = rel_linf(sim.download(lev, c), orc.download(lev, c))
            assert err < tol, (integrator, lev, c, err)


@pytest.mark.parametrize("interp", [1, 4])
@pytest.mark.parametrize("prec,tol", PRECS)
def test_subcycled_levels_other_interpolations(interp, prec, tol):
    inp = real_scalar(N=16, order=2, nghost=2, lam=1.0, L=16.0, cfl=0.25,
                      integrator=10, **{"amr.max_refinement_levels": 2,
                                        "amr.interpolation_type": interp})
    sim, orc = _setup(inp, prec, {1: LEV1, 2: LEV2_Q})
    for lev, bx in ((0, [[0, 0, 0, 15, 15, 15]]), (1, LEV1), (2, LEV2_Q)):
        h = 16.0 / (16 * 2 ** lev)
        for c, ph in ((0, 0.3), (1, 1.1)):
            d = _smooth(bx, h, 16.0, ph) * (1.0 if c == 0 else 0.2)
            sim.upload(lev, c, d)
            orc.upload(lev, c, d)
    sim.evolve(2)
    orc.evolve(2)
    for lev in range(3):
        for c in range(2):
            err = rel_linf(sim.download(lev, c), orc.download(lev, c))
            assert err < tol, (interp, lev, c, err)


@pytest.mark.parametrize("interp", [0, 1, 2, 4])
@pytest.mark.parametrize("prec", ALLP)
def test_fill_coarse_patch_matches_oracle(interp, prec):
    """FillCoarsePatch of level 1 from random level-0 data at the level's
    time; FP32-compensated states hold 32 significant bits (1e-8)."""
    inp = real_scalar(N=16, order=2, nghost=2, lam=1.0, L=16.0,
                      **{"amr.max_refinement_levels": 1,
                         "amr.interpolation_type": interp})
    sim, orc = _setup(inp, prec, {1: LEV1})
    rng = np.random.default_rng(7)
    for c in range(2):
        d = rng.standard_normal(sim.ncells(0))
        sim.upload(0, c, d)
        orc.upload(0, c, d)
    sim.fill_coarse_patch(1)
    orc.fill_coarse_patch(1)
    for c in range(2):
        err = rel_linf(sim.download(1, c), orc.download(1, c))
        assert err < (1e-13 if prec == shm.PREC_FP64 else 1e-8), (c, err)
    # The level steps from the interpolated state like the oracle.
    sim.advance(1, 1)
    orc.advance(1, 1)
    for c in range(2):
        err = rel_linf(sim.download(1, c), orc.download(1, c))
        assert err < (1e-12 if prec == shm.PREC_FP64 else 1e-6), (c, err)


def test_ghost_fill_changes_the_fine_solution():
    """Control: with the coarse level shifted by a constant, a fine level's
    step changes near its boundary (the ghosts come from the coarse level)
    and not far inside it."""
    inp = real_scalar(N=16, order=2, nghost=2, lam=0.0, L=16.0, integrator=21,
                      **{"amr.max_refinement_levels": 1})
    box = [[8, 8, 8, 23, 23, 23]]
    res = []
    for shift in (0.0, 1.0):
        sim = shm.Simulation("RealScalar", inp, precision=shm.PREC_FP64)
        sim.set_boxes(1, box)
        sim.upload(0, 0, _smooth([[0, 0, 0, 15, 15, 15]], 1.0, 16.0, 0.2) + shift)
        sim.upload(0, 1, np.zeros(16 ** 3))
        sim.upload(1, 0, _smooth(box, 0.5, 16.0, 0.2))
        sim.upload(1, 1, np.zeros(16 ** 3))
        sim.advance(1, 1)
        res.append(sim.download(1, 1).reshape(16, 16, 16))
    d = np.abs(res[1] - res[0])
    assert d[0].max() > 1e-3 and d[6:10, 6:10, 6:10].max() == 0.0


def test_unnested_level_is_refused():
    inp = real_scalar(N=16, order=1, nghost=2, lam=1.0, L=16.0,
                      **{"amr.max_refinement_levels": 2})
    sim = shm.Simulation("RealScalar", inp, precision=shm.PREC_FP64)
    sim.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    sim.set_boxes(2, [[16, 16, 16, 23, 23, 23]])
    with pytest.raises(shm.ShmError, match="not properly nested"):
        sim.advance(2, 1)


def test_interpolation_types_checked():
    for t in (3, 5, -1):
        with pytest.raises(shm.ShmError):
            shm.Simulation("RealScalar", real_scalar(
                N=16, **{"amr.interpolation_type": t}))
