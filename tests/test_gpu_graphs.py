"""Captured level steps (HIP graphs) against eager steps: every replay,
with host work (downloads) between replays, gives the eager result bit for
bit; so does a sub-cycled multi-level evolve, with and without coarse/fine
ghost cells (whose fill weights and coarse slot are refreshed per replay)."""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import (AMR_LEV1, AMR_LEV2, AMR_LEV2_Q, amr_boxes, complex_scalar,
                   gaussian_random_field, real_scalar, smooth_boxes)

pytestmark = pytest.mark.gpu


def _pair(project, inp, prec):
    a = shm.Simulation(project, inp, precision=prec)
    b = shm.Simulation(project, inp, precision=prec)
    b.set_graphs(True)
    return a, b


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
@pytest.mark.parametrize("integrator", [10, 21, 22, 11])
def test_graph_steps_match_eager(integrator, prec):
    if integrator == 10:
        inp, project = complex_scalar(N=32, integrator=10), "ComplexScalar"
    else:
        inp, project = real_scalar(N=32, order=2, integrator=integrator), "RealScalar"
    a, b = _pair(project, inp, prec)
    if project == "ComplexScalar":
        for s in (a, b):
            s.init_state(0, 1, 5, [0.05])
    else:
        phi = gaussian_random_field(32, seed=3, kcut=8, std=0.1)
        for s in (a, b):
            s.upload(0, 0, phi)
            s.upload(0, 1, np.zeros_like(phi))
    nc = 4 if project == "ComplexScalar" else 2
    for step in range(7):  # eager, capture of each slot, then replays
        a.advance(0, 1)
        b.advance(0, 1)
        for c in range(nc):
            assert np.array_equal(a.download(0, c), b.download(0, c)), (step, c)
        assert a.level_info(0)[1:] == b.level_info(0)[1:]
    assert a._get_info()["graph_replays"] == 0
    assert b._get_info()["graph_replays"] == 6  # all but the first (eager) step


def test_graph_recaptures_after_launch_change():
    inp = complex_scalar(N=32, integrator=10)
    a, b = _pair("ComplexScalar", inp, shm.PREC_FP32C16)
    for s in (a, b):
        s.init_state(0, 1, 9, [0.05])
    for step in range(6):
        if step == 3:
            for s in (a, b):
                s.set_launch(0, 8, 1, 0, 1)
        a.advance(0, 1)
        b.advance(0, 1)
        for c in range(4):
            assert np.array_equal(a.download(0, c), b.download(0, c)), (step, c)


def test_graph_subcycled_evolve_matches_eager():
    lev1, lev2 = amr_boxes(seed=2024)
    inp = complex_scalar(N=32, integrator=10, nghost=2, L=8.0, cfl=0.25, t0=0.125,
                         **{"amr.max_refinement_levels": 2, "amr.te_crit": 1e-3,
                            "amr.regrid_dt": 0.01})
    a, b = _pair("ComplexScalar", inp, shm.PREC_FP64)
    for s in (a, b):
        s.set_boxes(1, lev1)
        s.set_boxes(2, lev2)
        for lev in range(3):
            s.init_state(lev, 5, 77, [4, 5.0, 10.0, 1.5])
    for n in range(4):
        a.evolve(1)
        b.evolve(1)
        for lev in range(3):
            for c in range(4):
                assert np.array_equal(a.download(lev, c), b.download(lev, c)), (n, lev, c)
    for lev in range(3):
        assert np.array_equal(a.level_tag_counts(lev), b.level_tag_counts(lev))
    assert b._get_info()["graph_replays"] > 0


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
@pytest.mark.parametrize("integrator,interp", [(10, 2), (21, 2), (22, 2), (11, 2),
                                               (10, 4), (10, 1), (21, 0)])
def test_graph_coarse_fine_levels_match_eager(integrator, interp, prec):
    """Three levels whose fine levels fill coarse/fine ghost cells at every
    stage: three coarse steps (level 1 steps 6 times, level 2 12 times), all
    but each level's first step replayed from graphs, equal to eager steps
    bit for bit."""
    inp = real_scalar(N=16, order=2, nghost=2, lam=1.0, L=16.0, cfl=0.25,
                      integrator=integrator,
                      **{"amr.max_refinement_levels": 2,
                         "amr.interpolation_type": interp})
    lev2 = AMR_LEV2_Q if interp == 4 else AMR_LEV2
    a, b = _pair("RealScalar", inp, prec)
    for s in (a, b):
        s.set_boxes(1, AMR_LEV1)
        s.set_boxes(2, lev2)
        for lev, bx in ((0, [[0, 0, 0, 15, 15, 15]]), (1, AMR_LEV1), (2, lev2)):
            for c, ph in ((0, 0.3), (1, 1.1)):
                s.upload(lev, c, smooth_boxes(bx, 16.0 / (16 * 2 ** lev), 16.0, ph)
                         * (1.0 if c == 0 else 0.2))
    for n in range(3):
        a.evolve(1)
        b.evolve(1)
        for lev in range(3):
            for c in range(2):
                assert np.array_equal(a.download(lev, c), b.download(lev, c)), (n, lev, c)
    assert a._get_info()["graph_replays"] == 0
    assert b._get_info()["graph_replays"] == 2 + 5 + 11
