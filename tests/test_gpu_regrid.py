"""Local regrid on the GPU against the oracle (local_regrid/local_regrid.cpp):
block selection from tag masks (bit-exact blocks and distance), AddBoxes and
RemakeLevel (old data kept bit for bit, new boxes as FillPatch fills them),
and sub-cycled evolves that regrid locally at scheduled times (box lists
equal, fields to the usual bars).  The oracle's regrid is pinned in
test_oracle_regrid_pins.py.

Tolerances: relative L-infinity against max |oracle|; right after AddBoxes /
RemakeLevel 1e-13 (FP64), and for FP32-compensated states (FP64 stencil
tiles of the stored pair), whose level data were stepped before, the step
bar of test_gpu_amr.py, 1e-6; after further steps 1e-12 / 1e-6.  Evolves that regrid compare
FP64 only: regrids follow the tag masks, which are bit-exact in FP64
(FP32-compensated truncation errors may cross te_crit elsewhere).
"""
import itertools

import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import real_scalar, rel_linf, smooth_boxes
from oracle_py import Oracle, block_layout

pytestmark = pytest.mark.gpu

ALLP = [shm.PREC_FP64, shm.PREC_FP32C, shm.PREC_FP32C16, shm.PREC_FP32CS]
# Levels that step with coarse/fine ghost cells: the FP32-compensated path
# there forms its stencils on hi + lo (FP32CS).
FINE_PRECS = [shm.PREC_FP64, shm.PREC_FP32CS]


def _inputs(N=16, levels=1, bf=4, nbuf=1, **extra):
    d = real_scalar(N=N, order=2, nghost=2, lam=1.0, L=float(N), cfl=0.25,
                    **{"amr.max_refinement_levels": levels, "amr.blocking_factor": bf,
                       "amr.n_error_buf": nbuf})
    d.update(extra)
    return d


def _pair(inp, prec=shm.PREC_FP64):
    return shm.Simulation("RealScalar", inp, precision=prec), Oracle("RealScalar", inp)


@pytest.mark.parametrize("N,bf,nbuf,mgs", [(16, 2, 0, 16), (16, 4, 1, 8), (16, 8, 2, 16),
                                           (32, 4, 1, 16), (64, 8, 1, 32)])
def test_select_matches_oracle(N, bf, nbuf, mgs):
    rng = np.random.default_rng(N + bf + nbuf)
    sim, orc = _pair(_inputs(N=N, bf=bf, nbuf=nbuf, **{"amr.max_grid_size": mgs}))
    Np = 2 * N // bf
    blocks = [q for q in itertools.product(range(Np), repeat=3) if rng.random() < 0.3]
    boxes = block_layout("RealScalar", blocks, bf)
    sim.set_boxes(1, boxes)
    orc.set_boxes(1, boxes)
    for dens in (0.0, 0.003, 0.05):
        mask = (rng.random(sim.ncells(0)) < dens).astype(np.uint8)
        gb, gd = sim.regrid_select(0, mask)
        ob, od = orc.regrid_select(0, mask)
        assert np.array_equal(gb, ob), dens
        assert gd == od, (dens, gd, od)


def _bump(boxes, h, c, w, amp):
    out = []
    for b in boxes:
        k, j, i = np.meshgrid(*(np.arange(b[a], b[a + 3] + 1) for a in (2, 1, 0)),
                              indexing="ij")
        x, y, z = (i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h
        r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
        out.append((amp * np.exp(-r2 / w ** 2)).ravel())
    return np.concatenate(out)


L1 = [[8, 8, 8, 19, 19, 19]]
L2 = [[20, 20, 20, 31, 31, 31]]


def _load_bump(objs, levels):
    for o in objs:
        for lev, bx in levels:
            h = 16.0 / (16 * 2 ** lev)
            o.upload(lev, 0, _bump(bx, h, (7.0, 7.5, 8.0), 1.5, 1.0))
            o.upload(lev, 1, _bump(bx, h, (8.5, 7.0, 7.5), 2.0, 0.1))


def test_select_from_device_tags_matches_oracle():
    """Tags of level 1 from its truncation errors (TE against level 0) stay
    on the device; the selection reads them (mask None)."""
    inp = _inputs(levels=2, **{"amr.te_crit": 1e-3})
    sim, orc = _pair(inp)
    for o in (sim, orc):
        o.set_boxes(1, L1)
        o.set_boxes(2, L2)
    _load_bump((sim, orc), ((0, [[0, 0, 0, 15, 15, 15]]), (1, L1), (2, L2)))
    gm, gc, _ = sim.te_tag(1, shm.TE_COMPUTE | shm.TE_TAGS)
    om, oc, _ = orc.te_tag(1, shm.TE_COMPUTE | shm.TE_TAGS)
    assert np.array_equal(gm, om) and gm.sum() > 0
    gb, gd = sim.regrid_select(1)
    ob, od = orc.regrid_select(1, om)
    assert len(gb) > 0
    assert np.array_equal(gb, ob) and gd == od


@pytest.mark.parametrize("prec", FINE_PRECS)
def test_add_boxes_matches_oracle(prec):
    """After one coarse step (level 1 at t0 + dt0, its old state at t0 +
    dt1, level 0's states at t0 and t0 + dt0), level 1 gains two boxes:
    the new state from level 0 at its time, the old one interpolated halfway
    between level 0's states.  Then both step on."""
    inp = _inputs(levels=1, integrator=21)
    sim, orc = _pair(inp, prec)
    lev1 = [[8, 8, 8, 15, 15, 15]]
    for o in (sim, orc):
        o.set_boxes(1, lev1)
    for o in (sim, orc):
        for lev, bx in ((0, [[0, 0, 0, 15, 15, 15]]), (1, lev1)):
            h = 16.0 / (16 * 2 ** lev)
            o.upload(lev, 0, smooth_boxes(bx, h, 16.0, 0.3))
            o.upload(lev, 1, 0.2 * smooth_boxes(bx, h, 16.0, 1.1))
    sim.evolve(1)
    orc.evolve(1)
    before = [sim.download(1, c) for c in range(2)]
    nb = [[16, 8, 8, 23, 15, 15], [4, 20, 2, 11, 27, 9]]
    sim.add_boxes(1, nb)
    orc.add_boxes(1, nb)
    assert sim.boxes(1).tolist() == orc.boxes(1).tolist() == lev1 + nb
    # The level data were stepped: FP32-compensated states carry the step's
    # error into the interpolated boxes.
    for c in range(2):
        tol = 1e-13 if prec == shm.PREC_FP64 else 1e-6
        g = sim.download(1, c)
        assert np.array_equal(g[:512], before[c])
        assert rel_linf(g, orc.download(1, c)) < tol, c
        assert rel_linf(sim.download(1, c, old=True), orc.download(1, c, old=True)) < tol
    sim.evolve(1)
    orc.evolve(1)
    for lev in range(2):
        for c in range(2):
            err = rel_linf(sim.download(lev, c), orc.download(lev, c))
            bar = 1e-12 if prec == shm.PREC_FP64 else 1e-6
            assert err < bar, (lev, c, err)


@pytest.mark.parametrize("prec", FINE_PRECS)
def test_remake_level_matches_oracle(prec):
    inp = _inputs(levels=1, integrator=10)
    sim, orc = _pair(inp, prec)
    lev1 = [[8, 8, 8, 15, 15, 15], [16, 8, 8, 23, 15, 15]]
    for o in (sim, orc):
        o.set_boxes(1, lev1)
        for lev, bx in ((0, [[0, 0, 0, 15, 15, 15]]), (1, lev1)):
            h = 16.0 / (16 * 2 ** lev)
            o.upload(lev, 0, smooth_boxes(bx, h, 16.0, 0.7))
            o.upload(lev, 1, 0.1 * smooth_boxes(bx, h, 16.0, 0.2))
    sim.evolve(1)
    orc.evolve(1)
    new = [[12, 8, 8, 19, 15, 23], [4, 4, 4, 11, 7, 11]]
    sim.remake_level(1, new)
    orc.remake_level(1, new)
    for c in range(2):
        tol = 1e-13 if prec == shm.PREC_FP64 else 1e-6
        assert rel_linf(sim.download(1, c), orc.download(1, c)) < tol
    sim.evolve(1)
    orc.evolve(1)
    for c in range(2):
        bar = 1e-12 if prec == shm.PREC_FP64 else 1e-6
        assert rel_linf(sim.download(1, c), orc.download(1, c)) < bar


@pytest.mark.parametrize("te,integrator", [(1e-2, 21), (1e-3, 21), (1e-2, 10)])
@pytest.mark.parametrize("graphs", [False, True])
def test_evolve_with_local_regrid_matches_oracle(te, integrator, graphs):
    """Three levels, a Gaussian bump next to the finer boxes, truncation
    error tagging with regrids scheduled as often as possible and the local
    regrid on: after every coarse step the box lists, the regrid counters and
    the fields agree with the oracle; boxes were added (and at te_crit 1e-3
    the volume veto asked for global regrids, which keep the layout as the
    inputs allow, with the counters as they are)."""
    inp = _inputs(levels=2, integrator=integrator,
                  **{"amr.te_crit": te, "amr.regrid_dt": 0.01, "shm.local_regrid": 1,
                     "amr.volume_threshold_weak": 10.0,
                     "amr.volume_threshold_strong": 10.0,
                     "shm.keep_layout_on_global_regrid": 1})
    sim, orc = _pair(inp)
    sim.set_graphs(graphs)
    for o in (sim, orc):
        o.set_boxes(1, L1)
        o.set_boxes(2, L2)
    _load_bump((sim, orc), ((0, [[0, 0, 0, 15, 15, 15]]), (1, L1), (2, L2)))
    for n in range(3):
        sim.evolve(1)
        orc.evolve(1)
        for lev in range(3):
            assert sim.boxes(lev).tolist() == orc.boxes(lev).tolist(), (n, lev)
        gs, os_ = sim.regrid_stats(), orc.regrid_stats(3)
        for k in ("attempts", "global_wanted", "nregrids", "veto_level"):
            assert gs[k] == os_[k], (n, k)
        assert np.array_equal(gs["min_distance"], os_["min_distance"]), n
        for lev in range(3):
            for c in range(2):
                err = rel_linf(sim.download(lev, c), orc.download(lev, c))
                assert err < 1e-12, (n, lev, c, err)
    assert len(sim.boxes(1)) > 1 and len(sim.boxes(2)) > 1
    if te == 1e-3:
        assert sim.regrid_stats()["global_wanted"] > 0


def test_global_regrid_stops_the_evolve():
    """Without shm.keep_layout_on_global_regrid a wanted global regrid ends
    the evolve with SHM_ERR_UNSUPPORTED at the regrid, after the levels
    stepped to it, like the oracle; a following evolve continues."""
    inp = _inputs(levels=2, integrator=21,
                  **{"amr.te_crit": 1e-3, "amr.regrid_dt": 0.01, "shm.local_regrid": 1,
                     "amr.volume_threshold_weak": 10.0,
                     "amr.volume_threshold_strong": 10.0})
    sim, orc = _pair(inp)
    for o in (sim, orc):
        o.set_boxes(1, L1)
        o.set_boxes(2, L2)
    _load_bump((sim, orc), ((0, [[0, 0, 0, 15, 15, 15]]), (1, L1), (2, L2)))
    stopped = False
    for n in range(3):
        try:
            sim.evolve(1)
        except shm.ShmError as e:
            assert e.code == 4 and "global regrid" in str(e)
            with pytest.raises(RuntimeError):
                orc.evolve(1)
            stopped = True
        else:
            orc.evolve(1)
        for lev in range(3):
            assert sim.level_info(lev)[1] == orc.time(lev)
            assert sim.boxes(lev).tolist() == orc.boxes(lev).tolist()
        assert sim.regrid_stats()["global_wanted"] == orc.regrid_stats(3)["global_wanted"]
    assert stopped


def test_regrid_refusals():
    sim = shm.Simulation("RealScalar", _inputs(levels=1), precision=shm.PREC_FP64)
    sim.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    with pytest.raises(shm.ShmError, match="overlaps"):
        sim.add_boxes(1, [[12, 8, 8, 19, 15, 15]])
    with pytest.raises(shm.ShmError, match="even"):
        sim.add_boxes(1, [[17, 8, 8, 23, 15, 15]])
    with pytest.raises(shm.ShmError, match="no tags"):
        sim.regrid_select(0)
    with pytest.raises(ValueError):
        sim.regrid_select(0, np.zeros(5, np.uint8))
    # A box whose coarse stencil leaves level 0's domain cannot be nested
    # on level 2 without level 1 underneath: refused, the level unchanged.
    sim2 = shm.Simulation("RealScalar", _inputs(levels=2), precision=shm.PREC_FP64)
    sim2.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    sim2.set_boxes(2, [[20, 20, 20, 27, 27, 27]])
    with pytest.raises(shm.ShmError, match="nested"):
        sim2.add_boxes(2, [[40, 40, 40, 47, 47, 47]])
    assert sim2.boxes(2).tolist() == [[20, 20, 20, 27, 27, 27]]
    with pytest.raises(shm.ShmError):
        shm.Simulation("RealScalar", _inputs(bf=3))
