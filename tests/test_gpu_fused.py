"""Fused first two stages against separate launches: SSPRK3 (k1 kept on
chip) and RKN with a first node 0 (F0 written, the second stage's state
built on chip).  Identical bits, for every precision, tile shape and
stencil radius, on a periodic box, on ensembles of small independent boxes
(tiles wider than the box), and inside captured level steps; and against
the oracle."""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import (complex_scalar, ensemble_boxes, ensemble_fields, ensemble_inputs,
                   gaussian_random_field, real_scalar, rel_linf)
from oracle_py import Oracle

pytestmark = pytest.mark.gpu
ALLP = [shm.PREC_FP64, shm.PREC_FP32C, shm.PREC_FP32C16, shm.PREC_FP32CS]


def _pair(project, inp, prec, fv, graphs=False):
    a = shm.Simulation(project, inp, precision=prec)
    b = shm.Simulation(project, inp, precision=prec)
    a.set_launch(fused=0)
    b.set_launch(fused=1, fused_variant=fv)
    if graphs:
        b.set_graphs(True)
    return a, b


def _same(a, b, lev, nc, what):
    for c in range(nc):
        assert np.array_equal(a.download(lev, c), b.download(lev, c)), (what, c)


@pytest.mark.parametrize("fv", [0, 1, 2])
@pytest.mark.parametrize("prec", ALLP)
def test_fused_equals_two_launches_complex(prec, fv):
    a, b = _pair("ComplexScalar", complex_scalar(N=64), prec, fv)
    for s in (a, b):
        s.init_state(0, 1, 21, [0.05])
    a.advance(0, 4)
    b.advance(0, 4)
    assert b.get_launch(0)["fused"] == 1 and a.get_launch(0)["fused"] == 0
    _same(a, b, 0, 4, ("complex", prec, fv))


@pytest.mark.parametrize("tz,persistent,xcd", [(16, 0, 0), (20, 1, 0), (5, 1, 1), (7, 1, 0)])
@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16])
def test_fused_z_chunks_and_persistent_grids(prec, tz, persistent, xcd):
    """The fused pass unrolls its plane loop by the register-queue length
    (2R+2 = 6 here): z chunks of lengths that are not multiples of it, and
    persistent grids where a workgroup runs several tiles (128^3 in 64 x 16
    tiles and 5- or 7-plane chunks: 416 or 304 tiles for at most 256
    workgroups; the queues restart at each tile), give the two-launch
    result bit for bit."""
    inp = complex_scalar(N=128)
    a = shm.Simulation("ComplexScalar", inp, precision=prec)
    b = shm.Simulation("ComplexScalar", inp, precision=prec)
    # One call sets the whole launch shape.
    a.set_launch(tile_z=tz, persistent=persistent, xcd_split=xcd, fused=0)
    b.set_launch(tile_z=tz, persistent=persistent, xcd_split=xcd, fused=1,
                 fused_variant=2)
    for s in (a, b):
        s.init_state(0, 1, 22, [0.05])
    a.advance(0, 3)
    b.advance(0, 3)
    assert b.get_launch(0)["fused"] == 1 and a.get_launch(0)["fused"] == 0
    _same(a, b, 0, 4, ("chunks", prec, tz, persistent, xcd))


@pytest.mark.parametrize("order,tz", [(1, 5), (3, 12)])
def test_fused_chunks_radius(order, tz):
    """Queue lengths 4 and 8 (radius 1 and 3) on chunks of 5 and 12 planes."""
    inp = real_scalar(N=32, order=order, integrator=10, lam=1.0)
    a = shm.Simulation("RealScalar", inp, precision=shm.PREC_FP32C16)
    b = shm.Simulation("RealScalar", inp, precision=shm.PREC_FP32C16)
    a.set_launch(tile_z=tz, persistent=1, fused=0)
    b.set_launch(tile_z=tz, persistent=1, fused=1, fused_variant=2)
    phi = gaussian_random_field(32, seed=order + 7, kcut=8, std=0.2)
    for s in (a, b):
        s.upload(0, 0, phi)
        s.upload(0, 1, np.zeros_like(phi))
    a.advance(0, 3)
    b.advance(0, 3)
    assert b.get_launch(0)["fused"] == 1
    _same(a, b, 0, 2, ("chunks radius", order, tz))


@pytest.mark.parametrize("order", [1, 2, 3])
def test_fused_equals_two_launches_radius(order):
    inp = real_scalar(N=32, order=order, integrator=10, lam=1.0)
    a, b = _pair("RealScalar", inp, shm.PREC_FP32C16, 2)
    phi = gaussian_random_field(32, seed=order, kcut=8, std=0.2)
    for s in (a, b):
        s.upload(0, 0, phi)
        s.upload(0, 1, np.zeros_like(phi))
    a.advance(0, 3)
    b.advance(0, 3)
    _same(a, b, 0, 2, ("radius", order))


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
def test_fused_ensemble_small_boxes(prec):
    """Boxes of 16..48 cells: tiles wider and taller than the boxes wrap
    more than once."""
    boxes = ensemble_boxes(6, 16, 48, seed=12)
    inp = ensemble_inputs(boxes, integrator=10)
    a, b = _pair("RealScalar", inp, prec, 2)
    fields = ensemble_fields(boxes, seed=4)
    for s in (a, b):
        s.set_boxes(0, boxes)
        for c, f in enumerate(fields):
            s.upload(0, c, f)
    a.advance(0, 3)
    b.advance(0, 3)
    _same(a, b, 0, 2, "ensemble")
    orc = Oracle("RealScalar", inp)
    orc.set_boxes(0, boxes)
    for c, f in enumerate(fields):
        orc.upload(0, c, f)
    orc.advance(0, 3)
    tol = 1e-12 if prec == shm.PREC_FP64 else 1e-6
    for c in range(2):
        assert rel_linf(b.download(0, c), orc.download(0, c)) < tol


def test_fused_in_captured_steps():
    a, b = _pair("ComplexScalar", complex_scalar(N=32), shm.PREC_FP32C16, 2, graphs=True)
    for s in (a, b):
        s.init_state(0, 1, 8, [0.05])
    for step in range(5):
        a.advance(0, 1)
        b.advance(0, 1)
        _same(a, b, 0, 4, ("graph", step))
    assert b._get_info()["graph_replays"] == 4


def test_multibox_level_keeps_two_launches():
    """Boxes that do not wrap on themselves need ghost cells the fused
    kernel does not have: the level stays on the two-launch path."""
    inp = dict(complex_scalar(N=32), **{"amr.max_grid_size": 16})
    sim = shm.Simulation("ComplexScalar", inp, precision=shm.PREC_FP32C16)
    sim.set_launch(fused=1)
    assert sim.get_launch(0)["fused"] == 0


def _rkn_fields(N, seed):
    phi = gaussian_random_field(N, seed=seed, kcut=8, std=0.2)
    pi = gaussian_random_field(N, seed=seed + 1, kcut=8, std=0.05)
    return phi, pi


@pytest.mark.parametrize("fv", [0, 1, 2])
@pytest.mark.parametrize("prec", ALLP)
@pytest.mark.parametrize("integrator", [21, 22])
def test_fused_rkn_equals_separate_stages(integrator, prec, fv):
    """RKN4 (three stages: fused pair, then the final stage) and RKN5 (fused
    pair, a third stage that reads F0 and F1, the final stage); momenta
    nonzero, so the second stage's state carries c0 pi0."""
    inp = real_scalar(N=32, order=2, integrator=integrator, lam=1.0)
    a, b = _pair("RealScalar", inp, prec, fv)
    phi, pi = _rkn_fields(32, 5)
    for s in (a, b):
        s.upload(0, 0, phi)
        s.upload(0, 1, pi)
    a.advance(0, 4)
    b.advance(0, 4)
    assert b.get_launch(0)["fused"] == 1 and a.get_launch(0)["fused"] == 0
    _same(a, b, 0, 2, ("rkn", integrator, prec, fv))


@pytest.mark.parametrize("order", [1, 2, 3])
def test_fused_rkn_radius(order):
    inp = real_scalar(N=32, order=order, integrator=21, lam=1.0)
    a, b = _pair("RealScalar", inp, shm.PREC_FP32C16, 2)
    phi, pi = _rkn_fields(32, order)
    for s in (a, b):
        s.upload(0, 0, phi)
        s.upload(0, 1, pi)
    a.advance(0, 3)
    b.advance(0, 3)
    _same(a, b, 0, 2, ("rkn radius", order))


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
def test_fused_rkn_ensemble_small_boxes(prec):
    boxes = ensemble_boxes(6, 16, 48, seed=13)
    inp = ensemble_inputs(boxes)
    assert inp["integrator.type"] == 21
    a, b = _pair("RealScalar", inp, prec, 2)
    fields = ensemble_fields(boxes, seed=6)
    for s in (a, b):
        s.set_boxes(0, boxes)
        for c, f in enumerate(fields):
            s.upload(0, c, f)
    a.advance(0, 3)
    b.advance(0, 3)
    _same(a, b, 0, 2, "rkn ensemble")
    orc = Oracle("RealScalar", inp)
    orc.set_boxes(0, boxes)
    for c, f in enumerate(fields):
        orc.upload(0, c, f)
    orc.advance(0, 3)
    tol = 1e-12 if prec == shm.PREC_FP64 else 1e-6
    for c in range(2):
        assert rel_linf(b.download(0, c), orc.download(0, c)) < tol


def test_fused_rkn_in_captured_steps():
    inp = real_scalar(N=32, order=2, integrator=22, lam=1.0)
    a, b = _pair("RealScalar", inp, shm.PREC_FP32C16, 1, graphs=True)
    phi, pi = _rkn_fields(32, 9)
    for s in (a, b):
        s.upload(0, 0, phi)
        s.upload(0, 1, pi)
    for step in range(5):
        a.advance(0, 1)
        b.advance(0, 1)
        _same(a, b, 0, 2, ("rkn graph", step))
    assert b._get_info()["graph_replays"] == 4


def _user_tableau(nodes):
    # RKN4's tableau and weights with the given nodes.
    return {"integrator.rkn.nodes": nodes,
            "integrator.rkn.tableau": [0.0, 1. / 8., 0.0, 0.0, 0.5, 0.0],
            "integrator.rkn.weights_bar_b": [1. / 6., 1. / 3., 0.0],
            "integrator.rkn.weights_b": [1. / 6., 4. / 6., 1. / 6.]}


def test_fused_rkn_user_tableau():
    """A user tableau whose first node is 0 fuses (and equals RKN4, whose
    tableau it restates); one whose first node is not 0 keeps separate
    stages, since its first stage's state reads pi0 at neighbours."""
    inp = real_scalar(N=32, order=2, integrator=20, lam=1.0, **_user_tableau([0.0, 0.5, 1.0]))
    a, b = _pair("RealScalar", inp, shm.PREC_FP32C16, 2)
    ref = shm.Simulation("RealScalar", real_scalar(N=32, order=2, integrator=21, lam=1.0),
                         precision=shm.PREC_FP32C16)
    ref.set_launch(fused=0)
    phi, pi = _rkn_fields(32, 11)
    for s in (a, b, ref):
        s.upload(0, 0, phi)
        s.upload(0, 1, pi)
        s.advance(0, 3)
    assert b.get_launch(0)["fused"] == 1
    _same(a, b, 0, 2, "user tableau")
    _same(ref, b, 0, 2, "user tableau = RKN4")
    inp2 = real_scalar(N=32, order=2, integrator=20, lam=1.0,
                       **_user_tableau([0.1, 0.5, 1.0]))
    c = shm.Simulation("RealScalar", inp2, precision=shm.PREC_FP32C16)
    c.set_launch(fused=1)
    assert c.get_launch(0)["fused"] == 0


@pytest.mark.parametrize("prec", ALLP)
@pytest.mark.parametrize("integrator", [21, 22])
@pytest.mark.parametrize("layout", ["periodic", "boxes", "ensemble"])
def test_rkn_second_stage_state_equals_composed(layout, integrator, prec):
    """The first stage writing the second stage's field states (shm.rkn_state
    1, the default) against the second stage composing them from phi0, pi0
    and F0 at every stencil point (0): identical bits, on a periodic box, on
    a level chopped into boxes (same-level ghost copies of the states), and on
    an ensemble; separate stage launches throughout."""
    if layout == "ensemble":
        boxes = ensemble_boxes(5, 16, 40, seed=21)
        inp = ensemble_inputs(boxes, integrator=integrator)
        fields = ensemble_fields(boxes, seed=8)
    else:
        extra = {"amr.max_grid_size": 16} if layout == "boxes" else {}
        inp = real_scalar(N=32, order=2, integrator=integrator, lam=1.0, **extra)
        boxes, fields = None, _rkn_fields(32, 3)
    sims = []
    for state in (0, 1):
        s = shm.Simulation("RealScalar", dict(inp, **{"shm.rkn_state": state}),
                           precision=prec)
        s.set_launch(fused=0)
        if boxes is not None:
            s.set_boxes(0, boxes)
        for c, f in enumerate(fields):
            s.upload(0, c, f)
        s.advance(0, 3)
        sims.append(s)
    _same(sims[0], sims[1], 0, 2, ("rkn state", layout, integrator, prec))
