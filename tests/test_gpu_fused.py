"""Fused SSPRK3 stages one and two (k1 kept on chip) against the two-launch
path: identical bits, for every precision, tile shape and stencil radius,
on a periodic box, on ensembles of small independent boxes (tiles wider
than the box), and inside captured level steps; and against the oracle."""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import (complex_scalar, ensemble_boxes, ensemble_fields, ensemble_inputs,
                   gaussian_random_field, real_scalar, rel_linf)
from oracle_py import Oracle

pytestmark = pytest.mark.gpu
ALLP = [shm.PREC_FP64, shm.PREC_FP32C, shm.PREC_FP32C16]


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


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16])
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
