"""Kreiss-Oliger dissipation on the GPU against the oracle: KreissOligerDissipation
<2, 3> (kernels.h:56-106) added to the Rhs of every component
(macros.h:355-373), with per-component strengths and the order defaulting to
amr.nghost (sledgehamr.cpp:508-546).

Tolerances as in test_gpu_parity.py: relative L-infinity against max |oracle|,
1e-13 (FP64 Rhs) and 1e-5 (FP32-compensated Rhs, FP32 stencil inputs);
1e-12 / 1e-6 for fields after steps.  The oracle's dissipation term is
pinned by tests/test_oracle_pins.py::test_kreiss_oliger_plane_wave_symbol.
"""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import gaussian_random_field, random_fields, real_scalar, rel_linf
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

PRECS = [(shm.PREC_FP64, 1e-12), (shm.PREC_FP32C, 1e-6), (shm.PREC_FP32C16, 1e-6),
         (shm.PREC_FP32CS, 1e-6)]
ALLP = [shm.PREC_FP64, shm.PREC_FP32C, shm.PREC_FP32C16, shm.PREC_FP32CS]
KO = {
    2: {"sim.dissipation_order": 2, "sim.dissipation_strength_Phi": 0.3,
        "sim.dissipation_strength_dPhi": 0.7},
    3: {"sim.dissipation_strength": 0.2, "sim.dissipation_strength_dPhi": 0.5},
}


def _inputs(order, integrator=21, lap=1, **extra):
    d = real_scalar(N=32, order=lap, nghost=order, integrator=integrator,
                    lam=1.0, L=8.0)
    d.update(KO[order])
    d.update(extra)
    return d


def _load_same(sim, orc, fields):
    for c, f in enumerate(fields):
        sim.upload(0, c, f)
        orc.upload(0, c, f)


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("prec", ALLP)
def test_fill_rhs_with_dissipation(order, prec):
    inp = _inputs(order, lap=2)
    sim = shm.Simulation("RealScalar", inp, precision=prec)
    orc = Oracle("RealScalar", inp)
    fields = random_fields(sim.ncells(0), 2, seed=31 + order, scale=0.3)
    _load_same(sim, orc, fields)
    for add, w in [(False, 0.0), (True, 0.5)]:
        sim.fill_rhs(0, 0.2, w, add)
        orc.fill_rhs(0, 0.2, w, add)
        for c in range(2):
            err = rel_linf(sim.rhs(0, c), orc.rhs(0, c))
            rtol = {shm.PREC_FP64: 1e-13, shm.PREC_FP32CS: 1e-7}.get(prec, 1e-5)
            assert err < rtol, (c, add, err)


@pytest.mark.parametrize("integrator", [21, 22, 10])
@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("prec,tol", PRECS)
def test_advance_with_dissipation(integrator, order, prec, tol):
    inp = _inputs(order, integrator=integrator)
    sim = shm.Simulation("RealScalar", inp, precision=prec)
    orc = Oracle("RealScalar", inp)
    phi = gaussian_random_field(32, seed=77, kcut=16, std=0.1)
    pi = gaussian_random_field(32, seed=78, kcut=16, std=0.05)
    _load_same(sim, orc, [phi, pi])
    sim.advance(0, 12)
    orc.advance(0, 12)
    for c in range(2):
        err = rel_linf(sim.download(0, c), orc.download(0, c))
        assert err < tol, (integrator, order, c, err)


def test_dissipation_damps_high_modes():
    """A mode next to the Nyquist frequency keeps its discrete energy
    pi^2 + w^2 phi^2 without dissipation and loses most of it with order-2
    dissipation of strength 0.5: the momentum equation gains -g pi with
    g = 0.5 sin^4(th/2) / dx, exp(-g t) = 0.02 at t = 2 (RKN4, lambda = 0)."""
    N, dx = 32, 0.25
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    th = 2 * np.pi * 15 / N
    wave = np.cos(th * i).ravel()
    w2 = 4 * np.sin(th / 2) ** 2 / dx ** 2
    energy = {}
    for on in (False, True):
        inp = real_scalar(N=N, order=1, nghost=2, integrator=21, lam=0.0, L=8.0)
        if on:
            inp["sim.dissipation_strength"] = 0.5
        sim = shm.Simulation("RealScalar", inp, precision=shm.PREC_FP64)
        sim.upload(0, 0, wave)
        sim.upload(0, 1, np.zeros(N ** 3))
        sim.advance(0, 40)
        phi, pi = sim.download(0, 0), sim.download(0, 1)
        energy[on] = np.mean(pi ** 2) + w2 * np.mean(phi ** 2)
    e0 = w2 * np.mean(wave ** 2)
    assert abs(energy[False] / e0 - 1) < 1e-2, energy
    assert energy[True] < 0.1 * e0, energy


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
def test_multibox_dissipation_radius3(prec):
    """Order-3 dissipation over 8 boxes of 16^3 with 3 ghost cells equals
    the single box bit for bit."""
    base = _inputs(3, integrator=10)
    a = shm.Simulation("RealScalar", base, precision=prec)
    b = shm.Simulation("RealScalar", dict(base, **{"amr.max_grid_size": 16}),
                       precision=prec)
    glob = [gaussian_random_field(32, seed=s, kcut=12, std=0.2).reshape(32, 32, 32)
            for s in (5, 6)]
    boxes = b.boxes(0)
    assert len(boxes) == 8
    for c in range(2):
        a.upload(0, c, glob[c].ravel())
        b.upload(0, c, np.concatenate(
            [glob[c][bx[2]:bx[5] + 1, bx[1]:bx[4] + 1, bx[0]:bx[3] + 1].ravel()
             for bx in boxes]))
    a.advance(0, 4)
    b.advance(0, 4)
    for c in range(2):
        ga = a.download(0, c).reshape(32, 32, 32)
        gb = b.download(0, c)
        off = 0
        for bx in boxes:
            nb = (bx[3] - bx[0] + 1) * (bx[4] - bx[1] + 1) * (bx[5] - bx[2] + 1)
            blk = ga[bx[2]:bx[5] + 1, bx[1]:bx[4] + 1, bx[0]:bx[3] + 1].ravel()
            assert np.array_equal(blk, gb[off:off + nb]), (c, bx)
            off += nb


def test_dissipation_keeps_two_launch_ssprk3():
    """The fused stage-one-and-two kernel carries no dissipation term: an
    SSPRK3 level with dissipation stays on the two-launch path."""
    sim = shm.Simulation("RealScalar", _inputs(2, integrator=10),
                         precision=shm.PREC_FP32C16)
    sim.set_launch(fused=1)
    assert sim.get_launch(0)["fused"] == 0


def test_dissipation_order_checks():
    with pytest.raises(shm.ShmError):
        shm.Simulation("RealScalar", _inputs(2, **{"sim.dissipation_order": 4}))
    # Order 3 needs three ghost cells.
    with pytest.raises(shm.ShmError):
        shm.Simulation("RealScalar", _inputs(2, **{"sim.dissipation_order": 3}))
    with pytest.raises(shm.ShmError):
        shm.Simulation("RealScalar", _inputs(2, **{"sim.dissipation_strength_Phi": -1.0}))
    # Leapfrog's second kick would difference momenta it overwrites in place.
    with pytest.raises(shm.ShmError):
        shm.Simulation("RealScalar", _inputs(2, integrator=11))
