"""GPU parity: the gfx950 kernels (through the C ABI) against the FP64 oracle
on identical seeded inputs.

Tolerances (relative L-infinity per component, against max |oracle|):
  FP64 path            1e-12 for fields after multi-step runs, 1e-13 for Rhs
  FP32-compensated     1e-6 for fields (the target of the FP32 path),
                       1e-5 for a single Rhs evaluation with FP32 stencil
                       inputs, 1e-7 with FP64 stencil tiles (FP32CS: the
                       stored Rhs is FP32)
Tags and tag counters are compared exactly.
"""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import (complex_scalar, expected_tags, gaussian_random_field,
                   phase_transition, random_fields, real_scalar, rel_linf,
                   user_tag_case)
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

PRECS = [(shm.PREC_FP64, 1e-12), (shm.PREC_FP32C, 1e-6), (shm.PREC_FP32C16, 1e-6),
         (shm.PREC_FP32CS, 1e-6)]
ALLP = [shm.PREC_FP64, shm.PREC_FP32C, shm.PREC_FP32C16, shm.PREC_FP32CS]
# Representation error of the stored pair hi + lo.
PAIR_TOL = {shm.PREC_FP64: 1e-15, shm.PREC_FP32C: 1e-13, shm.PREC_FP32C16: 1e-9,
            shm.PREC_FP32CS: 1e-9}
# One Rhs evaluation: FP32 stencil inputs (FP32C, FP32C16), FP64 stencil
# tiles of the stored pair with an FP32 stored Rhs (FP32CS).
RHS_TOL = {shm.PREC_FP64: 1e-13, shm.PREC_FP32C: 1e-5, shm.PREC_FP32C16: 1e-5,
           shm.PREC_FP32CS: 1e-7}


def _load_same(sim, orc, lev, fields):
    for c, f in enumerate(fields):
        sim.upload(lev, c, f)
        orc.upload(lev, c, f)


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("prec", ALLP)
def test_fill_rhs_real_scalar(order, prec):
    inp = real_scalar(N=32, order=order, integrator=21)
    sim = shm.Simulation("RealScalar", inp, precision=prec)
    orc = Oracle("RealScalar", inp)
    n = sim.ncells(0)
    fields = random_fields(n, 2, seed=11 + order, scale=0.3)
    _load_same(sim, orc, 0, fields)
    for add, w in [(False, 0.0), (True, 0.25)]:
        sim.fill_rhs(0, 0.3, w, add)
        orc.fill_rhs(0, 0.3, w, add)
        for c in range(2):
            err = rel_linf(sim.rhs(0, c), orc.rhs(0, c))
            assert err < RHS_TOL[prec], (c, add, err)


@pytest.mark.parametrize("prec", ALLP)
def test_fill_rhs_complex_scalar(prec):
    inp = complex_scalar(N=32)
    sim = shm.Simulation("ComplexScalar", inp, precision=prec)
    orc = Oracle("ComplexScalar", inp)
    n = sim.ncells(0)
    fields = random_fields(n, 4, seed=5, scale=0.5)
    _load_same(sim, orc, 0, fields)
    sim.fill_rhs(0, 0.7)
    orc.fill_rhs(0, 0.7)
    for c in range(4):
        err = rel_linf(sim.rhs(0, c), orc.rhs(0, c))
        assert err < RHS_TOL[prec], (c, err)


@pytest.mark.parametrize("integrator", [21, 22, 10, 11])
@pytest.mark.parametrize("prec,tol", PRECS)
def test_advance_real_scalar(integrator, prec, tol):
    """Gaussian random field (|phi_k| ~ 1/k, k <= 16, std 0.1), pi = 0."""
    inp = real_scalar(N=32, order=1, integrator=integrator, cfl=0.2)
    sim = shm.Simulation("RealScalar", inp, precision=prec)
    orc = Oracle("RealScalar", inp)
    phi = gaussian_random_field(32, seed=1234, kcut=16, std=0.1)
    _load_same(sim, orc, 0, [phi, np.zeros_like(phi)])
    sim.advance(0, 20)
    orc.advance(0, 20)
    for c in range(2):
        err = rel_linf(sim.download(0, c), orc.download(0, c))
        assert err < tol, (integrator, c, err)
    assert abs(sim.level_info(0)[1] - orc.time(0)) < 1e-14


@pytest.mark.parametrize("prec,tol", PRECS)
def test_advance_complex_scalar_ssprk3(prec, tol):
    inp = complex_scalar(N=32, integrator=10)
    sim = shm.Simulation("ComplexScalar", inp, precision=prec)
    orc = Oracle("ComplexScalar", inp)
    orc.init_state(0, 1, seed=42, params=[0.05])
    fields = [orc.download(0, c) for c in range(4)]
    _load_same(sim, orc, 0, fields)
    sim.advance(0, 10)
    orc.advance(0, 10)
    for c in range(4):
        err = rel_linf(sim.download(0, c), orc.download(0, c))
        assert err < tol, (c, err)


def test_device_initial_state_matches_oracle():
    inp = complex_scalar(N=32, integrator=10)
    sim = shm.Simulation("ComplexScalar", inp, precision=shm.PREC_FP64)
    orc = Oracle("ComplexScalar", inp)
    sim.init_state(0, 1, seed=7, params=[0.05])
    orc.init_state(0, 1, seed=7, params=[0.05])
    for c in range(4):
        assert np.max(np.abs(sim.download(0, c) - orc.download(0, c))) < 1e-13


@pytest.mark.parametrize("prec", ALLP)
def test_multibox_equals_single_box(prec):
    """A level chopped into 8 boxes (box-to-box halos) gives the single-box
    result bit-for-bit."""
    base = complex_scalar(N=32, integrator=10)
    a = shm.Simulation("ComplexScalar", base, precision=prec)
    b = shm.Simulation("ComplexScalar", dict(base, **{"amr.max_grid_size": 16}),
                       precision=prec)
    orc = Oracle("ComplexScalar", base)
    orc.init_state(0, 1, seed=3, params=[0.05])
    glob = [orc.download(0, c).reshape(32, 32, 32) for c in range(4)]
    # Level data are box-major: reorder the global field for the 8 boxes.
    boxes = b.boxes(0)
    for c in range(4):
        a.upload(0, c, glob[c].ravel())
        parts = [glob[c][bx[2]:bx[5] + 1, bx[1]:bx[4] + 1, bx[0]:bx[3] + 1].ravel()
                 for bx in boxes]
        b.upload(0, c, np.concatenate(parts))
    a.advance(0, 5)
    b.advance(0, 5)
    for c in range(4):
        ga = a.download(0, c).reshape(32, 32, 32)
        gb = b.download(0, c)
        off = 0
        for bx in boxes:
            nb = (bx[3] - bx[0] + 1) * (bx[4] - bx[1] + 1) * (bx[5] - bx[2] + 1)
            blk = ga[bx[2]:bx[5] + 1, bx[1]:bx[4] + 1, bx[0]:bx[3] + 1].ravel()
            assert np.array_equal(blk, gb[off:off + nb]), (c, bx)
            off += nb


@pytest.mark.parametrize("prec", ALLP)
def test_te_tag_bit_exact(prec):
    """Tags and counters from identical fine and shadow data: the GPU mask
    equals the oracle's literal CPU tag loop bit-for-bit."""
    inp = complex_scalar(N=32, integrator=10, **{"amr.te_crit": 2e-3})
    sim = shm.Simulation("ComplexScalar", inp, precision=prec)
    orc = Oracle("ComplexScalar", inp)
    n = sim.ncells(0)
    rng = np.random.default_rng(99)
    fine = [1.0 + 0.01 * rng.standard_normal(n) for _ in range(4)]
    if prec != shm.PREC_FP64:  # the TE reads the FP32 values of the state
        fine = [f.astype(np.float32).astype(np.float64) for f in fine]
    _load_same(sim, orc, 0, fine)
    nc = n // 8
    crse = [1.0 + 0.01 * rng.standard_normal(nc) for _ in range(4)]
    if prec != shm.PREC_FP64:
        crse = [f.astype(np.float32).astype(np.float64) for f in crse]
    _load_same(sim, orc, -1, crse)
    gm, gc, gte = sim.te_tag(0, want_te=True)
    om, oc, ote = orc.te_tag(0, want_te=True)
    assert np.array_equal(gte, ote)
    assert np.array_equal(gm, om)
    assert np.array_equal(gc, oc)
    assert gm.sum() > 0 and gm.sum() < n
    # Averaged-down coarse data.
    for c in range(4):
        assert rel_linf(sim.download(-1, c), orc.download(-1, c)) < PAIR_TOL[prec]


def _box_major(dense, boxes):
    """A dense [z, y, x] array in the box-major order of `boxes`."""
    return np.concatenate([dense[b[2]:b[5] + 1, b[1]:b[4] + 1, b[0]:b[3] + 1].ravel()
                           for b in boxes])


@pytest.mark.parametrize("modifier,user_on", [(0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("prec", ALLP)
def test_te_tag_user_tags_and_modifiers_bit_exact(prec, modifier, user_on):
    """ComplexScalar's user criterion (string winding through the forward
    plaquettes, reading the neighbours at +1, across box faces and the
    periodic boundary) and its truncation-error modifiers (te / max|Psi_n|
    with the norm taken at the tagging, te * dt for the momenta): mask,
    counters (8 - user for blocks whose even cell is user-tagged) and
    truncation errors equal the oracle's literal CPU loop bit for bit, on 8
    boxes of 8^3, and agree with the closed-form expectation."""
    N = 16
    boxes = [[x, y, z, x + 7, y + 7, z + 7] for z in (0, 8) for y in (0, 8) for x in (0, 8)]
    cboxes = [[v // 2 for v in b[:3]] + [v // 2 for v in b[3:]] for b in boxes]
    fine, crse, delta, norms = user_tag_case(N, seed=17)
    crse = [c.astype(np.float32).astype(np.float64) for c in crse]
    inp = complex_scalar(N=N, integrator=10, L=N * 0.25, cfl=0.2,
                         **{"amr.te_crit": 1e-3, "project.tag_strings": user_on,
                            "project.te_modifier": modifier})
    sim = shm.Simulation("ComplexScalar", inp, precision=prec)
    orc = Oracle("ComplexScalar", inp)
    for o in (sim, orc):
        o.set_boxes(0, boxes)
        for c in range(4):
            o.upload(0, c, _box_major(fine[c], boxes))
            o.upload(-1, c, _box_major(crse[c], cboxes))
    gm, gc, gte = sim.te_tag(0, want_te=True)
    om, oc, ote = orc.te_tag(0, want_te=True)
    assert np.array_equal(gte, ote)
    assert np.array_equal(gm, om)
    assert np.array_equal(gc, oc), (gc, oc)
    want_mask, want_counts = expected_tags(fine, delta, norms, modifier, user_on)
    assert np.array_equal(gm.astype(bool), _box_major(want_mask, boxes))
    assert np.array_equal(gc, want_counts)
    if user_on:
        assert 0 < gc[1] < N ** 3


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_gw_spectrum(mode):
    inp = phase_transition(N=16)
    sim = shm.Simulation("PhaseTransitionGW", inp, precision=shm.PREC_FP64)
    orc = Oracle("PhaseTransitionGW", inp)
    n = sim.ncells(0)
    fields = random_fields(n, 14, seed=8, scale=0.2)
    _load_same(sim, orc, 0, fields)
    gb, gk = sim.gw_spectrum(mode=mode, nbins=24, projection_type=2)
    ob, ok = orc.gw_spectrum(mode=mode, nbins=24, projection_type=2)
    assert np.array_equal(gk, ok)
    assert rel_linf(gb, ob) < 1e-11


def test_gw_spectrum_modifiers_and_zero_padding():
    """Project spectrum modifiers (|k| u_ij; du_ij plus a translated copy
    with per-call parameters) and zero padding 1 and 2, in the three binning
    modes, against the oracle; calls alternate sizes, modifiers and
    projection types so the cached plans and k tables are exercised."""
    inp = phase_transition(N=16)
    sim = shm.Simulation("PhaseTransitionGW", inp, precision=shm.PREC_FP64)
    orc = Oracle("PhaseTransitionGW", inp)
    assert sim.gw_modifiers() == ["SpectrumUTimesK", "SpectrumPlusTranslatedCopy"]
    fields = random_fields(sim.ncells(0), 14, seed=31, scale=0.2)
    _load_same(sim, orc, 0, fields)
    cases = [(-1, (), 2, 0, 2), (0, (), 2, 0, 3), (1, (0.5, -1.25, 2.0), 1, 0, 2),
             (0, (), 1, 2, 1), (1, (3.0, 0.0, 0.0), 2, 1, 2), (-1, (), 1, 0, 2),
             (0, (), 2, 2, 2)]
    for mod, par, zp, mode, ptype in cases:
        gb, gk = sim.gw_spectrum(mode=mode, nbins=24, zero_padding=zp,
                                 projection_type=ptype, modifier=mod, params=par)
        ob, ok = orc.gw_spectrum(mode=mode, nbins=24, zero_padding=zp,
                                 projection_type=ptype, modifier=mod, params=par)
        assert np.array_equal(gk, ok), (mod, zp, mode)
        assert ob.max() > 0
        assert rel_linf(gb, ob) < 1e-11, (mod, zp, mode, rel_linf(gb, ob))
    with pytest.raises(shm.ShmError):
        sim.gw_spectrum(modifier=2)


@pytest.mark.parametrize("prec,tol", PRECS)
def test_advance_gw_rkn4(prec, tol):
    inp = phase_transition(N=16, integrator=21)
    sim = shm.Simulation("PhaseTransitionGW", inp, precision=prec)
    orc = Oracle("PhaseTransitionGW", inp)
    n = sim.ncells(0)
    fields = random_fields(n, 14, seed=21, scale=0.1)
    fields[0] = 0.5 + 0.3 * np.tanh(np.linspace(-3, 3, n))
    _load_same(sim, orc, 0, fields)
    sim.advance(0, 5)
    orc.advance(0, 5)
    for c in range(14):
        err = rel_linf(sim.download(0, c), orc.download(0, c))
        assert err < tol, (c, err)
