"""Every stage-kernel tile shape gives the same cells: the 64-wide tiles
(4, 8, 16 rows) and the 32-wide tile (8 rows, two rows per wavefront, for
boxes narrower than 64 cells) on boxes of 24-40 cells, bit for bit in FP64
and FP32-compensated, and the 32-wide tile against the oracle in FP64
(1e-12).  The FP32-compensated path equals the default tile bit for bit,
whose oracle parity the other suites hold; on this data (smooth fields,
dx = 1/4, four RKN4 steps) its momenta carry 3.7e-6 of FP32 stencil-input
rounding with every tile shape alike."""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import AMR_LEV1, real_scalar, rel_linf, smooth_boxes
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

NARROW = [[0, 0, 0, 23, 31, 15], [24, 0, 0, 47, 31, 15], [48, 0, 0, 63, 31, 15],
          [0, 32, 0, 39, 63, 15], [40, 32, 0, 63, 63, 15],
          [0, 0, 16, 63, 63, 63]]


def _run(integrator, prec, variant):
    inp = real_scalar(N=64, order=2, nghost=2, lam=1.0, L=16.0, integrator=integrator,
                      **{"shm.independent_boxes": 0})
    sim = shm.Simulation("RealScalar", inp, precision=prec)
    sim.set_boxes(0, NARROW)
    for c, ph in ((0, 0.4), (1, 1.3)):
        sim.upload(0, c, smooth_boxes(NARROW, 0.25, 16.0, ph) * (1.0 if c == 0 else 0.2))
    sim.set_launch(variant=variant, tile_z=16)
    sim.advance(0, 3)
    return sim, [sim.download(0, c) for c in range(2)]


@pytest.mark.parametrize("integrator", [10, 21, 11])
@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
def test_tile_variants_agree(integrator, prec):
    ref = None
    for v in (1, 0, 2, 3, 4):
        sim, out = _run(integrator, prec, v)
        assert sim.get_launch(0)["variant"] == v
        if ref is None:
            ref = out
        else:
            for c in range(2):
                assert np.array_equal(out[c], ref[c]), (v, c)
        sim.close()


@pytest.mark.parametrize("prec,tol", [(shm.PREC_FP64, 1e-12)])
def test_narrow_tile_matches_oracle(prec, tol):
    inp = real_scalar(N=64, order=2, nghost=2, lam=1.0, L=16.0, integrator=21)
    sim = shm.Simulation("RealScalar", inp, precision=prec)
    orc = Oracle("RealScalar", inp)
    for o in (sim, orc):
        o.set_boxes(0, NARROW)
        for c, ph in ((0, 0.4), (1, 1.3)):
            o.upload(0, c, smooth_boxes(NARROW, 0.25, 16.0, ph) * (1.0 if c == 0 else 0.2))
    sim.set_launch(variant=3)
    sim.advance(0, 4)
    orc.advance(0, 4)
    for c in range(2):
        assert rel_linf(sim.download(0, c), orc.download(0, c)) < tol


def test_narrow_tile_on_fine_levels():
    """Sub-cycled three-level steps with the 32-wide tile on every level
    (coarse/fine ghost fills included) equal the default tile bit for bit."""
    res = []
    for v in (1, 3):
        inp = real_scalar(N=16, order=2, nghost=2, lam=1.0, L=16.0, cfl=0.25,
                          integrator=10, **{"amr.max_refinement_levels": 1})
        sim = shm.Simulation("RealScalar", inp, precision=shm.PREC_FP32C16)
        sim.set_boxes(1, AMR_LEV1)
        for lev, bx in ((0, [[0, 0, 0, 15, 15, 15]]), (1, AMR_LEV1)):
            for c, ph in ((0, 0.3), (1, 1.1)):
                sim.upload(lev, c, smooth_boxes(bx, 16.0 / (16 * 2 ** lev), 16.0, ph))
        sim.set_launch(variant=v)
        sim.evolve(2)
        res.append([sim.download(lev, c) for lev in range(2) for c in range(2)])
    for a, b in zip(*res):
        assert np.array_equal(a, b)
