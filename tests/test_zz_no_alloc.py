"""The allocation-free tagging entry point (shm_te_tag_noalloc) against
shm_te_tag through the C ABI: masks and counters equal on two calls in a
row, the second on a smaller level, with user criteria, modifiers and
truncation errors; on a sub-cycled fine level (coarse/fine ghosts); and its
refusals (level not set up, level above the scratch array's maximum)."""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import AMR_LEV1, AMR_LEV2, complex_scalar, real_scalar, smooth_boxes

pytestmark = pytest.mark.gpu

FLAGS = shm.TE_COMPUTE | shm.TE_TAGS  # no average-down: the state stays


def _complex_level0(N, prec, seed):
    inp = complex_scalar(N=N, integrator=10, L=N * 0.25, cfl=0.2,
                         **{"amr.te_crit": 1e-3, "project.tag_strings": 1,
                            "project.te_modifier": 1})
    sim = shm.Simulation("ComplexScalar", inp, precision=prec)
    rng = np.random.default_rng(seed)
    n = sim.ncells(0)
    for c in range(4):
        f = rng.standard_normal(n).astype(np.float32).astype(np.float64)
        sim.upload(0, c, f)
        cr = (rng.standard_normal(n // 8) * 0.01).astype(np.float32).astype(np.float64)
        sim.upload(-1, c, cr)
    return sim


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16])
def test_noalloc_tagging_equals_te_tag_two_calls(prec):
    sims = [_complex_level0(32, prec, 5), _complex_level0(16, prec, 6)]
    for sim in sims:  # the second call: a smaller level on the same GPU
        m1, c1, _ = sim.te_tag(0, flags=FLAGS)
        m2, c2 = sim.te_tag_noalloc(0, flags=FLAGS)
        assert np.array_equal(m1, m2)
        assert np.array_equal(c1, c2)
        assert 0 < c1[1] and 0 < c1[0] < sim.ncells(0) * 8
    # And again on the larger level after the smaller one.
    m3, c3 = sims[0].te_tag_noalloc(0, flags=FLAGS)
    m4, c4, _ = sims[0].te_tag(0, flags=FLAGS)
    assert np.array_equal(m3, m4) and np.array_equal(c3, c4)


def test_noalloc_tagging_fine_level_with_coarse_fine_ghosts():
    inp = real_scalar(N=16, order=2, nghost=2, lam=1.0, L=16.0, cfl=0.25,
                      integrator=10, **{"amr.max_refinement_levels": 2,
                                        "amr.te_crit": 1e-4})
    sim = shm.Simulation("RealScalar", inp, precision=shm.PREC_FP64)
    sim.set_boxes(1, AMR_LEV1)
    sim.set_boxes(2, AMR_LEV2)
    for lev, bx in ((0, [[0, 0, 0, 15, 15, 15]]), (1, AMR_LEV1), (2, AMR_LEV2)):
        h = 16.0 / (16 * 2 ** lev)
        for c, ph in ((0, 0.3), (1, 1.1)):
            sim.upload(lev, c, smooth_boxes(bx, h, 16.0, ph) * (1.0 if c == 0 else 0.2))
    with pytest.raises(shm.ShmError) as e:
        sim.te_tag_noalloc(2, flags=FLAGS)
    assert e.value.code == 3
    m1, c1, _ = sim.te_tag(2, flags=FLAGS)
    m2, c2 = sim.te_tag_noalloc(2, flags=FLAGS)
    assert np.array_equal(m1, m2) and np.array_equal(c1, c2)
    assert c1[0] > 0


def test_noalloc_tagging_refuses_levels_above_the_scratch_size():
    inp = complex_scalar(N=512, integrator=10, **{"project.tag_strings": 1})
    sim = shm.Simulation("ComplexScalar", inp, precision=shm.PREC_FP32C16)
    with pytest.raises(shm.ShmError) as e:
        sim.te_tag_noalloc(0, flags=shm.TE_TAGS)
    assert e.value.code == 1 and "at most 33550336" in str(e.value)
