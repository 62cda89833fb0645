"""Initial-state arrays in the reference's orders: FillLevel::FromArray
(fill_level.cpp:205-223, the whole level with ind = i N^2 + j N + k) and
FillLevel::FromArrayChunks (fill_level.cpp:225-251, per box,
ind = (i - lo) ly lz + (j - lo) lz + (k - lo)).  The expected placement is
numpy's C order of an [i, j, k] array, compared bit for bit after a
download (FP64 storage)."""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import real_scalar

pytestmark = pytest.mark.gpu


def _box_major(glob, boxes):
    """[i, j, k] global array -> valid cells box after box, x fastest."""
    return np.concatenate([
        np.transpose(glob[b[0]:b[3] + 1, b[1]:b[4] + 1, b[2]:b[5] + 1], (2, 1, 0)).ravel()
        for b in boxes])


@pytest.mark.parametrize("mgs", [16, 8])
def test_from_array_whole_level(mgs):
    sim = shm.Simulation("RealScalar", real_scalar(N=16, **{"amr.max_grid_size": mgs}),
                         precision=shm.PREC_FP64)
    glob = np.random.default_rng(mgs).standard_normal((16, 16, 16))
    sim.upload_array(0, 0, glob.ravel(), 1)
    assert np.array_equal(sim.download(0, 0), _box_major(glob, sim.boxes(0)))


def test_from_array_chunks_per_box():
    sim = shm.Simulation("RealScalar", real_scalar(N=16, **{"amr.max_grid_size": 8}),
                         precision=shm.PREC_FP64)
    boxes = sim.boxes(0)
    glob = np.random.default_rng(1).standard_normal((16, 16, 16))
    chunks = np.concatenate([glob[b[0]:b[3] + 1, b[1]:b[4] + 1, b[2]:b[5] + 1].ravel()
                             for b in boxes])
    sim.upload_array(0, 1, chunks, 2)
    assert np.array_equal(sim.download(0, 1), _box_major(glob, boxes))


def test_from_array_fine_level():
    """A level-1 box: the whole-level array spans the level-1 domain."""
    sim = shm.Simulation("RealScalar", real_scalar(N=8, **{"amr.max_refinement_levels": 1}),
                         precision=shm.PREC_FP64)
    box = [4, 2, 6, 9, 7, 13]
    sim.set_boxes(1, [box])
    glob = np.random.default_rng(2).standard_normal((16, 16, 16))
    sim.upload_array(1, 0, glob.ravel(), 1)
    assert np.array_equal(sim.download(1, 0), _box_major(glob, [box]))
    with pytest.raises(shm.ShmError):
        sim.upload_array(1, 0, glob.ravel(), 3)
