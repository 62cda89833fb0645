"""Pins of the oracle's local regrid (local_regrid/local_regrid.cpp,
unique_layout.cpp), independent of its sequential restatement:

* block selection (DetermineNewBoxArray / CheckBorders) against a brute
  force over random masks: squared distances as the minimum over every cell
  of a block of the Euclidean distance to the tagged cell's fine index, and
  the reference's scan-order effect (a block unmarked once added is not
  measured again in that box) as "a cell measures a block iff its box-local
  index is at most that of the first cell adding it";
* a hand-built case where that effect changes the reported distance;
* FixNesting ranges worked out by hand (with the periodic wrap) and a
  nesting property on random block sets;
* the block layout covers exactly the blocks' cells;
* AddBoxes / RemakeLevel keep old data bit for bit and interpolate new boxes
  exactly on quadratic coarse data, linearly in time between coarse states;
* DoAttemptRegrid outcomes for the volume thresholds and the prechecks.
"""
import itertools

import numpy as np
import pytest

from cases import real_scalar
from oracle_py import Oracle, block_layout


def _inputs(N=16, levels=1, bf=4, nbuf=1, **extra):
    d = real_scalar(N=N, order=2, nghost=2, lam=1.0, L=float(N),
                    **{"amr.max_refinement_levels": levels, "amr.blocking_factor": bf,
                       "amr.n_error_buf": nbuf})
    d.update(extra)
    return d


def _blocks_to_boxes(blocks, bf):
    return block_layout("RealScalar", blocks, bf) if len(blocks) else np.zeros((0, 6), int)


def _cells_of_boxes(boxes):
    s = set()
    for b in boxes:
        for k in range(b[2], b[5] + 1):
            for j in range(b[1], b[4] + 1):
                for i in range(b[0], b[3] + 1):
                    s.add((i, j, k))
    return s


def _brute_select(boxes0, boxes1, mask, bf, N1, nbuf):
    """Blocks of level 1 to add and the smallest measured squared distance,
    by enumeration."""
    Np = N1 // bf
    thr = (nbuf + 1) ** 2
    cov = np.zeros((Np, Np, Np), bool)
    for q in itertools.product(range(Np), repeat=3):
        c = [qa * bf + bf // 2 for qa in q]
        cov[q] = any(all(b[a] <= c[a] <= b[a + 3] for a in range(3)) for b in boxes1)
    offs = list(itertools.product((-1, 0, 1), repeat=3))
    added, best, moff = set(), None, 0
    for b in boxes0:
        n = [b[a + 3] - b[a] + 1 for a in range(3)]
        tagged = []
        for local in range(n[0] * n[1] * n[2]):
            if mask[moff + local]:
                i = b[0] + local % n[0]
                j = b[1] + (local // n[0]) % n[1]
                k = b[2] + local // (n[0] * n[1])
                tagged.append((local, (2 * i, 2 * j, 2 * k)))
        dist = {}
        for local, f in tagged:
            for d in offs:
                u = tuple(f[a] // bf + d[a] for a in range(3))
                if cov[tuple(x % Np for x in u)]:
                    continue
                rng = [range(u[a] * bf, u[a] * bf + bf) for a in range(3)]
                d2 = min(sum((f[a] - c[a]) ** 2 for a in range(3))
                         for c in itertools.product(*rng))
                dist[(local, u)] = d2
        first = {}
        for (local, u), d2 in dist.items():
            if d2 < thr:
                first[u] = min(first.get(u, local), local)
                added.add(tuple(x % Np for x in u))
        for (local, u), d2 in dist.items():
            if local <= first.get(u, 1 << 40):
                best = d2 if best is None else min(best, d2)
        moff += n[0] * n[1] * n[2]
    return sorted(added), best


@pytest.mark.parametrize("case", range(8))
def test_select_blocks_brute_force(case):
    rng = np.random.default_rng(700 + case)
    bf = [2, 4, 4, 8][case % 4]
    nbuf = [0, 1, 2][case % 3]
    mgs = 8 if case % 2 else 16
    inp = _inputs(bf=bf, nbuf=nbuf, **{"amr.max_grid_size": mgs})
    o = Oracle("RealScalar", inp)
    Np = 32 // bf
    fine_blocks = [q for q in itertools.product(range(Np), repeat=3)
                   if rng.random() < 0.3]
    boxes1 = _blocks_to_boxes(fine_blocks, bf)
    if len(boxes1):
        o.set_boxes(1, boxes1)
    boxes0 = o.boxes(0)
    mask = (rng.random(o.ncells(0)) < 0.01).astype(np.uint8)
    got, dist = o.regrid_select(0, mask)
    want, best = _brute_select(boxes0, boxes1, mask, bf, 32, nbuf)
    assert [tuple(x) for x in got] == want
    assert dist == (-1.0 if best is None else np.sqrt(best))


def test_select_first_cell_unmarks_block():
    """Level 1 holds every 4-cell block but U = (3, 3, 3).  Cell A (fine
    index (16, 16, 12), d^2 = 2 to U) comes before cell B ((16, 14, 14),
    d^2 = 1) in k, j, i order; A adds U and unmarks it, so B's distance is
    never measured: sqrt(2), not 1."""
    inp = _inputs(bf=4, nbuf=1)
    o = Oracle("RealScalar", inp)
    blocks = [q for q in itertools.product(range(8), repeat=3) if q != (3, 3, 3)]
    o.set_boxes(1, _blocks_to_boxes(blocks, 4))
    mask = np.zeros(16 ** 3, np.uint8)
    for ci in ((8, 8, 6), (8, 7, 7)):
        mask[ci[0] + 16 * (ci[1] + 16 * ci[2])] = 1
    got, dist = o.regrid_select(0, mask)
    assert [tuple(x) for x in got] == [(3, 3, 3)]
    assert dist == pytest.approx(np.sqrt(2.0), abs=0)
    # Either cell alone: its own distance.
    for ci, d2 in (((8, 8, 6), 2.0), ((8, 7, 7), 1.0)):
        m = np.zeros(16 ** 3, np.uint8)
        m[ci[0] + 16 * (ci[1] + 16 * ci[2])] = 1
        assert o.regrid_select(0, m)[1] == np.sqrt(d2)


def test_select_own_block_and_threshold():
    inp = _inputs(bf=4, nbuf=1)
    o = Oracle("RealScalar", inp)
    o.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    m = np.zeros(16 ** 3, np.uint8)
    # Fine index 10 inside covered block 2; block 1 (cells 4..7) is 3 away.
    m[5 + 16 * (5 + 16 * 5)] = 1
    got, dist = o.regrid_select(0, m)
    assert len(got) == 0 and dist == 3.0
    # Fine index 6 inside uncovered block 1: distance 0, the block is added;
    # block 2 (8..11) is 2 away, 4 is not below (1 + 1)^2.
    m[:] = 0
    m[3 + 16 * (3 + 16 * 3)] = 1
    got, dist = o.regrid_select(0, m)
    assert [tuple(x) for x in got] == [(1, 1, 1)] and dist == 0.0
    # Periodic wrap: fine index 30 (block 7) next to block 0 across x = 32.
    o.set_boxes(1, [[28, 0, 0, 31, 31, 31]])
    m[:] = 0
    m[15 + 16 * (1 + 16 * 1)] = 1
    got, _ = o.regrid_select(0, m)
    assert (0, 0, 0) not in [tuple(x) for x in got]
    inp2 = _inputs(bf=4, nbuf=2)
    o2 = Oracle("RealScalar", inp2)
    o2.set_boxes(1, [[28, 0, 0, 31, 31, 31]])
    got2, _ = o2.regrid_select(0, m)
    assert (0, 0, 0) in [tuple(x) for x in got2]


def test_fix_nesting_ranges():
    """A level-2 block grown by nghost + 4 = 6 and coarsened by 2 spans
    level-1 cells 5..12 (block 4, cells 16..19); the reference adds blocks
    lo / 4 .. hi / 4 + 1 = 1..4 per axis that level 1 does not hold."""
    inp = _inputs(levels=2, bf=4)
    o = Oracle("RealScalar", inp)
    o.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    o.set_boxes(2, [[16, 16, 16, 23, 23, 23]])
    got = {tuple(x) for x in o.fix_nesting(2, [[4, 4, 4]])}
    want = set(itertools.product(range(1, 5), repeat=3)) - set(
        itertools.product((2, 3), repeat=3))
    assert got == want
    # Block 0: cells -6..9 coarsened to -3..4, cut into 29..31 and 0..4:
    # blocks 7..8 (8 wraps to 0) and 0..2 per axis.
    got = {tuple(x) for x in o.fix_nesting(2, [[0, 0, 0]])}
    axis = (7, 0, 1, 2)
    want = set(itertools.product(axis, repeat=3)) - set(itertools.product((2, 3), repeat=3))
    assert got == want


def test_fix_nesting_covers_grown_blocks():
    rng = np.random.default_rng(5)
    inp = _inputs(levels=2, bf=4)
    o = Oracle("RealScalar", inp)
    o.set_boxes(1, [[8, 8, 8, 23, 23, 23]])
    lev1 = _cells_of_boxes([[8, 8, 8, 23, 23, 23]])
    for _ in range(4):
        blocks = [q for q in itertools.product(range(16), repeat=3) if rng.random() < 0.004]
        add = {tuple(x) for x in o.fix_nesting(2, blocks)}
        have = lev1 | _cells_of_boxes(_blocks_to_boxes(sorted(add), 4) if add else [])
        for q in blocks:
            for c in itertools.product(*[range(4 * q[a] - 6, 4 * q[a] + 10) for a in range(3)]):
                cc = tuple((x // 2) % 32 for x in c)
                assert cc in have, (q, c)
        # Only blocks level 1 does not hold.
        assert all(not (2 <= b[0] <= 5 and 2 <= b[1] <= 5 and 2 <= b[2] <= 5) for b in add)


def test_block_layout_covers_blocks():
    rng = np.random.default_rng(11)
    for bf in (2, 4, 8):
        Np = 32 // bf
        blocks = [q for q in itertools.product(range(Np), repeat=3) if rng.random() < 0.2]
        boxes = block_layout("RealScalar", blocks, bf)
        cells = set()
        for b in boxes:
            assert all(b[a] % bf == 0 and (b[a + 3] + 1) % bf == 0 for a in range(3))
            c = _cells_of_boxes([b])
            assert not (cells & c)
            cells |= c
        want = set()
        for q in blocks:
            want |= _cells_of_boxes([[q[0] * bf, q[1] * bf, q[2] * bf,
                                      q[0] * bf + bf - 1, q[1] * bf + bf - 1,
                                      q[2] * bf + bf - 1]])
        assert cells == want
        assert len(boxes) <= len(blocks)


def _quad(boxes, h, coef):
    out = []
    for b in boxes:
        k, j, i = np.meshgrid(*(np.arange(b[a], b[a + 3] + 1) for a in (2, 1, 0)),
                              indexing="ij")
        x, y, z = (i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h
        a = coef
        out.append((a[0] + a[1] * x + a[2] * y * y + a[3] * x * z + a[4] * z * z
                    + a[5] * y * z).ravel())
    return np.concatenate(out)


def test_add_boxes_keeps_data_and_interpolates_quadratics():
    """Coarse states are quadratics in space (quadratic interpolation is
    exact), the old one at t = 0.5, the new one at t = 1.0; level 1 gains a
    box at its time 0.75 (and 0.5 for its old state): new cells hold the
    time-weighted quadratic, old boxes keep their data bit for bit."""
    inp = _inputs(bf=4)
    o = Oracle("RealScalar", inp)
    box0 = [[0, 0, 0, 15, 15, 15]]
    cn, co = [0.3, 0.1, 0.02, -0.01, 0.03, 0.015], [-0.2, 0.05, 0.01, 0.02, -0.02, 0.01]
    o.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    for c in range(2):
        o.upload(0, c, _quad(box0, 1.0, cn) * (c + 1))
        o.upload_old(0, c, _quad(box0, 1.0, co) * (c + 1), 0.5)
    o.set_time(0, 1.0, 2)
    rng = np.random.default_rng(3)
    old1 = [rng.standard_normal(512) for _ in range(2)]
    old1o = [rng.standard_normal(512) for _ in range(2)]
    for c in range(2):
        o.upload(1, c, old1[c])
        o.upload_old(1, c, old1o[c], 0.5)
    o.set_time(1, 0.75, 1)
    nb = [[16, 8, 8, 23, 15, 15], [4, 20, 2, 11, 27, 9]]
    o.add_boxes(1, nb)
    assert o.boxes(1).tolist() == [[8, 8, 8, 15, 15, 15]] + nb
    for c in range(2):
        new = o.download(1, c)
        assert np.array_equal(new[:512], old1[c])
        want = (0.5 * _quad(nb, 0.5, cn) + 0.5 * _quad(nb, 0.5, co)) * (c + 1)
        assert np.max(np.abs(new[512:] - want)) < 1e-13
        old = o.download(1, c, old=True)
        assert np.array_equal(old[:512], old1o[c])
        assert np.max(np.abs(old[512:] - _quad(nb, 0.5, co) * (c + 1))) < 1e-13


def test_remake_level_fills_from_level_and_coarse():
    inp = _inputs(bf=4)
    o = Oracle("RealScalar", inp)
    box0 = [[0, 0, 0, 15, 15, 15]]
    cn = [0.3, 0.1, 0.02, -0.01, 0.03, 0.015]
    o.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    for c in range(2):
        o.upload(0, c, _quad(box0, 1.0, cn))
        o.upload(1, c, np.full(512, 7.0 + c))
    new = [[8, 8, 8, 23, 15, 15]]
    o.remake_level(1, new)
    for c in range(2):
        v = o.download(1, c).reshape(8, 8, 16)
        assert np.all(v[:, :, :8] == 7.0 + c)
        want = _quad(new, 0.5, cn).reshape(8, 8, 16)[:, :, 8:]
        assert np.max(np.abs(v[:, :, 8:] - want)) < 1e-13


def _regrid_setup(**extra):
    inp = _inputs(bf=4, **extra)
    o = Oracle("RealScalar", inp)
    o.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    for lev in range(2):
        for c in range(2):
            o.upload(lev, c, np.zeros(o.ncells(lev)))
    o.set_time(0, 1.0, 1)
    o.set_time(1, 1.0, 2)
    return o


def test_local_regrid_adds_blocks_within_thresholds():
    o = _regrid_setup(**{"amr.volume_threshold_weak": 2.0,
                         "amr.volume_threshold_strong": 2.0})
    m = np.zeros(16 ** 3, np.uint8)
    # Fine index (6, 8, 8): its own block (1, 2, 2) at d^2 = 0, and blocks
    # (1, 1, 2), (1, 2, 1), (1, 1, 1) one and two cells away (y, z = 8 next
    # to cell 7); covered block (2, 2, 2) is not added.
    m[3 + 16 * (4 + 16 * 4)] = 1
    o.set_tags(0, m)
    assert o.local_regrid(0) == 1
    assert o.boxes(1).tolist() == [[8, 8, 8, 15, 15, 15], [4, 4, 4, 7, 11, 11]]
    st = o.regrid_stats(2)
    assert st["attempts"] == 1 and st["nregrids"] == 1 and st["min_distance"][1] == 0.0


def test_local_regrid_volume_veto_and_prechecks():
    # 8 new blocks of 64 cells on a 512-cell level: (512 + 512) / 512 = 2 >
    # 1.1, so a veto at level 0 asks for a global regrid.
    o = _regrid_setup()
    m = np.zeros(16 ** 3, np.uint8)
    for ci in itertools.product((2, 9), repeat=3):
        m[ci[0] + 16 * (ci[1] + 16 * ci[2])] = 1
    o.set_tags(0, m)
    assert o.local_regrid(0) == 0
    assert o.boxes(1).tolist() == [[8, 8, 8, 15, 15, 15]]
    assert o.regrid_stats(2)["veto_level"] == 0
    # The finest level cannot be regridded locally; max_local_regrids.
    o2 = _regrid_setup(**{"amr.max_local_regrids": 1})
    assert o2.local_regrid(1) == 0
    o2.set_tags(0, np.zeros(16 ** 3, np.uint8))
    assert o2.local_regrid(0) == 0
    assert o2.regrid_stats(2)["nregrids"] == 2
