"""Pins of the FP64 CPU oracle (no GPU).

Every check derives its expected value independently of the oracle's code:
polynomial exactness and plane-wave symbols of the central difference
stencils, closed-form amplification matrices and convergence orders of the
integrators, hand-built truncation-error blocks, numpy's FFT, the algebraic
identities of the transverse-traceless projector, plane-wave spectra, and
the reference's own k-bin tables (tests/golden/spectra_ks.json, extracted
from data/spectra_ks.hdf5).
"""
from __future__ import annotations

import itertools
import json
import os

import numpy as np
import pytest

from cases import (complex_scalar, expected_tags, phase_transition, real_scalar,
                   user_tag_case, winding_tags)
from oracle_py import (Oracle, oracle_lambda, oracle_r2c, oracle_unique_k2,
                       spectral_lib)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------------------
# Stencils (sledgehamr_utils.h:50-188)
# ---------------------------------------------------------------------------

def _grid(N, dx, shift):
    """Cell coordinates (x fastest, flattened [z][y][x])."""
    c = (np.arange(N) - shift) * dx
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    return x, y, z


def _lap_probe(N, order, phi, dx=1.0):
    """rhs(dPhi) of RealScalar with lambda = 0 is exactly Laplacian<order>."""
    orc = Oracle("RealScalar", real_scalar(N=N, order=order, lam=0.0, L=N * dx,
                                           nghost=order))
    orc.upload(0, 0, phi.ravel())
    orc.upload(0, 1, np.zeros(N ** 3))
    orc.fill_rhs(0, 0.0)
    return orc.rhs(0, 1).reshape(N, N, N)


def _interior(N, r):
    s = slice(r, N - r)
    return (s, s, s)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_laplacian_polynomial_exact(order):
    """A central 2p-order second difference is exact on polynomials whose
    restriction to each axis has degree <= 2p + 1; checked away from the
    periodic seam, with an analytic Laplacian."""
    N, deg = 20, 2 * order + 1
    x, y, z = _grid(N, 0.5, N / 2)
    rng = np.random.default_rng(order)
    phi = np.zeros_like(x)
    lap = np.zeros_like(x)
    for a, b, c in itertools.product(range(deg + 1), repeat=3):
        if max(a, b, c) > deg or a + b + c > deg + 2:
            continue
        w = rng.uniform(-1, 1)
        phi += w * x ** a * y ** b * z ** c
        if a >= 2:
            lap += w * a * (a - 1) * x ** (a - 2) * y ** b * z ** c
        if b >= 2:
            lap += w * b * (b - 1) * x ** a * y ** (b - 2) * z ** c
        if c >= 2:
            lap += w * c * (c - 1) * x ** a * y ** b * z ** (c - 2)
    got = _lap_probe(N, order, phi, dx=0.5)
    sl = _interior(N, order)
    err = np.max(np.abs(got[sl] - lap[sl])) / np.max(np.abs(lap[sl]))
    assert err < 1e-10, err


@pytest.mark.parametrize("order", [1, 2, 3])
def test_laplacian_polynomial_negative_control(order):
    """One degree higher the stencil is no longer exact (the check above
    would catch a stencil of lower order)."""
    N, deg = 20, 2 * order + 2
    x, y, z = _grid(N, 0.5, N / 2)
    phi = x ** deg
    lap = deg * (deg - 1) * x ** (deg - 2)
    got = _lap_probe(N, order, phi, dx=0.5)
    sl = _interior(N, order)
    assert np.max(np.abs(got[sl] - lap[sl])) > 1e-3


def _lap_symbol(order, th):
    """Per-axis eigenvalue (times dx^2, negated) of the central second
    difference of order 2p on exp(i th n)."""
    c = np.cos
    if order == 1:
        return 2 - 2 * c(th)
    if order == 2:
        return (30 - 32 * c(th) + 2 * c(2 * th)) / 12
    return (490 - 540 * c(th) + 54 * c(2 * th) - 4 * c(3 * th)) / 180


def _grad_symbol(order, th):
    """Eigenvalue (times dx, over i) of the central first difference."""
    s = np.sin
    if order == 1:
        return s(th)
    if order == 2:
        return (8 * s(th) - s(2 * th)) / 6
    return (45 * s(th) - 9 * s(2 * th) + s(3 * th)) / 30


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("m", [(1, 0, 0), (2, 3, 1), (5, 7, 3)])
def test_laplacian_plane_wave_symbol(order, m):
    N, dx = 16, 0.25
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    arg = 2 * np.pi * (m[0] * i + m[1] * j + m[2] * k) / N + 0.3
    phi = np.cos(arg)
    th = [2 * np.pi * q / N for q in m]
    lam = -sum(_lap_symbol(order, t) for t in th) / dx ** 2
    got = _lap_probe(N, order, phi, dx=dx)
    assert np.max(np.abs(got - lam * phi)) < 1e-11 * max(abs(lam), 1)


def test_gradient_plane_wave_symbol():
    """du_ij source of PhaseTransitionGW with u = 0 is d_i Phi d_j Phi
    through Gradient<3>: a plane wave gives the first-difference symbol."""
    N, dx = 16, 0.5
    m = (3, 1, 2)
    orc = Oracle("PhaseTransitionGW", phase_transition(N=N, L=N * dx))
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    arg = 2 * np.pi * (m[0] * i + m[1] * j + m[2] * k) / N + 0.1
    orc.upload(0, 0, np.sin(arg).ravel())
    for c in range(1, orc.ncomp):
        orc.upload(0, c, np.zeros(N ** 3))
    orc.fill_rhs(0, 0.0)
    g = [_grad_symbol(3, 2 * np.pi * q / N) / dx * np.cos(arg) for q in m]
    # du_xx, du_yy, du_zz, du_xy, du_xz, du_yz follow the six u_ij.
    want = {8: g[0] * g[0], 9: g[1] * g[1], 10: g[2] * g[2], 11: g[0] * g[1],
            12: g[0] * g[2], 13: g[1] * g[2]}
    for comp, w in want.items():
        got = orc.rhs(0, comp).reshape(N, N, N)
        assert np.max(np.abs(got - w)) < 1e-12, comp


def test_gradient_polynomial_exact():
    """Gradient<3> is exact on polynomials of degree <= 6 per axis."""
    N, dx = 20, 0.5
    x, y, z = _grid(N, dx, N / 2)
    orc = Oracle("PhaseTransitionGW", phase_transition(N=N, L=N * dx))
    phi = 0.01 * (x ** 6 - 2 * x ** 3 * y ** 2 + y ** 5 * z + z ** 4)
    gx = 0.01 * (6 * x ** 5 - 6 * x ** 2 * y ** 2)
    gy = 0.01 * (-4 * x ** 3 * y + 5 * y ** 4 * z)
    gz = 0.01 * (y ** 5 + 4 * z ** 3)
    orc.upload(0, 0, phi.ravel())
    for c in range(1, orc.ncomp):
        orc.upload(0, c, np.zeros(N ** 3))
    orc.fill_rhs(0, 0.0)
    sl = _interior(N, 3)
    for comp, w in {8: gx * gx, 11: gx * gy, 13: gy * gz, 10: gz * gz}.items():
        got = orc.rhs(0, comp).reshape(N, N, N)
        assert np.max(np.abs(got[sl] - w[sl])) < 1e-9 * np.max(np.abs(w[sl]))


# ---------------------------------------------------------------------------
# Kreiss-Oliger dissipation (kernels.h:56-106, applied per component at
# macros.h:355-373 with the strengths parsed at sledgehamr.cpp:508-546)
# ---------------------------------------------------------------------------

KO_CASES = {
    # order 2 from sim.dissipation_order, one strength per component
    2: {"sim.dissipation_order": 2, "sim.dissipation_strength_Phi": 0.3,
        "sim.dissipation_strength_dPhi": 0.7},
    # order 3 from amr.nghost (the default), default strength with an override
    3: {"sim.dissipation_strength": 0.2, "sim.dissipation_strength_dPhi": 0.5},
}
KO_STRENGTH = {2: (0.3, 0.7), 3: (0.2, 0.5)}


@pytest.mark.parametrize("order", [2, 3])
def test_kreiss_oliger_plane_wave_symbol(order):
    """The binomial 2p-th difference of cos(k.x) along axis a is
    (2i sin(th_a/2))^(2p) cos(k.x), so KO<p> of a plane wave is
    -s sum_a sin^(2p)(th_a/2) / dx times the wave, with s the strength of
    that component.  The dissipation term is the Rhs with KO minus without."""
    N, dx = 16, 0.25
    base = real_scalar(N=N, order=1, nghost=order, lam=0.0, L=N * dx)
    with_ko = Oracle("RealScalar", dict(base, **KO_CASES[order]))
    no_ko = Oracle("RealScalar", base)
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    waves, syms = [], []
    for m in [(3, 1, 0), (0, 5, 2)]:
        waves.append(np.cos(2 * np.pi * (m[0] * i + m[1] * j + m[2] * k) / N + 0.4))
        syms.append(sum(np.sin(np.pi * q / N) ** (2 * order) for q in m) / dx)
    for orc in (with_ko, no_ko):
        for c in range(2):
            orc.upload(0, c, waves[c].ravel())
        orc.fill_rhs(0, 0.0)
    for c in range(2):
        got = (with_ko.rhs(0, c) - no_ko.rhs(0, c)).reshape(N, N, N)
        want = -KO_STRENGTH[order][c] * syms[c] * waves[c]
        assert np.max(np.abs(got - want)) < 1e-11, (order, c)


# ---------------------------------------------------------------------------
# Coarse/fine ghost fill (FillPatchTwoLevels at level_synchronizer.cpp:92-209,
# GetLevelData at :613-632).  AMReX's interpolaters are not in the reference
# tree, so the semantics are defined here and pinned by what fixes them: exact
# reproduction of quadratics, the parent value for type 0, the time windows.
# ---------------------------------------------------------------------------

def _cf_oracle(fine_boxes, interp=2, ng=2, N=16, levels=1):
    inp = real_scalar(N=N, order=1, nghost=ng, lam=0.0, L=float(N),
                      **{"amr.max_refinement_levels": levels,
                         "amr.interpolation_type": interp})
    orc = Oracle("RealScalar", inp)
    orc.set_boxes(1, fine_boxes)
    return orc


def _box_cells(boxes, f):
    """f(x, y, z) at the cell centres of each box, box after box, x fastest;
    cell (i, j, k) of a level with spacing h has centre ((i + 1/2) h, ...)."""
    out = []
    for b, h in boxes:
        k, j, i = np.meshgrid(*(np.arange(b[a], b[a + 3] + 1) for a in (2, 1, 0)),
                              indexing="ij")
        out.append(f((i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h).ravel())
    return np.concatenate(out)


def _quad(a):
    return lambda x, y, z: (a[0] + a[1] * x + a[2] * y + a[3] * z + a[4] * x * x +
                            a[5] * y * y + a[6] * z * z + a[7] * x * y + a[8] * x * z +
                            a[9] * y * z)


def _ghosted(orc, boxes, g, comp=0):
    """Ghosted arrays of the level-1 boxes with their global index grids."""
    res = []
    for n, b in enumerate(boxes):
        shape = tuple(b[a + 3] - b[a] + 1 + 2 * g for a in (2, 1, 0))
        k, j, i = np.meshgrid(*(np.arange(b[a] - g, b[a + 3] + g + 1) for a in (2, 1, 0)),
                              indexing="ij")
        res.append((orc.fab(1, n, comp, shape), i, j, k))
    return res


def _inside(b, i, j, k):
    return ((i >= b[0]) & (i <= b[3]) & (j >= b[1]) & (j <= b[4]) &
            (k >= b[2]) & (k <= b[5]))


@pytest.mark.parametrize("ng", [2, 3])
def test_coarse_fine_fill_quadratic_exact(ng):
    """Two touching level-1 boxes over quadratic coarse states at t = 0 and
    t = 1, filled at t = 1/4: ghost cells another box covers hold that box's
    (cubic) data, every other ghost cell 3/4 q_old + 1/4 q_new at its centre."""
    boxes = [[8, 8, 8, 15, 15, 15], [16, 8, 8, 23, 15, 15]]
    orc = _cf_oracle(boxes, ng=ng)
    q_old = _quad([0.3, 0.2, -0.1, 0.05, 0.01, -0.02, 0.015, 0.007, -0.004, 0.003])
    q_new = _quad([-0.1, 0.05, 0.1, -0.2, -0.012, 0.006, 0.02, -0.005, 0.009, -0.008])
    cubic = lambda x, y, z: x ** 3 - y * z * z + 0.5
    coarse = [([0, 0, 0, 15, 15, 15], 1.0)]
    orc.upload_old(0, 0, _box_cells(coarse, q_old), 0.0)
    orc.upload(0, 0, _box_cells(coarse, q_new))
    orc.set_time(0, 1.0)
    orc.upload(1, 0, _box_cells([(b, 0.5) for b in boxes], cubic))
    orc.fill_patch(1, 0.25)
    for n, (a, i, j, k) in enumerate(_ghosted(orc, boxes, ng)):
        x, y, z = (i + 0.5) * 0.5, (j + 0.5) * 0.5, (k + 0.5) * 0.5
        other = _inside(boxes[1 - n], i, j, k)
        want = np.where(other, cubic(x, y, z),
                        0.75 * q_old(x, y, z) + 0.25 * q_new(x, y, z))
        ghost = ~_inside(boxes[n], i, j, k)
        assert other.any() and (ghost & ~other).any()
        assert np.max(np.abs(a - want)[ghost]) < 1e-12, n
        # Valid cells are untouched.
        assert np.array_equal(a[~ghost], cubic(x, y, z)[~ghost])


def test_coarse_fine_fill_piecewise_constant_is_not_quadratic():
    """Negative control of the quadratic check: type 0 misses the slope."""
    boxes = [[8, 8, 8, 15, 15, 15]]
    orc = _cf_oracle(boxes, interp=0)
    q = _quad([0.3, 0.2, -0.1, 0.05, 0.01, -0.02, 0.015, 0.007, -0.004, 0.003])
    orc.upload(0, 0, _box_cells([([0, 0, 0, 15, 15, 15], 1.0)], q))
    orc.fill_patch(1, orc.time(0))
    a, i, j, k = _ghosted(orc, boxes, 2)[0]
    ghost = ~_inside(boxes[0], i, j, k)
    err = np.abs(a - q((i + 0.5) * 0.5, (j + 0.5) * 0.5, (k + 0.5) * 0.5))[ghost]
    assert err.max() > 1e-2


def test_coarse_fine_fill_piecewise_constant_wraps():
    """Type 0 gives each ghost cell its parent's value, through the periodic
    wrap of the level-0 domain (the box touches z = 0)."""
    boxes = [[8, 8, 0, 15, 15, 7]]
    orc = _cf_oracle(boxes, interp=0, ng=3)
    c = np.random.default_rng(5).standard_normal(16 ** 3)
    orc.upload(0, 0, c)
    orc.fill_patch(1, orc.time(0))
    a, i, j, k = _ghosted(orc, boxes, 3)[0]
    ghost = ~_inside(boxes[0], i, j, k)
    par = c.reshape(16, 16, 16)[(k >> 1) % 16, (j >> 1) % 16, (i >> 1) % 16]
    assert np.array_equal(a[ghost], par[ghost])
    assert (k[ghost] < 0).any()


@pytest.mark.parametrize("time,want", [(0.0005, 1.0), (0.9995, 3.0), (0.5, 2.0),
                                       (0.25, 1.5)])
def test_coarse_fine_fill_time_windows(time, want):
    """GetLevelData: within 1e-3 |t_new - t_old| of a state's time that state
    alone, else linear in time between them."""
    boxes = [[8, 8, 8, 15, 15, 15]]
    orc = _cf_oracle(boxes)
    orc.upload_old(0, 0, np.full(16 ** 3, 1.0), 0.0)
    orc.upload(0, 0, np.full(16 ** 3, 3.0))
    orc.set_time(0, 1.0)
    orc.fill_patch(1, time)
    a, i, j, k = _ghosted(orc, boxes, 2)[0]
    assert np.all(a[~_inside(boxes[0], i, j, k)] == want)


def _children_in_ghosts(box, g):
    """Coarse cells whose 8 children are all ghost cells of the fine box."""
    lo = [(box[a] - g) // 2 for a in range(3)]
    hi = [(box[a + 3] + g + 1) // 2 - 1 for a in range(3)]
    out = []
    for c in itertools.product(*(range(lo[a], hi[a] + 1) for a in range(3))):
        kids = [(2 * c[0] + a, 2 * c[1] + b, 2 * c[2] + d)
                for a in (0, 1) for b in (0, 1) for d in (0, 1)]
        if all(not (box[0] <= x <= box[3] and box[1] <= y <= box[4] and box[2] <= z <= box[5])
               for x, y, z in kids):
            out.append(c)
    return out


def _fab_at(a, box, g, i, j, k):
    return a[k - box[2] + g, j - box[1] + g, i - box[0] + g]


@pytest.mark.parametrize("interp", [1, 4])
def test_coarse_fine_fill_conservative(interp):
    """Types 1 and 4 are conservative: the 8 children of a coarse cell
    average to it (random coarse data)."""
    box = [8, 8, 8, 15, 15, 15]
    orc = _cf_oracle([box], interp=interp)
    c = np.random.default_rng(9).standard_normal(16 ** 3)
    orc.upload(0, 0, c)
    orc.fill_patch(1, orc.time(0))
    a = _ghosted(orc, [box], 2)[0][0]
    cc = c.reshape(16, 16, 16)
    cells = _children_in_ghosts(box, 2)
    assert len(cells) > 100
    for ci, cj, ck in cells:
        kids = [_fab_at(a, box, 2, 2 * ci + p, 2 * cj + q, 2 * ck + r)
                for p in (0, 1) for q in (0, 1) for r in (0, 1)]
        assert abs(np.mean(kids) - cc[ck, cj, ci]) < 1e-14


def test_coarse_fine_fill_limited_linear():
    """Type 1 is exact on linear data and keeps a one-cell spike without
    overshoot: slopes vanish at the extremum and next to it."""
    box = [8, 8, 8, 15, 15, 15]
    orc = _cf_oracle([box], interp=1)
    lin = lambda x, y, z: 0.4 + 0.1 * x - 0.05 * y + 0.02 * z
    orc.upload(0, 0, _box_cells([([0, 0, 0, 15, 15, 15], 1.0)], lin))
    orc.fill_patch(1, orc.time(0))
    a, i, j, k = _ghosted(orc, [box], 2)[0]
    ghost = ~_inside(box, i, j, k)
    want = lin((i + 0.5) * 0.5, (j + 0.5) * 0.5, (k + 0.5) * 0.5)
    assert np.max(np.abs(a - want)[ghost]) < 1e-14
    spike = np.zeros((16, 16, 16))
    spike[5, 5, 3] = 1.0   # coarse (3, 5, 5): its children are ghost cells
    orc.upload(0, 0, spike.ravel())
    orc.fill_patch(1, orc.time(0))
    a = _ghosted(orc, [box], 2)[0][0]
    assert a[~_inside(box, i, j, k)].max() == 1.0
    for p, q, r in itertools.product((0, 1), repeat=3):
        assert _fab_at(a, box, 2, 6 + p, 10 + q, 10 + r) == 1.0
    assert _fab_at(a, box, 2, 6, 12, 12) == 0.0


def test_coarse_fine_fill_quartic_exact():
    """Type 4 maps the coarse cell averages of a tensor-product quartic to
    the fine cell averages of the same quartic."""
    box = [8, 8, 8, 15, 15, 15]
    orc = _cf_oracle([box], interp=4)
    coef = [(0.3, -0.2, 0.05, 0.01, -0.002), (1.0, 0.1, -0.03, 0.004, 0.0005),
            (0.5, 0.2, 0.02, -0.003, 0.0002)]

    def cell_avg(cf, lo, h):
        # average of sum cf[m] x^m over [lo, lo + h]
        return sum(c * ((lo + h) ** (m + 1) - lo ** (m + 1)) / ((m + 1) * h)
                   for m, c in enumerate(cf))

    n = np.arange(16)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    orc.upload(0, 0, (cell_avg(coef[0], i * 1.0, 1.0) * cell_avg(coef[1], j * 1.0, 1.0) *
                      cell_avg(coef[2], k * 1.0, 1.0)).ravel())
    orc.fill_patch(1, orc.time(0))
    a, i, j, k = _ghosted(orc, [box], 2)[0]
    ghost = ~_inside(box, i, j, k)
    want = (cell_avg(coef[0], i * 0.5, 0.5) * cell_avg(coef[1], j * 0.5, 0.5) *
            cell_avg(coef[2], k * 0.5, 0.5))
    assert np.max(np.abs(a - want)[ghost]) < 1e-12 * np.max(np.abs(want))
    # The quadratic type is not exact here (negative control).
    orc2 = _cf_oracle([box], interp=2)
    orc2.upload(0, 0, orc.download(0, 0))
    orc2.fill_patch(1, orc2.time(0))
    a2 = _ghosted(orc2, [box], 2)[0][0]
    assert np.max(np.abs(a2 - want)[ghost]) > 1e-6


def test_fill_coarse_patch():
    """FillCoarsePatch (level_synchronizer.cpp:56-88) interpolates every cell
    of the level, valid cells included (exact for a quadratic), and between
    the coarse level's two times takes the old state alone (the first of
    GetLevelData), where FillPatch would interpolate in time."""
    box = [8, 8, 8, 15, 15, 15]
    orc = _cf_oracle([box])
    q = _quad([0.3, 0.2, -0.1, 0.05, 0.01, -0.02, 0.015, 0.007, -0.004, 0.003])
    orc.upload(0, 0, _box_cells([([0, 0, 0, 15, 15, 15], 1.0)], q))
    orc.fill_coarse_patch(1)
    a, i, j, k = _ghosted(orc, [box], 2)[0]
    assert np.max(np.abs(a - q((i + 0.5) * 0.5, (j + 0.5) * 0.5, (k + 0.5) * 0.5))) < 1e-12
    orc.upload_old(0, 0, np.full(16 ** 3, 1.0), 0.0)
    orc.upload(0, 0, np.full(16 ** 3, 3.0))
    orc.set_time(0, 1.0)
    for t1, want in ((0.5, 1.0), (1.0, 3.0), (0.0, 1.0)):
        orc.set_time(1, t1)
        orc.fill_coarse_patch(1)
        assert np.all(_ghosted(orc, [box], 2)[0][0] == want), t1
    orc.set_time(1, 0.5)
    orc.fill_patch(1, 0.5)   # FillPatch interpolates in time
    a = _ghosted(orc, [box], 2)[0][0]
    assert np.all(a[~_inside(box, i, j, k)] == 2.0)


def test_coarse_fine_fill_requires_nesting():
    """A level-2 box at the corner of its level-1 box needs level-1 cells
    outside every level-1 box for its ghost cells."""
    inp = real_scalar(N=16, order=1, nghost=2, lam=0.0, L=16.0,
                      **{"amr.max_refinement_levels": 2})
    orc = Oracle("RealScalar", inp)
    orc.set_boxes(1, [[8, 8, 8, 15, 15, 15]])
    orc.set_boxes(2, [[16, 16, 16, 23, 23, 23]])
    with pytest.raises(RuntimeError, match="not properly nested"):
        orc.fill_patch(2, orc.time(2))


# ---------------------------------------------------------------------------
# Integrators (lsssprk3.cpp:14-31, rkn.cpp:25-205, leapfrog.cpp:13-57)
# ---------------------------------------------------------------------------

def _mode_case(integrator, cfl, N=16, m=2, order=1):
    """RealScalar with lambda = 0 and one Fourier mode along x: each cell
    obeys phi'' = -w^2 phi with w^2 the discrete Laplacian eigenvalue."""
    dx = 1.0
    orc = Oracle("RealScalar", real_scalar(N=N, order=order, lam=0.0, L=N * dx,
                                           cfl=cfl, integrator=integrator))
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    shape = np.cos(2 * np.pi * m * i / N + 0.2).ravel()
    w2 = _lap_symbol(order, 2 * np.pi * m / N) / dx ** 2
    return orc, shape, w2, cfl * dx


def _one_step(integrator, phi0, pi0, cfl):
    orc, shape, w2, h = _mode_case(integrator, cfl)
    orc.upload(0, 0, phi0 * shape)
    orc.upload(0, 1, pi0 * shape)
    orc.advance(0, 1)
    i0 = int(np.argmax(np.abs(shape)))
    return (orc.download(0, 0)[i0] / shape[i0], orc.download(0, 1)[i0] / shape[i0],
            w2, h)


def test_ssprk3_amplification_matrix():
    """Any three-stage third-order RK on y' = A y gives
    I + hA + (hA)^2/2 + (hA)^3/6."""
    for phi0, pi0 in ((1.0, 0.0), (0.0, 1.0)):
        p, q, w2, h = _one_step(10, phi0, pi0, cfl=0.3)
        A = np.array([[0, 1], [-w2, 0]]) * h
        M = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6
        want = M @ np.array([phi0, pi0])
        assert abs(p - want[0]) < 1e-14 and abs(q - want[1]) < 1e-14


def test_leapfrog_amplification_matrix():
    """Kick-drift-kick on phi'' = -w^2 phi."""
    for phi0, pi0 in ((1.0, 0.0), (0.0, 1.0)):
        p, q, w2, h = _one_step(11, phi0, pi0, cfl=0.3)
        ph = pi0 - h / 2 * w2 * phi0
        p1 = phi0 + h * ph
        q1 = ph - h / 2 * w2 * p1
        assert abs(p - p1) < 1e-14 and abs(q - q1) < 1e-14


@pytest.mark.parametrize("integrator,order", [(10, 3), (11, 2), (21, 4), (22, 5)])
def test_convergence_order(integrator, order):
    """Global error at a fixed time against the exact cos/sin solution
    falls as h^order."""
    errs = []
    for cfl in (0.2, 0.1):
        orc, shape, w2, h = _mode_case(integrator, cfl)
        w = np.sqrt(w2)
        orc.upload(0, 0, shape)
        orc.upload(0, 1, np.zeros_like(shape))
        steps = int(round(4.0 / h))
        orc.advance(0, steps)
        t = steps * h
        i0 = int(np.argmax(np.abs(shape)))
        e = abs(orc.download(0, 0)[i0] / shape[i0] - np.cos(w * t))
        e = max(e, abs(orc.download(0, 1)[i0] / shape[i0] + w * np.sin(w * t)))
        errs.append(e)
    rate = np.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.35, (errs, rate)


def test_ssprk3_matches_shu_osher_form():
    """Low-storage SSPRK3 equals u0 + dt/6 (k1 + k2 + 4 k3) with
    k1 = f(u0), k2 = f(u0 + dt k1), k3 = f(u0 + dt/4 (k1 + k2)), on a
    nonlinear problem (lambda = 1)."""
    N = 8
    inp = real_scalar(N=N, order=2, lam=1.0, L=N * 0.5, cfl=0.4, integrator=10)
    rng = np.random.default_rng(5)
    u0 = [rng.uniform(-1, 1, N ** 3), rng.uniform(-1, 1, N ** 3)]
    probe = Oracle("RealScalar", inp)

    def f(u):
        for c in range(2):
            probe.upload(0, c, u[c])
        probe.fill_rhs(0, 0.0)
        return [probe.rhs(0, c) for c in range(2)]

    dt = 0.4 * 0.5
    k1 = f(u0)
    k2 = f([u0[c] + dt * k1[c] for c in range(2)])
    k3 = f([u0[c] + dt / 4 * (k1[c] + k2[c]) for c in range(2)])
    want = [u0[c] + dt / 6 * (k1[c] + k2[c] + 4 * k3[c]) for c in range(2)]
    orc = Oracle("RealScalar", inp)
    for c in range(2):
        orc.upload(0, c, u0[c])
    orc.advance(0, 1)
    for c in range(2):
        assert np.max(np.abs(orc.download(0, c) - want[c])) < 1e-13


# ---------------------------------------------------------------------------
# Truncation errors and tagging (kernels.h:18-42, macros.h:489-566)
# ---------------------------------------------------------------------------

def test_te_tag_hand_built_blocks():
    """te = |coarse - mean of the 8 children|; a block passes when
    te >= te_crit for any component.  The reference's CPU loop SETs the 2x2x2
    block at the even cell and re-CLEARs the 7 odd cells when it visits them
    later (macros.h:504-527), so the mask holds the even corner only, while
    the counters add 8 per passing component and 8 once to the total."""
    N, crit = 8, 0.5
    inp = real_scalar(N=N, order=1, lam=0.0, L=float(N), **{"amr.te_crit": crit})
    orc = Oracle("RealScalar", inp)
    nc3 = (N // 2) ** 3
    crse_phi = np.zeros(nc3)
    crse_pi = np.zeros(nc3)
    eps = 2.0 ** -50
    crse_phi[0] = crit                      # at the threshold: tagged
    crse_phi[1] = crit * (1 - eps)          # just below: not tagged
    crse_phi[2] = -crit * (1 + eps)         # just above (negative): tagged
    crse_pi[2] = 2 * crit                   # second component, same block
    crse_pi[5] = crit                       # second component only
    crse_phi[7] = 0.75                      # children average 0.25 -> te 0.5
    # coarse cell 7 = (ci, cj, ck) = (3, 1, 0): fine children x 6..7,
    # y 2..3, z 0..1 (arrays are [z, y, x]).
    fine = np.zeros((N, N, N))
    fine[0:2, 2:4, 6:8] = 0.25
    orc.upload(0, 0, fine.ravel())
    orc.upload(0, 1, np.zeros(N ** 3))
    orc.upload(-1, 0, crse_phi)
    orc.upload(-1, 1, crse_pi)
    mask, counts, te = orc.te_tag(0, flags=1 | 4, want_te=True)

    assert te[0][7] == 0.5 and te[0][0] == crit and te[1][5] == crit
    tagged = {0, 2, 5, 7}
    want = np.zeros((N, N, N), dtype=np.uint8)
    for c in tagged:
        ci, cj, ck = c % 4, (c // 4) % 4, c // 16
        want[2 * ck, 2 * cj, 2 * ci] = 1
    assert np.array_equal(mask.reshape(N, N, N), want)
    assert counts[0] == 8 * len(tagged)            # total
    assert counts[1] == 0                          # user tags
    assert counts[2] == 8 * 3                      # Phi: blocks 0, 2, 7
    assert counts[3] == 8 * 2                      # dPhi: blocks 2, 5


def _complex_inputs(N, **extra):
    return complex_scalar(N=N, integrator=10, L=N * 0.25, cfl=0.2, **extra)


def test_user_string_tags_match_phase_winding():
    """ComplexScalar's user criterion (a string pierces one of the cell's
    three forward plaquettes) on a random field, where about half of the
    plaquettes wind: every cell, odd and even, is tagged as the winding
    computed from phase increments says; total and user counters both count
    the tagged cells (TagWithTruncationCpu without truncation errors)."""
    N = 8
    orc = Oracle("ComplexScalar", _complex_inputs(N, **{"project.tag_strings": 1}))
    rng = np.random.default_rng(3)
    re, im = rng.standard_normal((2, N, N, N))
    for c, f in enumerate([re, im, np.zeros_like(re), np.zeros_like(re)]):
        orc.upload(0, c, f.ravel())
    mask, counts, _ = orc.te_tag(0, flags=4)
    want = winding_tags(re + 1j * im)
    assert 0.2 < want.mean() < 0.9
    parity = (np.indices((N, N, N)).sum(axis=0) % 2).astype(bool)
    assert want[parity].any() and want[~parity].any()
    assert np.array_equal(mask.reshape(N, N, N).astype(bool), want)
    assert counts[0] == counts[1] == want.sum()
    # Switched off by the project input: no user tags.
    orc0 = Oracle("ComplexScalar", _complex_inputs(N))
    for c, f in enumerate([re, im, np.zeros_like(re), np.zeros_like(re)]):
        orc0.upload(0, c, f.ravel())
    m0, c0, _ = orc0.te_tag(0, flags=4)
    assert m0.sum() == 0 and c0[0] == c0[1] == 0


@pytest.mark.parametrize("modifier", [0, 1])
@pytest.mark.parametrize("user_on", [0, 1])
def test_te_tag_with_user_tags_and_modifiers_hand_built(modifier, user_on):
    """TagWithTruncationCpu (macros.h:489-566) with a user criterion and
    truncation-error modifiers: the mask is user(cell) or (even cell and
    some component passes), ntags_trunc adds 8 per passing component,
    ntags_user counts user tags and ntags_total adds them plus 8 - user(even
    cell) per passing block (macros.h:509-527).  Expected values follow from
    the chosen deltas by those rules, not from the oracle's loop."""
    N = 8
    fine, crse, delta, norms = user_tag_case(N)
    inp = _complex_inputs(N, **{"amr.te_crit": 1e-3, "project.tag_strings": user_on,
                                "project.te_modifier": modifier})
    orc = Oracle("ComplexScalar", inp)
    for c in range(4):
        orc.upload(0, c, fine[c].ravel())
        orc.upload(-1, c, crse[c].ravel())
    mask, counts, te = orc.te_tag(0, flags=1 | 4, want_te=True)
    assert np.allclose(te, np.abs(delta).reshape(4, -1), rtol=0, atol=1e-13)
    want_mask, want_counts = expected_tags(fine, delta, norms, modifier, user_on)
    assert np.array_equal(mask.reshape(N, N, N).astype(bool), want_mask)
    assert np.array_equal(counts, want_counts), (counts, want_counts)
    # Each decision class occurs, and user tags meet passing blocks.
    assert counts[2] > 0 and counts[4] > 0
    if user_on:
        assert 0 < counts[1] < N ** 3
        ev = winding_tags(fine[0] + 1j * fine[1])[0::2, 0::2, 0::2]
        assert ev.any() and (~ev).any()


def test_modifier_norm_is_taken_at_each_tagging():
    """SetParamsTruncationModifier reads max|Psi_n| of level 0 at every
    tagging (as FirstOrderPhaseTransition.cpp:133-157 does): with the same
    truncation errors, fields scaled by 4 fail the relative criterion that
    they passed before (a norm kept from the first tagging would pass them
    again)."""
    N = 8
    inp = _complex_inputs(N, **{"amr.te_crit": 1e-3, "project.te_modifier": 1})
    orc = Oracle("ComplexScalar", inp)
    fine, crse, delta, norms = user_tag_case(N)
    h = N // 2
    got = []
    for scale in (1.0, 4.0):
        for c in range(4):
            f = fine[c] * (scale if c < 2 else 1.0)
            mean = f.reshape(h, 2, h, 2, h, 2).mean(axis=(1, 3, 5))
            orc.upload(0, c, f.ravel())
            orc.upload(-1, c, (mean + delta[c]).ravel())
        _, counts, te = orc.te_tag(0, flags=1 | 4, want_te=True)
        assert np.allclose(te, np.abs(delta).reshape(4, -1), rtol=0, atol=1e-13)
        sn = [norms[0] * scale, norms[1] * scale, norms[2], norms[3]]
        _, want = expected_tags([fine[0] * scale, fine[1] * scale, fine[2], fine[3]],
                                 delta, sn, 1, 0)
        assert np.array_equal(counts, want), (scale, counts, want)
        got.append(counts)
    assert got[0][2] > 0 and got[1][2] == 0 and got[1][3] == 0
    assert np.array_equal(got[0][4:], got[1][4:])


# ---------------------------------------------------------------------------
# GW spectrum (gravitational_waves.cpp:52-322, gravitational_waves.h:101-133)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("N", [6, 8, 12])
def test_r2c_matches_numpy(N):
    rng = np.random.default_rng(N)
    a = rng.standard_normal((N, N, N))
    got = oracle_r2c(a)
    want = np.fft.rfftn(a)
    assert np.max(np.abs(got - want)) < 1e-11 * np.max(np.abs(want))


@pytest.mark.parametrize("ptype", [1, 2, 3])
def test_lambda_projector_identities(ptype):
    N = 16
    rng = np.random.default_rng(ptype)
    for _ in range(12):
        abc = tuple(int(v) for v in rng.integers(0, N, 3))
        if abc == (0, 0, 0):
            continue
        Lm = np.zeros((3, 3, 3, 3))
        for i, j, l, m in itertools.product(range(3), repeat=4):
            Lm[i, j, l, m] = oracle_lambda(i, j, l, m, abc, N, ptype)
        d = np.array([_grad_symbol(ptype, 2 * np.pi * (a - N if a > N // 2 - 1 else a) / N)
                      for a in abc])
        if np.linalg.norm(d) < 1e-12:
            continue
        tol = 1e-13
        assert np.max(np.abs(np.einsum("i,ijlm->jlm", d, Lm))) < tol   # transverse
        assert np.max(np.abs(np.einsum("iilm->lm", Lm))) < tol         # traceless
        assert np.max(np.abs(Lm - Lm.transpose(2, 3, 0, 1))) < tol     # symmetric
        P = np.einsum("ijlm,lmpq->ijpq", Lm, Lm)
        assert np.max(np.abs(P - Lm)) < 1e-12                          # projector
    assert oracle_lambda(0, 1, 0, 1, (0, 0, 0), N, ptype) == 0.0


def _plane_wave_spectrum(comp, axis, m, N=16, mode=1, nbins=8):
    orc = Oracle("PhaseTransitionGW", phase_transition(N=N, L=float(N)))
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    x = (i, j, k)[axis]
    for c in range(orc.ncomp):
        orc.upload(0, c, np.zeros(N ** 3))
    orc.upload(0, comp, np.cos(2 * np.pi * m * x / N).ravel())
    return orc.gw_spectrum(mode=mode, nbins=nbins, zero_padding=1, projection_type=3)


@pytest.mark.parametrize("comp,axis,want", [
    (11, 2, 1.0),    # du_xy, k along z: fully transverse-traceless
    (8, 2, 0.25),    # du_xx, k along z: TT part diag(1/2, -1/2, 0)
    (10, 2, 0.0),    # du_zz, k along z: longitudinal
    (13, 0, 1.0),    # du_yz, k along x (halved axis, multiplicity 2)
    (12, 1, 1.0),    # du_xz, k along y: transverse-traceless
    (11, 1, 0.0),    # du_xy, k along y: mixed longitudinal
])
def test_plane_wave_spectrum(comp, axis, want):
    m = 3
    spec, kv = _plane_wave_spectrum(comp, axis, m)
    assert abs(spec[m] - want) < 1e-12
    others = np.delete(spec, m)
    assert np.max(np.abs(others)) < 1e-12
    assert kv[m] == m * m


def test_spectrum_modes_conserve_total():
    spec1, _ = _plane_wave_spectrum(12, 1, 5, mode=1)
    spec0, k0 = _plane_wave_spectrum(12, 1, 5, mode=0)
    spec2, _ = _plane_wave_spectrum(12, 1, 5, mode=2, nbins=16)
    assert abs(spec0.sum() - 1.0) < 1e-12 and abs(spec2.sum() - 1.0) < 1e-12
    assert abs(spec0[int(np.searchsorted(k0, 25))] - 1.0) < 1e-12
    assert abs(spec1.sum() - 1.0) < 1e-12


# GW spectrum modifiers (GravitationalWavesSpectrumModifier,
# gravitational_waves.h:79-92; the project's SpectrumUTimesK and
# SpectrumPlusTranslatedCopy) and zero padding (utils/fft.h:143-261).

def _plane_wave_orc(comp, m, N=16, rand_seed=None):
    orc = Oracle("PhaseTransitionGW", phase_transition(N=N, L=float(N)))
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    rng = np.random.default_rng(rand_seed) if rand_seed is not None else None
    for c in range(orc.ncomp):
        f = np.zeros(N ** 3) if rng is None else rng.standard_normal(N ** 3)
        orc.upload(0, c, f)
    orc.upload(0, comp, np.cos(2 * np.pi * m * k / N).ravel())
    return orc


def test_gw_modifier_u_times_k_plane_wave():
    """SpectrumUTimesK selects u_ij (the du_ij, random here, do not enter)
    and scales every mode by |n| dk, dk = 2 pi / L: a transverse-traceless
    plane wave u_xy = cos(2 pi m z / N), whose du_ij spectrum would be 1 in
    bin m^2, gives (m dk)^2 there and nothing elsewhere."""
    N, m = 16, 3
    orc = _plane_wave_orc(5, m, N)          # u_xy
    for c in range(8, 14):                  # du_ij random
        orc.upload(0, c, np.random.default_rng(c).standard_normal(N ** 3))
    spec, kv = orc.gw_spectrum(mode=1, zero_padding=1, projection_type=3, modifier=0)
    dk = 2 * np.pi / N
    assert abs(spec[m] - (m * dk) ** 2) < 1e-12
    assert np.max(np.abs(np.delete(spec, m))) < 1e-12
    # Default: the du_ij (random) instead.
    spec_d, _ = orc.gw_spectrum(mode=1, zero_padding=1, projection_type=3)
    assert np.max(np.abs(np.delete(spec_d, m))) > 1e-3


@pytest.mark.parametrize("delta", [0.0, 1.0, 2.5])
def test_gw_modifier_translated_copy_plane_wave(delta):
    """SpectrumPlusTranslatedCopy multiplies every mode by 1 + exp(i k.d):
    for the TT plane wave du_xy = cos(2 pi m z / N) and d = (0, 0, delta)
    the spectrum in bin m^2 is |1 + exp(i m dk delta)|^2 = 2 + 2 cos(m dk
    delta) (both signs of k_z give the same)."""
    N, m = 16, 3
    orc = _plane_wave_orc(11, m, N)         # du_xy
    spec, _ = orc.gw_spectrum(mode=1, zero_padding=1, projection_type=3, modifier=1,
                              params=[0.0, 0.0, delta])
    dk = 2 * np.pi / N
    assert abs(spec[m] - (2 + 2 * np.cos(m * dk * delta))) < 1e-12
    assert np.max(np.abs(np.delete(spec, m))) < 1e-12


def _brute_force_gw_spectrum(fields, comps, N0, zp, L, ptype, utk):
    """Independent restatement for the pin below: zero-padded explicit DFT
    (no FFT), Lambda contracted in its trace form Tr(PAPA) - Tr(PA)^2 / 2
    with P = 1 - d d^T / |d|^2, d = the gradient symbol of the padded grid,
    multiplicity 2 off the x = 0 and x = N/2 planes, unique |n|^2 bins, over
    N^6; with utk every mode times |n| 2 pi / (L zp)."""
    N = N0 * zp
    W = np.exp(-2j * np.pi * np.outer(np.arange(N), np.arange(N)) / N)
    A = []
    for c in comps:
        pad = np.zeros((N, N, N))
        pad[:N0, :N0, :N0] = fields[c].reshape(N0, N0, N0)
        A.append(np.einsum("az,by,cx,zyx->abc", W, W, W, pad)[:, :, :N // 2 + 1])
    f = np.array([a - N if a >= N // 2 else a for a in range(N)])
    sym = np.array([_grad_symbol(ptype, 2 * np.pi * (a - N if a > N // 2 - 1 else a) / N)
                    for a in range(N)])
    k2 = sorted({int(x * x + y * y + z * z) for x in f for y in f for z in f})
    spec = np.zeros(len(k2))
    dk = 2 * np.pi / (L * zp)
    idx = [[0, 1, 2], [1, 3, 4], [2, 4, 5]]
    for cz in range(N):
        for by in range(N):
            for ax in range(N // 2 + 1):
                if ax == 0 and by == 0 and cz == 0:
                    continue
                d = np.array([sym[ax], sym[by], sym[cz]])
                P = np.eye(3) - np.outer(d, d) / d.dot(d)
                M = np.array([[A[idx[i][j]][cz, by, ax] for j in range(3)] for i in range(3)])
                if utk:
                    M = M * np.sqrt(f[ax] ** 2 + f[by] ** 2 + f[cz] ** 2) * dk
                val = 0.0
                for part in (M.real, M.imag):
                    B = P @ part
                    val += np.trace(B @ B) - 0.5 * np.trace(B) ** 2
                mult = 1.0 if ax in (0, N // 2) else 2.0
                spec[k2.index(int(f[ax] ** 2 + f[by] ** 2 + f[cz] ** 2))] += mult * val
    return spec / N ** 6, np.array(k2, dtype=float)


@pytest.mark.parametrize("utk", [False, True])
def test_gw_zero_padding_matches_brute_force_dft(utk):
    """zero_padding = 2 places the level at the corner of a 2N grid (the
    other blocks zero, utils/fft.h:152-190), transforms the padded grid
    with k tables of 2N and L -> 2L, and normalises by (2N)^6."""
    N0, zp = 4, 2
    L = N0 * 0.5
    orc = Oracle("PhaseTransitionGW", phase_transition(N=N0, L=L))
    rng = np.random.default_rng(11)
    fields = [rng.standard_normal(N0 ** 3) for _ in range(orc.ncomp)]
    for c, f in enumerate(fields):
        orc.upload(0, c, f)
    sel = [2, 5, 6, 3, 7, 4] if utk else [8, 11, 12, 9, 13, 10]
    spec, kv = orc.gw_spectrum(mode=0, zero_padding=zp, projection_type=2,
                               modifier=0 if utk else -1)
    want, wk = _brute_force_gw_spectrum(fields, sel, N0, zp, L, 2, utk)
    assert np.array_equal(kv, wk)
    assert np.max(np.abs(spec - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))


def _golden():
    with open(os.path.join(GOLDEN, "spectra_ks.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("N", [32, 64, 128])
def test_index_to_k_matches_reference_tables(N):
    g = _golden()[str(N)]
    lib = spectral_lib()
    for ptype, key in ((1, "k1"), (2, "k2")):
        got = np.array([lib.orc_index_to_k(a, N, ptype) for a in range(N)])
        assert np.max(np.abs(got - np.array(g[key]))) < 1e-15, key


@pytest.mark.parametrize("N", [32, 64, 128, 256])
def test_unique_k2_matches_reference_tables(N):
    g = _golden()[str(N)]
    got = oracle_unique_k2(N)
    assert len(got) == g["nks"] == g["nbins"]
    assert int(got.sum()) == g["bins_sum"] and int(got[-1]) == g["bins_last"]
    if "bins" in g:
        assert np.array_equal(got, np.array(g["bins"], dtype=np.int64))


# ---------------------------------------------------------------------------
# Independent boxes (ensemble of periodic realizations)
# ---------------------------------------------------------------------------

def test_ensemble_member_equals_single_periodic_box():
    """With shm.independent_boxes every box wraps onto itself: each member
    evolves bit-for-bit as the same box run alone as an ordinary periodic
    domain (whose ghost fill is the reference FillBoundary)."""
    from cases import ensemble_boxes, ensemble_fields, ensemble_inputs
    boxes = ensemble_boxes(4, 8, 24, seed=3)
    inp = ensemble_inputs(boxes)
    orc = Oracle("RealScalar", inp)
    orc.set_boxes(0, boxes)
    fields = ensemble_fields(boxes, seed=17)
    for c, f in enumerate(fields):
        orc.upload(0, c, f)
    orc.advance(0, 5)
    out = [orc.download(0, c) for c in range(2)]
    offs = np.cumsum([0] + [(b[3] + 1) ** 3 for b in boxes])
    for b, bx in enumerate(boxes):
        n = bx[3] + 1
        alone = real_scalar(N=n, order=2, integrator=21, L=n / 64.0, cfl=0.2, nghost=2)
        one = Oracle("RealScalar", alone)
        for c in range(2):
            one.upload(0, c, fields[c][offs[b]:offs[b + 1]])
        one.advance(0, 5)
        for c in range(2):
            assert np.array_equal(one.download(0, c), out[c][offs[b]:offs[b + 1]]), (b, c)


# ---------------------------------------------------------------------------
# Seeded object initial states (bubbles, string loops)
# ---------------------------------------------------------------------------

def _bubble_params(n, t_from, t_to):
    # count, phi_in, R0, wall width, nucleation times in [0, 1], window
    return [n, 1.0, 3.0, 2.0, 0.0, 1.0, t_from, t_to]


def test_bubble_injection_composes():
    """Vacuum plus the bubbles nucleated by t_mid, then the ones nucleated in
    (t_mid, 1], equals vacuum plus all bubbles nucleated by 1, bit for bit;
    the field stays within [0, phi_in] and reaches the bubble interior."""
    N = 24
    a = Oracle("PhaseTransitionGW", phase_transition(N=N, L=float(N)))
    b = Oracle("PhaseTransitionGW", phase_transition(N=N, L=float(N)))
    a.init_state(0, 3, 42, _bubble_params(10, -1.0, 1.0))
    b.init_state(0, 3, 42, _bubble_params(10, -1.0, 0.4))
    b.init_state(0, 4, 42, _bubble_params(10, 0.4, 1.0))
    pa, pb = a.download(0, 0), b.download(0, 0)
    assert np.array_equal(pa, pb)
    assert pa.min() >= 0.0 and pa.max() <= 1.0 and pa.max() > 0.9
    # Fewer bubbles nucleated by an earlier time: pointwise smaller field.
    c = Oracle("PhaseTransitionGW", phase_transition(N=N, L=float(N)))
    c.init_state(0, 3, 42, _bubble_params(10, -1.0, 0.4))
    assert np.all(c.download(0, 0) <= pa) and c.download(0, 0).sum() < pa.sum()


def _plaquette_winding(psi, axes):
    """Winding of arg(psi) around every unit plaquette spanned by two axes
    (periodic), in units of 2 pi."""
    a, b = axes
    p00 = psi
    p10 = np.roll(psi, -1, axis=a)
    p11 = np.roll(p10, -1, axis=b)
    p01 = np.roll(psi, -1, axis=b)
    d = lambda u, v: np.angle(v / u)
    w = d(p00, p10) + d(p10, p11) + d(p11, p01) + d(p01, p00)
    return np.rint(w / (2 * np.pi)).astype(int), w


def test_string_loops_carry_conserved_unit_winding():
    """psi = prod tanh(d/w) exp(i theta): every plaquette winds by 0 or
    +-1, strings are present, and the winding flux out of every cell cube
    vanishes (strings neither end nor branch)."""
    N = 32
    from cases import complex_scalar
    orc = Oracle("ComplexScalar", complex_scalar(N=N))
    orc.init_state(0, 5, 11, [3, 5.0, 9.0, 1.5])
    psi = (orc.download(0, 0) + 1j * orc.download(0, 1)).reshape(N, N, N)  # [z, y, x]
    assert np.max(np.abs(psi)) <= 1.0
    # Plaquette normals: x (axes y, z), y (axes z, x), z (axes x, y).
    wx, rx = _plaquette_winding(psi, (1, 0))
    wy, ry = _plaquette_winding(psi, (0, 2))
    wz, rz = _plaquette_winding(psi, (2, 1))
    for w, r in ((wx, rx), (wy, ry), (wz, rz)):
        assert np.max(np.abs(r / (2 * np.pi) - w)) < 1e-9
        assert set(np.unique(w)) <= {-1, 0, 1}
    assert np.count_nonzero(wx) + np.count_nonzero(wy) + np.count_nonzero(wz) > 20
    div = (np.roll(wx, -1, axis=2) - wx) + (np.roll(wy, -1, axis=1) - wy) + \
          (np.roll(wz, -1, axis=0) - wz)
    assert np.count_nonzero(div) == 0


def _three_level_oracle(ncoarse):
    from cases import amr_boxes, complex_scalar
    lev1, lev2 = amr_boxes(seed=2024)
    inp = complex_scalar(N=32, integrator=10, nghost=2, L=8.0, cfl=0.25,
                         t0=0.125, **{"amr.max_refinement_levels": 2,
                                      "amr.te_crit": 1e-3, "amr.regrid_dt": 0.01})
    orc = Oracle("ComplexScalar", inp)
    orc.set_boxes(1, lev1)
    orc.set_boxes(2, lev2)
    for lev in range(3):
        orc.init_state(lev, 5, 77, [4, 5.0, 10.0, 1.5])
    orc.evolve(ncoarse)
    return orc


# The order of level operations of two coarse steps of a 3-level hierarchy
# with a shadow level and regrids due as often as possible, worked out by
# hand from TimeStepper::Advance / ScheduleRegrid / SynchronizeLevels /
# DoRegridIfScheduled (time_stepper.cpp:90-171, 221-311) and the
# RegridScheduler (regrid_scheduler.cpp:9-74); times in units of the coarse
# dt after t0.
#   Step 1: no regrid is scheduled on level 0 (grid_old[0] holds no state
#   yet); level 1 schedules one at the end of its second step (odd istep),
#   so level 2's second step and level 1 compute truncation errors instead
#   of averaging down, level 1 regrids, level 0 then averages down.
#   Step 2: level 0 schedules for t0 + 2 (shadow level advanced at once);
#   level 1's own request merges into it (lowest level 0); every level
#   computes truncation errors at t0 + 2, and level 0 regrids.
HAND_LOG = [
    ("A", 0, 1.0), ("A", 1, 0.5), ("A", 2, 0.25), ("A", 2, 0.5), ("D", 1, 0.5),
    ("S", 1, 1.0), ("A", 1, 1.0), ("A", 2, 0.75), ("A", 2, 1.0), ("T", 2, 1.0),
    ("T", 1, 1.0), ("R", 1, 1.0), ("D", 0, 1.0),
    ("S", 0, 2.0), ("H", -1, 2.0), ("A", 0, 2.0), ("A", 1, 1.5), ("A", 2, 1.25),
    ("A", 2, 1.5), ("D", 1, 1.5), ("S", 1, 2.0), ("A", 1, 2.0), ("A", 2, 1.75),
    ("A", 2, 2.0), ("T", 2, 2.0), ("T", 1, 2.0), ("T", 0, 2.0), ("R", 0, 2.0),
]


def test_oracle_subcycled_event_order():
    """Sub-cycling on a static 3-level layout with the shadow level: the
    oracle's sequence of advances, average-downs, truncation errors,
    schedulings and regrids over two coarse steps is the hand-derived one
    (a wrong branch, e.g. averaging down where truncation errors are due, or
    level 0 scheduling in the first step, changes the log)."""
    orc = _three_level_oracle(2)
    dt0 = 0.25 * 8.0 / 32   # binary-exact times: the scheduler compares them exactly
    got = [(k, lev, (t - 0.125) / dt0) for k, lev, t in orc.events()]
    assert got == HAND_LOG
    for lev in range(3):
        assert orc.time(lev) == 0.125 + 2 * dt0
        assert orc.last_tags(lev)[0] > 0


# ---------------------------------------------------------------------------
# Spectra and projections of output functions (output_types/spectrum.cpp:12-113,
# output_types/projection.cpp:11-118); RealScalar declares ScaledPhi (params[0]
# times Phi, params = {project.output_scale}), dPhi2 and GradientEnergy.
# ---------------------------------------------------------------------------

def _out_oracle(N=16, L=8.0, t0=0.75, scale=1.5, levels=0):
    inp = real_scalar(N=N, order=1, nghost=2, lam=0.0, L=L,
                      **{"sim.t_start": t0, "project.output_scale": scale,
                         "amr.max_refinement_levels": levels})
    return Oracle("RealScalar", inp)


@pytest.mark.parametrize("m", [(3, 1, 0), (0, 2, 5), (8, 0, 0)])
def test_spectrum_plane_wave(m):
    """cos(2 pi m.n / N + p) has |F| = N^3 / 2 at +-m, one of which the
    x-halved layout holds (mult 2) unless m_x = 0 (both, mult 1): the bin of
    |m|^2 holds s^2 t / (2 dk).  At m_x = N/2 the two coincide: |F| = N^3 cos p,
    mult 1, s^2 t cos^2 p / dk."""
    N, L, t0, s, p = 16, 8.0, 0.75, 1.5, 0.3
    orc = _out_oracle(N, L, t0, s)
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    orc.upload(0, 0, np.cos(2 * np.pi * (m[0] * i + m[1] * j + m[2] * k) / N + p).ravel())
    orc.upload(0, 1, np.zeros(N ** 3))
    bins, ks = orc.spectrum(0)
    dk = 2 * np.pi / L
    want = s * s * t0 * (np.cos(p) ** 2 / dk if m[0] == N // 2 else 0.5 / dk)
    b = int(np.searchsorted(ks, sum(q * q for q in m)))
    assert ks[b] == sum(q * q for q in m)
    assert abs(bins[b] - want) < 1e-12 * want
    bins[b] = 0.
    assert np.max(np.abs(bins)) < 1e-12 * want


def test_spectrum_parseval():
    """Sum over bins = t / (dk N^3) sum_x f^2 (the halved layout with its
    multiplicities counts every mode once)."""
    N, L, t0, s = 16, 8.0, 0.75, 1.5
    orc = _out_oracle(N, L, t0, s)
    f = np.random.default_rng(3).standard_normal(N ** 3)
    orc.upload(0, 0, f)
    orc.upload(0, 1, f[::-1].copy())
    for fn, vals in ((0, s * f), (1, f[::-1] ** 2)):
        bins, _ = orc.spectrum(fn)
        want = t0 / (2 * np.pi / L) / N ** 3 * np.sum(vals ** 2)
        assert abs(bins.sum() - want) < 1e-11 * want, fn


def test_projection_orientation_single_level():
    """Single level, field = x: pixel [I, J] sums N cells of x_I (x index
    slow); field = y puts the variation on the second index."""
    N = 16
    orc = _out_oracle(N, L=16.0, scale=1.0)
    n = np.arange(N)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    for axis, g in ((0, i), (1, j)):
        orc.upload(0, 0, (g + 0.5).ravel())
        orc.upload(0, 1, np.zeros(N ** 3))
        d, c = orc.projection(0)
        want = N * (n + 0.5)
        assert np.allclose(d, want[:, None] if axis == 0 else want[None, :],
                           rtol=0, atol=1e-12)
        assert np.all(c == N)


def test_projection_two_levels_linear_in_z():
    """Field = z on both levels (cell centres): a coarse cell adds
    2 z_coarse = the sum of its two fine cells, so every pixel holds the
    fine column sum whatever the refinement; counts are 8 fine cells plus
    the 12 coarse cells the fine box leaves uncovered under it, else 16."""
    orc = _out_oracle(16, L=16.0, scale=1.0, levels=1)
    box = [8, 8, 4, 15, 15, 11]
    orc.set_boxes(1, [box])
    n = np.arange(16)
    k, j, i = np.meshgrid(n, n, n, indexing="ij")
    orc.upload(0, 0, ((k + 0.5) * 1.0).ravel())
    orc.upload(0, 1, np.zeros(16 ** 3))
    kk = np.arange(4, 12)
    orc.upload(1, 0, np.repeat((kk + 0.5) * 0.5, 64))
    orc.upload(1, 1, np.zeros(512))
    d, c = orc.projection(0, mode=0)
    assert d.shape == (32, 32)
    assert np.allclose(d, np.sum((np.arange(32) + 0.5) * 0.5), rtol=0, atol=1e-12)
    inside = np.zeros((32, 32), bool)
    inside[8:16, 8:16] = True
    assert np.all(c[inside] == 20) and np.all(c[~inside] == 16)
    dm, _ = orc.projection(0, mode=1)
    assert np.all(dm == 15.5)
    # Coarse level only.
    d0, c0 = orc.projection(0, mode=0, max_level=0)
    assert d0.shape == (16, 16) and np.all(c0 == 16)
    assert np.allclose(d0, np.sum(np.arange(16) + 0.5), rtol=0, atol=1e-12)
