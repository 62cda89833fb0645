"""Inputs files and seeded initial states shared by the parity tests."""
from __future__ import annotations

import numpy as np


def real_scalar(N=32, integrator=21, order=1, nghost=None, L=1.0, cfl=0.2,
                lam=1.0, **extra):
    d = {
        "project.name": "RealScalar",
        "project.laplacian_order": order,
        "project.lambda": lam,
        "sim.t_start": 0.0,
        "sim.t_end": 1.0,
        "sim.L": L,
        "sim.cfl": cfl,
        "amr.coarse_level_grid_size": N,
        "amr.nghost": nghost if nghost is not None else max(order, 1),
        "integrator.type": integrator,
    }
    d.update(extra)
    return d


def complex_scalar(N=32, integrator=10, nghost=2, L=None, cfl=0.2, t0=0.1,
                   **extra):
    d = {
        "project.name": "ComplexScalar",
        "project.lambda": 1.0,
        "sim.t_start": t0,
        "sim.t_end": 2.0,
        "sim.L": L if L is not None else N * 0.25,
        "sim.cfl": cfl,
        "amr.coarse_level_grid_size": N,
        "amr.nghost": nghost,
        "integrator.type": integrator,
    }
    d.update(extra)
    return d


def phase_transition(N=16, integrator=21, nghost=3, L=None, cfl=0.1, **extra):
    d = {
        "project.name": "PhaseTransitionGW",
        "project.delta": 1.5,
        "sim.t_start": 0.0,
        "sim.t_end": 1.0,
        "sim.L": L if L is not None else N * 0.5,
        "sim.cfl": cfl,
        "sim.gravitational_waves": 1,
        "amr.coarse_level_grid_size": N,
        "amr.nghost": nghost,
        "integrator.type": integrator,
    }
    d.update(extra)
    return d


def gaussian_random_field(N, seed, kcut=16, std=0.1):
    """phi with |phi_hat(n)| ~ |n|^-1 for 1 <= |n| <= kcut, random phases,
    Hermitian (real field), rescaled to the given standard deviation.
    Returned in x-fastest order (flattened [z][y][x])."""
    rng = np.random.default_rng(seed)
    f = np.fft.fftfreq(N) * N
    kz, ky, kx = np.meshgrid(f, f, f, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = np.where((kk >= 1) & (kk <= kcut), 1.0 / np.maximum(kk, 1), 0.0)
    phase = rng.uniform(0, 2 * np.pi, size=amp.shape)
    field = np.fft.ifftn(amp * np.exp(1j * phase)).real  # real part = Hermitian
    field *= std / field.std()
    return field.ravel()


def random_fields(n, ncomp, seed, scale=0.1, offset=0.0):
    rng = np.random.default_rng(seed)
    return [offset + scale * rng.standard_normal(n) for _ in range(ncomp)]


def rel_linf(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def ensemble_boxes(nbox, lo, hi, seed, mult=8):
    """Cubic boxes of seeded edges in [lo, hi] (multiples of mult), all at the
    index origin: an ensemble of independent periodic realizations."""
    rng = np.random.default_rng(seed)
    edges = rng.integers(lo // mult, hi // mult + 1, size=nbox) * mult
    return [[0, 0, 0, int(n) - 1, int(n) - 1, int(n) - 1] for n in edges]


def ensemble_inputs(boxes, integrator=21, order=2, cfl=0.2, dx=1.0 / 64):
    """configs[0] physics on every box, common dx (1/64), 4th-order
    Laplacian, ghost width 2."""
    # Index domain: the next power of two holding the largest box (the
    # coarse grid size must be a power of two).
    N = 1 << int(np.ceil(np.log2(max(b[3] + 1 for b in boxes))))
    return real_scalar(N=N, order=order, integrator=integrator, L=N * dx,
                       cfl=cfl, nghost=order, **{"shm.independent_boxes": 1})


def ensemble_fields(boxes, seed, std=0.1, kcut=16):
    """Per box a Gaussian random field (box-major, x fastest), pi = 0."""
    phi = np.concatenate([gaussian_random_field(b[3] + 1, seed + i, kcut=kcut, std=std)
                          for i, b in enumerate(boxes)])
    return [phi, np.zeros_like(phi)]


def bubble_params(n, t_from, t_to, t_lo=-2.0, t_hi=8.0, phi_in=1.0, R0=3.0, w=2.0):
    """Initial-state parameters of seeded bubbles (kinds 3 and 4): count,
    interior value, radius and wall width in level-0 cells, nucleation times
    uniform in [t_lo, t_hi], injection window (t_from, t_to]."""
    return [n, phi_in, R0, w, t_lo, t_hi, t_from, t_to]


def amr_boxes(seed, nbox=6, margin=4):
    """Nested static layout on a 32^3 base: level-1 boxes of edge 8/12/16 at
    16-cell slots of the 64^3 level-1 domain, each holding at most one
    level-2 box `margin` level-2 cells inside it on every side.  Coarse/fine
    ghost interpolation needs ceil(nghost / 2) + 1 coarse cells of margin:
    margin 4 serves nghost 2, margin 6 nghost 3."""
    rng = np.random.default_rng(seed)
    slots = [(x, y, z) for z in range(4) for y in range(4) for x in range(4)]
    lev1, lev2 = [], []
    for s in rng.choice(len(slots), nbox, replace=False):
        x, y, z = slots[s]
        e = int(rng.choice([8, 12, 16]))
        lo = (16 * x, 16 * y, 16 * z)
        lev1.append([*lo, *(v + e - 1 for v in lo)])
        e2 = int(rng.choice([8, 12, 16]))
        if 2 * margin + e2 <= 2 * e:
            lo2 = tuple(2 * v + margin for v in lo)
            lev2.append([*lo2, *(v + e2 - 1 for v in lo2)])
    return lev1, lev2


def smooth_boxes(boxes, h, L, phase):
    """Periodic smooth field at the cell centres of `boxes` (spacing h),
    box by box in the order of `boxes`, x fastest."""
    out = []
    for b in boxes:
        k, j, i = np.meshgrid(*(np.arange(b[a], b[a + 3] + 1) for a in (2, 1, 0)),
                              indexing="ij")
        x, y, z = ((i + 0.5) * h, (j + 0.5) * h, (k + 0.5) * h)
        w = 2 * np.pi / L
        f = (0.3 * np.sin(w * x + phase) * np.cos(2 * w * y) +
             0.2 * np.cos(w * z - 0.5 * phase) + 0.1 * np.sin(3 * w * (x + y + z)))
        out.append(f.ravel())
    return np.concatenate(out)


# Level-1 boxes of a 32^3 level-1 domain (coarse 16^3): two touching boxes
# (ghosts covered by the neighbour and interpolated next to it) and one at
# z = 0 whose ghosts wrap through the periodic boundary.
AMR_LEV1 = [[8, 8, 8, 15, 15, 15], [16, 8, 8, 23, 15, 15], [4, 20, 0, 11, 27, 7]]
# Level-2 boxes (64^3 domain), 4 cells inside level-1 boxes 0 and 1 (the
# second spans both); quadratic interpolation with 2 ghost cells needs 2
# level-1 cells of margin.
AMR_LEV2 = [[20, 20, 20, 27, 27, 27], [30, 20, 20, 43, 27, 27]]
# Level-2 boxes 6 level-2 cells inside level 1: the quartic reaches 2 coarse
# cells beyond the parent, so it needs 3 level-1 cells of margin.
AMR_LEV2_Q = [[22, 22, 22, 25, 25, 25], [28, 22, 22, 41, 25, 25]]


def amr_batched_boxes(seed, n1=2000, base=256, margin=6):
    """Static nested layout of the batched AMR workload.  Level 1 (2 base
    cells per axis) holds n1 boxes, one in each of n1 seeded 32-cell slots,
    even edges 16..32 at even offsets inside the slot; level 2 holds one box
    per level-1 box, even edges 16 .. min(64, 2 e1 - 2 margin), at least
    `margin` level-2 cells inside it (margin 6 serves 3 ghost cells)."""
    rng = np.random.default_rng(seed)
    ns = 2 * base // 32
    lev1, lev2 = [], []
    for s in sorted(rng.choice(ns ** 3, n1, replace=False)):
        slot = (s % ns, (s // ns) % ns, s // (ns * ns))
        e1 = int(rng.choice([16, 20, 24, 28, 32]))
        lo = [32 * c + 2 * int(rng.integers(0, (32 - e1) // 2 + 1)) for c in slot]
        lev1.append([*lo, *(v + e1 - 1 for v in lo)])
        emax = min(64, 2 * e1 - 2 * margin)
        e2 = 2 * int(rng.integers(8, emax // 2 + 1))
        room = 2 * e1 - 2 * margin - e2
        lo2 = [2 * v + margin + 2 * int(rng.integers(0, room // 2 + 1)) for v in lo]
        lev2.append([*lo2, *(v + e2 - 1 for v in lo2)])
    return lev1, lev2


def amr_batched_boxes_wide(seed, base=256, margin=6, frac=0.6, split=0.8):
    """Static nested layout of the batched AMR workload with level-1 edges
    16..64.  The level-1 domain (2 base cells per axis) is cut into 64-cell
    slots; a seeded fraction `frac` of them is used, each either holding one
    box of even edge 34..64 or (with probability `split`) cut into eight
    32-cell sub-slots holding one box of even edge 16..32 each, at even
    offsets.  Level 2 holds one box per level-1 box, even edge 16 ..
    min(64, 2 e1 - 2 margin), at least `margin` level-2 cells inside it
    (margin 6 serves 3 ghost cells).  About 2000 boxes per fine level at
    base 256."""
    rng = np.random.default_rng(seed)
    ns = 2 * base // 64
    lev1, lev2 = [], []
    slots = sorted(rng.choice(ns ** 3, int(frac * ns ** 3), replace=False))
    for s in slots:
        org = [64 * (s % ns), 64 * ((s // ns) % ns), 64 * (s // (ns * ns))]
        if rng.random() < split:
            subs = [(org[0] + 32 * a, org[1] + 32 * b, org[2] + 32 * c)
                    for c in (0, 1) for b in (0, 1) for a in (0, 1)]
            cell, elo, ehi = 32, 16, 32
        else:
            subs = [tuple(org)]
            cell, elo, ehi = 64, 34, 64
        for sub in subs:
            e1 = 2 * int(rng.integers(elo // 2, ehi // 2 + 1))
            lo = [v + 2 * int(rng.integers(0, (cell - e1) // 2 + 1)) for v in sub]
            lev1.append([*lo, *(v + e1 - 1 for v in lo)])
            emax = min(64, 2 * e1 - 2 * margin)
            e2 = 2 * int(rng.integers(8, emax // 2 + 1))
            room = 2 * e1 - 2 * margin - e2
            lo2 = [2 * v + margin + 2 * int(rng.integers(0, room // 2 + 1)) for v in lo]
            lev2.append([*lo2, *(v + e2 - 1 for v in lo2)])
    return lev1, lev2


# ---------------------------------------------------------------------------
# Tagging with user criteria and truncation-error modifiers (ComplexScalar
# with project.tag_strings / project.te_modifier).
# ---------------------------------------------------------------------------

def winding_tags(psi):
    """Winding of psi (complex [z, y, x], periodic) around the plaquettes
    with lower corner at each cell in the xy, xz and yz planes, from the
    principal-value phase increments along the four edges.  For straight
    edges that miss the origin this equals the signed ray-crossing count."""
    def wind(a, b):
        c0 = psi
        c1 = np.roll(psi, -1, axis=a)
        c2 = np.roll(c1, -1, axis=b)
        c3 = np.roll(psi, -1, axis=b)
        tot = sum(np.angle(y / x) for x, y in ((c0, c1), (c1, c2), (c2, c3), (c3, c0)))
        return np.rint(tot / (2 * np.pi)).astype(int)
    # axes of [z, y, x]: x = 2, y = 1, z = 0
    return (wind(2, 1) != 0) | (wind(2, 0) != 0) | (wind(1, 0) != 0)


def user_tag_case(N=8, scale=1.0, seed=5):
    """Fine data (FP32-representable, generic: no corner value or plaquette
    edge meets the origin of the complex plane), coarse data = child mean +
    delta with deltas chosen per coarse cell and component far from the
    decision thresholds of te_crit = 1e-3 with and without the modifiers
    (te / max|Psi_n| for the fields, te * dt for the momenta; dt = 0.05), so
    rounding of the child mean cannot move a decision."""
    rng = np.random.default_rng(seed)
    f32 = lambda a: a.astype(np.float32).astype(np.float64)
    fine = [scale * f32(rng.standard_normal((N, N, N))) for _ in range(2)]
    fine += [f32(0.3 * rng.standard_normal((N, N, N))) for _ in range(2)]
    h = N // 2
    mean = [f.reshape(h, 2, h, 2, h, 2).mean(axis=(1, 3, 5)) for f in fine]
    norms = [np.abs(f).max() for f in fine]
    kind = rng.integers(0, 3, size=(4, h, h, h))
    delta = np.zeros((4, h, h, h))
    for c in range(4):
        if c < 2:
            # 1.5e-3 passes alone, fails relative to the norm (about 3);
            # 3e-3 * norm passes both ways.
            vals = (0.0, 1.5e-3, 3e-3 * norms[c])
        else:
            # 5e-3 passes alone, fails times dt; 0.1 passes both ways.
            vals = (0.0, 5e-3, 0.1)
        for t, v in enumerate(vals):
            delta[c][kind[c] == t] = v * np.where(rng.random((h, h, h)) < 0.5, -1, 1)[kind[c] == t]
    crse = [mean[c] + delta[c] for c in range(4)]
    return fine, crse, delta, norms


def expected_tags(fine, delta, norms, modifier, user_on, dt=0.05, crit=1e-3):
    N = fine[0].shape[0]
    h = N // 2
    te = np.abs(delta)
    passes = np.zeros((4, h, h, h), dtype=bool)
    for c in range(4):
        m = te[c]
        if modifier:
            m = m / norms[c] if c < 2 else m * dt
        passes[c] = m >= crit
    user = winding_tags(fine[0] + 1j * fine[1]) if user_on else \
        np.zeros((N, N, N), dtype=bool)
    anyp = passes.any(axis=0)
    mask = user.copy()
    mask[0::2, 0::2, 0::2] |= anyp
    even_user = user[0::2, 0::2, 0::2]
    total = int(user.sum()) + int((8 - even_user[anyp]).sum())
    counts = [total, int(user.sum())] + [8 * int(passes[c].sum()) for c in range(4)]
    return mask, np.array(counts)
