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
