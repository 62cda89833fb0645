"""GPU parity at the measured sizes: the headline 512^3 complex-scalar level
(SSPRK3, 4th order) and the 384^3 phase-transition level with the GW tensor
(14 components, 6th order, RKN4, bubbles injected at their nucleation
times), two steps each from the same seeded initial states (both sides
generate them with the same counter-based draws), against the FP64 oracle.

Tolerances: relative L-infinity per component against max |oracle|; 1e-12
(FP64), 1e-6 (FP32-compensated: FP32 + bf16 storage with FP32 stencil
inputs, and with FP64 stencil tiles).  The spectrum is checked at the
reduced sizes (test_gpu_configs.py): the oracle's transform is a plain DFT.
"""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import bubble_params, complex_scalar, phase_transition, rel_linf
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def test_cfg1_512_two_steps_against_oracle():
    N = 512
    inp = complex_scalar(N=N, integrator=10, L=N * 0.25, cfl=0.2, t0=0.1)
    seed = 20240401
    orc = Oracle("ComplexScalar", inp)
    orc.init_state(0, 1, seed, [0.05])
    orc.advance(0, 2)
    ref = [orc.download(0, c) for c in range(4)]
    del orc
    for prec, tol in ((shm.PREC_FP64, 1e-12), (shm.PREC_FP32C16, 1e-6),
                      (shm.PREC_FP32CS, 1e-6)):
        sim = shm.Simulation("ComplexScalar", inp, precision=prec)
        sim.init_state(0, 1, seed, [0.05])
        sim.advance(0, 2)
        for c in range(4):
            err = rel_linf(sim.download(0, c), ref[c])
            assert err < tol, (prec, c, err)
        sim.close()


def test_cfg3_384_two_steps_against_oracle():
    N = 384
    inp = phase_transition(N=N, integrator=21, L=N * 0.5, cfl=0.2,
                           **{"shm.allow_any_grid_size": 1})
    seed = 20240401
    orc = Oracle("PhaseTransitionGW", inp)
    orc.init_state(0, 3, seed, bubble_params(64, -1e30, 0.0))
    t = orc.time(0)
    for _ in range(2):
        orc.advance(0, 1)
        orc.init_state(0, 4, seed, bubble_params(64, t, orc.time(0)))
        t = orc.time(0)
    ref = [orc.download(0, c) for c in range(14)]
    del orc
    assert np.max(np.abs(ref[0])) > 0.5          # bubbles present
    assert max(np.max(np.abs(r)) for r in ref[2:]) > 0  # tensor sourced
    for prec, tol in ((shm.PREC_FP64, 1e-12), (shm.PREC_FP32CS, 1e-6)):
        sim = shm.Simulation("PhaseTransitionGW", inp, precision=prec)
        sim.init_state(0, 3, seed, bubble_params(64, -1e30, 0.0))
        t = sim.level_info(0)[1]
        for _ in range(2):
            sim.advance(0, 1)
            sim.init_state(0, 4, seed, bubble_params(64, t, sim.level_info(0)[1]))
            t = sim.level_info(0)[1]
        for c in range(14):
            err = rel_linf(sim.download(0, c), ref[c])
            assert err < tol, (prec, c, err)
        sim.close()


def test_non_power_of_two_level_needs_opt_in():
    """The reference rejects a coarse level that is not a power of two
    (sledgehamr.cpp:438-444); shm.allow_any_grid_size admits even sizes."""
    inp = phase_transition(N=24, integrator=21, L=12.0, cfl=0.2)
    with pytest.raises(shm.ShmError) as e:
        shm.Simulation("PhaseTransitionGW", inp, precision=shm.PREC_FP64)
    assert e.value.code == 1
    inp["shm.allow_any_grid_size"] = 1
    sim = shm.Simulation("PhaseTransitionGW", inp, precision=shm.PREC_FP64)
    assert sim.ncells(0) == 24 ** 3
    sim.close()
    inp["amr.coarse_level_grid_size"] = 25
    with pytest.raises(shm.ShmError):
        shm.Simulation("PhaseTransitionGW", inp, precision=shm.PREC_FP64)
