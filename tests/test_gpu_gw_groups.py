"""GW levels whose stages run as three component groups (the project's
scalars; u_xx, u_yy, u_zz with their momenta; u_xy, u_xz, u_yz with theirs)
give the same fields as the single pass over all fourteen components, bit
for bit: each output is computed by the same cell functions from the same
inputs, and the groups only leave out the loads and outputs of the others.
The split applies because the dependency probe finds every output reading
its own group's components (the scalars and its own pairs) only.
"""
import numpy as np
import pytest

import sledgehamr_amd as shm
from cases import bubble_params, phase_transition

pytestmark = pytest.mark.gpu


def _run(prec, integrator, split, steps=3, N=32, rkn_state=1):
    inp = phase_transition(N=N, integrator=integrator,
                           **{"shm.gw_split": split, "shm.rkn_state": rkn_state})
    sim = shm.Simulation("PhaseTransitionGW", inp, precision=prec)
    sim.init_state(0, 3, 7, bubble_params(8, -1e30, 0.0))
    t = sim.level_info(0)[1]
    for _ in range(steps):
        sim.advance(0, 1)
        sim.init_state(0, 4, 7, bubble_params(8, t, sim.level_info(0)[1]))
        t = sim.level_info(0)[1]
    out = [sim.download(0, c) for c in range(14)]
    groups = sim.info["gw_groups"]
    sim.close()
    return out, groups


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
@pytest.mark.parametrize("integrator", [21, 22, 10])
def test_groups_equal_single_pass(prec, integrator):
    a, ga = _run(prec, integrator, 1)
    b, gb = _run(prec, integrator, 0)
    assert ga == 1 and gb == 0
    assert max(np.max(np.abs(x)) for x in a[2:]) > 0  # tensor sourced
    for c in range(14):
        assert np.array_equal(a[c], b[c]), c


@pytest.mark.parametrize("prec", [shm.PREC_FP64, shm.PREC_FP32C16, shm.PREC_FP32CS])
@pytest.mark.parametrize("split", [0, 1])
def test_rkn_stored_stage_states_equal_composed(prec, split):
    """RKN4 with the stage states stored by the stage before (the second
    stage's by the first, the third's by the second; shm.rkn_state 1)
    against states composed from phi0, pi0 and the forces at every stencil
    point (0): identical bits, with and without component groups."""
    a, _ = _run(prec, 21, split, rkn_state=1)
    b, _ = _run(prec, 21, split, rkn_state=0)
    for c in range(14):
        assert np.array_equal(a[c], b[c]), c
