"""A Newton iteration on the device captured in one hipGraph: Physics::UpdateState (with
UpdateVelAcc), Physics::Assemble, Physics::Constraint and CheckConvergence's norms
(physics.cpp:406-419, 255-303, 329-352; NonLinSolver.cpp:208-218) recorded by stream
capture and replayed.  Every replay, with eager work between replays, gives the same bits
as the eager sequence from the same state: K_con, f_con, the norms and the state vectors."""
import sys

import numpy as np
import pytest

from ice_fem_2025a_amd import abi, hip
from ice_fem_2025a_amd.config import make_problem
from ice_fem_2025a_amd.engine import Physics

pytestmark = pytest.mark.gpu

KINDS = [abi.IFE_STATE, abi.IFE_STATE_OLD, abi.IFE_DSTATE, abi.IFE_DDSTATE, abi.IFE_DSTATE_OLD, abi.IFE_DDSTATE_OLD]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,nel", [("hex3d_glen_at2", (3, 3, 2)), ("quad2d_thermal", (6, 4))])
def test_newton_iteration_graph_replays(require_gpu, name, nel):
    _say(f"[newton graph] {name} nel={nel}")
    prob = make_problem(name, nel=nel)
    e = Physics(prob)
    e.set_time_scheme(abi.IFE_TS_NEWMARK, 0.25, 0.5)
    e.assemble_all()
    m = prob.mesh
    zmin = m.coords[:, m.dim - 1].min()
    face = np.nonzero(np.abs(m.coords[:, m.dim - 1] - zmin) < 1e-12)[0]
    fields = [i for i, s in enumerate(prob.dof_steps) if s == 0]
    dofs = np.concatenate([e.dof_numbers(fi)[face] for fi in fields])
    x0 = e.get_state(0, abi.IFE_STATE)
    rng = np.random.default_rng(4)
    vals = x0[dofs] + rng.normal(scale=1e-6, size=dofs.size)
    e.set_constraints(0, dofs, vals)
    nfree = e.n_rows(0) - dofs.size
    dx = hip.DeviceArray.from_host(rng.normal(size=nfree) * 1e-7)
    out = hip.DeviceArray(3)
    saved = {(s, k): e.get_state(s, k) for s in range(e.nsteps) for k in KINDS}
    stream = hip.Stream()

    def restore():
        # states, the models' histories (the return map starts from the current plastic
        # strain, so it is part of the iteration's input) and c of the restored state
        for (s, k), v in saved.items():
            e.set_state(s, k, v)
        for mi in range(len(prob.models)):
            e.reset_step(mi)
        e.constrain(0)
        hip.synchronize()

    def iteration():
        e.update_state_ptr(0, dx.ptr, stream)
        e.assemble(0, stream)
        e.constrain(0, stream)
        e.residual_norms_async(0, dx.ptr, out.ptr, stream)

    def snapshot():
        stream.synchronize()
        _, _, Kc, fc = e.constrained(0)
        return [Kc, fc, out.get()] + [e.get_state(s, k) for s in range(e.nsteps) for k in KINDS]

    restore()
    iteration()
    ref = snapshot()
    assert np.isfinite(ref[2]).all() and ref[2][1] > 0
    restore()
    g = hip.Graph()
    with g.capture(stream):
        iteration()
    for rep in range(3):
        restore()
        g.launch(stream)
        got = snapshot()
        for i, (a, b) in enumerate(zip(got, ref)):
            assert np.array_equal(a, b), f"replay {rep}, item {i}"
        e.assemble_all()  # eager work between replays
        hip.synchronize()
    _say(f"  norms {ref[2]}")
