"""The device line search (ife_physics_line_search: NonLinSolver::HOLineSearch,
NonLinSolver.cpp:312-384, and the re-assembly of SolveStep_Standard, :169-175) on the Glen
u step of the 3D workload, with the increment from the device GMRES solve scaled so that
the search takes one trial, several trials, or runs out of trials:

* increments: the Newton step (accepted at the first trial), six times the Newton step (a
  multi-trial search) and an uphill direction nearly orthogonal to f_con with limits
  (1e-4, 1) (never accepted: the out-of-trials fallback to the best earlier trial);
* against a host restatement of HOLineSearch driving the engine's own primitives (update,
  Assemble, ConstraintFOnly) with the dot products in numpy: the same trial count and
  accepted eta (1e-12), step lengths and trial values within 1e-10, and the same final
  state, K_con and f_con within 1e-12;
* a whole Newton iteration -- line search, its trial assemblies and the re-assembly, the
  constraint elimination and the residual norms -- captured in one hipGraph by stream
  capture: every replay, with eager work between replays, gives the bits of the eager
  sequence from the same state."""
import sys

import numpy as np
import pytest

from ice_fem_2025a_amd import abi, hip
from ice_fem_2025a_amd.config import make_problem
from ice_fem_2025a_amd.engine import Physics

pytestmark = pytest.mark.gpu

KINDS = [abi.IFE_STATE, abi.IFE_STATE_OLD, abi.IFE_DSTATE, abi.IFE_DDSTATE, abi.IFE_DSTATE_OLD, abi.IFE_DDSTATE_OLD]
LIMS = (0.1, 1.0)  # nonlinSolver.linesearchLims of the shipped inputs


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sgn(x):
    return 1.0 if x > 0 else (-1.0 if x < 0 else 0.0)


def _setup(nel=(3, 3, 4)):
    prob = make_problem("hex3d_glen_at2", nel=nel)
    e = Physics(prob)
    e.set_time_scheme(abi.IFE_TS_NEWMARK, 0.25, 0.5)
    m = prob.mesh
    face = np.nonzero(np.abs(m.coords[:, 2] - m.coords[:, 2].min()) < 1e-12)[0]
    dofs = np.concatenate([e.dof_numbers(fi)[face] for fi in (0, 1, 2)])
    e.set_constraints(0, dofs, e.get_state(0, abi.IFE_STATE)[dofs])  # c = 0 from here on
    e.assemble(0)
    e.constrain(0)
    dx, info = e.solve(0)
    assert info["converged"] == 1
    saved = {(s, k): e.get_state(s, k) for s in range(e.nsteps) for k in KINDS}
    return prob, e, dx, saved


def _restore(e, prob, saved):
    for (s, k), v in saved.items():
        e.set_state(s, k, v)
    for mi in range(len(prob.models)):
        e.reset_step(mi)
    e.assemble(0)
    e.constrain(0)
    hip.synchronize()


def _host_line_search(e, step, dx, lims):
    """HOLineSearch (NonLinSolver.cpp:312-384) on the host over the engine's primitives."""
    f0 = e.constrained(step)[3]
    e0 = float(f0 @ dx)
    maxLS, L, ev, it, found = 10, [1.0] + [0.0] * 11, [0.0] * 12, 0, False
    Lcorr = None

    def update(v):
        d = hip.DeviceArray.from_host(v)
        e.update_state_ptr(step, d.ptr)
        hip.synchronize()

    while not found:
        dxc = L[it] * dx
        update(dxc)
        e.assemble(step)
        e.constrain_f(step)
        ev[it] = float(e.constrained(step)[3] @ dxc)
        if (abs(ev[it]) < abs(e0) and _sgn(ev[it]) == _sgn(e0)) or ev[it] == 0.0:
            found, Lcorr = True, L[it]
        else:
            if _sgn(ev[it]) != _sgn(e0):
                L[it + 1] = max(lims[0], min(lims[1], -e0 / (ev[it] - e0) * L[it]))
            else:
                L[it + 1] = 0.5 * L[it]
            if abs(L[it + 1] - L[it]) <= lims[0]:
                found, Lcorr = True, L[it]
        update((-L[it]) * dx)
        it += 1
        if it >= maxLS:
            found, eMin, Lcorr = True, e0, lims[0]
            for i in range(it - 1):
                if ev[i] < eMin:
                    eMin, Lcorr = ev[i], L[i]
    update(Lcorr * dx)
    if Lcorr < 1.0 - 1.0e-6:
        e.assemble(step)
        e.constrain(step)
    return {"trials": it, "eta": Lcorr, "L": L[:it], "e": ev[:it], "e0": e0}


def _increment(e, dx0, case):
    """(increment, limits) of a test case."""
    if case == "newton":
        return 1.0 * dx0, LIMS
    if case == "overshoot":
        return 6.0 * dx0, LIMS
    # uphill: a random direction with f_con . d = -1e-9 |f| |d|.  As f_con(x + L d) ~
    # f_con - L K d, e(L) ~ L e0 - L^2 d^T K d keeps the sign of e0 and exceeds it until
    # L < (|e0| / d^T K d)^1/2, far below the 2^-9 of ten halvings; the lower limit 1e-4 keeps
    # the |dL| <= lim stop from ending the search first
    f = e.constrained(0)[3]
    r = np.random.default_rng(5).normal(size=f.size) * np.max(np.abs(dx0))
    r -= (f @ r) / (f @ f) * f
    r -= 1e-9 * np.linalg.norm(r) / np.linalg.norm(f) * f
    return r, (1e-4, 1.0)


CASES = ["newton", "overshoot", "uphill"]


@pytest.mark.parametrize("case", CASES)
def test_line_search_matches_host_restatement(require_gpu, case):
    _say(f"[line search] hex3d_glen_at2 3x3x4, {case} increment")
    prob, e, dx0, saved = _setup()
    dxh, lims = _increment(e, dx0, case)
    ref = _host_line_search(e, 0, dxh, lims)
    xr = e.get_state(0, abi.IFE_STATE)
    _, _, Kr, fr = e.constrained(0)
    _restore(e, prob, saved)
    d = hip.DeviceArray.from_host(dxh)
    eta = hip.DeviceArray(1)
    e.line_search(0, d.ptr, *lims, eta.ptr)
    info = e.line_search_info(0)
    _say(f"  device: {info['trials']} trials, eta {info['eta']:.6f}, L {np.round(info['L'], 6)}; "
         f"host: {ref['trials']} trials, eta {ref['eta']:.6f}")
    assert info["trials"] == ref["trials"]
    assert abs(info["eta"] - ref["eta"]) <= 1e-12 and eta.get()[0] == info["eta"]
    assert np.allclose(info["L"], ref["L"], rtol=1e-10, atol=0)
    assert np.allclose(info["e"], ref["e"], rtol=1e-10, atol=1e-300)
    assert info["reassembled"] == (ref["eta"] < 1.0 - 1e-6)
    assert np.array_equal(d.get(), info["eta"] * dxh)  # dx scaled in place
    x = e.get_state(0, abi.IFE_STATE)
    assert np.max(np.abs(x - xr)) <= 1e-12 * np.max(np.abs(xr))
    _, _, K, f = e.constrained(0)
    assert np.max(np.abs(K - Kr)) <= 1e-12 * np.max(np.abs(Kr))
    assert np.max(np.abs(f - fr)) <= 1e-12 * np.max(np.abs(fr))


def test_line_search_cases_cover_the_branches(require_gpu):
    """Acceptance at the first trial, a multi-trial search, the out-of-trials fallback."""
    seen = {}
    for case in CASES:
        prob, e, dx0, saved = _setup()
        dxh, lims = _increment(e, dx0, case)
        d = hip.DeviceArray.from_host(dxh)
        e.line_search(0, d.ptr, *lims)
        seen[case] = e.line_search_info(0)["trials"]
    _say(f"  trials per case: {seen}")
    assert seen["newton"] == 1 and seen["overshoot"] > 1 and seen["uphill"] == 10


def test_newton_iteration_with_line_search_graph_replays(require_gpu):
    _say("[line search] Newton iteration in one hipGraph")
    prob, e, dx0, saved = _setup()
    dxh = 6.0 * dx0
    d = hip.DeviceArray(dxh.size)
    out = hip.DeviceArray(3)
    eta = hip.DeviceArray(1)
    stream = hip.Stream()

    def restore():
        _restore(e, prob, saved)
        hip.to_device(d.ptr, dxh)
        hip.synchronize()

    def iteration():
        e.line_search(0, d.ptr, *LIMS, eta.ptr, stream)
        e.residual_norms_async(0, d.ptr, out.ptr, stream)

    def snapshot():
        stream.synchronize()
        _, _, Kc, fc = e.constrained(0)
        return [Kc, fc, out.get(), eta.get(), d.get()] + [e.get_state(s, k) for s in range(e.nsteps) for k in KINDS] + \
            [e.history(0, nm) for nm in ("PlasticStrain", "ViscDiss")]

    restore()
    iteration()
    ref = snapshot()
    _say(f"  eager: eta {ref[3][0]:.6f}, norms {ref[2]}")
    assert ref[3][0] < 1.0
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
