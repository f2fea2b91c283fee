"""GPU checks of the differentiation forms.

- Per-element Gc / A: the primal assembly with element parameters equals the oracle's
  (K, f within 1e-12, row/step scaled).
- Adjoint (ife_physics_adjoint): g = d(lambda^T r)/d(Gc_e, A_e, n) against central finite
  differences (step 1e-6 relative) of the oracle's residuals on small meshes and of the
  engine's own residuals on the 32^3 adjoint workload; directional derivatives along seeded
  random directions and single-element coordinates agree within 1e-5 relative.
- JVP: with the exact tangent, K v (ife_physics_spmv) matches the central difference of
  each step's residual within 1e-6; the reference's tangent (which drops d2F/dsigma2 in the
  Glen map) is reported beside it.
- torch.autograd.Function (ice_fem_2025a_amd.torch_ops): backward() through the adjoint
  kernel against central differences of the Function's own forward pass.

Finite differences need the return map converged far below the reference's stop rule
(1e-9 relative on the squared residual), or its iteration-count jumps swamp a 1e-6
step: every FD check runs with the stop rule tightened on both sides (TIGHT).  The map
starts from the current history, so every evaluation first resets the history to the
committed one (ResetStep); the evaluations are then bitwise repeatable and the
differences carry no round-off from the starting point.  The difference is taken on the
residual vectors before the dot product with lambda (lambda^T (r+ - r-) / 2h): rows a
parameter does not reach then cancel exactly instead of leaving the round-off of the
whole lambda^T r in a small derivative."""
import sys

import numpy as np
import pytest

from ice_fem_2025a_amd.config import make_cotangent, make_problem
from ice_fem_2025a_amd.engine import Physics
from oracle.oracle import Oracle
from parity_util import RTOL_FP64, csr_err, vec_err

pytestmark = pytest.mark.gpu

TIGHT = [("rm_abs_tol", 1e-60), ("rm_rel_tol", 1e-28), ("rm_max_iterations", 250)]
RTOL_FD = 1e-5
H = 1e-6


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tighten(x):
    for k, v in TIGHT:
        x.set_param(0, k, v)


def _residuals(x, Gc, A, n):
    x.set_element_params(0, Gc, A)
    x.set_param(0, "n", n)
    x.reset_step(0)
    return x.residuals()


def _directions(prob, rng):
    E = prob.mesh.n_elems
    Gc, A = prob.elem_params[0]["Gc"], prob.elem_params[0]["A"]
    out = [("Gc random", rng.normal(size=E) * Gc, np.zeros(E), 0.0),
           ("A random", np.zeros(E), rng.normal(size=E) * A, 0.0),
           ("n", np.zeros(E), np.zeros(E), 1.0)]
    for k in (0, E - 1):
        dg = np.zeros(E); dg[k] = Gc[k]
        da = np.zeros(E); da[k] = A[k]
        out += [(f"Gc[{k}]", dg, np.zeros(E), 0.0), (f"A[{k}]", np.zeros(E), da, 0.0)]
    return out


def _check_directions(x, prob, lamcat, g, n0=3.0, seed=7):
    Gc, A = prob.elem_params[0]["Gc"], prob.elem_params[0]["A"]
    gGc, gA, gn = g
    for name, dG, dA, dn in _directions(prob, np.random.default_rng(seed)):
        rp = _residuals(x, Gc + H * dG, A + H * dA, n0 + H * dn)
        rm = _residuals(x, Gc - H * dG, A - H * dA, n0 - H * dn)
        fd = float(lamcat @ (rp - rm)) / (2 * H)
        ad = float(gGc @ dG + gA @ dA + gn.sum() * dn)
        err = abs(ad - fd) / max(abs(fd), 1e-300)
        _say(f"  {name:10s} adjoint {ad:+.12e}  fd {fd:+.12e}  rel {err:.2e}")
        assert err <= RTOL_FD, name
    x.set_element_params(0, Gc, A)
    x.set_param(0, "n", n0)


@pytest.mark.parametrize("nel", [(3, 3, 3), (5, 4, 6)])
def test_element_params_primal_matches_oracle(require_gpu, nel):
    _say(f"[adjoint] primal with per-element Gc/A, nel={nel}")
    prob = make_problem("hex3d_adjoint", nel=nel)
    o, e = Oracle(prob), Physics(prob)
    for s in range(o.nsteps):
        e.assemble(s)
        Ke, fe = e.Kf(s)
        Ko, fo = o.assemble(s)
        rp, _ = o.pattern(s)
        ek, ef = csr_err(Ke, Ko, rp), vec_err(fe, fo)
        _say(f"  step {s}: K err {ek:.2e}, f err {ef:.2e}")
        assert ek <= RTOL_FP64 and ef <= RTOL_FP64, f"step {s}"


@pytest.mark.parametrize("nel", [(2, 2, 2), (3, 3, 3)])
def test_adjoint_against_oracle_fd(require_gpu, nel):
    _say(f"[adjoint] vs central FD of the oracle residuals, nel={nel}")
    prob = make_problem("hex3d_adjoint", nel=nel)
    o, e = Oracle(prob), Physics(prob)
    _tighten(o)
    _tighten(e)
    lam = make_cotangent([e.n_rows(s) for s in range(e.nsteps)])
    lamcat = np.concatenate([lam[s] for s in range(e.nsteps)])
    e.reset_step(0)
    e.assemble_all()
    g = e.adjoint(lam)
    assert e.status()["nonfinite"] == 0
    assert np.any(g[0] != 0) and np.any(g[1] != 0) and g[2].sum() != 0
    _check_directions(o, prob, lamcat, g)


def test_adjoint_full_size_against_engine_fd(require_gpu):
    _say("[adjoint] hex3d_adjoint 32^3 vs central FD of the engine residuals")
    prob = make_problem("hex3d_adjoint")
    e = Physics(prob)
    _tighten(e)
    lam = make_cotangent([e.n_rows(s) for s in range(e.nsteps)])
    lamcat = np.concatenate([lam[s] for s in range(e.nsteps)])
    e.reset_step(0)
    e.assemble_all()
    g = e.adjoint(lam)
    assert e.status()["nonfinite"] == 0
    _check_directions(e, prob, lamcat, g)


def test_adjoint_same_in_mixed_mode(require_gpu):
    """The adjoint pass has no element contraction: mixed mode gives the FP64 numbers."""
    prob = make_problem("hex3d_adjoint", nel=(3, 3, 3))
    e64, emx = Physics(prob), Physics(prob, precision="mixed")
    lam = make_cotangent([e64.n_rows(s) for s in range(e64.nsteps)])
    for x in (e64, emx):
        x.assemble_all()
    g64, gmx = e64.adjoint(lam), emx.adjoint(lam)
    for a, b in zip(g64, gmx):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name,nel", [("hex3d_adjoint", (3, 3, 3)), ("hex3d_glen_at2", (4, 3, 5))])
def test_jvp_exact_tangent_against_fd(require_gpu, name, nel):
    _say(f"[jvp] {name} nel={nel}")
    from ice_fem_2025a_amd import abi
    prob = make_problem(name, nel=nel)
    e = Physics(prob)
    _tighten(e)
    rng = np.random.default_rng(11)
    for s in range(e.nsteps):
        x0 = e.get_state(s, abi.IFE_STATE)
        dx0 = e.get_state(s, abi.IFE_DSTATE)
        ddx0 = e.get_state(s, abi.IFE_DDSTATE)
        v = rng.normal(size=x0.size) * max(np.sqrt(np.mean(x0 ** 2)), 1e-30)
        h = H

        def f_at(sign):
            e.reset_step(0)
            e.set_state(s, abi.IFE_STATE, x0 + sign * h * v)
            e.set_state(s, abi.IFE_DSTATE, dx0 + sign * h * prob.du_dt * v)
            e.set_state(s, abi.IFE_DDSTATE, ddx0 + sign * h * prob.ddu_dt * v)
            e.assemble(s)
            return e.Kf(s)[1]

        fd = (f_at(+1) - f_at(-1)) / (2 * h)
        e.set_state(s, abi.IFE_STATE, x0)
        e.set_state(s, abi.IFE_DSTATE, dx0)
        e.set_state(s, abi.IFE_DDSTATE, ddx0)
        res = {}
        for exact in (1, 0):
            e.set_param(0, "exact_tangent", exact)
            e.reset_step(0)
            e.assemble(s)
            kv = e.spmv(s, v)
            res[exact] = np.linalg.norm(kv - fd) / np.linalg.norm(fd)
        e.set_param(0, "exact_tangent", 0)
        _say(f"  step {s}: |Kv - fd|/|fd| exact tangent {res[1]:.2e}, reference tangent {res[0]:.2e}")
        assert res[1] <= 1e-6, f"step {s}"


def test_autograd_function_against_fd(require_gpu):
    import torch

    from ice_fem_2025a_amd.torch_ops import assembly_residual
    _say("[autograd] AssemblyResidual backward vs central FD")
    prob = make_problem("hex3d_adjoint", nel=(3, 3, 3))
    e = Physics(prob)
    _tighten(e)
    lam = make_cotangent([e.n_rows(s) for s in range(e.nsteps)])
    lamt = torch.tensor(np.concatenate([lam[s] for s in range(e.nsteps)]), device="cuda")
    Gc0 = torch.tensor(prob.elem_params[0]["Gc"], device="cuda")
    A0 = torch.tensor(prob.elem_params[0]["A"], device="cuda")
    n0 = torch.tensor(3.0, dtype=torch.float64, device="cuda")
    Gc, A, n = (t.clone().requires_grad_(True) for t in (Gc0, A0, n0))
    e.reset_step(0)
    loss = (lamt * assembly_residual(e, Gc, A, n)).sum()
    loss.backward()

    def r_at(dG, dA, dn, sgn):
        e.reset_step(0)
        with torch.no_grad():
            return assembly_residual(e, Gc0 + sgn * H * dG, A0 + sgn * H * dA, n0 + sgn * H * dn)

    rng = np.random.default_rng(5)
    E = prob.mesh.n_elems
    zero = torch.zeros(E, dtype=torch.float64, device="cuda")
    for name, dG, dA, dn in [("Gc", torch.tensor(rng.normal(size=E), device="cuda") * Gc0, zero, 0.0),
                             ("A", zero, torch.tensor(rng.normal(size=E), device="cuda") * A0, 0.0),
                             ("n", zero, zero, 1.0)]:
        fd = float((lamt * (r_at(dG, dA, dn, +1) - r_at(dG, dA, dn, -1))).sum()) / (2 * H)
        ad = float((Gc.grad * dG).sum() + (A.grad * dA).sum() + n.grad * dn)
        err = abs(ad - fd) / max(abs(fd), 1e-300)
        _say(f"  {name}: autograd {ad:+.12e} fd {fd:+.12e} rel {err:.2e}")
        assert err <= RTOL_FD, name
