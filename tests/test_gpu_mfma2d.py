"""The 2D element kernel on MFMA (assemble2d_mfma: 4, 8 or 16 elements per wavefront on
v_mfma_f32_4x4x1_16b_f32, mixed precision) against the FP64 oracle:

* CSR values and residuals within RTOL_MIXED (row / step scaled) for the 64x32 LE/AT2 patch,
  the coupled u+phi+T domain of the batched workload, the lumped inertia / damping / gravity
  terms and the multiplier irreversibility (whose records carry the lumped weights);
* the three packings give the same bits: every element runs the same sequence of 4x4 products
  whatever the number of elements sharing an instruction;
* the batched form gives the same bits as the single-domain form, for every packing;
* the full seeded 64-domain batch (989,784 elements) after the autotuner: every domain's
  linear-elastic u step satisfies f = K U to the mixed bound (size-independent; the split-free
  linear-elastic residual is the tangent times the state), its K is symmetric, nothing is
  non-finite and every element is counted; two domains match the oracle."""
import sys

import numpy as np
import pytest
import scipy.sparse as sp

from ice_fem_2025a_amd import abi
from ice_fem_2025a_amd.config import add_multipliers, make_batch, make_problem
from ice_fem_2025a_amd.engine import Batch, Physics
from oracle.oracle import Oracle
from parity_util import RTOL_MIXED, csr_err, vec_err

pytestmark = pytest.mark.gpu

EPWS = (4, 8, 16)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _all_steps(e):
    e.reset_step(0)
    e.assemble_all()
    return [e.Kf(s) for s in range(e.nsteps)]


def _inertia(prob, seed=77):
    rng = np.random.default_rng(seed)
    NN = prob.mesh.n_nodes
    h = prob.inputs["mesh"]["lengths"][0] / prob.inputs["mesh"]["nel"][0]
    for c in ("ux", "uy"):
        prob.node_state[(abi.IFE_DSTATE, c)] = rng.uniform(-1, 1, NN) * 1e-3 * h
        prob.node_state[(abi.IFE_DDSTATE, c)] = rng.uniform(-1, 1, NN) * 1e2 * h
    m = prob.models[0]
    m.opts.Inertia, m.opts.Gravity = 1, 1
    m.material.Damping = 2.0e6
    m.pf.DensityDamage = m.pf.GravityDamage = abi.DAMAGE["Quadratic"]
    prob.ddu_dt = 2.5e3
    return prob


def _cases():
    yield "quad2d_le_at2 64x32", make_problem("quad2d_le_at2")
    yield "quad2d_thermal 13x7", make_problem("quad2d_thermal", nel=(13, 7))
    yield "inertia+damping+gravity 9x5", _inertia(make_problem("quad2d_le_at2", nel=(9, 5)))
    p = make_problem("quad2d_thermal", nel=(11, 6))
    p.models[0].opts.LumpedCap, p.models[0].opts.LumpedDriving = 1, 1
    yield "lumped capacity and driving 11x6", p


CASES = list(_cases())


@pytest.mark.parametrize("label,prob", CASES, ids=[c[0] for c in CASES])
def test_mfma_2d_against_oracle(require_gpu, label, prob):
    _say(f"[mfma2d] {label}")
    o = Oracle(prob)
    ref = [o.assemble(s) for s in range(o.nsteps)]
    pats = [o.pattern(s)[0] for s in range(o.nsteps)]
    first = None
    for epw in EPWS:
        e = Physics(prob, precision="mixed")
        e.set_param(0, "elements_per_wave", epw)
        got = _all_steps(e)
        st = e.status()
        assert st["nonfinite"] == 0 and st["elements"] == prob.mesh.n_elems
        for s, ((Ke, fe), (Ko, fo)) in enumerate(zip(got, ref)):
            ek, ef = csr_err(Ke, Ko, pats[s]), vec_err(fe, fo)
            _say(f"  epw {epw} step {s}: K err {ek:.2e} f err {ef:.2e}")
            assert ek <= RTOL_MIXED and ef <= RTOL_MIXED, (epw, s, ek, ef)
        if first is None:
            first = got
        else:
            assert all(np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) for a, b in zip(got, first)), epw


def test_mfma_2d_multipliers_against_oracle(require_gpu):
    """AugmentedLagrangeMult2: the phase step's multiplier rows come from the summed lumped
    weights the element records carry."""
    _say("[mfma2d] AugmentedLagrangeMult2")
    prob = add_multipliers(make_problem("quad2d_le_at2", nel=(10, 6)), "AugmentedLagrangeMult2")
    o = Oracle(prob)
    for epw in EPWS:
        e = Physics(prob, precision="mixed")
        e.set_param(0, "elements_per_wave", epw)
        got = _all_steps(e)
        for s in range(o.nsteps):
            Ko, fo = o.assemble(s)
            rp, _ = o.pattern(s)
            ek, ef = csr_err(got[s][0], Ko, rp), vec_err(got[s][1], fo)
            _say(f"  epw {epw} step {s}: K err {ek:.2e} f err {ef:.2e}")
            assert ek <= RTOL_MIXED and ef <= RTOL_MIXED


def test_mfma_2d_rejects_bad_packing(require_gpu):
    e = Physics(make_problem("quad2d_le_at2", nel=(3, 2)), precision="mixed")
    for bad in (0, 2, 3, 32, 4.5):
        with pytest.raises(abi.IfeError):
            e.set_param(0, "elements_per_wave", bad)
    e3 = Physics(make_problem("hex3d_glen_at2", nel=(2, 2, 2)), precision="mixed")
    with pytest.raises(abi.IfeError):
        e3.set_param(0, "elements_per_wave", 8)


def _snapshot(e):
    return [e.Kf(s) for s in range(e.nsteps)]


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


def test_mfma_2d_batch_matches_single_domain(require_gpu):
    _say("[mfma2d] batched vs single-domain, 7 domains of 6-400 elements")
    probs = make_batch(n_domains=7, nel_range=(6, 400))
    doms = [Physics(p, precision="mixed") for p in probs]
    b = Batch(doms)
    for epw in EPWS:
        ref = []
        for e in doms:
            e.set_param(0, "elements_per_wave", epw)
            ref.append(_all_steps(e))
        b.set_shape(elems_per_wave=epw)
        assert b.shape()["elems_per_wave"] == epw
        for e in doms:
            e.reset_step(0)
        b.assemble()
        for d, e in enumerate(doms):
            assert _same(_snapshot(e), ref[d]), (epw, d)
        b.capture()
        for rep in range(2):
            for e in doms:
                e.reset_step(0)
            b.launch()
            for d, e in enumerate(doms):
                assert _same(_snapshot(e), ref[d]), (epw, rep, d)
    with pytest.raises(abi.IfeError):
        b.set_shape(elems_per_wave=8, grid_assemble=256)
    bd = Batch([Physics(p) for p in probs[:2]])  # FP64: no MFMA packing
    with pytest.raises(abi.IfeError):
        bd.set_shape(elems_per_wave=16)


def test_full_64_domain_batch(require_gpu):
    _say("[mfma2d] full seeded 64-domain batch")
    probs = make_batch()
    E = sum(p.mesh.n_elems for p in probs)
    assert len(probs) == 64 and E == 989784
    doms = [Physics(p, precision="mixed") for p in probs]
    _say("  domains set up")
    b = Batch(doms)
    t = b.tune()
    _say(f"  tuned: {t}")
    assert t["candidates_timed"] >= 6 and t["elems_per_wave"] in (1, 4, 8, 16)
    for e in doms:
        e.reset_step(0)
    b.launch()
    worst_fku, worst_sym = 0.0, 0.0
    for d, (p, e) in enumerate(zip(probs, doms)):
        st = e.status()
        assert st["nonfinite"] == 0 and st["elements"] == p.mesh.n_elems, d
        rp, col = e.pattern(0)
        K, f = e.Kf(0)
        U = np.concatenate([p.node_state[(abi.IFE_STATE, c)] for c in ("ux", "uy")])
        # linear elastic, no split, no lumped terms: the u-step residual is K U
        rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
        KU = np.bincount(rows, weights=K * U[col], minlength=rp.size - 1)
        rowabs = np.bincount(rows, weights=np.abs(K * U[col]), minlength=rp.size - 1)
        worst_fku = max(worst_fku, np.max(np.abs(KU - f)) / (rowabs.max() + 1e-300))
        A = sp.csr_matrix((K, col, rp), shape=(rp.size - 1, rp.size - 1))
        worst_sym = max(worst_sym, abs(A - A.T).max() / (abs(A).max() + 1e-300))
    _say(f"  max |K U - f| / max row |K U|: {worst_fku:.2e}; max asymmetry: {worst_sym:.2e}")
    assert worst_fku <= 4 * RTOL_MIXED and worst_sym <= 4 * RTOL_MIXED
    for d in (int(np.argmin([p.mesh.n_elems for p in probs])), int(np.argmax([p.mesh.n_elems for p in probs]))):
        o = Oracle(probs[d])
        for s in range(o.nsteps):
            Ko, fo = o.assemble(s)
            rp, _ = o.pattern(s)
            Ke, fe = doms[d].Kf(s)
            ek, ef = csr_err(Ke, Ko, rp), vec_err(fe, fo)
            _say(f"  domain {d} ({probs[d].mesh.n_elems} elements) step {s}: K err {ek:.2e} f err {ef:.2e}")
            assert ek <= RTOL_MIXED and ef <= RTOL_MIXED
