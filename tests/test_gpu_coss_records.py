"""The Cosserat u step through element records and the ordered per-node gather
(coss_assemble_mfma + coss_gather_kernel, CossPhaseField.cpp:337-534), in element-range chunks:

* chunked (7 elements per chunk, so most nodes are completed across several chunks) and
  unchunked give the same bits, in FP64 and mixed precision, 2D and 3D: every entry is the sum
  of its element contributions in element order either way;
* the records path and the coloured read-modify-write path agree within 1e-14 (row scaled;
  they add the same contributions in a different order);
* the full 200x50x60 workload (600,000 elements): the u-step pattern has 2,790,895,104
  entries with 64-bit row offsets, and a seeded sample of node rows (interior, face, edge
  and corner nodes) matches the oracle on the sub-mesh of each node's elements within 1e-12
  (row/step scaled, the map converged as in test_gpu_coss's TIGHT cases): a node's rows
  involve only its elements, so the sub-mesh rows are the full mesh's rows."""
import copy
import ctypes as C
import sys

import numpy as np
import pytest

from ice_fem_2025a_amd import abi, hip
from ice_fem_2025a_amd.config import Problem, make_problem
from ice_fem_2025a_amd.engine import Physics
from ice_fem_2025a_amd.mesh import PatchMesh
from oracle.oracle import Oracle
from parity_util import RTOL_FP64, csr_err, vec_err

pytestmark = pytest.mark.gpu

TIGHT = [("rm_abs_tol", 1e-40), ("rm_rel_tol", 1e-30), ("rm_max_iterations", 60)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _u_step(prob, precision, chunk):
    e = Physics(prob, precision=precision)
    e.set_param(0, "coss_record_chunk", chunk)
    e.assemble(0)
    return e.Kf(0)


@pytest.mark.parametrize("name,nel", [("hex3d_coss", (4, 3, 5)), ("quad2d_coss", (9, 6))])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chunked_records_equal_unchunked(require_gpu, name, nel, precision):
    _say(f"[coss records] {name} nel={nel} {precision}")
    prob = make_problem(name, nel=nel)
    K0, f0 = _u_step(prob, precision, 0)
    K7, f7 = _u_step(prob, precision, 7)
    assert np.array_equal(K0, K7) and np.array_equal(f0, f7)
    if precision == "fp64":
        Kc, fc = _u_step(prob, precision, -1)  # the coloured read-modify-write path
        rp, _ = Physics(prob, device=False).pattern(0)
        ek, ef = csr_err(K0, Kc, rp), vec_err(f0, fc)
        _say(f"  records vs colours: K {ek:.2e}, f {ef:.2e}")
        assert ek <= 1e-14 and ef <= 1e-14


def _sub_problem(prob, elems):
    """The problem restricted to `elems` (ascending): nodes renumbered in ascending global id,
    so every field's columns keep their order."""
    m = prob.mesh
    nodes = np.unique(m.elems[elems])
    loc = -np.ones(m.n_nodes, dtype=np.int64)
    loc[nodes] = np.arange(nodes.size)
    sub = PatchMesh(dim=m.dim, nel=(len(elems),), lengths=(1.0,), coords=m.coords[nodes].copy(),
                    weights=m.weights[nodes].copy(), elems=loc[m.elems[elems]],
                    extraction_1d=m.extraction_1d[elems].copy(), ipcount1D=m.ipcount1D)
    nip = prob.nip
    ips = (np.asarray(elems)[:, None] * nip + np.arange(nip)[None, :]).ravel()
    sp = Problem(name=prob.name, inputs=prob.inputs, mesh=sub, dof_names=prob.dof_names, dof_steps=prob.dof_steps,
                 models=prob.models, t=prob.t, dt=prob.dt, du_dt=prob.du_dt, ddu_dt=prob.ddu_dt)
    sp.node_state = {k: v[nodes].copy() for k, v in prob.node_state.items()}
    sp.history = {k: v[:, ips].copy() for k, v in prob.history.items()}
    return sp, nodes, loc


def test_full_size_u_step_rows_against_the_oracle(require_gpu):
    _say("[coss records] hex3d_coss 200x50x60 (600,000 elements)")
    prob = make_problem("hex3d_coss")
    m = prob.mesh
    NN, NFU = m.n_nodes, 6
    e = Physics(prob)
    for k, v in TIGHT:
        e.set_param(0, k, v)
    _say("  engine set up")
    nr, nnz = C.c_int64(), C.c_int64()
    rp_h = C.POINTER(C.c_int64)()
    abi.check(e.L.ife_physics_pattern(e.h, 0, C.byref(nr), C.byref(nnz), C.byref(rp_h), None, None, None))
    assert nr.value == NFU * NN == 3907488 and nnz.value == 2790895104
    rp = np.ctypeslib.as_array(rp_h, shape=(nr.value + 1,))
    assert rp.dtype == np.int64 and rp[-1] == 2790895104 > 2 ** 31
    e.assemble(0)
    st = e.status()
    assert st["nonfinite"] == 0 and st["elements"] == 600000
    _say(f"  assembled: {st}")
    Kp, fp = e.Kf_ptrs(0)
    fe = np.empty(nr.value)
    hip.from_device(fe, fp)
    # nodes x-fastest on a 202 x 52 x 62 grid: interior, face, edge, corner samples
    gx, gy, gz = 202, 52, 62
    rng = np.random.default_rng(20261207)
    picks = [(0, 0, 0), (gx - 1, gy - 1, gz - 1), (0, gy // 2, gz // 2), (gx // 2, 0, gz - 1), (1, 1, 1)]
    picks += [tuple(int(rng.integers(2, g - 2)) for g in (gx, gy, gz)) for _ in range(7)]
    for (ix, iy, iz) in picks:
        node = ix + gx * (iy + gy * iz)
        elems = np.nonzero(np.any(m.elems == node, axis=1))[0]
        sp, nodes, loc = _sub_problem(prob, elems)
        o = Oracle(sp)
        for k, v in TIGHT:
            o.set_param(0, k, v)
        Ko, fo = o.assemble(0)
        rpo, _ = o.pattern(0)
        nl = nodes.size
        scale_f = np.max(np.abs(fe))
        worst = 0.0
        for al in range(NFU):
            rg, rl = al * NN + node, al * nl + loc[node]
            a, b = rp[rg], rp[rg + 1]
            Krow = np.empty(b - a)
            hip.from_device(Krow, Kp + 8 * int(a))
            Korow = Ko[rpo[rl]:rpo[rl + 1]]
            assert Krow.size == Korow.size
            sc = max(np.max(np.abs(Korow)), 1e-300)
            worst = max(worst, np.max(np.abs(Krow - Korow)) / sc)
            assert abs(fe[rg] - fo[rl]) <= RTOL_FP64 * scale_f, (node, al)
        _say(f"  node {node} ({ix},{iy},{iz}): {elems.size} elements, {b - a} columns per row, K err {worst:.2e}")
        assert worst <= RTOL_FP64, node
