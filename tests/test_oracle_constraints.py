"""The oracle's restatement of the Dirichlet constraint elimination (Physics::Constraint,
physics.cpp:329-352, with Constrainer::Assemble's selection matrices, Constrainer.cpp:68-121)
and of CheckConvergence's norms (NonLinSolver.cpp:208-218), pinned without the reference:

* K_con = U^T K U equals K restricted to the free rows and columns (selection PtAP), with
  the pattern of that restriction;
* the constrained solve is the free-row solve the reference's formula encodes: with
  dx_free solving K_con dx_free = f_con and dx_con = -c (c = value - x), the full increment
  satisfies the free rows of K dx = -f.  (The reference lifts the constraint with -c here
  while Physics::UpdateState, physics.cpp:408-411, then adds +c; from the second Newton
  iteration on c = 0.)  A wrong sign of either term of f_con, a wrong c, or a transposed
  selection fails this;
* duplicates with different values raise, equal duplicates are accepted;
* the norms on a hand-computed case."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle.oracle import constrain, residual_norms


def _random_system(n, seed, density=0.3):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=seed, format="csr") + sp.eye(n) * (n + 1.0)
    A = A.tocsr()
    A.sort_indices()
    return A, rng.normal(size=n), rng.normal(size=n)


@pytest.mark.parametrize("n,nc,seed", [(12, 4, 1), (40, 0, 2), (40, 39, 3), (57, 20, 4)])
def test_constrained_pattern_is_the_free_restriction(n, nc, seed):
    A, f, x = _random_system(n, seed)
    rng = np.random.default_rng(seed + 100)
    dofs = rng.choice(n, nc, replace=False)
    vals = rng.normal(size=nc)
    Kc, fcon = constrain(A.indptr, A.indices, A.data, f, x, dofs, vals)
    free = np.setdiff1d(np.arange(n), dofs)
    sub = A[free][:, free].tocsr()
    sub.sort_indices()
    Kc.sort_indices()
    assert Kc.shape == (free.size, free.size) and fcon.shape == (free.size,)
    np.testing.assert_array_equal(Kc.toarray(), sub.toarray())
    np.testing.assert_array_equal(Kc.indptr, sub.indptr)
    np.testing.assert_array_equal(Kc.indices, sub.indices)


@pytest.mark.parametrize("n,nc,seed", [(12, 4, 5), (33, 10, 6), (64, 31, 7)])
def test_constrained_step_solves_the_free_rows(n, nc, seed):
    A, f, x = _random_system(n, seed)
    rng = np.random.default_rng(seed + 200)
    dofs = rng.choice(n, nc, replace=False)
    vals = rng.normal(size=nc)
    Kc, fcon = constrain(A.indptr, A.indices, A.data, f, x, dofs, vals)
    free = np.setdiff1d(np.arange(n), dofs)
    dxf = np.linalg.solve(Kc.toarray(), fcon)
    dx = np.zeros(n)
    dx[free] = dxf
    dx[dofs] = -(vals - x[dofs])
    r = A @ dx + f
    assert np.abs(r[free]).max() <= 1e-10 * (np.abs(f).max() + np.abs(A @ dx).max())
    # the opposite lift (+c) leaves a residual of the size of 2 K_fc c on the free rows
    dx[dofs] = vals - x[dofs]
    if nc:
        assert np.abs((A @ dx + f)[free]).max() > 1e-3


def test_duplicate_constraints():
    A, f, x = _random_system(10, 8)
    constrain(A.indptr, A.indices, A.data, f, x, [3, 3, 5], [1.0, 1.0, 2.0])
    with pytest.raises(ValueError, match="doubly constrained"):
        constrain(A.indptr, A.indices, A.data, f, x, [3, 5, 3], [1.0, 2.0, 1.5])


def test_residual_norms_hand_case():
    dx = np.array([1.0, -2.0, 3.0])
    fc = np.array([-1.0, 1.0, 2.0])
    a, b, c = residual_norms(dx, fc)
    assert (a, b, c) == (6.0 / 3, 4.0 / 3, 9.0 / 3)
