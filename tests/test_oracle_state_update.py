"""The oracle's restatement of TimeScheme::UpdateVelAcc (TimeSchemes.cpp:36-62) and of
Physics::UpdateState (physics.cpp:408-411), pinned by what the schemes integrate exactly:

* Newmark (any beta, gamma) on uniformly accelerated motion x = x0 + v0 t + a t^2 / 2
  returns v0 + a dt and a;
* the Alpha scheme is the generalised trapezoid x1 = x0 + dt ((1 - alpha) v0 + alpha v1):
  v1 is recovered for any alpha, and Euler (alpha 1) is the backward difference;
* BDF2 as the reference codes it ("STILL TO CHECK"): the first-order differences;
* UpdateState puts the constrained dofs at their values and adds dx to the free ones."""
import numpy as np
import pytest

from oracle.oracle import update_state, update_vel_acc


@pytest.mark.parametrize("beta,gamma", [(0.25, 0.5), (0.3, 0.6), (0.5, 0.9)])
def test_newmark_exact_for_constant_acceleration(beta, gamma):
    rng = np.random.default_rng(1)
    x0, v0, a = rng.normal(size=(3, 50))
    dt = 0.37
    x1 = x0 + v0 * dt + 0.5 * a * dt * dt
    v1, a1 = update_vel_acc(0, beta, gamma, dt, x1, x0, v0, a)
    np.testing.assert_allclose(a1, a, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(v1, v0 + a * dt, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("alpha", [1.0, 0.5, 0.75])
def test_alpha_is_the_generalised_trapezoid(alpha):
    rng = np.random.default_rng(2)
    x0, v0, v1 = rng.normal(size=(3, 40))
    dt = 0.11
    x1 = x0 + dt * ((1 - alpha) * v0 + alpha * v1)
    v, a = update_vel_acc(1, alpha, 0.0, dt, x1, x0, v0, np.ones(40))
    np.testing.assert_allclose(v, v1, rtol=1e-12, atol=1e-12)
    assert not a.any()
    ve, _ = update_vel_acc(2, 0.0, 0.0, dt, x1, x0, v0, np.ones(40))
    np.testing.assert_allclose(ve, (x1 - x0) / dt, rtol=1e-15)


def test_bdf2_as_coded():
    x0, x1 = np.array([1.0, 2.0]), np.array([1.5, 1.0])
    v, a = update_vel_acc(3, 0.0, 0.0, 0.5, x1, x0, np.array([9.0, 9.0]), np.array([9.0, 9.0]))
    np.testing.assert_array_equal(v, [1.0, -2.0])
    np.testing.assert_array_equal(a, [2.0, -4.0])


def test_update_state_hits_the_constraints():
    rng = np.random.default_rng(3)
    x = rng.normal(size=20)
    dofs = np.array([2, 7, 11])
    vals = np.array([0.5, -1.0, 2.0])
    dx = rng.normal(size=17)
    y = update_state(x, dx, dofs, vals)
    np.testing.assert_array_equal(y[dofs], vals)
    free = np.setdiff1d(np.arange(20), dofs)
    np.testing.assert_allclose(y[free], x[free] + dx, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(update_state(x, np.ones(20)), x + 1)
