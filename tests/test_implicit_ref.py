"""Pins the dense implicit-differentiation reference (CPU) before it checks the GPU."""
import numpy as np
import pytest

from conftest import cloud_problem
from implicit_ref import implicit_grad, loss_value, plan
from oracle import oracle as O


def solved(seed, n=7, m=6, reg=0.3):
    M, a, b = cloud_problem(n, m, d=3, seed=seed)
    gamma = O.sinkhorn_bcd(M, a, b, reg, tol=1e-14, max_iter=200000)["dual"]
    return M, a, b, reg, gamma


def test_entropic_adjoint_is_the_envelope():
    # H w = H gamma  =>  w = gamma and G = T (the envelope theorem)
    M, a, b, reg, gamma = solved(1)
    G, w = implicit_grad(M, reg, gamma, "entropic")
    np.testing.assert_allclose(w, gamma, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(G, plan(M, reg, gamma), rtol=1e-9, atol=1e-15)
    # and the loss is the negated dual objective
    f = O.dual_obj(M, a, b, reg, gamma)
    assert abs(loss_value(M, reg, gamma, "entropic") + f) < 1e-12


@pytest.mark.parametrize("loss", ["entropic", "sharp"])
def test_implicit_grad_matches_finite_differences_in_M(loss):
    M, a, b, reg, gamma = solved(2)
    G, _ = implicit_grad(M, reg, gamma, loss)
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(6):
        E = rng.standard_normal(M.shape)
        vals = []
        for sgn in (1.0, -1.0):
            Mp = M + sgn * h * E
            gp = O.sinkhorn_bcd(Mp, a, b, reg, tol=1e-14, max_iter=200000)["dual"]
            vals.append(loss_value(Mp, reg, gp, loss))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - (G * E).sum()) < 1e-6 * max(1.0, abs(fd)), (loss, fd, (G * E).sum())


def test_dense_newton_refinement_reaches_the_unique_optimum():
    from implicit_ref import refine_dual
    M, a, b = cloud_problem(60, 45, d=5, seed=3)
    reg = 0.05
    rough = O.sinkhorn_bcd(M, a, b, reg, tol=1e-5, max_iter=100000)["dual"]
    tight = O.sinkhorn_bcd(M, a, b, reg, tol=1e-14, max_iter=1000000)["dual"]
    g, gn, it = refine_dual(M, a, b, reg, rough, tol=1e-14)
    assert gn < 1e-14 and 0 < it < 15
    np.testing.assert_allclose(g, tight, rtol=0, atol=1e-11)
