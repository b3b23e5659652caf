"""Pins the FP64 CPU oracle against checks that do not restate it.

The reference ships no tests or golden vectors (and cannot be built here: its
Eigen dependency is downloaded at build time), so each oracle function is
pinned by an independent route: an independent library routine
(scipy.special.logsumexp), closed forms, finite differences, dense linear
algebra, brute force on tiny inputs, invariants, and the uniqueness of the
entropic optimum.
"""
import numpy as np
import pytest
from scipy.special import logsumexp

from conftest import cloud_problem, readme_example
from oracle import oracle as O


def rand_problem(n=23, m=17, seed=0, reg=0.1):
    rng = np.random.default_rng(seed)
    M = rng.random((n, m))
    a = rng.random(n) + 0.1
    b = rng.random(m) + 0.1
    return M, a / a.sum(), b / b.sum(), reg


# --- dense passes (src/sinkhorn_problem_bcd.cpp, _obj.cpp) -----------------

@pytest.mark.parametrize("offset", [0.0, 50.0, -80.0])
def test_half_steps_match_scipy_logsumexp(offset):
    M, a, b, reg = rand_problem(reg=0.02)
    rng = np.random.default_rng(1)
    beta = rng.standard_normal(M.shape[1]) + offset
    alpha = rng.standard_normal(M.shape[0]) - offset
    al = reg * (np.log(a) - logsumexp((beta[None, :] - M) / reg, axis=1))
    be = reg * (np.log(b) - logsumexp((alpha[:, None] - M) / reg, axis=0))
    np.testing.assert_allclose(O.optimal_alpha(M, a, reg, beta), al, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.optimal_beta(M, b, reg, alpha), be, rtol=1e-13, atol=1e-13)


def test_half_step_marginal_identity():
    # after a beta update the column sums of T equal b; after an alpha update
    # the row sums equal a (the defining property of the block step)
    M, a, b, reg = rand_problem(seed=4)
    alpha = np.random.default_rng(2).standard_normal(M.shape[0])
    beta = O.optimal_beta(M, b, reg, alpha)
    T = np.exp((alpha[:, None] + beta[None, :] - M) / reg)
    np.testing.assert_allclose(T.sum(0), b, rtol=1e-13)
    alpha2 = O.optimal_alpha(M, a, reg, beta)
    T = np.exp((alpha2[:, None] + beta[None, :] - M) / reg)
    np.testing.assert_allclose(T.sum(1), a, rtol=1e-13)


def test_obj_grad_plan_brute_force_and_fd():
    M, a, b, reg = rand_problem(n=6, m=5, seed=3)
    n, m = M.shape
    gamma = np.random.default_rng(5).standard_normal(n + m - 1) * 0.1
    beta = np.append(gamma[n:], 0.0)
    T = np.exp((gamma[:n, None] + beta[None, :] - M) / reg)
    f_bf = reg * T.sum() - gamma[:n] @ a - gamma[n:] @ b[:-1]
    g_bf = np.concatenate([T.sum(1) - a, T.sum(0)[:-1] - b[:-1]])
    f, g = O.dual_obj_grad(M, a, b, reg, gamma)
    assert abs(f - f_bf) < 1e-13 and abs(O.dual_obj(M, a, b, reg, gamma) - f_bf) < 1e-13
    np.testing.assert_allclose(g, g_bf, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(O.get_plan(M, reg, gamma), T, rtol=1e-14)
    h = 1e-6
    fd = np.array([(O.dual_obj(M, a, b, reg, gamma + h * e) -
                    O.dual_obj(M, a, b, reg, gamma - h * e)) / (2 * h)
                   for e in np.eye(n + m - 1)])
    np.testing.assert_allclose(fd, g, rtol=1e-6, atol=1e-8)


# --- solvers (unique optimum, closed forms) --------------------------------

def test_separable_cost_closed_form():
    # M_ij = u_i + v_j  =>  T* = a b', alpha*_i = u_i + reg log a_i + c,
    # beta*_j = v_j + reg log b_j - c, with c fixing beta*_{m-1} = 0.
    rng = np.random.default_rng(7)
    n, m, reg = 12, 9, 0.3
    u, v = rng.random(n), rng.random(m)
    M = u[:, None] + v[None, :]
    a = rng.random(n) + 0.2
    b = rng.random(m) + 0.2
    a, b = a / a.sum(), b / b.sum()
    c = v[-1] + reg * np.log(b[-1])
    dual = np.concatenate([u + reg * np.log(a) + c, (v + reg * np.log(b) - c)[:-1]])
    r = O.sinkhorn_bcd(M, a, b, reg, tol=1e-13, max_iter=100)
    assert r["niter"] <= 2
    np.testing.assert_allclose(r["dual"], dual, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.get_plan(M, reg, r["dual"]), np.outer(a, b), rtol=1e-12)
    # the objective at the optimum: f* = reg sum(T*) - dual . [a; b_t]
    assert abs(r["obj_vals"][-1] - (reg - dual[:n] @ a - dual[n:] @ b[:-1])) < 1e-13


def independent_sinkhorn(M, a, b, reg, iters=20000):
    """Log-domain Sinkhorn in scipy, a different formulation of the same optimum."""
    f = np.zeros(M.shape[0])
    g = np.zeros(M.shape[1])
    la, lb = np.log(a), np.log(b)
    for _ in range(iters):
        f_new = reg * (la - logsumexp((g[None, :] - M) / reg, axis=1))
        g = reg * (lb - logsumexp((f_new[:, None] - M) / reg, axis=0))
        if np.max(np.abs(f_new - f)) < 1e-15:
            f = f_new
            break
        f = f_new
    return np.concatenate([f + g[-1], g[:-1] - g[-1]])


def test_all_drivers_reach_the_unique_optimum():
    M, a, b = cloud_problem(40, 30, d=4, seed=3)
    reg = 0.1
    ref = independent_sinkhorn(M, a, b, reg)
    plan_ref = O.get_plan(M, reg, ref)
    out = {
        "bcd": O.sinkhorn_bcd(M, a, b, reg, tol=1e-12, max_iter=100000),
        "ssns": O.sinkhorn_ssns(M, a, b, reg, tol=1e-12, max_iter=500),
        "ssns_cg": O.sinkhorn_ssns(M, a, b, reg, tol=1e-11, max_iter=500, method=0,
                                   cg_tol=1e-12, cg_max_iter=5000),
        # sparse Newton and SPLR shrink the Hessian density with the gradient
        # norm, so their line searches stall near 1e-9 on this problem
        "sn": O.sinkhorn_sparse_newton(M, a, b, reg, tol=1e-8, max_iter=3000, density=1.0),
        "splr": O.sinkhorn_splr(M, a, b, reg, tol=1e-8, max_iter=2000, density=0.5),
    }
    for name, r in out.items():
        err = r["mar_errs"][-1]
        assert err < (1e-8 if name in ("sn", "splr") else 1e-10), name
        # ||grad|| bounds the dual error through the smallest Hessian curvature
        tol = max(1e-8, 1e3 * err)
        np.testing.assert_allclose(r["dual"], ref, rtol=0, atol=tol, err_msg=name)
        np.testing.assert_allclose(O.get_plan(M, reg, r["dual"]), plan_ref, atol=10 * tol)
        assert len(r["obj_vals"]) == r["niter"] + 1


def test_bcd_history_semantics():
    M, a, b = readme_example(30, 20)
    r = O.sinkhorn_bcd(M, a, b, 0.1, tol=0.0, max_iter=7)
    assert r["niter"] == 7 and len(r["mar_errs"]) == 8
    assert all(x > 0 for x in np.diff(r["run_times"]) + 1e-12)
    r = O.sinkhorn_bcd(M, a, b, 0.1, tol=1e9, max_iter=7)
    assert r["niter"] == 0 and len(r["mar_errs"]) == 1
    # f decreases monotonically along BCD (it is exact block minimisation)
    r = O.sinkhorn_bcd(M, a, b, 0.1, tol=1e-10, max_iter=1000)
    f = np.asarray(r["obj_vals"])
    assert np.all(np.diff(f) <= 64 * np.finfo(float).eps * np.abs(f[1:]))


# --- sparsifiers (src/sinkhorn_sparsify3.cpp, src/sinkhorn_sparsify4.cpp) ---

def brute_threshold(T, delta, hint):
    """Value characterisation of the threshold rule (not the oracle's sort-and-scan):
    column j keeps T_ij >= tau_j, tau_j = min{v : sum_{T_kj <= v} T_kj > delta}
    (+inf if no such v with find_small, the column max with find_large); then each
    row whose dropped mass exceeds 2 delta re-keeps its dropped entries >= rho_i,
    rho_i = max{v : sum_{d_ik >= v} d_ik >= D_i - 2 delta}; exact zeros are kept."""
    n, m = T.shape
    keep = np.zeros((n, m - 1), bool)
    for j in range(m - 1):
        col = T[:, j]
        cands = [v for v in np.unique(col) if col[col <= v].sum() > delta]
        if cands:
            tau = min(cands)
        else:
            tau = np.inf if hint > 0.2 else col.max()
        keep[:, j] = col >= tau
    for i in range(n):
        d = np.where(keep[i], 0.0, T[i, :-1])
        D = d.sum()
        if D > 2 * delta:
            rho = max(v for v in np.unique(d) if d[d >= v].sum() >= D - 2 * delta)
            keep[i] |= d >= rho
    keep |= T[:, :-1] == 0.0
    return keep


@pytest.mark.parametrize("hint", [0.001, 0.5])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_threshold_rule_brute_force(hint, seed):
    rng = np.random.default_rng(seed)
    n, m = 9, 8
    T = rng.exponential(size=(n, m)) ** 3 * 0.01
    T[rng.random((n, m)) < 0.1] = 0.0
    for delta in (1e-4, 3e-3, 0.05, 10.0):
        cp, ri, va = O.sparsify_threshold(T, delta, hint)
        keep = np.zeros((n, m - 1), bool)
        for j in range(m - 1):
            keep[ri[cp[j]:cp[j + 1]], j] = True
        assert np.array_equal(keep, brute_threshold(T, delta, hint)), delta
        S = O.csc_to_dense(n, m, cp, ri, va)
        assert np.array_equal(S[keep], T[:, :-1][keep])
        dropped = np.where(keep, 0.0, T[:, :-1])
        assert np.all(dropped.sum(1) <= 2 * delta + 1e-15)


def test_topk_rule_brute_force():
    rng = np.random.default_rng(3)
    n, m = 11, 9
    T = rng.random((n, m))
    for density in (0.0, 0.05, 0.3, 1.0, 1.5):
        cp, ri, va = O.sparsify_density(T, density, True)
        S = O.csc_to_dense(n, m, cp, ri, va)
        k = int(n * (m - 1) * min(max(density, 0.0), 1.0))
        body = T[:, :-1]
        thr = np.sort(body.ravel())[::-1][k - 1] if k > 0 else np.inf
        keep = body >= thr
        keep[0, :] = True
        keep[:, 0] = True
        assert np.array_equal(S != 0, keep)
        np.testing.assert_array_equal(S[keep], body[keep])
        cp2, ri2, va2 = O.sparsify_density(T, density, False)
        assert len(ri2) == k


# --- Hessian operators, CG, BFGS (src/sinkhorn_hess.cpp, _cg.cpp, _linsolve.cpp)

def test_hessian_ops_cg_and_direct_against_dense():
    M, a, b = cloud_problem(25, 19, d=3, seed=8)
    reg = 0.1
    n, m = M.shape
    gamma = independent_sinkhorn(M, a, b, reg, 50)
    T, _ = O.compute_T(M, reg, gamma)
    f, g = O.dual_obj_grad(M, a, b, reg, gamma)
    cp, ri, va = O.sparsify_density(T, 0.3, True)
    va = va / reg
    h1 = (g[:n] + a) / reg
    h2 = (g[n:] + b[:-1]) / reg
    S = O.csc_to_dense(n, m, cp, ri, va)
    Hs = np.block([[np.diag(h1), S], [S.T, np.diag(h2)]])
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n + m - 1)
    np.testing.assert_allclose(O.apply_hsx(n, m, h1, h2, (cp, ri, va), x), Hs @ x, rtol=1e-12)
    shift = 1e-2
    z = rng.standard_normal(m - 1)
    Delta = np.diag(h2 + shift) - S.T @ np.diag(1.0 / (h1 + shift)) @ S
    np.testing.assert_allclose(O.apply_deltax(n, m, h1, h2, (cp, ri, va), z, shift), Delta @ z,
                               rtol=1e-10, atol=1e-12)
    ref = np.linalg.solve(Hs + shift * np.eye(n + m - 1), x)
    sol, it, err = O.hess_cg(n, m, h1, h2, (cp, ri, va), x, shift, tol=1e-12, max_iter=1000)
    assert err < 1e-12 and it < 1000
    np.testing.assert_allclose(sol, ref, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(O.hess_direct(n, m, h1, h2, (cp, ri, va), x, shift), ref,
                               rtol=1e-9, atol=1e-12)


def test_bfgs_inverse_update_identity():
    # Nocedal & Wright (6.17): H+ = (I - rho s y')H(I - rho y s') + rho s s',
    # rho = 1/(y's), with H = A^-1; the combination must equal H+ r.
    rng = np.random.default_rng(9)
    N = 7
    B = rng.standard_normal((N, N))
    A = B @ B.T + N * np.eye(N)
    r, y, s = rng.standard_normal((3, N))
    if y @ s < 0:
        y = -y
    H = np.linalg.inv(A)
    rho = 1.0 / (y @ s)
    I = np.eye(N)
    Hp = (I - rho * np.outer(s, y)) @ H @ (I - rho * np.outer(y, s)) + rho * np.outer(s, s)
    got = O.low_rank_combine(H @ r, H @ y, r, y, s)
    np.testing.assert_allclose(got, Hp @ r, rtol=1e-12)


# --- line-search interpolation (src/sinkhorn_problem_ls.cpp:72-223) -----------

def test_cubic_minimizer_against_polynomial_fit():
    rng = np.random.default_rng(11)
    for _ in range(50):
        c = rng.standard_normal(4)
        c[3] = abs(c[3]) + 0.1        # positive cubic coefficient: a local min exists
        p = np.polynomial.Polynomial(c)
        dp = p.deriv()
        roots = [r.real for r in dp.roots() if abs(r.imag) < 1e-12]
        if not roots:
            continue
        xmin = max(roots, key=lambda t: p.deriv(2)(t))
        if p.deriv(2)(xmin) <= 0:
            continue
        a_, b_ = xmin - 1.3, xmin + 0.7
        x, ok = O.cubic_minimizer(a_, b_, p(a_), p(b_), dp(a_), dp(b_))
        assert ok
        assert abs(x - xmin) < 1e-8 * max(1, abs(xmin))
    # a quadratic through the data reduces to the vertex
    q = np.polynomial.Polynomial([1.0, -2.0, 0.5])
    x, ok = O.cubic_minimizer(0.0, 3.0, q(0.0), q(3.0), q.deriv()(0.0), q.deriv()(3.0))
    assert ok and abs(x - 2.0) < 1e-12


def test_step_selection_cases():
    # Case 1 (ft > fl): the new step lies strictly inside (al, at)
    t = O.step_selection(0.0, np.inf, 1.0, 0.0, np.inf, 1.0, -1.0, np.inf, 2.0)
    assert 0.0 < t < 1.0
    # non-finite trial value: midpoint
    assert O.step_selection(0.0, np.inf, 1.0, 0.0, np.inf, np.inf, -1.0, np.inf, 1.0) == 0.5
    # degenerate interval
    assert O.step_selection(0.3, 0.3, 1.0, 0.0, 0.0, 0.0, -1.0, -1.0, -1.0) == 0.3


# --- golden fixture (independent optimum of the C1 recipe) --------------------

def test_golden_c1_readme_recipe():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "c1_readme_n1000_m800.npz")
    z = np.load(path)
    M, a, b = readme_example(1000, 800)
    np.testing.assert_array_equal(a, z["a"])
    r = O.sinkhorn_bcd(M, a, b, float(z["reg"]), tol=1e-12, max_iter=10000)
    assert r["niter"] < 10000
    # the fixture's own marginal error is ~1e-15; 1e-12 pins the dual to ~1e-9
    np.testing.assert_allclose(r["dual"], z["dual"], rtol=0, atol=1e-9)
    assert abs(r["obj_vals"][-1] - float(z["obj"])) < 1e-12


# --- dense Newton, L-BFGS dual, APDAGD (src/sinkhorn_{newton,lbfgs_dual,apdagd}.cpp) ---

def test_newton_direction_against_dense_numpy_solve():
    # H assembled from an independently computed T (numpy exp), then
    # numpy.linalg.solve: the oracle's Cholesky direction must agree
    M, a, b, reg = rand_problem(n=9, m=7, seed=11)
    n, m = M.shape
    gamma = np.random.default_rng(12).standard_normal(n + m - 1) * 0.05
    beta = np.append(gamma[n:], 0.0)
    T = np.exp((gamma[:n, None] + beta[None, :] - M) / reg)
    g = np.concatenate([T.sum(1) - a, T.sum(0)[:-1] - b[:-1]])
    H = np.zeros((n + m - 1, n + m - 1))
    H[:n, :n] = np.diag(T.sum(1))
    H[n:, n:] = np.diag(T.sum(0)[:-1])
    H[:n, n:] = T[:, :-1]
    H[n:, :n] = T[:, :-1].T
    H /= reg
    np.testing.assert_allclose(O.newton_direction(M, a, b, reg, gamma),
                               np.linalg.solve(H, -g), rtol=1e-10, atol=1e-12)


def test_lbfgs_two_loop_against_dense_bfgs_recursion():
    # The two-loop product equals the dense inverse-BFGS recursion
    # H+ = (I - rho s y')H(I - rho y s') + rho s s' started at H0 = I / theta,
    # theta = y'y / s'y of the newest pair, over the last mm pairs.
    rng = np.random.default_rng(21)
    N, mm = 11, 6
    A = rng.standard_normal((N, N))
    A = A @ A.T + N * np.eye(N)
    for npairs in (1, 4, 9):
        S = rng.standard_normal((npairs, N))
        Y = S @ A          # curvature pairs of a convex quadratic: s'y > 0
        v = rng.standard_normal(N)
        keep = range(max(0, npairs - mm), npairs)
        theta = Y[-1] @ Y[-1] / (S[-1] @ Y[-1])
        H = np.eye(N) / theta
        I = np.eye(N)
        for k in keep:
            rho = 1.0 / (Y[k] @ S[k])
            H = (I - rho * np.outer(S[k], Y[k])) @ H @ (I - rho * np.outer(Y[k], S[k])) \
                + rho * np.outer(S[k], S[k])
        np.testing.assert_allclose(O.lbfgs_apply(S, Y, v, mm), H @ v, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_nocedal_wright_search_returns_a_strong_wolfe_point(seed):
    M, a, b, reg = rand_problem(n=15, m=12, seed=30 + seed, reg=0.05)
    n, m = M.shape
    rng = np.random.default_rng(seed)
    gamma = rng.standard_normal(n + m - 1) * 0.05
    f0, g0 = O.dual_obj_grad(M, a, b, reg, gamma)
    d = -g0 + 0.3 * rng.standard_normal(n + m - 1) * np.abs(g0).mean()
    if g0 @ d > 0:
        d = -g0
    for init in (1e-3, 1.0, 50.0):
        rc, step, f, g, lf = O.nw_line_search(M, a, b, reg, gamma, d, init)
        assert rc == 0
        # accepted point is where the oracle says, with its own f and gradient
        f_chk, g_chk = O.dual_obj_grad(M, a, b, reg, gamma + step * d)
        assert abs(f - f_chk) <= 1e-14 * max(1.0, abs(f))
        np.testing.assert_allclose(g, g_chk, rtol=0, atol=1e-14)
        # sufficient decrease (ftol 1e-4) and strong curvature (wolfe 0.9)
        assert f - f0 <= 1e-4 * step * (g0 @ d) + 1e-15
        assert abs(g @ d) <= 0.9 * abs(g0 @ d) + 1e-15
        assert lf == f
    # an ascent direction is refused (LBFGSpp's logic_error)
    rc = O.nw_line_search(M, a, b, reg, gamma, g0, 1.0)[0]
    assert rc == -3


def test_newton_lbfgs_apdagd_reach_the_unique_optimum():
    M, a, b = cloud_problem(40, 30, d=4, seed=3)
    reg = 0.1
    ref = independent_sinkhorn(M, a, b, reg)
    plan_ref = O.get_plan(M, reg, ref)
    nw = O.sinkhorn_newton(M, a, b, reg, tol=1e-11, max_iter=100)
    lb = O.sinkhorn_lbfgs_dual(M, a, b, reg, tol=1e-8, max_iter=5000)
    for name, r, bound in (("newton", nw, 1e-11), ("lbfgs", lb, 1e-8)):
        err = r["mar_errs"][-1]
        assert err <= bound, name
        tol = max(1e-8, 1e3 * err)
        np.testing.assert_allclose(r["dual"], ref, rtol=0, atol=tol, err_msg=name)
        np.testing.assert_allclose(O.get_plan(M, reg, r["dual"]), plan_ref, atol=10 * tol)
        assert len(r["obj_vals"]) == r["niter"] + 1
    # Newton converges quadratically: few iterations, ||g|| squares at the end
    assert nw["niter"] <= 15
    e = nw["mar_errs"]
    assert e[-2] < 1e-3 and e[-1] < 1e3 * e[-2] ** 2
    # APDAGD: the averaged primal plan converges to T*, with gap and
    # (full) marginal error below tol at the stop
    # (first-order: about 4400 iterations to 1e-6 here)
    ap = O.sinkhorn_apdagd(M, a, b, reg, tol=1e-6, max_iter=20000)
    assert ap["niter"] < 20000
    P = ap["plan"]
    full_err = np.sqrt(((P.sum(1) - a) ** 2).sum() + ((P.sum(0) - b) ** 2).sum())
    assert abs(full_err - ap["mar_errs"][-1]) < 1e-15 and full_err <= 1e-6
    np.testing.assert_allclose(P, plan_ref, rtol=0, atol=1e-4)
    primal = (M * P).sum() - reg * (P * (1 - np.log(P))).sum()
    assert primal - (-ap["obj_vals"][-1]) <= 1e-6 + 1e-12
    # the stop test follows the history push (src/sinkhorn_apdagd.cpp:160-174):
    # a converged run records niter + 2 entries
    assert len(ap["obj_vals"]) == ap["niter"] + 2


def test_separable_cost_newton_and_lbfgs():
    rng = np.random.default_rng(7)
    n, m, reg = 12, 9, 0.3
    u, v = rng.random(n), rng.random(m)
    M = u[:, None] + v[None, :]
    a = rng.random(n) + 0.2
    b = rng.random(m) + 0.2
    a, b = a / a.sum(), b / b.sum()
    c = v[-1] + reg * np.log(b[-1])
    dual = np.concatenate([u + reg * np.log(a) + c, (v + reg * np.log(b) - c)[:-1]])
    r = O.sinkhorn_newton(M, a, b, reg, tol=1e-13, max_iter=50)
    np.testing.assert_allclose(r["dual"], dual, rtol=1e-12, atol=1e-12)
    # L-BFGS: f differences drop below FP64 resolution near ||g|| ~ 5e-9
    # here, where the line search (like LBFGSpp's) gives up with an error
    r = O.sinkhorn_lbfgs_dual(M, a, b, reg, tol=1e-8, max_iter=200)
    assert r["mar_errs"][-1] <= 1e-8
    np.testing.assert_allclose(O.get_plan(M, reg, r["dual"]), np.outer(a, b), rtol=1e-6)
    with pytest.raises(RuntimeError, match="unable to sufficiently decrease"):
        O.sinkhorn_lbfgs_dual(M, a, b, reg, tol=1e-12, max_iter=200)


# --- feasible rounding and primal value (src/approx_proj.h, src/sinkhorn_problem_obj.cpp:300-319) ---

def test_approx_proj_feasibility_and_bound():
    # Altschuler, Weed and Rigollet (2017), Lemma 7: the rounded plan has
    # marginals exactly (a, b), is nonnegative, and moves by at most
    # 2 (|r - a|_1 + |c - b|_1) in l1 from the input.
    rng = np.random.default_rng(41)
    for n, m in ((7, 5), (30, 44), (1, 9)):
        a = rng.random(n) + 0.1
        b = rng.random(m) + 0.1
        a, b = a / a.sum(), b / b.sum()
        T = rng.random((n, m)) * rng.uniform(0.5, 2.0) / (n * m)
        P = O.approx_proj(T, a, b)
        np.testing.assert_allclose(P.sum(1), a, rtol=1e-13, atol=1e-16)
        np.testing.assert_allclose(P.sum(0), b, rtol=1e-13, atol=1e-16)
        assert P.min() >= -1e-17
        bound = 2 * (np.abs(T.sum(1) - a).sum() + np.abs(T.sum(0) - b).sum())
        assert np.abs(P - T).sum() <= bound * (1 + 1e-12)
    # an exactly feasible input leaves |err_a|_1 = 0: the reference's 0/0
    T = np.full((2, 2), 0.25)
    assert np.isnan(O.approx_proj(T, np.array([0.5, 0.5]), np.array([0.5, 0.5]))).all()


def test_primal_value_meets_the_dual_at_the_optimum():
    # strong duality of the entropic problem: at the dual optimum the primal
    # value of the (nearly feasible) plan equals -f*; away from it the primal
    # value of a feasible plan bounds -f from above (weak duality)
    M, a, b = cloud_problem(25, 19, d=3, seed=8)
    reg = 0.1
    r = O.sinkhorn_bcd(M, a, b, reg, tol=1e-11, max_iter=100000)
    pv = O.primal_val(M, a, b, reg, r["dual"])
    assert abs(pv + r["obj_vals"][-1]) < 1e-9
    g = r["dual"] + 0.05 * np.random.default_rng(3).standard_normal(len(r["dual"]))
    f = O.dual_obj(M, a, b, reg, g)
    assert O.primal_val(M, a, b, reg, g) >= -f - 1e-12


# --------------------------------------------------------------------------
# QROT (quadratically regularised OT)
# --------------------------------------------------------------------------
def test_qrot_obj_grad_brute_force_and_fd():
    M, a, b = cloud_problem(6, 5, d=2, seed=31)
    rng = np.random.default_rng(3)
    gamma = 0.3 * rng.standard_normal(11)
    for tau in (0.0, 0.7):
        f, g = O.qrot_dual_obj_grad(M, a, b, 0.2, gamma, tau)
        al, be = gamma[:6], gamma[6:]
        D = np.maximum(al[:, None] + be[None, :] - M, 0.0)
        d = al.sum() - be.sum()
        f_ref = 0.5 * (D * D).sum() - 0.2 * (al @ a + be @ b) + 0.5 * tau * d * d
        g_ref = np.concatenate([D.sum(1) - 0.2 * a + tau * d, D.sum(0) - 0.2 * b - tau * d])
        assert abs(f - f_ref) < 1e-14 and np.max(np.abs(g - g_ref)) < 1e-14
        h = 1e-6
        for k in range(11):
            e = np.zeros(11)
            e[k] = h
            fd = (O.qrot_dual_obj_grad(M, a, b, 0.2, gamma + e, tau)[0]
                  - O.qrot_dual_obj_grad(M, a, b, 0.2, gamma - e, tau)[0]) / (2 * h)
            assert abs(fd - g[k]) < 1e-7


def test_qrot_piecewise_root():
    rng = np.random.default_rng(4)
    for trial in range(30):
        y = rng.standard_normal(rng.integers(1, 40))
        c = float(rng.exponential(0.5 if trial % 3 else 50.0))
        x = O.qrot_pl_root(y, c)
        assert abs(np.maximum(x - y, 0.0).sum() - c) < 1e-12 * max(1.0, c)
        lo, hi = y.min(), y.max() + c          # independent bisection
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if np.maximum(mid - y, 0.0).sum() < c else (lo, mid)
        assert abs(x - 0.5 * (lo + hi)) < 1e-12 * max(1.0, abs(x))
    # every point active: x = (c + sum y) / len
    y = np.array([0.0, 0.1, 0.2])
    assert abs(O.qrot_pl_root(y, 10.0) - (10.0 + 0.3) / 3) < 1e-15


def test_qrot_optimal_alpha_beta_zero_their_gradient_blocks():
    M, a, b = cloud_problem(12, 9, d=3, seed=32)
    reg = 0.05
    alpha = 0.1 * np.random.default_rng(5).standard_normal(12)
    beta = O.qrot_optimal_beta(M, b, reg, alpha)
    _, g = O.qrot_dual_obj_grad(M, a, b, reg, np.concatenate([alpha, beta]))
    assert np.max(np.abs(g[12:])) < 1e-15
    alpha2 = O.qrot_optimal_alpha(M, a, reg, beta)
    _, g = O.qrot_dual_obj_grad(M, a, b, reg, np.concatenate([alpha2, beta]))
    assert np.max(np.abs(g[:12])) < 1e-15


def test_qrot_drivers_reach_the_qp_optimum():
    # The primal min <P, M> + reg/2 ||P||^2 over the transport polytope,
    # solved independently by scipy's SLSQP on a 5 x 4 problem.
    from scipy.optimize import minimize
    M, a, b = cloud_problem(5, 4, d=2, seed=33)
    reg = 0.1
    n, m = M.shape
    cons = [{"type": "eq", "fun": lambda p: p.reshape(n, m).sum(1) - a},
            {"type": "eq", "fun": lambda p: p.reshape(n, m).sum(0)[:-1] - b[:-1]}]
    qp = minimize(lambda p: p @ M.ravel() + 0.5 * reg * p @ p, np.outer(a, b).ravel(),
                  jac=lambda p: M.ravel() + reg * p, bounds=[(0, None)] * (n * m),
                  constraints=cons, method="SLSQP", options={"ftol": 1e-15, "maxiter": 1000})
    P_qp = qp.x.reshape(n, m)
    # (L-BFGS at tol 1e-10 ends in LBFGS++'s line-search error: f no longer
    # decreases at FP64 resolution)
    for solve, tol in ((O.qrot_bcd, 1e-10), (O.qrot_gd, 1e-10), (O.qrot_lbfgs_dual, 1e-9),
                       (O.qrot_lbfgs_semi_dual, 1e-9)):
        r = solve(M, a, b, reg, tol=tol, max_iter=100000)
        assert np.max(np.abs(r["plan"] - P_qp)) < 1e-6, solve.__name__
        # strong duality: primal value = -f / reg at the optimum
        assert abs(r["prim_vals"][-1] + r["obj_vals"][-1] / reg) < 1e-8
        assert abs(r["prim_vals"][-1] - qp.fun) < 1e-8


def test_qrot_driver_histories():
    M, a, b = cloud_problem(20, 15, d=3, seed=34)
    reg, tol = 0.05, 1e-7
    for solve in (O.qrot_bcd, O.qrot_gd):
        r = solve(M, a, b, reg, tol=tol, max_iter=20000)
        assert len(r["obj_vals"]) == len(r["prim_vals"]) == r["niter"] + 1
        g = np.array(r["mar_errs"]) * reg          # mar_errs = ||grad|| / reg
        assert g[-1] < tol and np.all(g[:-1] >= tol)
        _, grad = O.qrot_dual_obj_grad(M, a, b, reg, r["dual"])
        assert abs(np.linalg.norm(grad) - g[-1]) < 1e-14
        # weak duality along the way: primal value of the rounded plan >= -f / reg
        assert np.all(np.array(r["prim_vals"]) >= -np.array(r["obj_vals"]) / reg - 1e-12)
    assert O.qrot_bcd(M, a, b, reg, tol=1e9)["niter"] == 0
    r = O.qrot_gd(M, a, b, reg, tol=0.0, max_iter=7)
    assert r["niter"] == 7 and len(r["obj_vals"]) == 8
    # BCD's objective decreases monotonically (exact block minimisation)
    r = O.qrot_bcd(M, a, b, reg, tol=tol, max_iter=20000)
    assert np.all(np.diff(r["obj_vals"]) <= 1e-15)


def test_qrot_semi_dual_gradient_is_the_envelope():
    # f(alpha) = dual(alpha, beta*(alpha)); its gradient is D 1 - reg a (the
    # beta block vanishes at beta*): central differences of the semi-dual
    # objective against the gradient L-BFGS sees at its first iterate
    M, a, b = cloud_problem(9, 7, d=2, seed=35)
    reg = 0.1
    rng = np.random.default_rng(6)
    alpha = 0.05 * rng.standard_normal(9)

    def semi(al):
        be = O.qrot_optimal_beta(M, b, reg, al)
        return O.qrot_dual_obj_grad(M, a, b, reg, np.concatenate([al, be]))

    f, g = semi(alpha)
    h = 1e-7
    for k in range(9):
        e = np.zeros(9)
        e[k] = h
        fd = (semi(alpha + e)[0] - semi(alpha - e)[0]) / (2 * h)
        assert abs(fd - g[k]) < 1e-6
    r = O.qrot_lbfgs_semi_dual(M, a, b, reg, tol=1e-9, max_iter=5000)
    # the result's beta is optimal for its alpha, and the histories line up
    assert np.max(np.abs(r["dual"][9:] - O.qrot_optimal_beta(M, b, reg, r["dual"][:9]))) == 0.0
    assert len(r["obj_vals"]) == len(r["prim_vals"]) == len(r["mar_errs"])


def test_qrot_apdagd_reaches_the_qp_optimum():
    # the averaged primal plan of APDAGD against scipy's SLSQP solution of the
    # primal QP; the stop rule (gap and full marginal error <= tol) holds
    from scipy.optimize import minimize
    M, a, b = cloud_problem(5, 4, d=2, seed=33)
    reg, tol = 0.1, 1e-8
    n, m = M.shape
    cons = [{"type": "eq", "fun": lambda p: p.reshape(n, m).sum(1) - a},
            {"type": "eq", "fun": lambda p: p.reshape(n, m).sum(0)[:-1] - b[:-1]}]
    qp = minimize(lambda p: p @ M.ravel() + 0.5 * reg * p @ p, np.outer(a, b).ravel(),
                  jac=lambda p: M.ravel() + reg * p, bounds=[(0, None)] * (n * m),
                  constraints=cons, method="SLSQP", options={"ftol": 1e-15, "maxiter": 1000})
    r = O.qrot_apdagd(M, a, b, reg, tol=tol, max_iter=2000000)
    P = r["plan"]
    assert len(r["obj_vals"]) == r["niter"] + 2          # initial entry + every iteration
    full = np.sqrt(((P.sum(1) - a) ** 2).sum() + ((P.sum(0) - b) ** 2).sum())
    assert full <= tol and abs(full - r["mar_errs"][-1]) < 1e-15
    assert np.max(np.abs(P - qp.x.reshape(n, m))) < 1e-5
    primal = (P * M).sum() + 0.5 * reg * (P * P).sum()
    assert abs(primal - qp.fun) < 1e-7


def _qrot_dense_hess(M, gamma, shift, tau):
    n, m = M.shape
    sig = (gamma[:n, None] + gamma[None, n:] - M > 0).astype(float)
    H = np.zeros((n + m, n + m))
    H[:n, :n] = np.diag(sig.sum(1))
    H[n:, n:] = np.diag(sig.sum(0))
    H[:n, n:] = sig
    H[n:, :n] = sig.T
    v = np.concatenate([np.ones(n), -np.ones(m)])
    return H + shift * np.eye(n + m) + tau * np.outer(v, v)


def test_qrot_hessian_is_the_gradient_jacobian_and_solves():
    # Away from the kinks the QROT dual is piecewise quadratic: central
    # differences of the gradient give H + tau vv' exactly (to rounding).
    # The block-elimination solve (Schur complement CG) against numpy's dense
    # solve, with and without tau, and the CG step count against the bound m.
    M, a, b = cloud_problem(8, 6, d=2, seed=41)
    reg = 0.1
    rng = np.random.default_rng(8)
    alpha = 0.02 * rng.standard_normal(8)
    gamma = np.concatenate([alpha, O.qrot_optimal_beta(M, b, reg, alpha)])
    gamma += 0.01 * rng.standard_normal(14)
    for tau in (0.0, 0.3):
        J = np.zeros((14, 14))
        h = 1e-7
        for k in range(14):
            e = np.zeros(14)
            e[k] = h
            J[:, k] = (O.qrot_dual_obj_grad(M, a, b, reg, gamma + e, tau)[1] -
                       O.qrot_dual_obj_grad(M, a, b, reg, gamma - e, tau)[1]) / (2 * h)
        Hd = _qrot_dense_hess(M, gamma, 0.0, tau)
        assert np.max(np.abs(J - Hd)) < 1e-6
        nact = np.sum(_qrot_dense_hess(M, gamma, 0.0, 0.0)[:8, 8:])
        assert 0 < nact < 48                         # some, not all, entries active
        rhs = rng.standard_normal(14)
        for shift in (1e-3, 0.5):
            x, it = O.qrot_hess_solve(M, gamma, rhs, shift, tau, tol=1e-14, max_iter=500)
            ref = np.linalg.solve(_qrot_dense_hess(M, gamma, shift, tau), rhs)
            assert np.max(np.abs(x - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))
            assert 0 < it <= 6 + 1                   # CG on a 6 x 6 SPD system
    # zero right-hand side: no CG step, zero solution
    x, it = O.qrot_hess_solve(M, gamma, np.zeros(14), 1e-3)
    assert it == 0 and not np.any(x)


def test_qrot_grssn_reaches_the_qp_optimum():
    from scipy.optimize import minimize
    M, a, b = cloud_problem(5, 4, d=2, seed=33)
    reg = 0.1
    n, m = M.shape
    cons = [{"type": "eq", "fun": lambda p: p.reshape(n, m).sum(1) - a},
            {"type": "eq", "fun": lambda p: p.reshape(n, m).sum(0)[:-1] - b[:-1]}]
    qp = minimize(lambda p: p @ M.ravel() + 0.5 * reg * p @ p, np.outer(a, b).ravel(),
                  jac=lambda p: M.ravel() + reg * p, bounds=[(0, None)] * (n * m),
                  constraints=cons, method="SLSQP", options={"ftol": 1e-15, "maxiter": 1000})
    for method in (0, 1):
        r = O.qrot_grssn(M, a, b, reg, tol=1e-11, max_iter=500, method=method)
        assert r["niter"] < 100 and len(r["obj_vals"]) == r["niter"] + 1
        assert np.max(np.abs(r["plan"] - qp.x.reshape(n, m))) < 1e-6
        assert abs(r["prim_vals"][-1] - qp.fun) < 1e-8
        # Armijo steps along a descent direction: f never increases
        assert np.all(np.diff(r["obj_vals"]) <= 1e-15)
        assert r["mar_errs"][-1] * reg < 1e-11 and len(r["cg_iters"]) == r["niter"]


def test_qrot_grssn_history_rules():
    M, a, b = cloud_problem(30, 20, d=3, seed=42)
    reg = 0.05
    r = O.qrot_grssn(M, a, b, reg, tol=1e-9, max_iter=200)
    g = np.array(r["mar_errs"]) * reg
    assert g[-1] < 1e-9 and np.all(g[:-1] >= 1e-9)
    _, grad = O.qrot_dual_obj_grad(M, a, b, reg, r["dual"])
    assert abs(np.linalg.norm(grad) - g[-1]) < 1e-14
    # the semismooth Newton method ends superlinearly: far fewer steps than GD
    assert r["niter"] < 60
    assert O.qrot_grssn(M, a, b, reg, tol=1e9)["niter"] == 0
    r = O.qrot_grssn(M, a, b, reg, tol=0.0, max_iter=3)
    assert r["niter"] == 3 and len(r["obj_vals"]) == 4


def test_qrot_assn_reaches_the_qp_optimum_and_history_rules():
    from scipy.optimize import minimize
    M, a, b = cloud_problem(5, 4, d=2, seed=33)
    reg = 0.1
    n, m = M.shape
    cons = [{"type": "eq", "fun": lambda p: p.reshape(n, m).sum(1) - a},
            {"type": "eq", "fun": lambda p: p.reshape(n, m).sum(0)[:-1] - b[:-1]}]
    qp = minimize(lambda p: p @ M.ravel() + 0.5 * reg * p @ p, np.outer(a, b).ravel(),
                  jac=lambda p: M.ravel() + reg * p, bounds=[(0, None)] * (n * m),
                  constraints=cons, method="SLSQP", options={"ftol": 1e-15, "maxiter": 1000})
    for method in (0, 1):
        r = O.qrot_assn(M, a, b, reg, tol=1e-11, max_iter=5000, method=method)
        assert r["niter"] < 5000 and len(r["obj_vals"]) == r["niter"] + 1
        assert np.max(np.abs(r["plan"] - qp.x.reshape(n, m))) < 1e-6
        assert abs(r["prim_vals"][-1] - qp.fun) < 1e-8
    # the same optimum as GRSSN on a larger problem (plans: the dual is
    # unique only up to alpha + c, beta - c when sum a = sum b)
    M, a, b = cloud_problem(30, 20, d=3, seed=42)
    r = O.qrot_assn(M, a, b, 0.1, tol=1e-9, max_iter=2000)
    g = np.array(r["mar_errs"]) * 0.1
    assert g[-1] < 1e-9 and np.all(g[:-1] >= 1e-9)
    q = O.qrot_grssn(M, a, b, 0.1, tol=1e-9, max_iter=2000)
    assert np.max(np.abs(r["plan"] - q["plan"])) < 1e-7
    assert O.qrot_assn(M, a, b, 0.1, tol=1e9)["niter"] == 0


def test_qrot_s5n_reaches_the_qp_optimum_and_history_rules():
    from scipy.optimize import minimize
    M, a, b = cloud_problem(5, 4, d=2, seed=33)
    reg = 0.1
    n, m = M.shape
    cons = [{"type": "eq", "fun": lambda p: p.reshape(n, m).sum(1) - a},
            {"type": "eq", "fun": lambda p: p.reshape(n, m).sum(0)[:-1] - b[:-1]}]
    qp = minimize(lambda p: p @ M.ravel() + 0.5 * reg * p @ p, np.outer(a, b).ravel(),
                  jac=lambda p: M.ravel() + reg * p, bounds=[(0, None)] * (n * m),
                  constraints=cons, method="SLSQP", options={"ftol": 1e-15, "maxiter": 1000})
    for method in (0, 1):
        r = O.qrot_s5n(M, a, b, reg, tol=1e-11, max_iter=500, method=method)
        assert r["niter"] < 100 and len(r["obj_vals"]) == r["niter"] + 1
        assert np.max(np.abs(r["plan"] - qp.x.reshape(n, m))) < 1e-6
        assert abs(r["prim_vals"][-1] - qp.fun) < 1e-8
    # steps are taken only when the model ratio is positive: f never
    # increases; the primal value of the rounded plan never drops below -f /
    # reg (the gap that sets the shift stays >= 0); same optimum as GRSSN
    M, a, b = cloud_problem(60, 45, d=4, seed=7)
    r = O.qrot_s5n(M, a, b, 0.05, tol=1e-9, max_iter=500)
    assert np.all(np.diff(r["obj_vals"]) <= 1e-15)
    assert np.all(np.array(r["prim_vals"]) >= -np.array(r["obj_vals"]) / 0.05 - 1e-12)
    g = np.array(r["mar_errs"]) * 0.05
    assert g[-1] < 1e-9 and np.all(g[:-1] >= 1e-9)
    q = O.qrot_grssn(M, a, b, 0.05, tol=1e-9, max_iter=500)
    assert np.max(np.abs(r["plan"] - q["plan"])) < 1e-6


def test_qrot_pdaam_reaches_the_qp_optimum():
    # the averaged primal plan of PDAAM against scipy's SLSQP solution of the
    # primal QP; the stop rule (gap and full marginal error <= tol) holds; the
    # same plan as APDAGD on a larger problem
    from scipy.optimize import minimize
    M, a, b = cloud_problem(5, 4, d=2, seed=33)
    reg, tol = 0.1, 1e-8
    n, m = M.shape
    cons = [{"type": "eq", "fun": lambda p: p.reshape(n, m).sum(1) - a},
            {"type": "eq", "fun": lambda p: p.reshape(n, m).sum(0)[:-1] - b[:-1]}]
    qp = minimize(lambda p: p @ M.ravel() + 0.5 * reg * p @ p, np.outer(a, b).ravel(),
                  jac=lambda p: M.ravel() + reg * p, bounds=[(0, None)] * (n * m),
                  constraints=cons, method="SLSQP", options={"ftol": 1e-15, "maxiter": 1000})
    r = O.qrot_pdaam(M, a, b, reg, tol=tol, max_iter=2000000)
    P = r["plan"]
    assert len(r["obj_vals"]) == r["niter"] + 2
    full = np.sqrt(((P.sum(1) - a) ** 2).sum() + ((P.sum(0) - b) ** 2).sum())
    assert full <= tol and abs(full - r["mar_errs"][-1]) < 1e-15
    assert np.max(np.abs(P - qp.x.reshape(n, m))) < 1e-5
    primal = (P * M).sum() + 0.5 * reg * (P * P).sum()
    assert abs(primal - qp.fun) < 1e-7
    M, a, b = cloud_problem(40, 30, d=4, seed=3)
    r = O.qrot_pdaam(M, a, b, 0.1, tol=1e-5, max_iter=200000)
    q = O.qrot_apdagd(M, a, b, 0.1, tol=1e-5, max_iter=200000)
    assert np.max(np.abs(r["plan"] - q["plan"])) < 1e-4 * np.max(q["plan"])
