"""Pins of the oracle's line searches and Newton-type driver logic by what the
mathematics fixes, not by restating the code.

* ``line_selection`` (src/sinkhorn_problem_ls.cpp:13-41) and
  ``line_search_wolfe`` (:226-470) run on analytic one-dimensional objectives
  whose minimisers, Wolfe points and exits are known in closed form, and on
  the EROT dual with an independent numpy objective.
* The SSNS, sparse-Newton and SPLR drivers (src/sinkhorn_ssns.cpp:16-234,
  src/sinkhorn_sparse_newton.cpp:24-170, src/sinkhorn_splr.cpp:24-199) record
  every iteration (the oracle's trace); each record is recomputed here from
  its own gamma with numpy: the gradient from T, the sparsified Hessian by a
  brute-force kept set, the direction by a dense solve (the inverse-BFGS
  matrix for SPLR's low-rank step), the line-selection choice and rho from
  objective values, the strong Wolfe conditions at the accepted step, and
  the point whose T the next iteration uses.
"""
import numpy as np
import pytest

from conftest import cloud_problem
from oracle import oracle as O
from test_oracle import brute_threshold

C1, C2 = 1e-4, 0.9
STEP_MIN, STEP_MAX = 1e-10, 2.0


# --- numpy EROT dual, independent of the oracle -----------------------------

def T_at(M, reg, gamma):
    n, m = M.shape
    beta = np.r_[gamma[n:], 0.0]
    return np.exp((gamma[:n, None] + beta[None, :] - M) / reg)


def f_mixed(T, gamma, a, b, reg):
    """reg sum(T) - alpha'a - beta~'b~ with T formed elsewhere (the reference's
    dual_obj_grad(gamma, g, T, false))."""
    n = a.shape[0]
    return reg * T.sum() - gamma[:n] @ a - gamma[n:] @ b[:-1]


def grad_of(T, a, b):
    return np.r_[T.sum(1) - a, T.sum(0)[:-1] - b[:-1]]


def f_at(M, a, b, reg, gamma):
    return f_mixed(T_at(M, reg, gamma), gamma, a, b, reg)


def hess_dense(T, keep, reg):
    n, m = T.shape
    N = n + m - 1
    H = np.zeros((N, N))
    H[np.arange(n), np.arange(n)] = T.sum(1) / reg
    H[n + np.arange(m - 1), n + np.arange(m - 1)] = T.sum(0)[:-1] / reg
    S = np.where(keep, T[:, :-1], 0.0) / reg
    H[:n, n:] = S
    H[n:, :n] = S.T
    return H


def topk_keep(T, density):
    n, m = T.shape
    body = T[:, :-1]
    k = int(n * (m - 1) * min(max(density, 0.0), 1.0))
    thr = np.sort(body.ravel())[::-1][k - 1] if k > 0 else np.inf
    keep = body >= thr
    keep[0, :] = True
    keep[:, 0] = True
    return keep


def select_rule(fvals, f0):
    """First candidate lowering f0, else the first of the smallest values;
    also the index of the last candidate evaluated."""
    best, fbest = 0, np.inf
    for k, fv in enumerate(fvals):
        if fv < fbest:
            best, fbest = k, fv
        if fbest < f0:
            return best, k
    return best, len(fvals) - 1


# --- line_selection on analytic objectives ----------------------------------

def test_line_selection_first_decrease_else_best():
    cand = (1.0, 0.5, 0.25, 0.1)
    x0, d = np.zeros(1), np.ones(1)
    # f(t) = (t - 0.4)^2 - 0.1: f(1) = 0.26 > f(0) = 0.06, f(0.5) = -0.09 < f(0)
    calls = []

    def f1(x):
        calls.append(x[0])
        return (x[0] - 0.4) ** 2 - 0.1

    st, fv, last = O.line_selection(f1, x0, d, f1(x0))
    calls.pop(0)
    assert st == 0.5 and last == 0.5 and calls == [1.0, 0.5]
    assert fv == pytest.approx(-0.09)
    # ascent direction: nothing lowers f, every candidate is tried, the best
    # (smallest step for a convex f) is returned and is the last one tried
    st, fv, last = O.line_selection(lambda x: x[0] ** 2 + x[0], x0, d, 0.0)
    assert st == 0.1 and last == 0.1 and fv == pytest.approx(0.11)
    # non-convex values, none below f0 = 0: the first strict minimum wins,
    # while the last candidate tried is still 0.1
    vals = {1.0: 0.3, 0.5: 0.2, 0.25: 0.2, 0.1: 0.5}
    st, fv, last = O.line_selection(lambda x: vals[round(x[0], 6)], x0, d, 0.0)
    assert st == 0.5 and fv == 0.2 and last == 0.1
    for f0, want in ((0.25, 0.5), (0.35, 1.0), (0.0, 0.5)):
        st, _, _ = O.line_selection(lambda x: vals[round(x[0], 6)], x0, d, f0)
        assert st == want


# --- line_search_wolfe on analytic objectives --------------------------------

def wolfe_1d(phi, dphi, init=1.0, max_iter=20):
    trials = []

    def fg(x):
        t = x[0]
        trials.append(t)
        return phi(t), np.array([dphi(t)])

    out = O.line_search_wolfe(fg, np.zeros(1), np.ones(1), phi(0.0), np.array([dphi(0.0)]),
                              init_step=init, max_iter=max_iter)
    return out, trials


def strong_wolfe(phi, dphi, t):
    return (phi(t) <= phi(0.0) + C1 * t * dphi(0.0)
            and abs(dphi(t)) <= C2 * abs(dphi(0.0)))


@pytest.mark.parametrize("tstar,init", [(0.3, 1.0), (1.7, 1.0), (0.05, 1.0), (0.9, 0.2)])
def test_wolfe_quadratic_returns_a_strong_wolfe_step(tstar, init):
    phi = lambda t: 0.5 * (t - tstar) ** 2
    dphi = lambda t: t - tstar
    out, trials = wolfe_1d(phi, dphi, init)
    assert not out["recompute_T"]
    assert out["step"] == trials[-1] == out["last_step"] and out["trials"] == len(trials)
    assert strong_wolfe(phi, dphi, out["step"])
    assert len(trials) <= 20


def test_wolfe_exits_at_step_max_on_a_linear_descent():
    # phi(t) = -t: sufficient decrease everywhere, curvature never met; the
    # search extrapolates (Case 2) until step_max and stops there
    out, trials = wolfe_1d(lambda t: -t, lambda t: -1.0, init=0.5)
    assert out["step"] == STEP_MAX == trials[-1] and not out["recompute_T"]
    assert all(t2 > t1 for t1, t2 in zip(trials, trials[1:]))


def test_wolfe_exits_at_step_min_when_every_step_is_too_long():
    # minimiser 5e-13 < step_min: psi > 0 at every admissible step, the
    # step-min safeguard drives the trials down to step_min, which is returned
    phi = lambda t: -t + 1e12 * t * t
    dphi = lambda t: -1.0 + 2e12 * t
    out, trials = wolfe_1d(phi, dphi)
    assert out["step"] == STEP_MIN == trials[-1] and not out["recompute_T"]
    assert all(t2 < t1 for t1, t2 in zip(trials, trials[1:]))


def test_wolfe_refuses_an_ascent_direction():
    out, trials = wolfe_1d(lambda t: t * t + t, lambda t: 2 * t + 1.0, init=0.7)
    assert out["recompute_T"] and out["step"] == 0.7 and trials == [] and out["trials"] == 0
    out, _ = wolfe_1d(lambda t: t * t + t, lambda t: 2 * t + 1.0, init=5.0)
    assert out["step"] == STEP_MAX
    out, _ = wolfe_1d(lambda t: t * t + t, lambda t: 2 * t + 1.0, init=0.0)
    assert out["step"] == STEP_MIN


def psi(phi, dphi, t):
    return phi(t) - phi(0.0) - t * C1 * dphi(0.0)


@pytest.mark.parametrize("max_iter", [1, 2, 3, 20])
def test_wolfe_exhausted_trials_return_the_best_point(max_iter):
    # |phi'| = 1 everywhere but at the kink, so the curvature test never
    # holds; with the trials used up the search returns either the step of its
    # last update (kept only when psi formed from the last trial's f at that
    # step is no worse than the best psi seen, recompute_T false) or the trial
    # with the smallest psi (recompute_T true)
    phi = lambda t: abs(t - 0.37) - 0.37
    dphi = lambda t: 1.0 if t > 0.37 else -1.0
    out, trials = wolfe_1d(phi, dphi, init=1.0, max_iter=max_iter)
    assert out["trials"] == len(trials) == max_iter
    psis = [psi(phi, dphi, t) for t in trials]
    best = min(range(len(trials)), key=lambda k: (psis[k], -k))
    if out["recompute_T"]:
        lo = trials[best] if psis[best] <= 0.0 else 0.0
        assert out["step"] == (lo if lo > 0 else 1.0)
    else:
        last_f = phi(trials[-1])
        p_new = last_f - phi(0.0) - out["step"] * C1 * dphi(0.0)
        assert p_new <= min(0.0, min(psis)) + 1e-15
        assert out["last_step"] == trials[-1]


def test_wolfe_on_the_erot_dual():
    M, a, b = cloud_problem(30, 22, d=4, seed=11)
    reg = 0.05
    rng = np.random.default_rng(4)
    n, m = M.shape
    gamma = np.r_[0.02 * rng.standard_normal(n), 0.02 * rng.standard_normal(m - 1)]
    f0, g0 = O.dual_obj_grad(M, a, b, reg, gamma)
    # steepest descent and a Newton direction (dense Hessian, numpy solve)
    T = T_at(M, reg, gamma)
    H = hess_dense(T, np.ones((n, m - 1), bool), reg)
    for d in (-g0, -np.linalg.solve(H, g0), -30.0 * g0):
        out = O.erot_line_search_wolfe(M, a, b, reg, gamma, d)
        t = out["step"]
        assert not out["recompute_T"] and out["last_step"] == t
        fx = f_at(M, a, b, reg, gamma + t * d)
        gx = grad_of(T_at(M, reg, gamma + t * d), a, b)
        assert fx <= f0 + C1 * t * (g0 @ d) + 1e-15 * abs(f0)
        assert abs(gx @ d) <= C2 * abs(g0 @ d)
        np.testing.assert_allclose(out["grad"], gx, rtol=1e-12, atol=1e-15)


# --- driver traces ------------------------------------------------------------

def small_problem(seed=5, n=28, m=21, reg=0.05):
    M, a, b = cloud_problem(n, m, d=4, seed=seed)
    return M, a, b, reg


def check_ssns_trace(M, a, b, reg, res, recs, mu0, dir_rtol):
    n, m = M.shape
    assert len(recs) == res["niter"]
    mu, hint = mu0, 0.001
    for k, r in enumerate(recs):
        h = r["hdr"]
        gamma, tp, d = r["gamma"], r["tpoint"], r["direc"]
        T = T_at(M, reg, tp)
        g = grad_of(T, a, b)
        f = f_mixed(T, gamma, a, b, reg)
        assert h[0] == 1 and h[1] == k
        assert h[2] == pytest.approx(f, rel=1e-12) and h[2] == res["obj_vals"][k]
        assert h[3] == pytest.approx(np.linalg.norm(g), rel=1e-10)
        assert h[5] == mu and h[4] == pytest.approx(mu * np.linalg.norm(g), rel=1e-10)
        delta = 0.01 * np.linalg.norm(g)
        assert h[6] == pytest.approx(delta, rel=1e-10)
        keep = brute_threshold(T, h[6], hint)
        Hs = hess_dense(T, keep, reg)
        N = n + m - 1
        dd = -np.linalg.solve(Hs + h[4] * np.eye(N), g)
        np.testing.assert_allclose(d, dd, rtol=dir_rtol, atol=dir_rtol * np.abs(dd).max())
        cand = (1.0, 0.5, 0.25, 0.1)
        fv = [f_at(M, a, b, reg, gamma + c * d) for c in cand]
        best, last = select_rule(fv, h[2])
        assert h[7] == cand[best] and h[12] == cand[last]
        assert h[8] == pytest.approx(fv[best], rel=1e-12)
        s = cand[best] * d
        numer = h[2] - fv[best]
        denom = -g @ s - 0.5 * s @ (Hs @ s)
        assert h[11] == pytest.approx(denom, rel=1e-8)
        # f - f_new cancels near the optimum: compare on the scale of f
        fres = 1e-13 * abs(h[2])
        assert abs(h[10] - numer) <= fres
        rho = h[9]
        assert abs(rho - numer / denom) <= fres / abs(denom) + 1e-7 * abs(rho)
        mu = 4.0 * mu if rho < 0.25 else (max(0.5 * mu, 1e-3) if rho > 0.75 else mu)
        assert h[14] == mu
        if k + 1 < len(recs):
            nxt = recs[k + 1]
            if rho > 0.0:
                np.testing.assert_allclose(nxt["gamma"], gamma + s, rtol=1e-15, atol=1e-16)
                # the next T is the last candidate's
                np.testing.assert_allclose(nxt["tpoint"], gamma + cand[last] * d,
                                           rtol=1e-14, atol=1e-15)
                hint = h[13]
            else:
                assert np.array_equal(nxt["gamma"], gamma)
                assert np.array_equal(nxt["tpoint"], tp)


def test_ssns_trace_recomputed_in_numpy():
    M, a, b, reg = small_problem()
    res, recs = O.traced("sinkhorn_ssns", M, a, b, reg, tol=1e-9, max_iter=200, method=1)
    assert res["mar_errs"][-1] < 1e-9 and len(recs) >= 3
    check_ssns_trace(M, a, b, reg, res, recs, 1.0, 1e-9)


def test_ssns_accepts_a_step_that_raises_f():
    # mu0 < 0 makes Hs + mu ||g|| I negative definite, so the first direction
    # ascends: no candidate lowers f, yet both halves of rho are negative and
    # the step is accepted (rho > 0).  For a convex dual the smallest
    # candidate is then the best and the last evaluated, so the T reused
    # afterwards is the accepted point's.  Plain CG (method 0) solves the
    # indefinite system; the reference's default LDLT would too.
    M, a, b, reg = small_problem(seed=6)
    res, recs = O.traced("sinkhorn_ssns", M, a, b, reg, tol=1e-8, max_iter=200, method=0,
                         mu0=-200.0)
    h = recs[0]["hdr"]
    assert h[8] > h[2] and h[9] > 0.0 and h[7] == 0.1 == h[12]
    assert h[14] == 1e-3
    assert res["mar_errs"][-1] < 1e-8
    check_ssns_trace(M, a, b, reg, res, recs, -200.0, 1e-5)


def check_newton_trace(M, a, b, reg, res, recs, splr, density_max, shift_max, dir_rtol):
    n, m = M.shape
    N = n + m - 1
    assert len(recs) == res["niter"]
    gn0 = None
    density = 0.1 * density_max
    g_prev, gamma_prev = np.zeros(N), np.zeros(N)
    for k, r in enumerate(recs):
        h = r["hdr"]
        gamma, tp, d = r["gamma"], r["tpoint"], r["direc"]
        T = T_at(M, reg, tp)
        g = grad_of(T, a, b)
        gn = np.linalg.norm(g)
        gn0 = gn if k == 0 else gn0
        assert h[0] == (3 if splr else 2) and h[1] == k
        assert h[2] == pytest.approx(f_mixed(T, gamma, a, b, reg), rel=1e-12)
        assert h[3] == pytest.approx(gn, rel=1e-10)
        assert h[4] == pytest.approx(min(gn, shift_max), rel=1e-10)
        assert h[5] == pytest.approx(density, rel=1e-12)
        A = hess_dense(T, topk_keep(T, h[5]), reg) + h[4] * np.eye(N)
        y, s = g - g_prev, gamma - gamma_prev
        low_rank = splr and k > 0 and (y @ s) > 1e-6 * (y @ y)
        assert bool(h[9]) == low_rank
        if low_rank:
            np.testing.assert_allclose(r["y"], y, rtol=1e-9, atol=1e-14 * max(a.max(), b.max()))
            np.testing.assert_allclose(r["s"], s, rtol=1e-12, atol=1e-15)
            rho = 1.0 / (y @ s)
            Ai = np.linalg.inv(A)
            E = np.eye(N) - rho * np.outer(s, y)
            Hp = E @ Ai @ E.T + rho * np.outer(s, s)
            dd = -Hp @ g
        else:
            Hp = np.linalg.inv(A)
            dd = -np.linalg.solve(A, g)
        # g carries an absolute rounding error of ~1e-16 of the marginals,
        # which the solve amplifies by ||A^-1|| (it dominates once ||g|| is
        # at the noise level)
        gerr = 1e-14 * max(a.max(), b.max()) * np.sqrt(N)
        assert np.linalg.norm(d - dd) <= (dir_rtol * np.linalg.norm(dd)
                                          + np.linalg.norm(Hp, 2) * gerr)
        t, rec, last = h[7], bool(h[8]), h[11]
        dg0 = g @ d
        fx = f_at(M, a, b, reg, gamma + t * d)
        gx = grad_of(T_at(M, reg, gamma + t * d), a, b)
        if not rec and last == t:
            psit = fx - h[2] - t * C1 * dg0
            converged = psit <= 0.0 and abs(gx @ d) <= C2 * abs(dg0)
            at_min = t <= STEP_MIN and (psit > 0.0 or gx @ d - C1 * dg0 >= 0.0)
            at_max = t >= STEP_MAX and psit <= 0.0 and gx @ d - C1 * dg0 < 0.0
            assert converged or at_min or at_max or h[10] == 0 or abs(psit) < 1e-13 * abs(h[2])
        g_prev, gamma_prev = g, gamma
        if k + 1 < len(recs):
            nxt = recs[k + 1]
            np.testing.assert_allclose(nxt["gamma"], gamma + t * d, rtol=1e-15, atol=1e-16)
            # T (and so f, g and H) of the next iteration: the returned step
            # when it was not the last trial evaluated and recompute_T is set,
            # else the last trial, which is the returned step except when the
            # trials ran out
            want_tp = nxt["gamma"] if rec else gamma + last * d
            np.testing.assert_allclose(nxt["tpoint"], want_tp, rtol=1e-14, atol=1e-15)
            gn_next = nxt["hdr"][3]
            if splr:
                bad = gn < gn0 and gn_next > gn
                density *= 1.1 if bad else 0.99
            else:
                bad = gn < gn0 and gn_next > 1.1 * gn
                density *= 2.0 if bad else 0.9
            density = min(density_max, max(0.01 * density_max, density))


@pytest.mark.parametrize("splr", [False, True])
@pytest.mark.parametrize("density", [1.0, 0.3])
@pytest.mark.parametrize("tol,max_iter", [(1e-8, 300), (1e-11, 40)])
def test_newton_family_trace_recomputed_in_numpy(splr, density, tol, max_iter):
    # tol 1e-11 lies below what FP64 objective differences resolve on this
    # problem (||g|| stalls near 4e-9): the searches then run on rounding
    # noise, exhaust their trials and take the fallback exits, which the same
    # checks cover
    M, a, b, reg = small_problem(seed=7)
    name = "sinkhorn_splr" if splr else "sinkhorn_sparse_newton"
    res, recs = O.traced(name, M, a, b, reg, tol=tol, max_iter=max_iter, method=1,
                         density=density, shift=1e-3)
    assert len(recs) >= 3
    if tol >= 1e-8:
        assert res["mar_errs"][-1] < tol
    if splr:
        assert any(r["hdr"][9] for r in recs)        # the low-rank step ran
    check_newton_trace(M, a, b, reg, res, recs, splr, density, 1e-3, 1e-8)


def test_sparse_newton_first_direction_with_every_entry_kept():
    # density 1 keeps the whole Hessian (top-k with k = n(m-1) from the
    # first iteration on, since 0.1 * density_max is clamped to... the
    # schedule starts at 0.1, so use density_max = 10: 0.1 * 10 = 1)
    M, a, b, reg = small_problem(seed=8)
    n, m = M.shape
    res, recs = O.traced("sinkhorn_sparse_newton", M, a, b, reg, tol=1e-9, max_iter=5,
                         method=1, density=10.0, shift=1e-3)
    r = recs[0]
    T = T_at(M, reg, r["gamma"])
    g = grad_of(T, a, b)
    H = hess_dense(T, np.ones((n, m - 1), bool), reg)
    shift = min(np.linalg.norm(g), 1e-3)
    d = -np.linalg.solve(H + shift * np.eye(n + m - 1), g)
    np.testing.assert_allclose(r["direc"], d, rtol=1e-9, atol=1e-9 * np.abs(d).max())
    assert r["hdr"][5] == 1.0


@pytest.mark.parametrize("splr,ls_max_iter,want", [(False, 1, "recompute"), (False, 2, "kept"),
                                                   (True, 2, "kept")])
def test_newton_family_exhausted_wolfe_trials(splr, ls_max_iter, want):
    # With one or two Wolfe trials the search runs out often, which drives the
    # exits of the exhausted branch: I_lo with recompute_T ("recompute"), or
    # the unevaluated step of the last update with the last trial's T
    # ("kept").  Every record is recomputed as above.
    M, a, b, reg = small_problem(seed=9)
    name = "sinkhorn_splr" if splr else "sinkhorn_sparse_newton"
    res, recs = O.traced(name, M, a, b, reg, tol=1e-8, max_iter=80, method=1, density=0.3,
                         shift=1e-3, ls_max_iter=ls_max_iter)
    kinds = set()
    for r in recs:
        h = r["hdr"]
        if h[10] == ls_max_iter and not h[8] and h[11] != h[7]:
            kinds.add("kept")
        elif h[8]:
            kinds.add("recompute")
    assert want in kinds
    check_newton_trace(M, a, b, reg, res, recs, splr, 0.3, 1e-3, 1e-8)
