"""Dense FP64 implicit-differentiation reference for the EROT losses (CPU).

Given the dual optimum gamma of min f (the oracle's solvers), form the plan T,
the dense dual Hessian H ((n+m-1) square, beta_{m-1} fixed to 0), solve
H w = dL/dgamma with numpy, and return G = dL/dM = dL/dM|gamma - w' dg/dM.
Test infrastructure: the checker of the GPU adjoint.
"""
import numpy as np


def plan(M, reg, gamma):
    n, m = M.shape
    beta = np.append(gamma[n:], 0.0)
    return np.exp((gamma[:n, None] + beta[None, :] - M) / reg)


def dual_hessian(T, reg):
    n, m = T.shape
    N = n + m - 1
    H = np.zeros((N, N))
    H[:n, :n] = np.diag(T.sum(1))
    H[n:, n:] = np.diag(T.sum(0)[: m - 1])
    H[:n, n:] = T[:, : m - 1]
    H[n:, :n] = T[:, : m - 1].T
    return H / reg


def loss_value(M, reg, gamma, loss):
    T = plan(M, reg, gamma)
    if loss == "sharp":
        return float((T * M).sum())
    # <T, M> - reg h(T), h(T) = sum T (1 - log T), with 0 log 0 = 0 for
    # entries that underflow
    with np.errstate(divide="ignore", invalid="ignore"):
        ent = np.where(T > 0.0, T * (1.0 - np.log(T)), 0.0)
    return float((T * M).sum() - reg * ent.sum())


def implicit_grad(M, reg, gamma, loss):
    n, m = M.shape
    T = plan(M, reg, gamma)
    H = dual_hessian(T, reg)
    alpha = gamma[:n]
    beta = np.append(gamma[n:], 0.0)
    if loss == "entropic":
        rhs = H @ gamma
        direct = T * (1.0 - (alpha[:, None] + beta[None, :]) / reg)
    else:
        TM = T * M
        rhs = np.concatenate([TM.sum(1), TM.sum(0)[: m - 1]]) / reg
        direct = T * (1.0 - M / reg)
    w = np.linalg.solve(H, rhs)
    wn = w[:n]
    wm = np.append(w[n:], 0.0)
    G = direct + T * (wn[:, None] + wm[None, :]) / reg
    return G, w


def dual_obj_grad(M, a, b, reg, gamma):
    """FP64 dual objective f and gradient g (gamma = [alpha; beta_0..m-2])."""
    n, m = M.shape
    T = plan(M, reg, gamma)
    r = T.sum(1)
    c = T.sum(0)[: m - 1]
    f = reg * r.sum() - gamma[:n] @ a - gamma[n:] @ b[: m - 1]
    return f, np.concatenate([r - a, c - b[: m - 1]]), T


def refine_dual(M, a, b, reg, gamma, tol=1e-13, max_iter=30):
    """Newton's method on the dual with the dense FP64 Hessian (numpy solve)
    and Armijo backtracking, from a nearby gamma, until ||g|| < tol.  The
    dual optimum is unique, so the result does not depend on the start.
    Returns (gamma, ||g||, iterations)."""
    gamma = np.array(gamma, dtype=np.float64)
    f, g, T = dual_obj_grad(M, a, b, reg, gamma)
    gn = np.linalg.norm(g)
    it = 0
    while gn >= tol and it < max_iter:
        H = dual_hessian(T, reg)
        d = -np.linalg.solve(H, g)
        del H
        step, slope = 1.0, g @ d
        for _ in range(40):
            f1, g1, T1 = dual_obj_grad(M, a, b, reg, gamma + step * d)
            if f1 <= f + 1e-4 * step * slope or np.linalg.norm(g1) < gn:
                break
            step *= 0.5
        gamma = gamma + step * d
        f, g, T = f1, g1, T1
        gn = np.linalg.norm(g)
        it += 1
    return gamma, gn, it
