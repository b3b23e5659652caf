"""Newton-system solves of the device CG against direct factorisations.

The reference solves its sparsified Newton systems (Hs + shift I) d = rhs with
a sparse LDL^T factorisation (src/sinkhorn_linsolve.cpp:20-56,94-104, the
matrix assembled by to_spmat at src/sinkhorn_hess.cpp:57-96).  This build
solves the same systems by block elimination and CG on the Schur complement
Delta = diag(h2) + shift - sigma' diag(1 / (h1 + shift)) sigma
(src/sinkhorn_cg.cpp:20-49).  In exact arithmetic both give the unique
solution of an SPD system; these tests build the Hessians the drivers build
(top-k at density 1% and 5%, the threshold rule) at 4096^2 and compare the
device CG, run to a relative residual of 1e-12, with a Cholesky factorisation
of Delta (LAPACK), and at 1000^2 with a sparse LU of the whole
(n + m - 1)-square matrix (SuperLU), which is what an LDL^T solve returns up
to rounding.  The error bound is the CG tolerance times the condition number
of Delta, which the shift bounds by (max h2 + shift) / shift.
"""
import numpy as np
import pytest

from conftest import cloud_problem

pytestmark = pytest.mark.gpu

scipy_linalg = pytest.importorskip("scipy.linalg")
scipy_sparse = pytest.importorskip("scipy.sparse")
from scipy.sparse.linalg import spsolve  # noqa: E402


def sigma_dense(H, n, m):
    S = np.zeros((n, m - 1))
    for j in range(m - 1):
        sl = slice(H["colptr"][j], H["colptr"][j + 1])
        S[H["rowind"][sl], j] = H["cval"][sl]
    return S


def schur_solve(H, S, rhs, shift):
    """Block elimination with a Cholesky factorisation of Delta."""
    n = H["h1"].size
    d1 = H["h1"] + shift
    C = np.diag(H["h2"] + shift) - S.T @ (S / d1[:, None])
    cf = scipy_linalg.cho_factor(C)
    r1, r2 = rhs[:n], rhs[n:]
    y = scipy_linalg.cho_solve(cf, r2 - S.T @ (r1 / d1))
    x = (r1 - S @ y) / d1
    return np.concatenate([x, y])


def solved_point(P):
    """BCD to ||g|| < 1e-4: the kind of point the Newton drivers reach."""
    from regot import _lib
    r = P.solve(_lib.SOLVER_BCD, tol=1e-4, max_iter=5000, compute_plan=False)
    return r.dual


@pytest.mark.parametrize("rule", ["topk1", "topk5", "threshold"])
def test_newton_system_4096_matches_cholesky(gpu_lib, rule):
    n = m = 4096
    reg = 0.01
    print(f"direct solve 4096: {rule}: problem", flush=True)
    M, a, b = cloud_problem(n, m, d=16, seed=41)
    M32 = M.astype(np.float32)
    with gpu_lib.Problem(M32, a, b, reg) as P:
        P.set_hess_storage(0)
        gamma = solved_point(P)
        _, g = P.dual_obj_grad(gamma)
        print(f"direct solve 4096: {rule}: Hessian", flush=True)
        if rule == "topk1":
            P.sparsified_hess_with_density(gamma, 0.01)
        elif rule == "topk5":
            P.sparsified_hess_with_density(gamma, 0.05)
        else:
            P.sparsified_hess(gamma, 1e-3 * np.linalg.norm(g), 0.5)
        H = P.hess_export()
        shift = 1e-3 * H["h2"].max()
        rhs = -g
        print(f"direct solve 4096: {rule}: CG", flush=True)
        sol, it, err = P.hess_cg(rhs, shift, tol=1e-12, max_iter=20000)
    assert err < 1e-12
    print(f"direct solve 4096: {rule}: Cholesky ({it} CG iterations)", flush=True)
    S = sigma_dense(H, n, m)
    ref = schur_solve(H, S, rhs, shift)
    # the factorisation solves the system it was given: residual check
    Hv = np.concatenate([H["h1"] * ref[:n] + S @ ref[n:], S.T @ ref[:n] + H["h2"] * ref[n:]])
    assert np.linalg.norm(Hv + shift * ref - rhs) <= 1e-10 * np.linalg.norm(rhs)
    cond = (H["h2"].max() + shift) / shift
    rel = np.linalg.norm(sol - ref) / np.linalg.norm(ref)
    print(f"direct solve 4096: {rule}: relative difference {rel:.3e}, cond bound {cond:.1f}",
          flush=True)
    assert rel <= 10 * 1e-12 * cond


def test_newton_system_1000_matches_sparse_lu_of_full_matrix(gpu_lib):
    n, m, reg = 1000, 900, 0.02
    M, a, b = cloud_problem(n, m, d=8, seed=43)
    with gpu_lib.Problem(M, a, b, reg) as P:
        P.set_hess_storage(0)
        gamma = solved_point(P)
        _, g = P.dual_obj_grad(gamma)
        P.sparsified_hess_with_density(gamma, 0.02)
        H = P.hess_export()
        shift = 1e-4 * H["h2"].max()
        sol, it, err = P.hess_cg(-g, shift, tol=1e-12, max_iter=20000)
    assert err < 1e-12
    # the full matrix as to_spmat assembles it: [[diag h1, sigma], [sigma', diag h2]] + shift I
    S = scipy_sparse.csc_matrix((H["cval"], H["rowind"], H["colptr"]), shape=(n, m - 1))
    A = scipy_sparse.bmat([[scipy_sparse.diags(H["h1"] + shift), S],
                           [S.T, scipy_sparse.diags(H["h2"] + shift)]], format="csc")
    ref = spsolve(A, -g)
    cond = (H["h2"].max() + shift) / shift
    rel = np.linalg.norm(sol - ref) / np.linalg.norm(ref)
    assert rel <= 10 * 1e-12 * cond, (rel, cond, it)
