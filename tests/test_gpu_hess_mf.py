"""Matrix-free sparsified Hessian (FP32 bulk) against the stored CSR / CSC.

The matrix-free build keeps the kept entries of the threshold rule (or every
entry, the full Hessian) as a dense FP32 matrix with zeros where an entry is
dropped, written by the pass that decides the kept set, and every product reads
it once; so on one handle and one dual point both storages must give the same
counts and the same products up to the order of the FP64 sums; the CG solves
and the drivers that use them must agree to their own tolerances.  The stored
form is checked against the oracle in test_gpu_hessian.py.
"""
import numpy as np
import pytest

from conftest import cloud_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (n, m): one column chunk partial / full, two chunks, six, eight full
# (m - 1 = 16384, the largest matrix-free width); n below and above the
# compute-unit count (one row or several per workgroup).
SHAPES = [(700, 1300), (200, 2049), (150, 2050), (64, 12001), (48, 16385)]


def near_point(M, a, b, reg, seed):
    """Two BCD sweeps from zero, perturbed: a point whose kept set is dense."""
    alpha = np.zeros(M.shape[0])
    beta = O.optimal_beta(M, b, reg, alpha)
    for _ in range(2):
        alpha = O.optimal_alpha(M, a, reg, beta)
        beta = O.optimal_beta(M, b, reg, alpha)
    g = np.concatenate([alpha + beta[-1], beta[:-1] - beta[-1]])
    return g + 0.05 * reg * np.random.default_rng(seed).standard_normal(g.shape)


def products(P, n, m, x, z, rhs, shift):
    d = P.hess_apply_delta(x, shift)
    h = P.hess_apply_hs(z)
    sol, it, err = P.hess_cg(rhs, shift, tol=1e-10, max_iter=10 * m)
    return d, h, sol, it, err


@pytest.mark.parametrize("n,m", SHAPES)
def test_matrix_free_products_match_stored(gpu_lib, n, m):
    reg = 0.05
    M, a, b = cloud_problem(n, m, seed=n + m)
    M32 = M.astype(np.float32)
    gamma = near_point(M32.astype(np.float64), a, b, reg, seed=n)
    rng = np.random.default_rng(m)
    x = rng.standard_normal(m - 1)
    z = rng.standard_normal(n + m - 1)
    rhs = rng.standard_normal(n + m - 1)
    with gpu_lib.Problem(M32, a, b, reg) as P:
        _, g = P.dual_obj_grad(gamma)
        for rule in ("threshold", "full"):
            print(f"matrix-free {n}x{m} {rule}", flush=True)
            P.set_hess_storage(0)
            if rule == "threshold":
                nnz0 = P.sparsified_hess(gamma, 1e-3 * np.linalg.norm(g), 0.5)
            else:
                nnz0 = P.sparsified_hess_with_density(gamma, 1.0)
            info0 = P.hess_info()
            H = P.hess_export()
            h1, h2 = H["h1"], H["h2"]
            shift = 0.05 * h1.mean()
            ref = products(P, n, m, x, z, rhs, shift)
            P.set_hess_storage(1)
            if rule == "threshold":
                nnz1 = P.sparsified_hess(gamma, 1e-3 * np.linalg.norm(g), 0.5)
                # the same kept set: the same counts and density
                assert nnz1 == nnz0
                assert P.hess_info() == info0
                assert info0[0] > 0.1 * n * (m - 1)   # a dense kept set
            else:
                P.sparsified_hess_with_density(gamma, 1.0)
            with pytest.raises(gpu_lib.RegotError):
                P.hess_export()
            got = products(P, n, m, x, z, rhs, shift)
            # |sigma' ((sigma x) / (h1 + s))|_j <= h2_j max|x| (row and column
            # sums of sigma are h1 and h2): the terms' scale
            sc_d = (h2.max() + shift) * np.abs(x).max()
            assert np.abs(got[0] - ref[0]).max() <= 1e-12 * sc_d
            sc_h = max(h1.max(), h2.max()) * np.abs(z).max()
            assert np.abs(got[1] - ref[1]).max() <= 1e-12 * sc_h
            # CG to 1e-10 on (Hs + s I), cond <= (max h + s) / s
            assert got[4] < 1e-10 and ref[4] < 1e-10
            assert abs(got[3] - ref[3]) <= 2
            cond = (max(h1.max(), h2.max()) * 2 + shift) / shift
            assert np.linalg.norm(got[2] - ref[2]) <= 1e-10 * cond * np.linalg.norm(ref[2])
            P.set_hess_storage(-1)


@pytest.mark.parametrize("solver,kw", [("sinkhorn_ssns", {}),
                                       ("sinkhorn_ssns", dict(method=0)),
                                       ("sinkhorn_newton", {})])
def test_drivers_matrix_free_match_stored(gpu_lib, solver, kw):
    # SSNS's kept sets are dense here (tens of percent): the default storage
    # goes matrix-free, so run both storages explicitly and compare.
    n, m, reg = 900, 1100, 0.01
    if solver == "sinkhorn_newton":
        n, m = 300, 420
    M, a, b = cloud_problem(n, m, seed=33)
    M32 = M.astype(np.float32)
    runs = {}
    for mode in (0, 1):
        print(f"{solver} {kw} {n}x{m} storage {mode}", flush=True)
        runs[mode] = getattr(gpu_lib, solver)(M32, a, b, reg, tol=1e-8, max_iter=300,
                                              verify=True, hess_storage=mode, **kw)
    st, mf = runs[0], runs[1]
    for r in (st, mf):
        assert r.mar_errs[-1] < 1e-8 and r.mar_err_fp64 < 1e-8
    assert abs(st.niter - mf.niter) <= max(1, st.niter // 10)
    if solver == "sinkhorn_ssns":
        # the first build sees the same point: the same kept set
        assert st.densities[0] == mf.densities[0]
    # both duals meet ||g|| < 1e-8: they agree to ~1e-8 / (smallest curvature)
    rel = np.linalg.norm(st.dual - mf.dual) / np.linalg.norm(st.dual)
    assert rel < 5e-6
    assert abs(st.obj_vals[-1] - mf.obj_vals[-1]) < 1e-9 * abs(st.obj_vals[-1])


def test_implicit_grad_matrix_free_matches_stored(gpu_lib):
    torch = pytest.importorskip("torch")
    n, m, reg = 200, 260, 0.1
    M, a, b = cloud_problem(n, m, d=4, seed=12)
    M32 = M.astype(np.float32)
    gamma = O.sinkhorn_bcd(M32.astype(np.float64), a, b, reg, tol=1e-12, max_iter=100000)["dual"]
    out = {}
    for mode in (0, 1):
        for loss in ("entropic", "sharp"):
            print(f"implicit grad {loss} storage {mode}", flush=True)
            with gpu_lib.Problem(M32, a, b, reg) as P:
                P.set_hess_storage(mode)
                G = torch.empty(n, m, dtype=torch.float64, device="cuda:0")
                torch.cuda.synchronize()
                value, w, it, err = P.implicit_grad(gamma, G.data_ptr(), np.float64, m,
                                                    loss=loss, cg_tol=1e-12)
                assert err < 1e-12
                out[mode, loss] = (value, w, G.cpu().numpy())
    for loss in ("entropic", "sharp"):
        v0, w0, G0 = out[0, loss]
        v1, w1, G1 = out[1, loss]
        assert abs(v0 - v1) <= 1e-12 * abs(v0)
        np.testing.assert_allclose(w1, w0, rtol=1e-7, atol=1e-9 * np.abs(w0).max())
        np.testing.assert_allclose(G1, G0, rtol=1e-7, atol=1e-9 * np.abs(G0).max())
