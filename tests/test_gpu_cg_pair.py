"""Two right-hand sides in the same CG launches against two separate solves.

SPLR's low-rank solve (src/sinkhorn_linsolve.cpp:160-225) solves the Newton
system for two right-hand sides.  The GPU runs both CGs in the same kernel
launches (grid row 0 and 1), each with its own vectors, scalars and stop
test, so each must return exactly what a CG of its own returns: the same
solution bit for bit, the same iteration count and the same residual, also
when one side stops long before the other, hits the cap, has a zero
right-hand side or is already converged at the start.  Matrix-free Hessians
run the two solves one after the other.
"""
import numpy as np
import pytest

from conftest import cloud_problem

pytestmark = pytest.mark.gpu


def near_optimum(P):
    from regot import _lib
    return P.solve(_lib.SOLVER_BCD, tol=1e-5, max_iter=5000, compute_plan=False).dual


def same(single, pair, k):
    x, it, err = single
    (x0, x1), its, errs = pair
    xs = (x0, x1)[k]
    assert its[k] == it
    assert errs[k] == err
    assert np.array_equal(xs, x)


@pytest.mark.parametrize("bulk", ["fp32", "fp64"])
def test_cg_pair_matches_single_solves(gpu_lib, bulk):
    n, m, reg = 700, 900, 0.02
    M, a, b = cloud_problem(n, m, d=8, seed=81)
    if bulk == "fp32":
        M = M.astype(np.float32)
    rng = np.random.default_rng(8)
    N = n + m - 1
    with gpu_lib.Problem(M, a, b, reg) as P:
        P.set_hess_storage(0)
        gamma = near_optimum(P)
        P.sparsified_hess_with_density(gamma, 0.02)
        h2max = P.hess_export()["h2"].max()
        # (shift scale, tol, cap, rhs1 kind): both long; different lengths
        # (the second right-hand side nearly constant); a cap that ends both;
        # a zero right-hand side; both converged at the start (tol > 1)
        cases = [(1e-4, 1e-12, 5000, "rand"), (1e-3, 1e-10, 5000, "smooth"),
                 (1e-5, 1e-12, 37, "rand"), (1e-3, 1e-10, 5000, "zero"),
                 (1e-3, 2.0, 5000, "rand")]
        for q, (sc, tol, cap, kind) in enumerate(cases):
            shift = sc * h2max
            r0 = rng.standard_normal(N)
            if kind == "rand":
                r1 = rng.standard_normal(N)
            elif kind == "smooth":
                r1 = np.ones(N) + 1e-3 * rng.standard_normal(N)
            else:
                r1 = np.zeros(N)
            print(f"cg pair {bulk} case {q}: {kind}", flush=True)
            s0 = P.hess_cg(r0, shift, tol=tol, max_iter=cap)
            s1 = P.hess_cg(r1, shift, tol=tol, max_iter=cap)
            pr = P.hess_cg2(r0, r1, shift, tol=tol, max_iter=cap)
            print(f"  iterations {s0[1]} / {s1[1]}", flush=True)
            same(s0, pr, 0)
            same(s1, pr, 1)
            if q == 2:
                assert s0[1] == cap and s1[1] == cap
            if kind == "zero":
                assert s1[1] == 0 and np.all(pr[0][1] == 0.0)
            if tol > 1:
                assert s0[1] == 0 and s1[1] == 0
            elif kind != "zero":
                assert s0[1] > 16 and s1[1] > 16   # past the first batch


def test_cg_pair_matrix_free_runs_two_solves(gpu_lib):
    n, m, reg = 400, 600, 0.05
    M, a, b = cloud_problem(n, m, seed=82)
    M32 = M.astype(np.float32)
    rng = np.random.default_rng(9)
    N = n + m - 1
    with gpu_lib.Problem(M32, a, b, reg) as P:
        P.set_hess_storage(1)
        gamma = near_optimum(P)
        P.sparsified_hess_with_density(gamma, 1.0)
        r0, r1 = rng.standard_normal(N), rng.standard_normal(N)
        shift = 1e-3
        s0 = P.hess_cg(r0, shift, tol=1e-10, max_iter=5000)
        s1 = P.hess_cg(r1, shift, tol=1e-10, max_iter=5000)
        pr = P.hess_cg2(r0, r1, shift, tol=1e-10, max_iter=5000)
    same(s0, pr, 0)
    same(s1, pr, 1)
