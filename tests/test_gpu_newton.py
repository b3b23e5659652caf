"""GPU parity of the Newton-type drivers (SSNS, sparse Newton, SPLR).

The oracle runs the same algorithms in FP64 on the CPU.  Its default linear
solver is a dense Cholesky (the mathematical equivalent of the reference's
SimplicialLDLT default); the GPU uses CG at a 1e-10 relative residual, so
trajectories agree to the CG tolerance and iteration counts within a few.
The EROT optimum is unique, so every solver must land on the same dual.
"""
import numpy as np
import pytest

from conftest import cloud_problem, readme_example
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(x, ref):
    return np.max(np.abs(np.asarray(x) - ref)) / max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("reg", [0.1, 0.01])
def test_ssns_readme_fp64(gpu_lib, reg):
    M, a, b = readme_example()
    r = gpu_lib.sinkhorn_ssns(M, a, b, reg, tol=1e-10, max_iter=500)
    o = O.sinkhorn_ssns(M, a, b, reg, tol=1e-10, max_iter=500, method=1)
    ref = O.sinkhorn_bcd(M, a, b, reg, tol=1e-12, max_iter=200000)["dual"]
    assert r.niter < 500 and r.mar_errs[-1] < 1e-10
    assert abs(r.niter - o["niter"]) <= 3
    k = min(6, len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-9)
    np.testing.assert_allclose(r.densities[:k], o["densities"][:k], rtol=1e-12)
    assert len(r.densities) == len(r.obj_vals) == r.niter + 1
    # where the marginals pin the dual (b_j not tiny), all solvers agree
    well = np.concatenate([a > 1e-8, b[:-1] > 1e-8])
    assert rel(r.dual[well], ref[well]) < 1e-6
    np.testing.assert_allclose(r.plan, O.get_plan(M, reg, r.dual), rtol=1e-10, atol=1e-300)


@pytest.mark.parametrize("solver", ["sinkhorn_sparse_newton", "sinkhorn_splr"])
def test_newton_family_readme_fp64(gpu_lib, solver):
    M, a, b = readme_example()
    reg = 0.1
    kw = dict(density=0.5, shift=1e-3)
    r = getattr(gpu_lib, solver)(M, a, b, reg, tol=1e-9, max_iter=300, **kw)
    o = getattr(O, solver)(M, a, b, reg, tol=1e-9, max_iter=300, method=1, **kw)
    assert r.niter < 300 and r.mar_errs[-1] < 1e-9
    assert abs(r.niter - o["niter"]) <= max(3, 0.1 * o["niter"])
    k = min(5, len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-8)
    np.testing.assert_allclose(r.densities[:k], o["densities"][:k], rtol=1e-12)
    # The converged dual is pinned only up to a near-flat direction (alpha - c,
    # beta + c moves mass in the last column only, whose b is ~1e-5 here), so
    # compare the objective and the plan, which that direction barely moves.
    assert abs(r.obj_vals[-1] - o["obj_vals"][-1]) < 1e-10 * abs(o["obj_vals"][-1])
    np.testing.assert_allclose(r.plan, O.get_plan(M, reg, o["dual"]), rtol=0, atol=1e-8)


@pytest.mark.parametrize("solver,kw", [("sinkhorn_ssns", {}),
                                       ("sinkhorn_sparse_newton", dict(density=1.0)),
                                       ("sinkhorn_splr", dict(density=0.2))])
def test_newton_fp32_cloud(gpu_lib, solver, kw):
    # FP32 bulk: M in FP32; the Newton drivers' objective and gradient passes
    # accumulate in FP64, the sparsified Hessian stays FP32.
    n, m, reg = 400, 350, 0.02
    M, a, b = cloud_problem(n, m, seed=21)
    M32 = M.astype(np.float32)
    r = getattr(gpu_lib, solver)(M32, a, b, reg, tol=1e-8, max_iter=400, verify=True, **kw)
    o = getattr(O, solver)(M32.astype(np.float64), a, b, reg, tol=1e-8, max_iter=400, method=1,
                           **kw)
    ref = O.sinkhorn_bcd(M32.astype(np.float64), a, b, reg, tol=1e-11, max_iter=100000)["dual"]
    assert o["niter"] < 400
    assert r.mar_errs[-1] < 1e-8
    assert r.mar_err_fp64 < 1e-8
    assert abs(r.niter - o["niter"]) <= max(3, 0.1 * o["niter"])
    # mar_err <= 1e-8 pins the dual to ~1e-8 / (smallest curvature)
    assert rel(r.dual, ref) < 5e-6


def test_dropin_contract(gpu_lib, capfd):
    M, a, b = readme_example(40, 30)
    # verbose output goes to the C-level stdout, as the reference's std::cout does
    r = gpu_lib.sinkhorn_ssns(M, a, b, 0.1, 1e-8, 100, 1, method=0, unknown_kwarg=3)
    out = capfd.readouterr().out
    assert out.startswith("iter = 0, objval = ")
    assert ", ||grad|| = " in out
    assert isinstance(r.niter, int)
    d = r.dual
    d[0] = 1e9                      # fresh copy on every read
    assert r.dual[0] != 1e9
    assert isinstance(r.obj_vals, list) and isinstance(r.densities, list)
    assert r.plan.dtype == np.float64 and r.plan.shape == (40, 30)
    # max_iter exhaustion and immediate convergence
    r = gpu_lib.sinkhorn_splr(M, a, b, 0.1, tol=0.0, max_iter=4)
    assert r.niter == 4 and len(r.mar_errs) == 5
    r = gpu_lib.sinkhorn_sparse_newton(M, a, b, 0.1, tol=1e9)
    assert r.niter == 0 and len(r.obj_vals) == 1
