"""GPU parity of the dense-Newton, L-BFGS-dual and APDAGD drivers.

The oracle runs the same algorithms in FP64 on the CPU (dense Cholesky for
Newton's direction, the two-loop recursion and Nocedal-Wright search for
L-BFGS, the plan averaging for APDAGD).  The GPU forms Newton's direction by a
preconditioned CG to a 1e-12 relative residual instead of the Cholesky solve,
and sums in a different order, so histories agree to ~1e-9 relative and
iteration counts within a few (first-order methods amplify rounding slowly
over thousands of steps).
"""
import numpy as np
import pytest

from conftest import cloud_problem, readme_example
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel(x, ref):
    return np.max(np.abs(np.asarray(x) - ref)) / max(1.0, np.max(np.abs(ref)))


# Newton's Armijo test compares objective values; once ||g|| ~ 1e-9 the
# decrease it asks for (~ g'H^-1 g / 2 ~ 1e-18) is below the objective's FP64
# resolution, and acceptance depends on summation order (the reference's
# own outcome there depends on Eigen's).  The tolerances below stop the
# quadratic phase before that regime.
@pytest.mark.parametrize("reg", [0.1, 0.02])
def test_newton_dense_fp64(gpu_lib, reg):
    M, a, b = cloud_problem(60, 45, d=4, seed=5)
    print("newton fp64 60x45 reg=%g" % reg, flush=True)
    r = gpu_lib.sinkhorn_newton(M, a, b, reg, tol=1e-8, max_iter=100)
    o = O.sinkhorn_newton(M, a, b, reg, tol=1e-8, max_iter=100)
    assert r.niter == o["niter"] and r.niter < 100
    assert len(r.obj_vals) == r.niter + 1 and r.densities == []
    np.testing.assert_allclose(r.obj_vals, o["obj_vals"], rtol=1e-11)
    e_r, e_o = np.array(r.mar_errs), np.array(o["mar_errs"])
    big = e_o > 1e-7
    np.testing.assert_allclose(e_r[big], e_o[big], rtol=1e-5)
    assert e_r[-1] < 1e-8
    assert rel(r.dual, o["dual"]) < 1e-8
    np.testing.assert_allclose(r.plan, O.get_plan(M, reg, r.dual), rtol=1e-10, atol=1e-300)


def test_newton_dense_readme_and_fp32(gpu_lib):
    M, a, b = readme_example(100, 80)
    print("newton readme 100x80", flush=True)
    r = gpu_lib.sinkhorn_newton(M, a, b, 0.1, tol=1e-8, max_iter=100)
    o = O.sinkhorn_newton(M, a, b, 0.1, tol=1e-8, max_iter=100)
    # With theta = 0.5 and the exact Newton direction, a quadratic model meets
    # the sufficient-decrease test with equality, so close to the optimum the
    # unit step's acceptance is decided by rounding (in the reference too):
    # here at ||g|| = 5.7e-7 a rejected unit step costs one or two halvings.
    assert abs(r.niter - o["niter"]) <= 3 and r.mar_errs[-1] < 1e-8
    k = 13
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-10)
    assert abs(r.obj_vals[-1] - o["obj_vals"][-1]) < 1e-12 * abs(o["obj_vals"][-1])
    # FP32 bulk: M in FP32, objective/gradient in FP64 sums, Hessian FP32
    M, a, b = cloud_problem(300, 260, seed=8)
    M32 = M.astype(np.float32)
    print("newton fp32 300x260", flush=True)
    r = gpu_lib.sinkhorn_newton(M32, a, b, 0.02, tol=1e-7, max_iter=100, verify=True)
    o = O.sinkhorn_newton(M32.astype(np.float64), a, b, 0.02, tol=1e-7, max_iter=100)
    assert r.mar_errs[-1] < 1e-7 and r.mar_err_fp64 < 1e-7
    assert abs(r.niter - o["niter"]) <= 2
    assert rel(r.dual, o["dual"]) < 1e-7


@pytest.mark.parametrize("reg", [0.1, 0.02])
def test_lbfgs_dual_fp64(gpu_lib, reg):
    M, a, b = cloud_problem(60, 45, d=4, seed=6)
    print("lbfgs fp64 60x45 reg=%g" % reg, flush=True)
    r = gpu_lib.sinkhorn_lbfgs_dual(M, a, b, reg, tol=1e-7, max_iter=2000)
    o = O.sinkhorn_lbfgs_dual(M, a, b, reg, tol=1e-7, max_iter=2000)
    assert r.mar_errs[-1] <= 1e-7 and r.niter < 2000
    assert abs(r.niter - o["niter"]) <= max(2, 0.05 * o["niter"])
    assert len(r.obj_vals) == r.niter + 1 and r.densities == []
    k = min(15, len(o["obj_vals"]), len(r.obj_vals))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-10)
    np.testing.assert_allclose(r.mar_errs[:k], o["mar_errs"][:k], rtol=1e-6)
    ref = O.sinkhorn_newton(M, a, b, reg, tol=1e-11, max_iter=100)["dual"]
    assert rel(r.dual, ref) < 1e-4


def test_lbfgs_dual_edges(gpu_lib):
    # the initial point already passes: minimize() returns 1 with one history
    # entry (src/LBFGSpp/LBFGS.h:84-89)
    M, a, b = cloud_problem(30, 20, d=3, seed=9)
    r = gpu_lib.sinkhorn_lbfgs_dual(M, a, b, 0.1, tol=10.0, max_iter=50)
    o = O.sinkhorn_lbfgs_dual(M, a, b, 0.1, tol=10.0, max_iter=50)
    assert r.niter == o["niter"] == 1 and len(r.obj_vals) == 1
    # max_iter caps the iterations (history niter + 1)
    r = gpu_lib.sinkhorn_lbfgs_dual(M, a, b, 0.1, tol=1e-14, max_iter=5)
    o = O.sinkhorn_lbfgs_dual(M, a, b, 0.1, tol=1e-14, max_iter=5)
    assert r.niter == o["niter"] == 5 and len(r.obj_vals) == 6
    np.testing.assert_allclose(r.obj_vals, o["obj_vals"], rtol=1e-11)
    # an unreachable tolerance ends in LBFGSpp's line-search error
    with pytest.raises(RuntimeError, match="line search"):
        gpu_lib.sinkhorn_lbfgs_dual(M, a, b, 0.1, tol=1e-15, max_iter=3000)


def test_apdagd_fp64(gpu_lib):
    M, a, b = cloud_problem(40, 30, d=4, seed=3)
    reg = 0.1
    print("apdagd fp64 40x30", flush=True)
    r = gpu_lib.sinkhorn_apdagd(M, a, b, reg, tol=1e-5, max_iter=5000)
    o = O.sinkhorn_apdagd(M, a, b, reg, tol=1e-5, max_iter=5000)
    assert r.niter < 5000 and abs(r.niter - o["niter"]) <= max(2, 0.02 * o["niter"])
    assert len(r.obj_vals) == r.niter + 2
    k = min(50, len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-10)
    np.testing.assert_allclose(r.mar_errs[:k], o["mar_errs"][:k], rtol=1e-8)
    # the result's plan is the averaged primal plan, not T(dual)
    P = r.plan
    full_err = np.sqrt(((P.sum(1) - a) ** 2).sum() + ((P.sum(0) - b) ** 2).sum())
    assert full_err <= 1e-5 and abs(full_err - r.mar_errs[-1]) < 1e-12
    # 1300 steps of a first-order method carry summation-order rounding into
    # the averaged plan at the 1e-6 relative level
    np.testing.assert_allclose(P, o["plan"], rtol=1e-5, atol=1e-12)
    # the duals drift apart along the near-flat direction (alpha - c, beta + c),
    # which only rescales the last column's mass
    assert rel(r.dual, o["dual"]) < 1e-5


def test_apdagd_fp32_ragged(gpu_lib):
    # FP32 M, n and m not multiples of the 256-column strips or 128-row blocks
    M, a, b = cloud_problem(301, 517, seed=12)
    M32 = M.astype(np.float32)
    print("apdagd fp32 301x517", flush=True)
    r = gpu_lib.sinkhorn_apdagd(M32, a, b, 0.05, tol=1e-4, max_iter=3000)
    o = O.sinkhorn_apdagd(M32.astype(np.float64), a, b, 0.05, tol=1e-4, max_iter=3000)
    assert r.niter < 3000 and abs(r.niter - o["niter"]) <= max(2, 0.02 * o["niter"])
    k = min(30, len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-9)
    np.testing.assert_allclose(r.plan, o["plan"], rtol=1e-5, atol=1e-12)


def test_first_order_dropin_contract(gpu_lib, capfd):
    M, a, b = readme_example(40, 30)
    # positional arguments as in the reference binding, unknown keywords ignored
    r = gpu_lib.sinkhorn_newton(M, a, b, 0.1, 1e-8, 50, 1, unknown_kwarg=3)
    out = capfd.readouterr().out
    assert out.startswith("iter = 0, objval = ") and ", ||grad|| = " in out
    assert isinstance(r.niter, int) and r.plan.shape == (40, 30)
    r = gpu_lib.sinkhorn_lbfgs_dual(M, a, b, 0.1, 1e-6, 500, 1)
    out = capfd.readouterr().out
    assert out.startswith("iter = 0, objval = ")
    r = gpu_lib.sinkhorn_apdagd(M, a, b, 0.1, 1e-3, 50, 1)
    out = capfd.readouterr().out
    assert out.startswith("iter = 0, primal_obj = ") and ", mar_err = " in out
    assert r.plan.dtype == np.float64 and r.plan.flags["C_CONTIGUOUS"]
    p1 = r.plan
    p1[0, 0] = -1.0                 # fresh copy on every read
    assert r.plan[0, 0] != -1.0
    # max_iter exhaustion: APDAGD records max_iter + 1 entries
    r = gpu_lib.sinkhorn_apdagd(M, a, b, 0.1, tol=0.0, max_iter=4)
    assert r.niter == 4 and len(r.mar_errs) == 5
    r = gpu_lib.sinkhorn_newton(M, a, b, 0.1, tol=1e9)
    assert r.niter == 0 and len(r.obj_vals) == 1
