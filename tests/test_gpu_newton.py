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


def test_ssns_accepts_a_step_that_raises_f(gpu_lib):
    # mu0 < 0 makes Hs + mu ||g|| I negative definite, so the first direction
    # ascends: no candidate lowers f, both halves of rho are negative, the
    # step is accepted, and the next gradient and Hessian come from the last
    # candidate's T (src/sinkhorn_ssns.cpp:185-187).  Plain CG on both sides.
    M, a, b = cloud_problem(28, 21, d=4, seed=6)
    reg = 0.05
    print("ssns mu0=-200 28x21", flush=True)
    r = gpu_lib.sinkhorn_ssns(M, a, b, reg, tol=1e-8, max_iter=200, method=0, mu0=-200.0)
    o = O.sinkhorn_ssns(M, a, b, reg, tol=1e-8, max_iter=200, method=0, mu0=-200.0)
    assert r.obj_vals[1] > r.obj_vals[0] and o["obj_vals"][1] > o["obj_vals"][0]
    assert abs(r.niter - o["niter"]) <= 1 and r.mar_errs[-1] < 1e-8
    k = min(6, len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-9)
    np.testing.assert_allclose(r.mar_errs[:k], o["mar_errs"][:k], rtol=1e-6)
    np.testing.assert_allclose(r.densities[:k], o["densities"][:k], rtol=1e-12)
    assert rel(r.dual, o["dual"]) < 1e-6


@pytest.mark.parametrize("solver,ls_max_iter", [("sinkhorn_sparse_newton", 2),
                                                ("sinkhorn_splr", 2),
                                                ("sinkhorn_sparse_newton", 1)])
def test_newton_family_exhausted_wolfe_trials(gpu_lib, solver, ls_max_iter):
    # One or two Wolfe trials: the first search runs out of trials and keeps
    # the unevaluated step of its last update with the last trial's T (f from
    # that T and the new potentials), or falls back to I_lo and recomputes.
    # The direct solve of the oracle against a 1e-13 PCG on the GPU.
    M, a, b = cloud_problem(28, 21, d=4, seed=9)
    reg = 0.05
    kw = dict(density=0.3, shift=1e-3, ls_max_iter=ls_max_iter)
    print("%s ls_max_iter=%d 28x21" % (solver, ls_max_iter), flush=True)
    r = getattr(gpu_lib, solver)(M, a, b, reg, tol=1e-8, max_iter=80, cg_tol=1e-13, **kw)
    o = getattr(O, solver)(M, a, b, reg, tol=1e-8, max_iter=80, method=1, **kw)
    k = min(12, len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-11)
    np.testing.assert_allclose(r.mar_errs[:k], o["mar_errs"][:k], rtol=1e-7)
    np.testing.assert_allclose(r.densities[:k], o["densities"][:k], rtol=1e-13)
    if o["niter"] < 80:
        assert abs(r.niter - o["niter"]) <= max(2, 0.1 * o["niter"])
        assert abs(r.obj_vals[-1] - o["obj_vals"][-1]) < 1e-10 * abs(o["obj_vals"][-1])


def test_splr_one_wolfe_trial_stops_on_a_nonfinite_objective(gpu_lib):
    # With a single trial SPLR's low-rank steps run away on this problem and
    # the objective overflows; both stop on the non-finite f at the same
    # iteration (src/sinkhorn_splr.cpp:101).
    M, a, b = cloud_problem(28, 21, d=4, seed=9)
    reg = 0.05
    kw = dict(density=0.3, shift=1e-3, ls_max_iter=1)
    print("splr ls_max_iter=1 28x21", flush=True)
    r = gpu_lib.sinkhorn_splr(M, a, b, reg, tol=1e-8, max_iter=80, cg_tol=1e-13, **kw)
    o = O.sinkhorn_splr(M, a, b, reg, tol=1e-8, max_iter=80, method=1, **kw)
    assert o["niter"] == r.niter < 10
    assert not np.isfinite(o["obj_vals"][-1]) and not np.isfinite(r.obj_vals[-1])
    k = len(o["obj_vals"]) - 1
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-10)
