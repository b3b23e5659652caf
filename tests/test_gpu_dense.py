"""GPU parity of the dense passes and the BCD solver against the FP64 oracle.

Tolerances: FP64 bulk 1e-10 relative, FP32 bulk 1e-6 relative (scale =
max(1, |reference|_inf)), as the port's correctness bar states.
"""
import numpy as np
import pytest

from conftest import cloud_problem, readme_example
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-10, "fp32": 1e-6}


def rel_err(x, ref):
    x, ref = np.asarray(x), np.asarray(ref)
    return np.max(np.abs(x - ref)) / max(1.0, np.max(np.abs(ref)))


def problem_as(M, bulk):
    """The oracle solves the same problem: for FP32 bulk, M rounded to FP32."""
    if bulk == "fp32":
        Mc = M.astype(np.float32)
        return Mc, Mc.astype(np.float64)
    return M, M


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
@pytest.mark.parametrize("shape,reg", [((300, 257), 0.1), ((129, 700), 0.01), ((64, 64), 1e-3)])
def test_half_steps(gpu_lib, bulk, shape, reg):
    n, m = shape
    M, a, b = cloud_problem(n, m, seed=n + m)
    Mg, Mo = problem_as(M, bulk)
    rng = np.random.default_rng(1)
    with gpu_lib.Problem(Mg, a, b, reg) as P:
        beta = rng.standard_normal(m) * 0.3
        alpha = P.optimal_alpha(beta)
        assert rel_err(alpha, O.optimal_alpha(Mo, a, reg, beta)) < TOL[bulk]
        alpha0 = np.zeros(n)
        beta0 = P.optimal_beta(alpha0)
        assert rel_err(beta0, O.optimal_beta(Mo, b, reg, alpha0)) < TOL[bulk]
        # huge offsets: exercises the adaptive shift in both directions
        beta_off = beta + 1e3 * reg
        alpha2 = P.optimal_alpha(beta_off)
        assert rel_err(alpha2, O.optimal_alpha(Mo, a, reg, beta_off)) < TOL[bulk]
        alpha_off = alpha - 5e2 * reg
        beta2 = P.optimal_beta(alpha_off)
        assert rel_err(beta2, O.optimal_beta(Mo, b, reg, alpha_off)) < TOL[bulk]


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
def test_obj_grad_plan(gpu_lib, bulk):
    n, m, reg = 200, 333, 0.05
    M, a, b = cloud_problem(n, m, seed=7)
    Mg, Mo = problem_as(M, bulk)
    with gpu_lib.Problem(Mg, a, b, reg) as P:
        alpha = np.zeros(n)
        beta = O.optimal_beta(Mo, b, reg, alpha)
        gamma = np.concatenate([alpha + beta[-1], beta[:-1] - beta[-1]])
        gamma += np.random.default_rng(3).standard_normal(n + m - 1) * 0.01
        f = P.dual_obj(gamma)
        fo = O.dual_obj(Mo, a, b, reg, gamma)
        assert abs(f - fo) / max(1.0, abs(fo)) < TOL[bulk]
        f2, g = P.dual_obj_grad(gamma)
        fo2, go = O.dual_obj_grad(Mo, a, b, reg, gamma)
        assert abs(f2 - fo2) / max(1.0, abs(fo2)) < TOL[bulk]
        # gradient entries are O(a_i): compare against the marginal scale
        assert np.max(np.abs(g - go)) < TOL[bulk] * max(a.max(), b.max()) * 10
        plan = P.get_plan(gamma)
        po = O.get_plan(Mo, reg, gamma)
        assert np.max(np.abs(plan - po) / np.maximum(po, 1e-300)) < 1e-12


def test_bcd_readme_fp64_matches_oracle_history(gpu_lib):
    M, a, b = readme_example()
    r = gpu_lib.sinkhorn_bcd(M, a, b, 0.1, tol=1e-9, max_iter=2000)
    o = O.sinkhorn_bcd(M, a, b, 0.1, tol=1e-9, max_iter=2000)
    assert abs(r.niter - o["niter"]) <= 1
    assert rel_err(r.dual, o["dual"]) < 1e-10
    k = min(len(r.obj_vals), len(o["obj_vals"]))
    assert len(r.obj_vals) == r.niter + 1 == len(r.mar_errs) == len(r.run_times)
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r.mar_errs[:k], o["mar_errs"][:k], rtol=1e-6, atol=1e-14)
    assert r.mar_errs[-1] < 1e-9
    plan = r.plan
    assert plan.shape == (100, 80) and plan.flags["C_CONTIGUOUS"]
    np.testing.assert_allclose(plan, O.get_plan(M, 0.1, o["dual"]), rtol=1e-8, atol=1e-15)
    assert r.densities == []
    assert all(np.diff(r.run_times) >= 0)


def test_bcd_fp32_cloud(gpu_lib):
    n, m, reg = 1000, 900, 0.05
    M, a, b = cloud_problem(n, m, seed=11)
    M32 = M.astype(np.float32)
    r = gpu_lib.sinkhorn_bcd(M32, a, b, reg, tol=1e-8, max_iter=5000, verify=True)
    o = O.sinkhorn_bcd(M32.astype(np.float64), a, b, reg, tol=1e-8, max_iter=5000)
    assert r.niter < 5000
    assert abs(r.niter - o["niter"]) <= max(3, 0.02 * o["niter"])
    assert rel_err(r.dual, o["dual"]) < 1e-6
    assert r.mar_err_fp64 < 1e-7
    po = O.get_plan(M32.astype(np.float64), reg, o["dual"])
    assert np.max(np.abs(r.plan - po)) < 1e-6 * po.max()


def test_bcd_edge_cases(gpu_lib):
    M, a, b = readme_example(7, 5)
    # immediate convergence: huge tol -> niter 0, one history entry
    r = gpu_lib.sinkhorn_bcd(M, a, b, 0.1, tol=1e9)
    assert r.niter == 0 and len(r.obj_vals) == 1
    # max_iter exhaustion
    r = gpu_lib.sinkhorn_bcd(M, a, b, 0.1, tol=0.0, max_iter=17)
    o = O.sinkhorn_bcd(M, a, b, 0.1, tol=0.0, max_iter=17)
    assert r.niter == 17 and len(r.mar_errs) == 18
    assert rel_err(r.dual, o["dual"]) < 1e-10
    # single row / single column problems
    for shape in [(1, 6), (6, 1), (1, 1)]:
        Mx, ax, bx = readme_example(*shape)
        r = gpu_lib.sinkhorn_bcd(Mx, ax, bx, 0.1, tol=1e-10, max_iter=50)
        o = O.sinkhorn_bcd(Mx, ax, bx, 0.1, tol=1e-10, max_iter=50)
        assert r.niter == o["niter"]
        assert rel_err(r.dual, o["dual"]) < 1e-10


def test_device_fp64_exp2_accuracy(gpu_lib):
    # The table-driven FP64 2^x of the one-pass objective / gradient and
    # APDAGD kernels, against exp2 in 80-bit extended precision rounded to
    # double (a near-correctly-rounded reference), on random and special
    # arguments: at most one ulp, and rarely that.
    from regot import _lib
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-1080.0, 1030.0, 500000), rng.uniform(-2.0, 2.0, 500000),
                        np.arange(-1075.0, 1025.0, 0.25),
                        [0.0, -0.0, 1023.999, 1024.0, -1074.0, -1074.5, -1075.0, -1100.5,
                         1e300, -1e300, np.inf, -np.inf]])
    y = np.empty_like(x)
    _lib.check(_lib.lib().regot_check_exp2(_lib.dptr(x), _lib.dptr(y), len(x), 0))
    with np.errstate(over="ignore"):
        ref = np.exp2(x.astype(np.longdouble)).astype(np.float64)
    fin = np.isfinite(ref) & (ref > 0)
    ulp = np.abs(y[fin].view(np.int64) - ref[fin].view(np.int64))
    assert ulp.max() <= 1
    assert np.mean(ulp > 0) < 0.01
    assert np.all(y[~fin] == ref[~fin])
    z = np.array([np.nan])
    w = np.empty(1)
    _lib.check(_lib.lib().regot_check_exp2(_lib.dptr(z), _lib.dptr(w), 1, 0))
    assert np.isnan(w[0])


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
def test_round_plan_and_primal_value(gpu_lib, bulk):
    # approx_proj of T(gamma) and the primal value against the oracle, on a
    # ragged shape (strips and row blocks not full), at a point away from the
    # optimum (large rounding correction) and at the optimum (primal = -f*)
    from conftest import cloud_problem
    M, a, b = cloud_problem(301, 517, seed=4)
    if bulk == "fp32":
        M = M.astype(np.float32).astype(np.float64)
        Mp = M.astype(np.float32)
    else:
        Mp = M
    reg = 0.05
    o = O.sinkhorn_bcd(M, a, b, reg, tol=1e-11, max_iter=20000)
    rng = np.random.default_rng(1)
    with gpu_lib.Problem(Mp, a, b, reg) as P:
        for g in (o["dual"], o["dual"] + 0.02 * rng.standard_normal(len(o["dual"]))):
            R = P.round_plan(g)
            Ro = O.approx_proj(O.get_plan(M, reg, g), a, b)
            # the correction err_a err_b' / |err_a|_1 carries the cancellation
            # of a - rowsum: absolute agreement at 1e-16 of a unit-mass plan
            np.testing.assert_allclose(R, Ro, rtol=1e-12, atol=1e-16)
            np.testing.assert_allclose(R.sum(1), a, rtol=1e-12)
            np.testing.assert_allclose(R.sum(0), b, rtol=1e-12)
            pv = P.primal_val(g)
            assert abs(pv - O.primal_val(M, a, b, reg, g)) < 1e-12 * max(1.0, abs(pv))
        assert abs(P.primal_val(o["dual"]) + o["obj_vals"][-1]) < 1e-9


def test_device_fp64_log_accuracy(gpu_lib):
    # The table-driven FP64 log of the APDAGD plan pass (entropy term), against
    # log in 80-bit extended precision rounded to double, on log-uniform
    # arguments over the whole double range, subnormals, arguments next to 1
    # and special values: at most one ulp, and rarely that.
    from regot import _lib
    rng = np.random.default_rng(4)
    x = np.concatenate([np.exp2(rng.uniform(-1074.0, 1023.9, 500000)), rng.uniform(0.5, 2.0, 300000),
                        1.0 + rng.uniform(-1e-3, 1e-3, 100000), 1.0 + rng.uniform(-1e-12, 1e-12, 10000),
                        rng.uniform(5e-324, 2.3e-308, 50000),
                        [1.0, 2.0, 0.5, 0.75, 1.5, np.nextafter(1.0, 0.0), np.nextafter(1.0, 2.0),
                         5e-324, 2.2250738585072014e-308, 1.7976931348623157e308]])
    y = np.empty_like(x)
    _lib.check(_lib.lib().regot_check_log(_lib.dptr(x), _lib.dptr(y), len(x), 0))
    ref = np.log(x.astype(np.longdouble)).astype(np.float64)
    assert np.all(y[x == 1.0] == 0.0)
    nz = ref != 0
    ulp = np.abs(y[nz].view(np.int64) - ref[nz].view(np.int64))
    assert ulp.max() <= 1
    assert np.mean(ulp > 0) < 0.01
    z = np.array([0.0, -0.0, np.inf, -1.0, -np.inf, np.nan])
    w = np.empty_like(z)
    _lib.check(_lib.lib().regot_check_log(_lib.dptr(z), _lib.dptr(w), len(z), 0))
    assert w[0] == -np.inf and w[1] == -np.inf and w[2] == np.inf
    assert np.all(np.isnan(w[3:]))
