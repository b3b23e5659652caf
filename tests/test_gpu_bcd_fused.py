"""GPU parity of the single-read BCD iteration (one pass over M per iteration).

The FP32 BCD driver runs the single-read form wherever it is eligible
(``bcd_fused``, default 1); ``bcd_fused=0`` keeps the two-pass form.  Checked here against the FP64 oracle (same iteration
semantics, same stop test) and against the two-pass GPU path, on shapes that
cover every column-vector count (m up to 16384, full-width and ragged), the
rows whose range check falls back to an exact row maximum, a warm start, the
eligibility fallback and workgroups holding more rows than threads.
"""
import numpy as np
import pytest

from conftest import cloud_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_err(x, ref):
    return np.max(np.abs(np.asarray(x) - ref)) / max(1.0, np.max(np.abs(ref)))


def fused_expected(r):
    """Single-read passes a solve runs: one per iteration up to the iterate
    where an FP32 solve hands over to its FP64 finish, that one included."""
    pf = r.stats["precise_from"]
    return (pf if pf >= 0 else r.niter) + 1


def solve(gpu_lib, M, a, b, reg, tol, max_iter, fused, x0=None):
    """BCD through a problem handle with kernel profiling on; returns the
    result and the number of single-read passes that ran."""
    kw = {"bcd_fused": fused}
    if x0 is not None:
        kw["x0"] = x0
    with gpu_lib.Problem(M, a, b, reg) as P:
        P.profile(True)
        r = P.solve(0, tol=tol, max_iter=max_iter, compute_plan=False, verify=True, **kw)
        fused_launches = P.profile_stats(2)[0]
    return r, fused_launches


# (n, m, reg): m covers NC = ceil(m / 2048) = 1, 3, 7, 8, full-width and ragged
SHAPES = [(50, 1500, 0.05), (130, 2048, 0.05), (257, 5000, 0.02), (301, 12289, 0.05),
          (64, 16384, 0.05), (90, 14500, 0.05)]


@pytest.mark.parametrize("n,m,reg", SHAPES)
def test_fused_bcd_matches_oracle(gpu_lib, n, m, reg):
    M, a, b = cloud_problem(n, m, seed=n + m)
    M32 = M.astype(np.float32)
    print("fused bcd %dx%d reg=%g" % (n, m, reg), flush=True)
    r, nf = solve(gpu_lib, M32, a, b, reg, 1e-8, 5000, fused=1)
    o = O.sinkhorn_bcd(M32.astype(np.float64), a, b, reg, tol=1e-8, max_iter=5000)
    assert nf == fused_expected(r)    # every FP32 iteration used the single read
    assert r.niter < 5000 and abs(r.niter - o["niter"]) <= max(2, 0.02 * o["niter"])
    assert len(r.obj_vals) == r.niter + 1
    assert r.mar_errs[-1] < 1e-8 and all(e >= 1e-8 for e in r.mar_errs[:-1])
    k = min(10, len(o["obj_vals"]), len(r.obj_vals))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-9)
    assert rel_err(r.dual, o["dual"]) < 1e-6
    assert r.mar_err_fp64 < 1.05e-8


def test_fused_matches_two_pass(gpu_lib):
    M, a, b = cloud_problem(400, 9000, seed=4)
    M32 = M.astype(np.float32)
    print("fused vs two-pass 400x9000", flush=True)
    rf, nf = solve(gpu_lib, M32, a, b, 0.02, 1e-8, 5000, fused=1)
    rt, nt = solve(gpu_lib, M32, a, b, 0.02, 1e-8, 5000, fused=0)
    assert nf > 0 and nt == 0
    assert abs(rf.niter - rt.niter) <= 1
    assert abs(rf.obj_fp64 - rt.obj_fp64) < 1e-12 * abs(rt.obj_fp64)
    assert rel_err(rf.dual, rt.dual) < 1e-7
    assert rf.mar_err_fp64 < 1.05e-8 and rt.mar_err_fp64 < 1.05e-8


def test_fused_row_range_fallback(gpu_lib):
    # Row 0 is free (M = 0) but carries a_0 = 1e-7: in the first iteration its
    # row sum is ~1e7 a_0.  Row 1 costs 1 everywhere at reg 0.01: its terms
    # are ~2^-144, below FP32 range.  Both rows leave [2^-20, 2^20] and take
    # the exact-maximum path.
    n, m, reg = 40, 3000, 0.01
    M, a, b = cloud_problem(n, m, seed=21)
    M[0, :] = 0.0
    M[1, :] = 1.0
    a[0] = 1e-7
    a /= a.sum()
    M32 = M.astype(np.float32)
    print("fused bcd range fallback", flush=True)
    r, nf = solve(gpu_lib, M32, a, b, reg, 1e-6, 20000, fused=1)
    o = O.sinkhorn_bcd(M32.astype(np.float64), a, b, reg, tol=1e-6, max_iter=20000)
    assert nf > 0 and np.all(np.isfinite(r.dual))
    assert abs(r.niter - o["niter"]) <= max(2, 0.02 * o["niter"])
    # the in-loop objective sums FP32 terms: ~1e-7 relative of reg sum r
    np.testing.assert_allclose(r.obj_vals[:5], o["obj_vals"][:5], rtol=0, atol=1e-7 * reg)
    assert rel_err(r.dual, o["dual"]) < 1e-6
    assert r.mar_err_fp64 < 1.05e-6


def test_fused_warm_start(gpu_lib):
    # with x0 the first iteration keeps the two-pass form (its beta is not
    # the result of a beta update), the rest run single-read
    M, a, b = cloud_problem(200, 6000, seed=5)
    M32 = M.astype(np.float32)
    r0, _ = solve(gpu_lib, M32, a, b, 0.03, 1e-3, 5, fused=0)
    x0 = np.asarray(r0.dual) + 0.01
    rf, nf = solve(gpu_lib, M32, a, b, 0.03, 1e-8, 5000, fused=1, x0=x0)
    rt, _ = solve(gpu_lib, M32, a, b, 0.03, 1e-8, 5000, fused=0, x0=x0)
    assert nf == rf.niter           # all but iteration 0
    assert abs(rf.niter - rt.niter) <= 1 and rel_err(rf.dual, rt.dual) < 1e-7
    assert rf.mar_err_fp64 < 1.05e-8


def test_fused_ineligible_marginals_fall_back(gpu_lib):
    # min a min b below 2^-60 sum a sum b: the single-read pass is not used,
    # and the result is the two-pass one bit for bit
    M, a, b = cloud_problem(100, 3000, seed=6)
    a[3] = 1e-40
    a /= a.sum()
    M32 = M.astype(np.float32)
    rf, nf = solve(gpu_lib, M32, a, b, 0.05, 1e-8, 5000, fused=1)
    rt, _ = solve(gpu_lib, M32, a, b, 0.05, 1e-8, 5000, fused=0)
    assert nf == 0
    assert rf.niter == rt.niter and np.array_equal(rf.dual, rt.dual)


def test_fused_many_rows_per_workgroup(gpu_lib):
    # 140000 rows over at most 256 workgroups: > 512 rows each, more rows than
    # threads, so every thread finishes several rows in the epilogue
    n, m, reg = 140000, 1100, 0.05
    rng = np.random.default_rng(7)
    X = rng.standard_normal((n, 8)).astype(np.float32)
    Y = rng.standard_normal((m, 8)).astype(np.float32)
    M = (X * X).sum(1)[:, None] + (Y * Y).sum(1)[None, :] - 2.0 * (X @ Y.T)
    np.maximum(M, 0.0, out=M)
    M /= M.max()
    a = 0.9 * rng.dirichlet(np.ones(n)) + 0.1 / n
    b = 0.9 * rng.dirichlet(np.ones(m)) + 0.1 / m
    a /= a.sum()
    b /= b.sum()
    print("fused bcd 140000x1100", flush=True)
    rf, nf = solve(gpu_lib, M, a, b, reg, 1e-7, 2000, fused=1)
    rt, _ = solve(gpu_lib, M, a, b, reg, 1e-7, 2000, fused=0)
    assert nf == rf.niter + 1
    assert abs(rf.niter - rt.niter) <= 1
    assert rel_err(rf.dual, rt.dual) < 1e-6
    assert rf.mar_err_fp64 < 1.05e-7


def test_fused_edge_shapes(gpu_lib):
    # one row; three rows over the full 16384 columns (one row per workgroup);
    # max_iter = 0 and an immediately met tolerance: the single-read pass
    # still runs the stop test of iteration 0.  (Rows carry a_i ~ 1 / n here:
    # FP32-bulk row sums resolve their marginal to ~1e-7, hence tol 1e-6.)
    for n, m in ((1, 3000), (3, 16384)):
        M, a, b = cloud_problem(n, m, seed=n)
        M32 = M.astype(np.float32)
        print("fused edge %dx%d" % (n, m), flush=True)
        r, nf = solve(gpu_lib, M32, a, b, 0.05, 1e-6, 200, fused=1)
        o = O.sinkhorn_bcd(M32.astype(np.float64), a, b, 0.05, tol=1e-6, max_iter=200)
        assert nf == r.niter + 1 and abs(r.niter - o["niter"]) <= 1
        assert rel_err(r.dual, o["dual"]) < 1e-6
    M, a, b = cloud_problem(50, 2500, seed=3)
    M32 = M.astype(np.float32)
    for tol, max_iter in ((1e-8, 0), (1e9, 100)):
        r, nf = solve(gpu_lib, M32, a, b, 0.05, tol, max_iter, fused=1)
        o = O.sinkhorn_bcd(M32.astype(np.float64), a, b, 0.05, tol=tol, max_iter=max_iter)
        assert r.niter == o["niter"] == 0 and nf == 1 and len(r.obj_vals) == 1
        np.testing.assert_allclose(r.obj_vals, o["obj_vals"], rtol=1e-9)
        assert rel_err(r.dual, o["dual"]) < 1e-6


@pytest.mark.parametrize("n,m,reg,tol,uniform", [(1, 3000, 0.05, 1e-9, False),
                                                 (3, 16384, 0.05, 1e-9, False),
                                                 (50, 2500, 0.05, 1e-10, False),
                                                 (1200, 1100, 0.01, 1e-10, True)])
def test_fp32_bulk_finishes_below_its_resolution(gpu_lib, n, m, reg, tol, uniform):
    # FP32 sums resolve each marginal to ~2^-23 of its size; below that the
    # FP32 loop hands over to FP64 passes over the same FP32 M and goes on to
    # tol, so the iteration count and the dual match the FP64 oracle's.
    M, a, b = cloud_problem(n, m, d=4 if uniform else 16, seed=n + 3)
    if uniform:
        a, b = np.full(n, 1.0 / n), np.full(m, 1.0 / m)
    M32 = M.astype(np.float32)
    print("fp32 finish %dx%d tol %g" % (n, m, tol), flush=True)
    r, nf = solve(gpu_lib, M32, a, b, reg, tol, 5000, fused=1)
    o = O.sinkhorn_bcd(M32.astype(np.float64), a, b, reg, tol=tol, max_iter=5000)
    pf = r.stats["precise_from"]
    assert 0 <= pf <= r.niter and nf == fused_expected(r)
    assert abs(r.niter - o["niter"]) <= 1
    assert r.mar_errs[-1] < tol and r.mar_err_fp64 < 1.05 * tol
    assert len(r.obj_vals) == len(r.mar_errs) == len(r.run_times) == r.niter + 1
    assert all(e >= tol for e in r.mar_errs[:-1])
    # FP32-bulk objectives carry ~1e-9 relative sum error; from the handover
    # on they are FP64 sums along a trajectory within ~1e-8 of the oracle's
    k = min(len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-8)
    np.testing.assert_allclose(r.obj_vals[pf:k], o["obj_vals"][pf:k], rtol=1e-11)
    assert rel_err(r.dual, o["dual"]) < 1e-8
