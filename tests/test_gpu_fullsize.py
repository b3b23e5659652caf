"""Full-size GPU runs of the benchmark configurations, checked through
size-independent properties.

At n = m = 8192 and 16384 the FP64 oracle is too slow to run, so these tests
check what holds at any size: the returned dual meets the marginals to the
tolerance when its plan is re-evaluated in FP64 on the CPU (numpy, in row
chunks, independent of the library and of the oracle); every solver lands on
the same unique optimum (same objective, same plan up to the tolerance); the
histories keep their reference semantics (niter + 1 entries, the stop test
strict at the last entry).
"""
import numpy as np
import pytest

import bench

pytestmark = pytest.mark.gpu


def cpu_marginal_error(M, a, b, reg, gamma, chunk=2048):
    """||[T1 - a; (T'1 - b)[:-1]]||, T = exp((alpha + beta - M) / reg) in FP64."""
    n, m = M.shape
    alpha = gamma[:n]
    beta = np.append(gamma[n:], 0.0)
    r = np.empty(n)
    c = np.zeros(m)
    for i0 in range(0, n, chunk):
        T = np.exp((alpha[i0:i0 + chunk, None] + beta[None, :]
                    - M[i0:i0 + chunk].astype(np.float64)) / reg)
        r[i0:i0 + chunk] = T.sum(1)
        c += T.sum(0)
    g = np.concatenate([r - a, c[:-1] - b[:-1]])
    return float(np.linalg.norm(g))


@pytest.mark.parametrize("reg", [0.1, 0.01])
def test_c2_bcd_and_ssns_8192(gpu_lib, reg):
    # sinkhorn_bcd and sinkhorn_ssns, n = m = 8192, X, Y ~ N(0, I_16),
    # M = sqeuclid / max in FP32, a, b = 0.9 Dirichlet(1) + 0.1/n, tol 1e-8
    n = 8192
    M, a, b = bench.make_problem(n, n, 16, 0)
    print("c2 n=%d reg=%g" % (n, reg), flush=True)
    with gpu_lib.Problem(M, a, b, reg) as P:
        rb = P.solve(0, tol=1e-8, max_iter=20000, compute_plan=False, verify=True)
        rs = P.solve(1, tol=1e-8, max_iter=500, compute_plan=False)
    for r in (rb, rs):
        assert r.mar_errs[-1] < 1e-8 and len(r.obj_vals) == r.niter + 1
        assert all(e >= 1e-8 for e in r.mar_errs[:-1])
        # the FP32-bulk in-loop ||g|| and its FP64 re-evaluation differ by the
        # FP32 summation error (~1e-11 here)
        err = cpu_marginal_error(M, a, b, reg, r.dual)
        assert err < 1.01e-8, err
    # the unique optimum: same objective (BCD's in-loop objective is an
    # FP32-bulk sum, ~1e-10 relative; compare its FP64 re-evaluation), and
    # duals that agree where the marginals pin them
    assert abs(rb.obj_fp64 - rs.obj_vals[-1]) < 1e-12 * max(1.0, abs(rb.obj_fp64))
    assert abs(rb.obj_vals[-1] - rs.obj_vals[-1]) < 1e-9 * max(1.0, abs(rb.obj_fp64))
    assert np.max(np.abs(rb.dual - rs.dual)) < 1e-4 * max(1.0, np.max(np.abs(rb.dual)))


def test_c3_sparse_newton_and_splr_16384(gpu_lib):
    # sinkhorn_sparse_newton and sinkhorn_splr, n = m = 16384, mixture clouds
    # in R^32, reg 0.001, density 1% and 5%, tol 1e-8
    n = 16384
    M, a, b = bench.c3_problem(n)
    reg = 0.001
    out = {}
    with gpu_lib.Problem(M, a, b, reg) as P:
        for name, sid in (("sn", 2), ("splr", 3)):
            for dens in (0.01, 0.05):
                print("c3 %s density %g" % (name, dens), flush=True)
                r = P.solve(sid, tol=1e-8, max_iter=500, compute_plan=False, density=dens)
                assert r.niter < 500 and r.mar_errs[-1] < 1e-8
                assert len(r.densities) == len(r.obj_vals) == r.niter + 1
                out[(name, dens)] = r
    f = [r.obj_vals[-1] for r in out.values()]
    assert max(f) - min(f) < 1e-10 * max(1.0, abs(f[0]))
    # FP64 re-check of one dual per solver on the CPU
    for key in (("sn", 0.01), ("splr", 0.05)):
        err = cpu_marginal_error(M, a, b, reg, out[key].dual)
        assert err < 1e-8, (key, err)
