"""The workloads of the five benchmark configurations on the GPU, each checked
against the FP64 oracle or, where the oracle cannot finish at full size, an
independent FP64 evaluation:

* README recipe, n=1000, m=800, FP64 (reference README.md:111-122 scaled):
  the drop-in sinkhorn_bcd against the oracle's run and the golden optimum;
* n=m=8192 Gaussian clouds, FP32 M, reg 0.1 and 0.01: sinkhorn_bcd against
  the oracle's FP64 run on the same (cast-back) M, dual and plan to 1e-6
  relative; sinkhorn_ssns against that optimum;
* 512 batched problems, n, m ~ U{64..4096}: every problem's marginal error
  re-evaluated in FP64 on the host, every 32nd problem against the oracle;
* 4096 x 64 BF16 clouds, reg 0.05: the loss and dL/dX of the differentiable
  pipeline against an FP64 reference built on the same (matrix-core) cost
  matrix: the certified FP64 optimum and its plan (the envelope form of the
  implicit derivative; the dense solve is singular to FP64 at this size).
"""
import os

import numpy as np
import pytest

from conftest import readme_example
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def rel(x, ref):
    return np.max(np.abs(np.asarray(x) - ref)) / max(1e-300, np.max(np.abs(ref)))


def gaussian_problem(n, d, seed):
    """X, Y ~ N(0, I_d), M = sqeuclid / max in FP32, a, b = 0.9 Dirichlet(1) +
    0.1 / n normalised."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d))
    Y = rng.standard_normal((n, d))
    M = X @ Y.T
    M *= -2.0
    M += (X * X).sum(1)[:, None]
    M += (Y * Y).sum(1)[None, :]
    np.maximum(M, 0.0, out=M)
    M /= M.max()
    M32 = M.astype(np.float32)
    del M
    a = 0.9 * rng.dirichlet(np.ones(n)) + 0.1 / n
    b = 0.9 * rng.dirichlet(np.ones(n)) + 0.1 / n
    return M32, a / a.sum(), b / b.sum()


def fp64_marginal_error(M, reg, dual, a, b, rows=1024):
    """||[T 1 - a; (T' 1 - b)_{0..m-2}]|| in FP64 numpy, in row chunks."""
    n, m = M.shape
    alpha, beta = dual[:n], np.append(dual[n:], 0.0)
    r = np.empty(n)
    c = np.zeros(m)
    for i0 in range(0, n, rows):
        T = np.exp((alpha[i0:i0 + rows, None] + beta[None, :]
                    - M[i0:i0 + rows].astype(np.float64)) / reg)
        r[i0:i0 + rows] = T.sum(1)
        c += T.sum(0)
    return np.sqrt(np.sum((r - a) ** 2) + np.sum((c[:-1] - b[:-1]) ** 2))


def gauge_rel(x, ref, n):
    """Relative difference of two duals after removing the best shift along
    v = [1_n; -1_{m-1}] (alpha + c, beta - c).  That direction moves only the
    plan's last column, whose marginal the stop test leaves out, so two duals
    at ||g|| < tol can differ along it by O(tol reg / b_{m-1})."""
    d = np.asarray(x) - ref
    v = np.concatenate([np.ones(n), -np.ones(len(ref) - n)])
    c = (d @ v) / (v @ v)
    return np.max(np.abs(d - c * v)) / np.max(np.abs(ref))


def max_plan_diff(M, reg, d1, d2, rows=1024, last_col=True):
    """max |T(d1) - T(d2)| / max T(d2), FP64, in row chunks; last_col=False
    leaves out column m-1 (see gauge_rel)."""
    n, m = M.shape
    a1, b1 = d1[:n], np.append(d1[n:], 0.0)
    a2, b2 = d2[:n], np.append(d2[n:], 0.0)
    diff = top = 0.0
    for i0 in range(0, n, rows):
        Mc = M[i0:i0 + rows].astype(np.float64)
        T1 = np.exp((a1[i0:i0 + rows, None] + b1[None, :] - Mc) / reg)
        T2 = np.exp((a2[i0:i0 + rows, None] + b2[None, :] - Mc) / reg)
        if not last_col:
            T1, T2 = T1[:, :-1], T2[:, :-1]
        diff = max(diff, np.abs(T1 - T2).max())
        top = max(top, T2.max())
    return diff / top


# --- README recipe, n = 1000, m = 800, FP64 ------------------------------------

def test_readme_recipe_n1000_m800_fp64(gpu_lib):
    M, a, b = readme_example(1000, 800)
    print("readme recipe 1000x800 fp64", flush=True)
    r = gpu_lib.sinkhorn_bcd(M, a, b, 0.1, tol=1e-8, max_iter=10000)
    o = O.sinkhorn_bcd(M, a, b, 0.1, tol=1e-8, max_iter=10000)
    assert abs(r.niter - o["niter"]) <= 1 and r.mar_errs[-1] < 1e-8
    k = min(len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-12)
    np.testing.assert_allclose(r.mar_errs[:k], o["mar_errs"][:k], rtol=1e-7)
    assert rel(r.dual, o["dual"]) < 1e-10
    # to the golden optimum (an independent scipy Sinkhorn, converged)
    z = np.load(os.path.join(HERE, "golden", "c1_readme_n1000_m800.npz"))
    r = gpu_lib.sinkhorn_bcd(M, a, b, 0.1, tol=1e-12, max_iter=10000)
    assert r.mar_errs[-1] < 1e-12
    np.testing.assert_allclose(r.dual, z["dual"], rtol=0, atol=1e-9)
    assert abs(r.obj_vals[-1] - float(z["obj"])) < 1e-12


# --- n = m = 8192 Gaussian clouds, FP32 M ------------------------------------

@pytest.fixture(scope="module")
def c2_problem():
    return gaussian_problem(8192, 16, 0)


@pytest.mark.parametrize("reg", [0.1, 0.01])
def test_gaussian_n8192_bcd_and_ssns_against_the_oracle(gpu_lib, c2_problem, reg):
    M32, a, b = c2_problem
    M64 = M32.astype(np.float64)
    print("gaussian 8192^2 reg %g: oracle bcd" % reg, flush=True)
    o = O.sinkhorn_bcd(M64, a, b, reg, tol=1e-8, max_iter=1000)
    assert o["mar_errs"][-1] < 1e-8
    print("gaussian 8192^2 reg %g: gpu bcd (niter oracle %d)" % (reg, o["niter"]), flush=True)
    r = gpu_lib.sinkhorn_bcd(M32, a, b, reg, tol=1e-8, max_iter=1000, compute_plan=False,
                             verify=True)
    assert abs(r.niter - o["niter"]) <= 1
    assert r.mar_errs[-1] < 1e-8 and r.mar_err_fp64 < 1.05e-8
    # in-loop FP32-bulk objectives: ~1e-9 relative sum error
    k = min(len(r.obj_vals), len(o["obj_vals"]))
    np.testing.assert_allclose(r.obj_vals[:k], o["obj_vals"][:k], rtol=1e-8)
    assert rel(r.dual, o["dual"]) < 1e-6
    assert max_plan_diff(M32, reg, r.dual, o["dual"]) < 1e-6
    assert fp64_marginal_error(M32, reg, r.dual, a, b) < 1.05e-8
    print("gaussian 8192^2 reg %g: gpu ssns" % reg, flush=True)
    s = gpu_lib.sinkhorn_ssns(M32, a, b, reg, tol=1e-8, max_iter=500, compute_plan=False)
    assert s.niter < 500 and s.mar_errs[-1] < 1e-8
    assert fp64_marginal_error(M32, reg, s.dual, a, b) < 1.05e-8
    # The unique optimum, both at ||g|| < 1e-8.  Two different solvers stop
    # at different points of that tolerance ball: the duals differ by up to
    # ||g|| / lambda_min(H) (lambda_min ~ min a / reg; freely along the
    # near-flat shift [1; -1], which moves only the last column, see
    # gauge_rel) and the plans by that over reg, while the objective, flat at
    # the optimum, agrees to second order.  (SSNS against its own FP64 oracle
    # run, trajectory by trajectory, is in test_gpu_newton.py.)
    n = M32.shape[0]
    pd = max_plan_diff(M32, reg, s.dual, o["dual"], last_col=False)
    print("ssns against bcd: dual %.2e, after the shift %.2e, plan %.2e"
          % (rel(s.dual, o["dual"]), gauge_rel(s.dual, o["dual"], n), pd), flush=True)
    assert pd < 1e-6 / reg
    assert gauge_rel(s.dual, o["dual"], n) < 1e-4
    assert abs(s.obj_vals[-1] - o["obj_vals"][-1]) < 1e-9 * abs(o["obj_vals"][-1])


# --- 512 batched problems, n, m ~ U{64..4096} ---------------------------------

def batched_recipe(P, seed=0):
    """P problems, n, m ~ U{64..4096}, X, Y ~ N(0, I_8), M = sqeuclid / max in
    FP32, a, b = 0.9 Dirichlet(1) + 0.1 uniform, reg ~ logU[0.005, 0.1]."""
    rng = np.random.default_rng(seed)
    ns = rng.integers(64, 4097, size=P)
    ms = rng.integers(64, 4097, size=P)
    regs = np.exp(rng.uniform(np.log(0.005), np.log(0.1), size=P))
    Ms, As, Bs = [], [], []
    for p in range(P):
        n, m = int(ns[p]), int(ms[p])
        X = rng.standard_normal((n, 8))
        Y = rng.standard_normal((m, 8))
        M = (X * X).sum(1)[:, None] + (Y * Y).sum(1)[None, :] - 2.0 * (X @ Y.T)
        np.maximum(M, 0.0, out=M)
        M /= M.max()
        Ms.append(M.astype(np.float32))
        a = 0.9 * rng.dirichlet(np.ones(n)) + 0.1 / n
        b = 0.9 * rng.dirichlet(np.ones(m)) + 0.1 / m
        As.append(a / a.sum())
        Bs.append(b / b.sum())
    return ns, ms, regs, Ms, As, Bs


def test_batched_512_problems(gpu_lib):
    P, tol = 512, 1e-6
    print("batched 512: generating", flush=True)
    ns, ms, regs, Ms, As, Bs = batched_recipe(P)
    print("batched 512: solving (%.2e elements)" % float((ns * ms).sum()), flush=True)
    out = gpu_lib.sinkhorn_bcd_batch(Ms, As, Bs, regs, tol=tol, max_iter=100000)
    assert np.all(out["mar_err"] < tol)
    print("batched 512: FP64 re-evaluation", flush=True)
    worst = 0.0
    for p in range(P):
        e = fp64_marginal_error(Ms[p], regs[p], out["dual"][p], As[p], Bs[p])
        worst = max(worst, e / tol)
        # the FP32 sums resolve each marginal to ~1e-7 of its size
        assert e < tol + 1e-7 * np.sqrt(np.sum(As[p] ** 2) + np.sum(Bs[p] ** 2)), (p, e)
    print("batched 512: worst FP64 marginal error / tol = %.3f" % worst, flush=True)
    # every 32nd problem against the oracle (about 15 s of FP64 CPU work)
    compared = 0
    for p in range(0, P, 32):
        M64 = Ms[p].astype(np.float64)
        o = O.sinkhorn_bcd(M64, As[p], Bs[p], regs[p], tol=tol, max_iter=100000)
        assert abs(int(out["niter"][p]) - o["niter"]) <= max(1, 0.01 * o["niter"]), p
        # FP32-bulk objective: reg sum T in FP32 sums (~1e-8 of reg)
        assert abs(out["obj"][p] - o["obj_vals"][-1]) < 1e-7 * regs[p], p
        assert max_plan_diff(Ms[p], regs[p], out["dual"][p], o["dual"]) < 1e-4, p
        compared += 1
    assert compared == P // 32


# --- 4096 x 64 BF16 clouds, reg 0.05: differentiable loss ---------------------

def test_differentiable_bf16_4096x64(gpu_lib):
    torch = pytest.importorskip("torch")
    from implicit_ref import dual_obj_grad, loss_value, plan
    from regot.autograd import sinkhorn_loss, sqeuclid
    n = m = 4096
    d, reg = 64, 0.05
    g = torch.Generator().manual_seed(0)
    X = torch.randn(n, d, generator=g).to(torch.bfloat16).to("cuda:0").requires_grad_(True)
    Y = torch.randn(m, d, generator=g).to(torch.bfloat16).to("cuda:0").requires_grad_(True)
    print("differentiable 4096x64 bf16: forward + backward", flush=True)
    info = {}
    L = sinkhorn_loss(X, Y, reg=reg, tol=1e-8, max_iter=3000, solver="ssns", info=info)
    L.backward()
    assert info["mar_err"] < 1e-8
    Mt = sqeuclid(X, Y)
    torch.cuda.synchronize()
    M = Mt.cpu().numpy().astype(np.float64)
    a = np.full(n, 1.0 / n)
    b = np.full(m, 1.0 / m)
    # FP64 optimum on the same cost matrix: FP64-bulk SSNS, certified by an
    # FP64 numpy evaluation of its marginals (the optimum is unique)
    print("differentiable 4096x64 bf16: FP64 optimum", flush=True)
    ref = gpu_lib.sinkhorn_ssns(M, a, b, reg, tol=1e-11, max_iter=3000, compute_plan=False)
    f_ref, g_ref, _ = dual_obj_grad(M, a, b, reg, ref.dual)
    assert np.linalg.norm(g_ref) < 1e-10
    # dL/dM of the entropic loss at the optimum is the plan itself (the
    # envelope theorem; the dense implicit-diff solve reproduces it at sizes
    # where the dual Hessian is well conditioned -- test_implicit_ref.py).
    # Here the plan is nearly a permutation (reg 0.05 against squared
    # distances ~128), the (n+m-1)^2 Hessian is singular to FP64, and a dense
    # solve returns rounding noise, so the reference uses the envelope form.
    L_ref = loss_value(M, reg, ref.dual, "entropic")
    G_ref = plan(M, reg, ref.dual)
    assert abs(L.item() - L_ref) < 1e-7 * abs(L_ref)
    assert abs(L_ref + f_ref) < 1e-9 * abs(L_ref)            # L* = -f* at the optimum
    Xd = X.detach().double().cpu().numpy()
    Yd = Y.detach().double().cpu().numpy()
    gX_ref = 2.0 * (G_ref.sum(1)[:, None] * Xd - G_ref @ Yd)
    # the GPU adjoint dL/dM in FP64, at the forward's dual (off the optimum by
    # the forward's tolerance) and at the reference optimum (the adjoint alone)
    P = gpu_lib.Problem.from_device(Mt.data_ptr(), n, m, np.float32, a, b, reg)
    G = torch.empty(n, m, dtype=torch.float64, device="cuda:0")
    errs = []
    for gamma in (info["gamma"], ref.dual):
        torch.cuda.synchronize()
        P.implicit_grad(gamma, G.data_ptr(), np.float64, m, loss="entropic", cg_tol=1e-10)
        Gh = G.cpu().numpy()
        errs.append(rel(2.0 * (Gh.sum(1)[:, None] * Xd - Gh @ Yd), gX_ref))
    P.close()
    err_bf16 = rel(X.grad.double().cpu().numpy(), gX_ref)
    print("differentiable: loss %.10g (ref %.10g), dL/dX rel err: forward dual %.2e, reference "
          "dual %.2e, BF16 output %.2e" % (L.item(), L_ref, errs[0], errs[1], err_bf16),
          flush=True)
    assert errs[1] < 1e-9            # the adjoint itself
    assert errs[0] < 1e-3            # plus the forward's stop at ||g|| < 1e-8
    assert err_bf16 < 1e-2           # plus BF16 rounding of the gradient
