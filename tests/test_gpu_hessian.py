"""GPU parity of the sparsified Hessian, its products and the CG solve.

FP64 bulk: the kept sets must equal the oracle's exactly (random inputs have
no ties) and values agree to 1e-12.  FP32 bulk: the rule's invariants hold and
Hessian products agree with the oracle's to the FP32 value accuracy.
"""
import numpy as np
import pytest

from conftest import cloud_problem
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def point(M, a, b, reg, iters=3, seed=0, noise=0.02):
    """A dual point near a few BCD iterations, perturbed."""
    n, m = M.shape
    alpha = np.zeros(n)
    beta = O.optimal_beta(M, b, reg, alpha)
    for _ in range(iters):
        alpha = O.optimal_alpha(M, a, reg, beta)
        beta = O.optimal_beta(M, b, reg, alpha)
    g = np.concatenate([alpha + beta[-1], beta[:-1] - beta[-1]])
    return g + noise * reg * np.random.default_rng(seed).standard_normal(g.shape)


def csc_dense(n, m, colptr, rowind, val):
    return O.csc_to_dense(n, m, colptr, rowind, val)


@pytest.mark.parametrize("hint", [0.001, 0.5])
@pytest.mark.parametrize("reg", [0.05, 0.001])
def test_threshold_rule_fp64_matches_oracle(gpu_lib, hint, reg):
    n, m = 150, 121
    M, a, b = cloud_problem(n, m, seed=5)
    gamma = point(M, a, b, reg)
    f, g = O.dual_obj_grad(M, a, b, reg, gamma)
    delta = 0.01 * np.linalg.norm(g)
    T, _ = O.compute_T(M, reg, gamma)
    cp, ri, va = O.sparsify_threshold(T, delta, hint)
    with gpu_lib.Problem(M, a, b, reg, bulk="fp64") as P:
        nnz = P.sparsified_hess(gamma, delta, hint)
        stored, zeros, dens = P.hess_info()
        H = P.hess_export()
    # the reference counts kept exact zeros of T as stored entries
    assert nnz == len(ri)
    assert stored + zeros == len(ri)
    Sg = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"])
    So = csc_dense(n, m, cp, ri, va) / reg
    kept_o = np.zeros((n, m - 1), bool)
    for j in range(m - 1):
        kept_o[ri[cp[j]:cp[j + 1]], j] = True
    kept_g = Sg != 0
    assert np.array_equal(kept_g, kept_o & (So != 0))
    np.testing.assert_allclose(Sg, So, rtol=1e-12, atol=0)
    N = n + m - 1
    assert abs(dens - (2.0 * len(ri) + N) / N / N) < 1e-15
    np.testing.assert_allclose(H["h1"], (g[:n] + a) / reg, rtol=1e-12)
    np.testing.assert_allclose(H["h2"], (g[n:] + b[:-1]) / reg, rtol=1e-12)
    # CSR holds the same matrix
    Sr = np.zeros((n, m - 1))
    for i in range(n):
        Sr[i, H["colind"][H["rowptr"][i]:H["rowptr"][i + 1]]] = H["rval"][H["rowptr"][i]:H["rowptr"][i + 1]]
    assert np.array_equal(Sr, Sg)


def test_threshold_rule_fp32_invariants(gpu_lib):
    n, m, reg = 300, 257, 0.02
    M, a, b = cloud_problem(n, m, seed=9)
    M32 = M.astype(np.float32)
    Mo = M32.astype(np.float64)
    gamma = point(Mo, a, b, reg)
    f, g = O.dual_obj_grad(Mo, a, b, reg, gamma)
    delta = 0.01 * np.linalg.norm(g)
    T, _ = O.compute_T(Mo, reg, gamma)
    with gpu_lib.Problem(M32, a, b, reg) as P:
        P.sparsified_hess(gamma, delta, 0.001)
        H = P.hess_export()
    S = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"]) * reg
    kept = S != 0
    dropped = np.where(kept, 0.0, T[:, :-1])
    # per-row dropped mass <= 2 delta (row rule), per column the threshold rule
    assert np.all(dropped.sum(1) <= 2 * delta * (1 + 1e-5) + 1e-300)
    cp, ri, va = O.sparsify_threshold(T, delta, 0.001)
    So = csc_dense(n, m, cp, ri, va)
    diff = np.abs(S - So)
    # entries may switch sides only at the threshold boundary: small mass
    assert diff.sum() <= 4 * delta * (m - 1) * 1e-3 + 1e-6 * So.sum()
    assert np.max(np.abs(S[kept] - T[:, :-1][kept]) / T[:, :-1][kept]) < 2e-6


def column_tau(S, delta):
    """Per column the smallest value v with sum_{S_ij <= v} S_ij > delta
    (+inf when the column's mass never passes delta)."""
    St = np.sort(S, axis=0)
    C = np.cumsum(St, axis=0)
    hit = C > delta
    idx = hit.argmax(axis=0)
    return np.where(hit.any(axis=0), St[idx, np.arange(S.shape[1])], np.inf)


@pytest.mark.parametrize("n,m,side", [(300, 257, "le"), (200, 2100, "gt"), (64, 12001, "gt"),
                                      (40, 16500, "gt")])
@pytest.mark.parametrize("hint", [0.001, 0.5])
def test_threshold_rule_fp32_exact_on_device_values(gpu_lib, n, m, side, hint):
    # FP32 bulk builds tau from row-major M and the row side in one pass (m -
    # 1 <= 16384; wider: the per-row kernel).  The rule is checked exactly on
    # the device's own T: every nonzero sigma_ij = T_ij / reg comes from the
    # full Hessian (top-k at density 1), and the oracle's threshold rule runs
    # on that matrix with delta / reg.
    reg = 0.02
    M, a, b = cloud_problem(n, m, seed=n + m)
    M32 = M.astype(np.float32)
    gamma = point(M32.astype(np.float64), a, b, reg, iters=2, seed=n)
    print(f"fp32 threshold {n}x{m} hint {hint}", flush=True)
    with gpu_lib.Problem(M32, a, b, reg) as P:
        _, g = P.dual_obj_grad(gamma)
        delta = 0.01 * np.linalg.norm(g)
        P.sparsified_hess_with_density(gamma, 1.0)
        F = P.hess_export()
        Sf = np.zeros((n, m - 1))
        for i in range(n):
            sl = slice(F["rowptr"][i], F["rowptr"][i + 1])
            Sf[i, F["colind"][sl]] = F["rval"][sl]
        nnz = P.sparsified_hess(gamma, delta, hint)
        stored, zeros, _ = P.hess_info()
        H = P.hess_export()
    Sg = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"])
    Sp = np.concatenate([Sf, np.zeros((n, 1))], axis=1)
    cp, ri, va = O.sparsify_threshold(Sp, delta / reg, hint)
    kept_o = np.zeros((n, m - 1), bool)
    for j in range(m - 1):
        kept_o[ri[cp[j]:cp[j + 1]], j] = True
    kept_o &= Sf != 0
    assert np.array_equal(Sg != 0, kept_o)
    assert np.array_equal(Sg[kept_o], Sf[kept_o])   # same values, bit for bit
    assert stored == int(kept_o.sum()) and nnz == stored + zeros
    # the cases cover both sides of the row rule: rows whose dropped mass
    # under the column thresholds alone stays <= 2 delta (rho_i = +inf), and
    # rows above it (rho_i from the radix select)
    tau = column_tau(Sf, delta / reg)
    D = np.where(Sf < tau[None, :], Sf, 0.0).sum(1)
    if hint < 0.2:
        assert np.any(D <= 2 * delta / reg) if side == "le" else np.any(D > 2 * delta / reg)


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
@pytest.mark.parametrize("density", [0.0, 0.01, 0.1, 1.0])
def test_topk_rule(gpu_lib, bulk, density):
    n, m, reg = 97, 131, 0.03
    M, a, b = cloud_problem(n, m, seed=3)
    Mg = M.astype(np.float32) if bulk == "fp32" else M
    Mo = Mg.astype(np.float64)
    gamma = point(Mo, a, b, reg)
    T, _ = O.compute_T(Mo, reg, gamma)
    cp, ri, va = O.sparsify_density(T, density, True)
    with gpu_lib.Problem(Mg, a, b, reg) as P:
        nnz = P.sparsified_hess_with_density(gamma, density)
        H = P.hess_export()
    So = csc_dense(n, m, cp, ri, va)
    Sg = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"]) * reg
    k = int(n * (m - 1) * density)
    if bulk == "fp64":
        assert nnz == len(ri)
        assert np.array_equal(Sg != 0, So != 0)
        np.testing.assert_allclose(Sg, So, rtol=1e-12)
    else:
        # FP32 values may reorder near the k-th value; the count is exact
        assert nnz == len(ri)
        assert np.sum((Sg != 0) != (So != 0)) <= max(4, 0.002 * k)
    # preserve: row 0 and column 0 always stored
    assert np.all(Sg[0, :] != 0) and np.all(Sg[:, 0] != 0)


def topk_ties_expected(Mq, reg, density):
    """Kept set of the top-k rule with ties at the k-th value taken in
    row-major order (preserve on), for T = exp(-M / reg) at gamma = 0 with M
    on a few levels, so that whole levels tie exactly."""
    n, m = Mq.shape
    B = Mq[:, :-1]
    k = int(n * (m - 1) * density)
    kept = np.zeros(B.shape, bool)
    if k > 0:
        levels = np.unique(B)            # ascending M = descending T
        before = 0
        for lv in levels:
            cnt = int(np.sum(B == lv))
            if before + cnt >= k:
                kept |= B < lv
                take = k - before
                tie = np.flatnonzero((B == lv).ravel())[:take]   # row-major ranks
                kept.ravel()[tie] = True
                break
            before += cnt
    kept[0, :] = True
    kept[:, 0] = True
    return kept


@pytest.mark.parametrize("top_log2", [-124.0, -130.0])
def test_topk_rule_values_near_fp32_subnormals(gpu_lib, top_log2):
    # FP32 bulk, four cost levels scaled so the largest T is 2^top_log2: at
    # -124 the levels are 2^-124, 2^-124.7 (the k-th value), 2^-125.4 (the
    # lowest normal binade) and 2^-126.2 (subnormal), so the CSR writer may
    # use the raw exponential, while the preserved row 0 and column 0 keep
    # subnormal entries; at -130 every value is subnormal and the writer runs
    # with the exact exponential.  m > 4096 puts columns past the first step
    # of every pass.
    n, m, reg, density = 40, 5000, 0.5, 0.4
    rng = np.random.default_rng(71)
    Mq = 0.25 * rng.integers(0, 4, size=(n, m)).astype(np.float64)
    a = np.full(n, 1.0 / n)
    b = np.full(m, 1.0 / m)
    c = -reg * top_log2 * np.log(2.0)       # alpha_i = -c: T = 2^top_log2 exp(-M / reg)
    gamma = np.concatenate([np.full(n, -c), np.zeros(m - 1)])
    with gpu_lib.Problem(Mq.astype(np.float32), a, b, reg) as P:
        P.set_hess_storage(0)
        print(f"top-k near subnormals: top 2^{top_log2}", flush=True)
        nnz = P.sparsified_hess_with_density(gamma, density)
        H = P.hess_export()
    kept = topk_ties_expected(Mq, reg, density)
    Sg = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"])
    assert nnz == int(kept.sum())
    assert np.array_equal(Sg != 0, kept)
    T = np.exp(-c / reg - Mq[:, :-1] / reg) / reg
    # the FP32 exponent argument near -128 is rounded to 2^-17: up to
    # ln2 * 2^-17 = 5.3e-6 relative; FP32 subnormals carry fewer bits (2^-149
    # absolute against values >= 2^-133)
    np.testing.assert_allclose(Sg[kept], T[kept], rtol=3e-3 if top_log2 < -126 else 1e-5)


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
@pytest.mark.parametrize("n,m,density", [(150, 121, 0.3), (400, 300, 0.005), (64, 50, 1.0)])
def test_topk_rule_ties(gpu_lib, bulk, n, m, density):
    # four cost levels: the k-th value falls inside a class of exact ties;
    # the sparse case builds the CSC by the counting sort, the dense ones by
    # column strips that look ties up in the CSR
    rng = np.random.default_rng(n + m)
    Mq = 0.25 * rng.integers(0, 4, size=(n, m)).astype(np.float64)
    a = np.full(n, 1.0 / n)
    b = np.full(m, 1.0 / m)
    reg = 0.5
    Mg = Mq.astype(np.float32) if bulk == "fp32" else Mq
    gamma = np.zeros(n + m - 1)
    with gpu_lib.Problem(Mg, a, b, reg) as P:
        nnz = P.sparsified_hess_with_density(gamma, density)
        H = P.hess_export()
    kept = topk_ties_expected(Mq, reg, density)
    Sg = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"])
    assert nnz == int(kept.sum())
    assert np.array_equal(Sg != 0, kept)
    T = np.exp(-Mq[:, :-1] / reg) / reg
    np.testing.assert_allclose(Sg[kept], T[kept], rtol=2e-6 if bulk == "fp32" else 1e-13)
    Sr = np.zeros((n, m - 1))
    for i in range(n):
        sl = slice(H["rowptr"][i], H["rowptr"][i + 1])
        assert np.all(np.diff(H["colind"][sl]) > 0)
        Sr[i, H["colind"][sl]] = H["rval"][sl]
    assert np.array_equal(Sr, Sg)
    for j in range(m - 1):
        assert np.all(np.diff(H["rowind"][H["colptr"][j]:H["colptr"][j + 1]]) > 0)


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
def test_hessian_products_and_cg(gpu_lib, bulk):
    n, m, reg = 180, 140, 0.05
    M, a, b = cloud_problem(n, m, seed=13)
    Mg = M.astype(np.float32) if bulk == "fp32" else M
    gamma = point(Mg.astype(np.float64), a, b, reg)
    rng = np.random.default_rng(2)
    with gpu_lib.Problem(Mg, a, b, reg) as P:
        P.sparsified_hess_with_density(gamma, 0.05)
        H = P.hess_export()
        csc = (H["colptr"], H["rowind"], H["cval"])
        x = rng.standard_normal(m - 1)
        shift = 1e-3
        d_g = P.hess_apply_delta(x, shift)
        d_o = O.apply_deltax(n, m, H["h1"], H["h2"], csc, x, shift)
        np.testing.assert_allclose(d_g, d_o, rtol=1e-12, atol=1e-12 * np.abs(d_o).max())
        v = rng.standard_normal(n + m - 1)
        h_g = P.hess_apply_hs(v)
        h_o = O.apply_hsx(n, m, H["h1"], H["h2"], csc, v)
        np.testing.assert_allclose(h_g, h_o, rtol=1e-12, atol=1e-12 * np.abs(h_o).max())
        # dense check of the operator itself
        S = csc_dense(n, m, *csc)
        Hd = np.block([[np.diag(H["h1"]), S], [S.T, np.diag(H["h2"])]])
        np.testing.assert_allclose(h_g, Hd @ v, rtol=1e-10, atol=1e-10 * np.abs(h_g).max())
        rhs = rng.standard_normal(n + m - 1)
        sol, it, err = P.hess_cg(rhs, shift, tol=1e-10)
        sol_o, it_o, err_o = O.hess_cg(n, m, H["h1"], H["h2"], csc, rhs, shift, tol=1e-10)
        ref = np.linalg.solve(Hd + shift * np.eye(n + m - 1), rhs)
        assert err < 1e-10
        assert abs(it - it_o) <= 2
        np.testing.assert_allclose(sol, ref, rtol=1e-6, atol=1e-8 * np.abs(ref).max())
        np.testing.assert_allclose(sol, sol_o, rtol=1e-7, atol=1e-9 * np.abs(ref).max())
        # iteration cap is honoured
        _, it2, err2 = P.hess_cg(rhs, shift, tol=1e-14, max_iter=3)
        assert it2 == 3 and err2 > 1e-14


@pytest.mark.parametrize("bulk", ["fp64", "fp32"])
def test_topk_rule_repeated_builds(gpu_lib, bulk):
    # Successive builds on one handle (its candidate and tally buffers reused):
    # a nearby point, a shifted point, jumps in density and the drivers'
    # schedule (density down 10% per build at a drifting point).  Every build
    # must keep the oracle's set.
    n, m, reg = 300, 257, 0.03
    M, a, b = cloud_problem(n, m, seed=29)
    Mg = M.astype(np.float32) if bulk == "fp32" else M
    Mo = Mg.astype(np.float64)
    gamma0 = point(Mo, a, b, reg)
    rng = np.random.default_rng(5)
    steps = [(gamma0, 0.01),
             (gamma0 + 1e-6 * reg * rng.standard_normal(gamma0.shape), 0.01),
             (gamma0 + np.concatenate([np.full(n, 5 * reg), np.zeros(m - 1)]), 0.01),
             (gamma0, 0.1),
             (gamma0 + 1e-5 * reg * rng.standard_normal(gamma0.shape), 0.1),
             (gamma0, 0.003)]
    g = gamma0
    for q in range(8):
        g = g + 0.02 * reg * rng.standard_normal(g.shape)
        steps.append((g, 0.05 * 0.9 ** q))
    with gpu_lib.Problem(Mg, a, b, reg) as P:
        for q, (g, density) in enumerate(steps):
            print(f"top-k repeated build {q}: density {density:.4f}", flush=True)
            T, _ = O.compute_T(Mo, reg, g)
            cp, ri, va = O.sparsify_density(T, density, True)
            nnz = P.sparsified_hess_with_density(g, density)
            H = P.hess_export()
            So = csc_dense(n, m, cp, ri, va)
            Sg = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"]) * reg
            assert nnz == len(ri), q
            if bulk == "fp64":
                assert np.array_equal(Sg != 0, So != 0), q
                np.testing.assert_allclose(Sg, So, rtol=1e-12)
            else:
                # FP32 values may reorder near the k-th value; the count is exact
                k = int(n * (m - 1) * density)
                assert np.sum((Sg != 0) != (So != 0)) <= max(4, 0.002 * k), q


def test_topk_rule_ties_repeated_build(gpu_lib):
    # the same tied build twice on one handle: both take the same ties in
    # row-major order
    n, m, density = 150, 121, 0.3
    rng = np.random.default_rng(7)
    Mq = 0.25 * rng.integers(0, 4, size=(n, m)).astype(np.float64)
    a = np.full(n, 1.0 / n)
    b = np.full(m, 1.0 / m)
    gamma = np.zeros(n + m - 1)
    kept = topk_ties_expected(Mq, 0.5, density)
    with gpu_lib.Problem(Mq, a, b, 0.5) as P:
        for _ in range(2):
            nnz = P.sparsified_hess_with_density(gamma, density)
            H = P.hess_export()
            Sg = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"])
            assert nnz == int(kept.sum())
            assert np.array_equal(Sg != 0, kept)


@pytest.mark.parametrize("n,m", [(40000, 50), (50, 40000)])
@pytest.mark.parametrize("rule", ["topk", "threshold"])
def test_rules_tall_and_wide(gpu_lib, n, m, rule):
    # Row counts, column counts and tie offsets are scanned in passes of
    # 16384 by one workgroup: 40000 rows or 39999 columns take three passes.
    # FP64 bulk keeps exactly the oracle's set, and CSR and CSC agree.
    reg = 0.05
    M, a, b = cloud_problem(n, m, seed=n + 3 * m)
    gamma = point(M, a, b, reg, iters=2)
    T, _ = O.compute_T(M, reg, gamma)
    print(f"{rule} {n}x{m}", flush=True)
    with gpu_lib.Problem(M, a, b, reg, bulk="fp64") as P:
        if rule == "topk":
            cp, ri, va = O.sparsify_density(T, 0.02, True)
            nnz = P.sparsified_hess_with_density(gamma, 0.02)
            scale = reg
        else:
            _, g = O.dual_obj_grad(M, a, b, reg, gamma)
            delta = 0.01 * np.linalg.norm(g)
            cp, ri, va = O.sparsify_threshold(T, delta, 0.001)
            nnz = P.sparsified_hess(gamma, delta, 0.001)
            scale = 1.0
        H = P.hess_export()
    assert nnz == len(ri)
    C = csc_dense(n, m, H["colptr"], H["rowind"], H["cval"])
    Sg = C * scale
    So = csc_dense(n, m, cp, ri, va) / (1.0 if rule == "topk" else reg)
    assert np.array_equal(Sg != 0, So != 0)
    np.testing.assert_allclose(Sg, So, rtol=1e-12, atol=0)
    rp = H["rowptr"]
    assert rp[0] == 0 and rp[-1] == H["colptr"][-1] and np.all(np.diff(rp) >= 0)
    rows = np.repeat(np.arange(n), np.diff(rp))
    Sr = np.zeros((n, m - 1))
    Sr[rows, H["colind"][:rp[-1]]] = H["rval"][:rp[-1]]
    assert np.array_equal(Sr, C)
