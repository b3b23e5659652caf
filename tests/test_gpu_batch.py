"""GPU parity of the batched BCD launch against the FP64 oracle, problem by problem.

Every problem of a batch must reproduce what sinkhorn_bcd gives it alone:
same niter, same dual, same final objective and marginal error.
"""
import numpy as np
import pytest

from conftest import cloud_problem, readme_example
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def mixed_batch(seed, count, lo=24, hi=300, dtype=np.float64):
    rng = np.random.default_rng(seed)
    Ms, As, Bs, regs = [], [], [], []
    for q in range(count):
        n, m = (int(x) for x in rng.integers(lo, hi, size=2))
        M, a, b = cloud_problem(n, m, d=8, seed=1000 * seed + q, dtype=dtype)
        Ms.append(M)
        As.append(a)
        Bs.append(b)
        regs.append(float(np.exp(rng.uniform(np.log(0.02), np.log(0.3)))))
    return Ms, As, Bs, regs


def test_batch_fp64_matches_oracle_per_problem(gpu_lib):
    Ms, As, Bs, regs = mixed_batch(1, 10)
    # degenerate and ragged shapes next to ordinary ones
    for n, m in [(1, 7), (9, 1), (1, 1), (3, 257), (257, 3)]:
        M, a, b = cloud_problem(n, m, d=8, seed=n * 31 + m)
        Ms.append(M)
        As.append(a)
        Bs.append(b)
        regs.append(0.1)
    tol, max_iter = 1e-10, 3000
    out = gpu_lib.sinkhorn_bcd_batch(Ms, As, Bs, regs, tol=tol, max_iter=max_iter, bulk="fp64")
    for p in range(len(Ms)):
        o = O.sinkhorn_bcd(Ms[p], As[p], Bs[p], regs[p], tol=tol, max_iter=max_iter)
        assert abs(int(out["niter"][p]) - o["niter"]) <= 1, p
        ref = o["dual"]
        err = np.max(np.abs(out["dual"][p] - ref)) / max(1.0, np.max(np.abs(ref)))
        assert err < 1e-10, (p, err)
        if int(out["niter"][p]) == o["niter"]:
            assert abs(out["obj"][p] - o["obj_vals"][-1]) < 1e-12 * max(1.0, abs(o["obj_vals"][-1]))
            assert abs(out["mar_err"][p] - o["mar_errs"][-1]) < 1e-13 + 1e-6 * o["mar_errs"][-1]
        assert out["mar_err"][p] < tol or out["niter"][p] == max_iter


def test_batch_fp32_cloud(gpu_lib):
    Ms, As, Bs, regs = mixed_batch(2, 12, lo=64, hi=700, dtype=np.float32)
    tol = 1e-7
    out = gpu_lib.sinkhorn_bcd_batch(Ms, As, Bs, regs, tol=tol, max_iter=5000)
    assert out["niter"].dtype == np.int32
    for p in range(len(Ms)):
        M64 = Ms[p].astype(np.float64)
        o = O.sinkhorn_bcd(M64, As[p], Bs[p], regs[p], tol=tol, max_iter=5000)
        assert out["mar_err"][p] < tol
        assert abs(int(out["niter"][p]) - o["niter"]) <= max(2, 0.05 * o["niter"]), p
        # FP64 re-check of the returned dual on the CPU
        _, g = O.dual_obj_grad(M64, As[p], Bs[p], regs[p], out["dual"][p])
        assert np.linalg.norm(g) < 2 * tol, (p, np.linalg.norm(g))


def test_batch_matches_single_problem_path(gpu_lib):
    # same arithmetic as regot.sinkhorn_bcd on each problem (FP64 bulk)
    Ms, As, Bs, regs = mixed_batch(3, 4, lo=100, hi=400)
    out = gpu_lib.sinkhorn_bcd_batch(Ms, As, Bs, regs, tol=1e-9, max_iter=2000, bulk="fp64")
    # (the tilings differ, so FP64 sums differ in the last bits)
    for p in range(len(Ms)):
        r = gpu_lib.sinkhorn_bcd(Ms[p], As[p], Bs[p], regs[p], tol=1e-9, max_iter=2000)
        assert abs(int(out["niter"][p]) - r.niter) <= 1
        np.testing.assert_allclose(out["dual"][p], r.dual, rtol=0, atol=1e-10)
        if int(out["niter"][p]) == r.niter:
            assert abs(out["mar_err"][p] - r.mar_errs[-1]) <= 1e-14 + 1e-6 * r.mar_errs[-1]


def test_batch_uneven_convergence_and_limits(gpu_lib):
    # one fast problem, one slow one, one that hits max_iter, one already converged:
    # the fast ones leave the tile tables while the slow one keeps running
    M, a, b = readme_example(120, 90)
    Ms = [M, M, M, M]
    As = [a] * 4
    Bs = [b] * 4
    regs = [1.0, 0.05, 0.01, 0.2]
    B = gpu_lib.BatchProblem(Ms, As, Bs, regs, bulk="fp64")
    out = B.solve_bcd(tol=1e-9, max_iter=150)
    for p in range(4):
        o = O.sinkhorn_bcd(M, a, b, regs[p], tol=1e-9, max_iter=150)
        assert int(out["niter"][p]) == o["niter"], p
        np.testing.assert_allclose(out["dual"][p], o["dual"], rtol=0, atol=1e-10)
    assert int(out["niter"][2]) == 150
    assert out["rounds"] >= 150
    # a second solve on the same handle starts again from the reference init
    out2 = B.solve_bcd(tol=1e9, max_iter=10)
    assert np.all(out2["niter"] == 0)
    out3 = B.solve_bcd(tol=1e-9, max_iter=150)
    np.testing.assert_array_equal(out3["niter"], out["niter"])
    np.testing.assert_array_equal(np.concatenate(out3["dual"]), np.concatenate(out["dual"]))
    B.close()


def test_batch_device_inputs(gpu_lib):
    torch = pytest.importorskip("torch")
    Ms, As, Bs, regs = mixed_batch(4, 3, lo=50, hi=200, dtype=np.float32)
    dev = [torch.from_numpy(M).to("cuda:0") for M in Ms]
    torch.cuda.synchronize()
    descs = [(t.data_ptr(), t.shape[0], t.shape[1], np.float32) for t in dev]
    a = gpu_lib.sinkhorn_bcd_batch(descs, As, Bs, regs, tol=1e-7, max_iter=3000)
    b = gpu_lib.sinkhorn_bcd_batch(Ms, As, Bs, regs, tol=1e-7, max_iter=3000)
    np.testing.assert_array_equal(a["niter"], b["niter"])
    np.testing.assert_array_equal(np.concatenate(a["dual"]), np.concatenate(b["dual"]))


def test_batch_graph_replay_matches_eager_launches(gpu_lib):
    # The rounds as hipGraph replays (default) and as one-by-one launches
    # give bitwise the same results: same kernels, same order.  The solve
    # replays each captured graph many times, rebuilds it when the tables
    # shrink, and a second solve replays again after eager work in between.
    Ms, As, Bs, regs = mixed_batch(5, 9, lo=40, hi=260)
    B = gpu_lib.BatchProblem(Ms, As, Bs, regs, bulk="fp64")
    e = B.solve_bcd(tol=1e-10, max_iter=2000, graph=False)
    g = B.solve_bcd(tol=1e-10, max_iter=2000, graph=True)
    assert g["graph_launches"] >= 2 and g["graph_builds"] >= 2
    np.testing.assert_array_equal(g["niter"], e["niter"])
    np.testing.assert_array_equal(g["obj"], e["obj"])
    np.testing.assert_array_equal(g["mar_err"], e["mar_err"])
    np.testing.assert_array_equal(np.concatenate(g["dual"]), np.concatenate(e["dual"]))
    r = gpu_lib.sinkhorn_bcd(Ms[0], As[0], Bs[0], regs[0], tol=1e-10, max_iter=2000)
    assert abs(r.niter - int(g["niter"][0])) <= 1
    g2 = B.solve_bcd(tol=1e-10, max_iter=2000, graph=True)
    np.testing.assert_array_equal(np.concatenate(g2["dual"]), np.concatenate(e["dual"]))
    assert g2["graph_launches"] > g["graph_launches"]
    B.close()


def test_batch_autotuned_tiles(gpu_lib):
    # The autotuner times the coarse-tile candidates once per handle and keeps
    # the fastest; the solve then starts from the reference init, so every
    # problem still matches the oracle, and later solves reuse the choice.
    # enough rows for coarse tiles (>= 2048 tiles of 16 rows)
    Ms, As, Bs, regs = mixed_batch(6, 150, lo=200, hi=900, dtype=np.float32)
    B = gpu_lib.BatchProblem(Ms, As, Bs, regs)
    out = B.solve_bcd(tol=1e-7, max_iter=5000, autotune=True)
    assert out["tuned"] and out["rows_per_wave"] in (2, 4)
    assert out["block_rows"] in (256, 512, 1024)
    assert sum(t > 0 for t in out["tune_ms"]) == 6
    for p in range(0, len(Ms), 15):
        o = O.sinkhorn_bcd(Ms[p].astype(np.float64), As[p], Bs[p], regs[p], tol=1e-7,
                           max_iter=5000)
        assert abs(int(out["niter"][p]) - o["niter"]) <= max(2, 0.05 * o["niter"]), p
        assert out["mar_err"][p] < 1e-7
    again = B.solve_bcd(tol=1e-7, max_iter=5000, autotune=True)
    assert (again["rows_per_wave"], again["block_rows"]) == (out["rows_per_wave"],
                                                              out["block_rows"])
    np.testing.assert_array_equal(np.concatenate(again["dual"]), np.concatenate(out["dual"]))
    B.close()


def single_read_batch(seed):
    # FP32 problems around the single-read range 1024 < m <= 4096: ragged
    # widths (one to four column chunks, partial last vectors), tiny and
    # ragged n, and neighbours that keep two-pass rounds (m = 1024, 4097)
    shapes = [(300, 1025), (17, 1027), (1, 2049), (3, 1500), (260, 2048), (129, 3001),
              (64, 4095), (40, 4096), (90, 1024), (33, 4097), (200, 600)]
    Ms, As, Bs, regs = [], [], [], []
    rng = np.random.default_rng(seed)
    for q, (n, m) in enumerate(shapes):
        M, a, b = cloud_problem(n, m, d=8, seed=7000 * seed + q, dtype=np.float32)
        Ms.append(M)
        As.append(a)
        Bs.append(b)
        regs.append(float(np.exp(rng.uniform(np.log(0.02), np.log(0.2)))))
    return shapes, Ms, As, Bs, regs


def test_batch_single_read_rounds_match_oracle_and_two_pass(gpu_lib):
    shapes, Ms, As, Bs, regs = single_read_batch(8)
    tol = 1e-7
    B = gpu_lib.BatchProblem(Ms, As, Bs, regs)
    print("single-read batch: fused rounds", flush=True)
    f = B.solve_bcd(tol=tol, max_iter=5000)
    print("single-read batch: two-pass rounds", flush=True)
    t = B.solve_bcd(tol=tol, max_iter=5000, single_read=False)
    want = np.array([1024 < m <= 4096 for _, m in shapes])
    np.testing.assert_array_equal(f["single_read"], want)
    assert not t["single_read"].any()
    for p, (n, m) in enumerate(shapes):
        M64 = Ms[p].astype(np.float64)
        o = O.sinkhorn_bcd(M64, As[p], Bs[p], regs[p], tol=tol, max_iter=5000)
        for r in (f, t):
            assert r["mar_err"][p] < tol, (p, r["mar_err"][p])
            assert abs(int(r["niter"][p]) - o["niter"]) <= max(2, 0.05 * o["niter"]), (p, n, m)
            # FP64 re-check of the returned dual on the CPU
            _, g = O.dual_obj_grad(M64, As[p], Bs[p], regs[p], r["dual"][p])
            assert np.linalg.norm(g) < 2 * tol, (p, n, m, np.linalg.norm(g))
        # both round kinds stop within a crossing of tol of each other
        assert abs(int(f["niter"][p]) - int(t["niter"][p])) <= max(1, 0.02 * o["niter"]), p
        if not want[p]:
            # problems outside the range run the same two-pass rounds either way
            assert int(f["niter"][p]) == int(t["niter"][p])
            np.testing.assert_array_equal(f["dual"][p], t["dual"][p])
    B.close()


def test_batch_single_read_tail_and_graph_replays(gpu_lib):
    # One problem converges slowly: the others leave the tables and its tiles
    # shrink to the minimum row count; graph replays give bitwise the same
    # results as launches, and a second solve repeats the first.
    M, a, b = cloud_problem(400, 2500, d=8, seed=91, dtype=np.float32)
    Ms, As, Bs = [M, M, M], [a, a, a], [b, b, b]
    regs = [0.2, 0.05, 0.01]
    B = gpu_lib.BatchProblem(Ms, As, Bs, regs)
    e = B.solve_bcd(tol=1e-7, max_iter=4000)
    assert e["single_read"].all()
    g = B.solve_bcd(tol=1e-7, max_iter=4000, graph=True)
    np.testing.assert_array_equal(g["niter"], e["niter"])
    np.testing.assert_array_equal(np.concatenate(g["dual"]), np.concatenate(e["dual"]))
    e2 = B.solve_bcd(tol=1e-7, max_iter=4000)
    np.testing.assert_array_equal(np.concatenate(e2["dual"]), np.concatenate(e["dual"]))
    M64 = M.astype(np.float64)
    for p in range(3):
        o = O.sinkhorn_bcd(M64, a, b, regs[p], tol=1e-7, max_iter=4000)
        assert abs(int(e["niter"][p]) - o["niter"]) <= max(2, 0.05 * o["niter"]), p
        assert e["mar_err"][p] < 1e-7
    # max_iter cuts the slow problem: niter = max_iter exactly
    c = B.solve_bcd(tol=1e-12, max_iter=37)
    assert int(c["niter"][2]) == 37
    B.close()
