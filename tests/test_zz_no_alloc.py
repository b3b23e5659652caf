"""The no-allocation entry point regot_dual_obj_grad_noalloc against
regot_dual_obj_grad (same kernels, same results) through the C-ABI, on two
calls in a row on one GPU, the second with a smaller problem, in both bulk
precisions; and its size limit."""
import ctypes

import numpy as np
import pytest

from conftest import cloud_problem

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def padded(M, dtype, align_elems):
    """Device copy of M with a row pitch rounded up to align_elems."""
    n, m = M.shape
    ld = -(-m // align_elems) * align_elems
    buf = np.zeros((n, ld), dtype=dtype)
    buf[:, :m] = M
    return torch.from_numpy(buf).to("cuda:0"), ld


def call_noalloc(L, Md, bulk, ld, n, m, a, b, reg, gamma):
    da = torch.from_numpy(a).to("cuda:0")
    db = torch.from_numpy(b).to("cuda:0")
    dg = torch.from_numpy(gamma).to("cuda:0")
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda:0")
    grad = torch.full((n + m - 1,), float("nan"), dtype=torch.float64, device="cuda:0")
    stream = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    rc = L.regot_dual_obj_grad_noalloc(
        ctypes.c_void_p(Md.data_ptr()), bulk, ld, n, m, ctypes.c_void_p(da.data_ptr()),
        ctypes.c_void_p(db.data_ptr()), reg, ctypes.c_void_p(dg.data_ptr()),
        ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
        ctypes.c_void_p(stream), 0)
    torch.cuda.synchronize()
    assert rc == 0, L.regot_last_error()
    return out.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.parametrize("bulk", ["fp32", "fp64"])
def test_noalloc_matches_dual_obj_grad(gpu_lib, bulk):
    from regot import _lib
    L = _lib.lib()
    code = _lib.REGOT_F32 if bulk == "fp32" else _lib.REGOT_F64
    dt = np.float32 if bulk == "fp32" else np.float64
    reg = 0.05
    for n, m in ((700, 1300), (300, 257)):          # the second call is smaller
        M, a, b = cloud_problem(n, m, d=6, seed=n)
        rng = np.random.default_rng(m)
        with gpu_lib.Problem(M.astype(dt), a, b, reg, bulk=bulk) as P:
            alpha = P.optimal_alpha(np.zeros(m))
            gamma = np.concatenate([alpha, np.zeros(m - 1)]) + 1e-3 * rng.standard_normal(
                n + m - 1)
            f_ref, g_ref = P.dual_obj_grad(gamma)
        print("noalloc %s %dx%d" % (bulk, n, m), flush=True)
        Md, ld = padded(M.astype(dt), dt, 16 // np.dtype(dt).itemsize)
        out, grad = call_noalloc(L, Md, code, ld, n, m, a, b, reg, gamma)
        np.testing.assert_allclose(out[0], f_ref, rtol=1e-14)
        np.testing.assert_allclose(out[1], np.sum(g_ref ** 2), rtol=1e-12)
        np.testing.assert_allclose(grad, g_ref, rtol=1e-13, atol=1e-16)


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu_lib):
    # n = 5e6 rows need 2n + 2m doubles for the split potentials and log
    # marginals alone (> 4 Mi doubles = 32 MiB); every buffer passed is
    # allocated at its full size, so even an accepted call would stay in bounds
    from regot import _lib
    L = _lib.lib()
    n, m = 5_000_000, 2
    M = torch.zeros(n, m, dtype=torch.float64, device="cuda:0")
    a = torch.full((n,), 1.0 / n, dtype=torch.float64, device="cuda:0")
    b = torch.full((m,), 0.5, dtype=torch.float64, device="cuda:0")
    g = torch.zeros(n + m - 1, dtype=torch.float64, device="cuda:0")
    grad = torch.zeros(n + m - 1, dtype=torch.float64, device="cuda:0")
    out = torch.zeros(2, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    rc = L.regot_dual_obj_grad_noalloc(vp(M), _lib.REGOT_F64, m, n, m, vp(a), vp(b), 0.1, vp(g),
                                       vp(out), vp(grad), None, 0)
    torch.cuda.synchronize()
    assert rc < 0 and b"scratch" in L.regot_last_error()
