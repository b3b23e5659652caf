"""GPU parity of the differentiable loss: the BF16 matrix-core cost, the
implicit-function adjoint against the dense FP64 reference, and the autograd
gradient against central finite differences."""
import numpy as np
import pytest

from conftest import cloud_problem
from implicit_ref import implicit_grad, loss_value, plan
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def bf16_bits_to_float(t):
    return t.float().cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("n,m,d", [(300, 200, 64), (97, 131, 20), (64, 64, 8), (130, 70, 100), (65, 129, 128)])
def test_sqeuclid_bf16_matches_fp64(gpu_lib, n, m, d):
    from regot.autograd import sqeuclid
    g = torch.Generator().manual_seed(n + m + d)
    X = torch.randn(n, d, generator=g).to(torch.bfloat16).to("cuda:0")
    Y = torch.randn(m, d, generator=g).to(torch.bfloat16).to("cuda:0")
    M = sqeuclid(X, Y)
    torch.cuda.synchronize()
    Xd, Yd = bf16_bits_to_float(X), bf16_bits_to_float(Y)
    ref = (Xd * Xd).sum(1)[:, None] + (Yd * Yd).sum(1)[None, :] - 2.0 * Xd @ Yd.T
    got = M.cpu().numpy().astype(np.float64)
    assert M.dtype == torch.float32 and got.shape == (n, m)
    # FP32 accumulation of d exact BF16 products, then an FP32 store
    tol = 2.0 ** -20 * (np.abs(Xd) @ np.abs(Yd).T) + 2.0 ** -23 * np.abs(ref) + 1e-12
    assert np.all(np.abs(got - ref) <= tol)


@pytest.mark.parametrize("loss", ["entropic", "sharp"])
def test_implicit_grad_matches_dense_reference(gpu_lib, loss):
    n, m, reg = 70, 50, 0.2
    M, a, b = cloud_problem(n, m, d=4, seed=11)
    gamma = O.sinkhorn_bcd(M, a, b, reg, tol=1e-14, max_iter=100000)["dual"]
    G_ref, w_ref = implicit_grad(M, reg, gamma, loss)
    P = gpu_lib.Problem(M, a, b, reg, bulk="fp64")
    G = torch.empty(n, m, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    value, w, it, err = P.implicit_grad(gamma, G.data_ptr(), np.float64, m, loss=loss,
                                        cg_tol=1e-13)
    P.close()
    # the entropic adjoint is warm-started at its exact solution (the dual)
    assert err < 1e-12 and (it > 0 or loss == "entropic")
    np.testing.assert_allclose(w, w_ref, rtol=1e-8, atol=1e-10)
    np.testing.assert_allclose(G.cpu().numpy(), G_ref, rtol=1e-8, atol=1e-12)
    assert abs(value - loss_value(M, reg, gamma, loss)) < 1e-10 * max(1.0, abs(value))
    if loss == "entropic":
        np.testing.assert_allclose(G.cpu().numpy(), plan(M, reg, gamma), rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("loss", ["entropic", "sharp"])
def test_autograd_matches_central_differences(gpu_lib, loss):
    from regot.autograd import sinkhorn_loss
    n, m, d, reg = 48, 40, 6, 0.3
    g = torch.Generator().manual_seed(5)
    X = torch.randn(n, d, generator=g, dtype=torch.float64).to("cuda:0").requires_grad_(True)
    Y = torch.randn(m, d, generator=g, dtype=torch.float64).to("cuda:0").requires_grad_(True)
    kw = dict(reg=reg, tol=1e-13, max_iter=200000, loss=loss, cg_tol=1e-13)
    L = sinkhorn_loss(X, Y, **kw)
    L.backward()
    gX = X.grad.detach().cpu().numpy()
    gY = Y.grad.detach().cpu().numpy()
    h = 1e-3
    rng = np.random.default_rng(7)
    with torch.no_grad():
        for _ in range(32):
            DX = rng.standard_normal((n, d))
            DY = rng.standard_normal((m, d))
            nrm = np.sqrt((DX * DX).sum() + (DY * DY).sum())   # unit directions
            DX /= nrm
            DY /= nrm
            tX = torch.from_numpy(DX).to("cuda:0")
            tY = torch.from_numpy(DY).to("cuda:0")
            lp = sinkhorn_loss(X + h * tX, Y + h * tY, **kw).item()
            lm = sinkhorn_loss(X - h * tX, Y - h * tY, **kw).item()
            fd = (lp - lm) / (2 * h)
            an = (gX * DX).sum() + (gY * DY).sum()
            # central differences are exact to O(h^2) = 1e-6 here
            assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (fd, an)


def test_bf16_pipeline_against_dense_reference(gpu_lib):
    # BF16 clouds -> matrix-core cost -> FP32-bulk solve -> adjoint; checked
    # against the FP64 dense reference on the same cost matrix
    from regot.autograd import sinkhorn_loss, sqeuclid
    n = m = 256
    d, reg = 64, 2.0
    g = torch.Generator().manual_seed(9)
    X = torch.randn(n, d, generator=g).to(torch.bfloat16).to("cuda:0").requires_grad_(True)
    Y = torch.randn(m, d, generator=g).to(torch.bfloat16).to("cuda:0").requires_grad_(True)
    info = {}
    L = sinkhorn_loss(X, Y, reg=reg, tol=1e-9, max_iter=20000, info=info)
    L.backward()
    M = sqeuclid(X, Y).cpu().numpy().astype(np.float64)
    a = np.full(n, 1.0 / n)
    b = np.full(m, 1.0 / m)
    gamma = O.sinkhorn_bcd(M, a, b, reg, tol=1e-12, max_iter=200000)["dual"]
    G_ref, _ = implicit_grad(M, reg, gamma, "entropic")
    Xd, Yd = bf16_bits_to_float(X.detach()), bf16_bits_to_float(Y.detach())
    gX_ref = 2.0 * (G_ref.sum(1)[:, None] * Xd - G_ref @ Yd)
    gX = X.grad.float().cpu().numpy()
    assert abs(L.item() - loss_value(M, reg, gamma, "entropic")) < 1e-6 * abs(L.item())
    # BF16 gradient output: ~3 significant digits
    err = np.abs(gX - gX_ref).max() / np.abs(gX_ref).max()
    assert err < 1e-2, err
