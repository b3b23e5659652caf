"""Differentiable entropic OT loss between point clouds for PyTorch.

``sinkhorn_loss(X, Y, a, b, reg)`` returns L = <T*, M> - reg h(T*) (or the
sharp <T*, M>) for the squared-Euclidean cost M = |x|^2 + |y|^2 - 2 X Y' and
the optimal plan T* of the EROT problem, as a scalar tensor differentiable in
X and Y:

* forward: M is formed on the GPU -- on the matrix cores from BF16 clouds
  (``regot_sqeuclid_bf16``), or in FP64 by torch for FP32/FP64 clouds -- and
  the EROT problem is solved by this package's HIP solvers;
* backward: dL/dM = G by the implicit function theorem on the dual optimality
  conditions (``regot_implicit_grad``: one device CG solve with the dual
  Hessian), then dL/dx_i = 2 (sum_j G_ij x_i - (G Y)_i) and
  dL/dy_j = 2 (sum_i G_ij y_j - (G' X)_j) (library GEMMs).

PyTorch is plumbing here: the tensors live in its allocator, the solver and
the adjoint run in libregot_hip.so.  Because PyTorch-ROCm carries its own HIP
runtime, the two are synchronised at the boundary (``torch.cuda.synchronize``
before the library reads a tensor; the library returns only when done).
"""
import numpy as np
import torch

from . import Problem, sqeuclid_bf16
from . import _lib

_SOLVERS = {"bcd": _lib.SOLVER_BCD, "ssns": _lib.SOLVER_SSNS,
            "sparse_newton": _lib.SOLVER_SPARSE_NEWTON, "splr": _lib.SOLVER_SPLR}


def sqeuclid(X, Y):
    """Squared-Euclidean cost of two point clouds on the GPU.

    BF16 inputs run the matrix-core kernel (FP32 accumulation, FP32 M); FP32 or
    FP64 inputs are formed in FP64 by torch and returned in FP64."""
    if X.dtype == torch.bfloat16 and Y.dtype == torch.bfloat16:
        Xc = X.detach().contiguous()
        Yc = Y.detach().contiguous()
        n, d = Xc.shape
        m = Yc.shape[0]
        M = torch.empty(n, m, dtype=torch.float32, device=X.device)
        torch.cuda.synchronize(X.device)
        sqeuclid_bf16(Xc.data_ptr(), Xc.stride(0), Yc.data_ptr(), Yc.stride(0), n, m, d,
                      M.data_ptr(), M.stride(0), device=X.device.index or 0)
        return M
    Xd = X.detach().double()
    Yd = Y.detach().double()
    M = (Xd * Xd).sum(1)[:, None] + (Yd * Yd).sum(1)[None, :] - 2.0 * (Xd @ Yd.T)
    return M.contiguous()


class _SinkhornLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X, Y, a, b, reg, tol, max_iter, loss, solver, cg_tol):
        M = sqeuclid(X, Y)
        n, m = M.shape
        dt = np.float32 if M.dtype == torch.float32 else np.float64
        torch.cuda.synchronize(M.device)
        P = Problem.from_device(M.data_ptr(), n, m, dt, a, b, reg,
                                bulk="fp32" if dt == np.float32 else "fp64",
                                device=M.device.index or 0)
        res = P.solve(_SOLVERS[solver], tol=tol, max_iter=max_iter, compute_plan=False)
        gamma = res.dual
        G = torch.empty(n, m, dtype=M.dtype, device=M.device)
        value, w, it, err = P.implicit_grad(gamma, G.data_ptr(), dt, G.stride(0), loss=loss,
                                            cg_tol=cg_tol)
        P.close()
        ctx.save_for_backward(X, Y, G)
        ctx.info = {"niter": res.niter, "mar_err": res.mar_errs[-1], "cg_iters": it,
                    "cg_err": err, "gamma": gamma, "w": w}
        return torch.tensor(value, dtype=torch.float64, device=X.device)

    @staticmethod
    def backward(ctx, gout):
        X, Y, G = ctx.saved_tensors
        Xf = X.detach().to(G.dtype)
        Yf = Y.detach().to(G.dtype)
        s = gout.to(G.dtype)
        gX = gY = None
        if ctx.needs_input_grad[0]:
            gX = (2.0 * s) * (G.sum(1, keepdim=True) * Xf - G @ Yf)
            gX = gX.to(X.dtype)
        if ctx.needs_input_grad[1]:
            gY = (2.0 * s) * (G.sum(0)[:, None] * Yf - G.T.contiguous() @ Xf)
            gY = gY.to(Y.dtype)
        return gX, gY, None, None, None, None, None, None, None, None


def sinkhorn_loss(X, Y, a=None, b=None, reg=0.05, tol=1e-6, max_iter=10000, loss="entropic",
                  solver="bcd", cg_tol=1e-10, info=None):
    """Entropic (``loss="entropic"``) or sharp (``"sharp"``) OT loss between the
    point clouds X (n x d) and Y (m x d) on a GPU, differentiable in X and Y.
    a, b default to uniform weights.  If ``info`` is a dict it receives the
    solver's iteration count and marginal error and the adjoint's CG figures."""
    n, m = X.shape[0], Y.shape[0]
    a = np.full(n, 1.0 / n) if a is None else np.asarray(a, dtype=np.float64)
    b = np.full(m, 1.0 / m) if b is None else np.asarray(b, dtype=np.float64)
    if loss not in ("entropic", "sharp"):
        raise ValueError("loss must be 'entropic' or 'sharp'")
    if solver not in _SOLVERS:
        raise ValueError("solver must be one of %s" % sorted(_SOLVERS))
    out = _SinkhornLoss.apply(X, Y, a, b, float(reg), float(tol), int(max_iter), loss, solver,
                              float(cg_tol))
    if info is not None and out.grad_fn is not None:
        info.update(out.grad_fn.info if hasattr(out.grad_fn, "info") else {})
    return out
