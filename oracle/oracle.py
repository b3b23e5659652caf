"""ctypes front end of the FP64 CPU oracle (oracle/erot_oracle.c).

TEST INFRASTRUCTURE ONLY: the parity checker for the HIP build.  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module; the ``regot`` package never does.

Every wrapper takes row-major float64 numpy arrays and names the reference
function it restates (yixuan/regot-python v0.0.3, ``src/sinkhorn_*``).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liberot_oracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class _Opts(ctypes.Structure):
    _fields_ = [
        ("lin_method", ctypes.c_int),
        ("mu0", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("density", ctypes.c_double),
        ("cg_tol", ctypes.c_double),
        ("cg_max_iter", ctypes.c_int),
        ("lowrank_cg", ctypes.c_int),
        ("dense_T", ctypes.c_int),
        ("ls_max_iter", ctypes.c_int),
    ]


def build():
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        for name in ("orc_dual_obj", "orc_dual_obj_grad", "orc_compute_T", "orc_bcd_iterations",
                     "orc_line_selection", "orc_line_search_wolfe", "orc_erot_line_search_wolfe",
                     "orc_cubic_minimizer", "orc_step_selection", "orc_primal_val",
                     "orc_qrot_dual_obj_grad", "orc_qrot_pl_root", "orc_qrot_primal_val"):
            getattr(L, name).restype = ctypes.c_double
        for name in ("orc_sparsify_mat3", "orc_sparsify_density", "orc_trace_used"):
            getattr(L, name).restype = ctypes.c_longlong
        L.orc_trace_set.argtypes = [ctypes.c_void_p, ctypes.c_long]
        _lib = L
    return _lib


def _d(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def _p(x):
    return x.ctypes.data_as(_dp)


# --------------------------------------------------------------------------
# Problem dense passes
# --------------------------------------------------------------------------
def optimal_beta(M, b, reg, alpha):
    """src/sinkhorn_problem_bcd.cpp:149-174"""
    M, b, alpha = _d(M), _d(b), _d(alpha)
    n, m = M.shape
    beta = np.empty(m)
    lib().orc_optimal_beta(n, m, _p(M), _p(b), ctypes.c_double(reg), _p(alpha), _p(beta))
    return beta


def optimal_alpha(M, a, reg, beta):
    """src/sinkhorn_problem_bcd.cpp:177-201"""
    M, a, beta = _d(M), _d(a), _d(beta)
    n, m = M.shape
    alpha = np.empty(n)
    lib().orc_optimal_alpha(n, m, _p(M), _p(a), ctypes.c_double(reg), _p(beta), _p(alpha))
    return alpha


def compute_T(M, reg, gamma):
    """src/sinkhorn_problem_obj.cpp:139-188 -> (T, sum T)"""
    M, gamma = _d(M), _d(gamma)
    n, m = M.shape
    T = np.empty((n, m))
    s = lib().orc_compute_T(n, m, _p(M), ctypes.c_double(reg), _p(gamma), _p(T))
    return T, s


def dual_obj(M, a, b, reg, gamma):
    """src/sinkhorn_problem_obj.cpp:324-342"""
    M, a, b, gamma = _d(M), _d(a), _d(b), _d(gamma)
    n, m = M.shape
    return lib().orc_dual_obj(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), _p(gamma))


def dual_obj_grad(M, a, b, reg, gamma):
    """src/sinkhorn_problem_obj.cpp:363-393 -> (f, grad)"""
    M, a, b, gamma = _d(M), _d(a), _d(b), _d(gamma)
    n, m = M.shape
    g = np.empty(n + m - 1)
    f = lib().orc_dual_obj_grad(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), _p(gamma),
                                _p(g), None, 1)
    return f, g


def get_plan(M, reg, gamma):
    """src/sinkhorn_result.h:32-50"""
    M, gamma = _d(M), _d(gamma)
    n, m = M.shape
    plan = np.empty((n, m))
    lib().orc_get_plan(n, m, _p(M), ctypes.c_double(reg), _p(gamma), _p(plan))
    return plan


# --------------------------------------------------------------------------
# Sparsifiers -> (colptr, rowind, val) CSC of the n x (m-1) block
# --------------------------------------------------------------------------
def _csc_from(call, n, m):
    colptr = np.empty(m, dtype=np.int32)
    ri = ctypes.POINTER(ctypes.c_int)()
    va = ctypes.POINTER(ctypes.c_double)()
    nnz = call(colptr.ctypes.data_as(_ip), ctypes.byref(ri), ctypes.byref(va))
    rowind = np.ctypeslib.as_array(ri, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(va, shape=(max(nnz, 1),))[:nnz].copy()
    lib().orc_free(ri)
    lib().orc_free(va)
    return colptr, rowind, val


def sparsify_threshold(T, delta, density_hint):
    """SSNS threshold rule, src/sinkhorn_sparsify3.cpp:276-405"""
    T = _d(T)
    n, m = T.shape
    return _csc_from(lambda cp, ri, va: lib().orc_sparsify_mat3(
        n, m, _p(T), ctypes.c_double(delta), ctypes.c_double(density_hint), cp, ri, va), n, m)


def sparsify_density(T, density, preserve=True):
    """Global top-k rule, src/sinkhorn_sparsify4.cpp:77-174"""
    T = _d(T)
    n, m = T.shape
    return _csc_from(lambda cp, ri, va: lib().orc_sparsify_density(
        n, m, _p(T), ctypes.c_double(density), int(preserve), cp, ri, va), n, m)


def csc_to_dense(n, m, colptr, rowind, val):
    S = np.zeros((n, m - 1))
    for j in range(m - 1):
        for k in range(colptr[j], colptr[j + 1]):
            S[rowind[k], j] = val[k]
    return S


# --------------------------------------------------------------------------
# Hessian operators (CSC sigma), src/sinkhorn_hess.cpp:99-164, src/sinkhorn_cg.cpp
# --------------------------------------------------------------------------
def _csc_args(colptr, rowind, val):
    colptr = np.ascontiguousarray(colptr, dtype=np.int32)
    rowind = np.ascontiguousarray(rowind, dtype=np.int32)
    val = _d(val)
    return colptr, rowind, val


def apply_deltax(n, m, h1, h2, csc, x, shift):
    cp, ri, va = _csc_args(*csc)
    h1, h2, x = _d(h1), _d(h2), _d(x)
    res = np.empty(m - 1)
    lib().orc_apply_deltax(n, m, _p(h1), _p(h2), cp.ctypes.data_as(_ip), ri.ctypes.data_as(_ip),
                           _p(va), _p(x), ctypes.c_double(shift), _p(res))
    return res


def apply_hsx(n, m, h1, h2, csc, x):
    cp, ri, va = _csc_args(*csc)
    h1, h2, x = _d(h1), _d(h2), _d(x)
    res = np.empty(n + m - 1)
    lib().orc_apply_hsx(n, m, _p(h1), _p(h2), cp.ctypes.data_as(_ip), ri.ctypes.data_as(_ip),
                        _p(va), _p(x), _p(res))
    return res


def hess_cg(n, m, h1, h2, csc, rhs, shift, tol=1e-8, max_iter=0):
    """(Hs + shift I)^{-1} rhs via CG on the Schur complement -> (x, iters, rel. residual)"""
    cp, ri, va = _csc_args(*csc)
    h1, h2, rhs = _d(h1), _d(h2), _d(rhs)
    res = np.empty(n + m - 1)
    err = ctypes.c_double(0)
    it = lib().orc_hess_cg(n, m, _p(h1), _p(h2), cp.ctypes.data_as(_ip), ri.ctypes.data_as(_ip),
                           _p(va), _p(rhs), ctypes.c_double(shift), ctypes.c_double(tol),
                           int(max_iter), _p(res), ctypes.byref(err))
    return res, it, err.value


def hess_direct(n, m, h1, h2, csc, rhs, shift):
    cp, ri, va = _csc_args(*csc)
    h1, h2, rhs = _d(h1), _d(h2), _d(rhs)
    res = np.empty(n + m - 1)
    rc = lib().orc_hess_direct(n, m, _p(h1), _p(h2), cp.ctypes.data_as(_ip),
                               ri.ctypes.data_as(_ip), _p(va), _p(rhs), ctypes.c_double(shift),
                               _p(res))
    if rc != 0:
        raise RuntimeError("dense Cholesky failed (%d)" % rc)
    return res


def low_rank_combine(iAr, iAy, r, y, s):
    """src/sinkhorn_linsolve.cpp:207-213"""
    iAr, iAy, r, y, s = (_d(v) for v in (iAr, iAy, r, y, s))
    out = np.empty_like(r)
    lib().orc_low_rank_combine(ctypes.c_long(len(r)), _p(iAr), _p(iAy), _p(r), _p(y), _p(s),
                               _p(out))
    return out


def cubic_minimizer(a, b, fa, fb, ga, gb):
    ex = ctypes.c_int(0)
    r = lib().orc_cubic_minimizer(*(ctypes.c_double(v) for v in (a, b, fa, fb, ga, gb)),
                                  ctypes.byref(ex))
    return r, bool(ex.value)


def step_selection(al, au, at, fl, fu, ft, gl, gu, gt):
    return lib().orc_step_selection(*(ctypes.c_double(v) for v in (al, au, at, fl, fu, ft, gl, gu, gt)))


# --------------------------------------------------------------------------
# Drivers -> dict with the SinkhornResult fields
# --------------------------------------------------------------------------
def _run(fn, M, a, b, reg, tol, max_iter, verbose, opts, with_dens):
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    dual = np.empty(n + m - 1)
    h = [np.empty(max_iter + 1) for _ in range(4)]
    hl = ctypes.c_int(0)
    args = [n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), ctypes.c_double(tol),
            int(max_iter), int(verbose)]
    if opts is not None:
        args.append(ctypes.byref(opts))
    args += [_p(dual), _p(h[0]), _p(h[1]), _p(h[2])]
    if with_dens:
        args.append(_p(h[3]))
    args.append(ctypes.byref(hl))
    niter = fn(*args)
    k = hl.value
    out = dict(niter=niter, dual=dual, obj_vals=list(h[0][:k]), mar_errs=list(h[1][:k]),
               run_times=list(h[2][:k]), densities=list(h[3][:k]) if with_dens else [])
    return out


def _opts(method=1, mu0=1.0, shift=1e-3, density=0.01, cg_tol=0.0, cg_max_iter=0,
          lowrank_cg=False, ls_max_iter=0):
    return _Opts(int(method), float(mu0), float(shift), float(density), float(cg_tol),
                 int(cg_max_iter), int(lowrank_cg), 1, int(ls_max_iter))


def sinkhorn_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0):
    """src/sinkhorn_bcd.cpp:14-124"""
    return _run(lib().orc_sinkhorn_bcd, M, a, b, reg, tol, max_iter, verbose, None, False)


def sinkhorn_ssns(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kw):
    """src/sinkhorn_ssns.cpp:16-234 (method defaults to 1, src/solvers.cpp:425-441)"""
    return _run(lib().orc_sinkhorn_ssns, M, a, b, reg, tol, max_iter, verbose, _opts(**kw), True)


def sinkhorn_sparse_newton(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kw):
    """src/sinkhorn_sparse_newton.cpp:24-170"""
    return _run(lib().orc_sinkhorn_sparse_newton, M, a, b, reg, tol, max_iter, verbose,
                _opts(**kw), True)


def sinkhorn_splr(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kw):
    """src/sinkhorn_splr.cpp:24-199"""
    return _run(lib().orc_sinkhorn_splr, M, a, b, reg, tol, max_iter, verbose, _opts(**kw), True)


# --------------------------------------------------------------------------
# Line searches over generic objectives, and the driver trace
# --------------------------------------------------------------------------
_OBJ_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, _dp)
_OBJGRAD_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, _dp, _dp)


def line_selection(fobj, gamma, direc, curobj, cand=(1.0, 0.5, 0.25, 0.1)):
    """src/sinkhorn_problem_ls.cpp:13-41 on a Python objective f(x) ->
    (step, f at the step, last candidate evaluated)."""
    gamma, direc = _d(gamma), _d(direc)
    N = gamma.shape[0]
    cb = _OBJ_FN(lambda ctx, x: float(fobj(np.ctypeslib.as_array(x, (N,)).copy())))
    c = _d(cand)
    fv = ctypes.c_double()
    last = ctypes.c_double()
    work = np.empty(N)
    st = lib().orc_line_selection(ctypes.c_long(N), cb, None, _p(c), len(c), _p(gamma),
                                  _p(direc), ctypes.c_double(curobj), ctypes.byref(fv),
                                  _p(work), ctypes.byref(last))
    return st, fv.value, last.value


def _wolfe_out(st, rec, last, nt, x, g):
    return dict(step=st, recompute_T=bool(rec.value), last_step=last.value,
                trials=nt.value, x=x, grad=g)


def line_search_wolfe(fg, gamma, direc, curobj, curgrad, init_step=1.0, max_iter=20):
    """src/sinkhorn_problem_ls.cpp:226-470 on a Python objective
    fg(x) -> (f, grad).  Returns a dict: step, recompute_T, last_step,
    trials, and x / grad of the last trial."""
    gamma, direc, curgrad = _d(gamma), _d(direc), _d(curgrad)
    N = gamma.shape[0]

    def call(ctx, x, g):
        f, gr = fg(np.ctypeslib.as_array(x, (N,)).copy())
        np.ctypeslib.as_array(g, (N,))[:] = gr
        return float(f)

    cb = _OBJGRAD_FN(call)
    rec, nt, last = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    x, g = np.zeros(N), np.zeros(N)
    st = lib().orc_line_search_wolfe(ctypes.c_long(N), cb, None, ctypes.c_double(init_step),
                                      _p(gamma), _p(direc), ctypes.c_double(curobj), _p(curgrad),
                                      int(max_iter), ctypes.byref(rec), _p(x), _p(g),
                                      ctypes.byref(last), ctypes.byref(nt))
    return _wolfe_out(st, rec, last, nt, x, g)


def erot_line_search_wolfe(M, a, b, reg, gamma, direc, init_step=1.0, max_iter=20):
    """line_search_wolfe on the EROT dual with the oracle's own f and gradient."""
    M, a, b, gamma, direc = _d(M), _d(a), _d(b), _d(gamma), _d(direc)
    n, m = M.shape
    f, g0 = dual_obj_grad(M, a, b, reg, gamma)
    rec, nt, last = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    x, g = np.zeros(n + m - 1), np.zeros(n + m - 1)
    st = lib().orc_erot_line_search_wolfe(
        n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), ctypes.c_double(init_step), _p(gamma),
        _p(direc), ctypes.c_double(f), _p(g0), int(max_iter), ctypes.byref(rec), _p(x), _p(g),
        ctypes.byref(last), ctypes.byref(nt))
    return _wolfe_out(st, rec, last, nt, x, g)


TRACE_HDR = 16


def traced(solver, M, a, b, reg, tol=1e-6, max_iter=1000, **kw):
    """Run sinkhorn_ssns / sinkhorn_sparse_newton / sinkhorn_splr with the
    driver trace on -> (result dict, list of per-iteration records).  Each
    record: ``hdr`` (the scalars listed in erot_oracle.h) and the vectors
    ``gamma``, ``tpoint``, ``direc``, ``y``, ``s``."""
    n, m = np.shape(M)
    N = n + m - 1
    rec_len = TRACE_HDR + 5 * N
    buf = np.zeros(rec_len * (max_iter + 1))
    lib().orc_trace_set(buf.ctypes.data, buf.shape[0])
    try:
        res = globals()[solver](M, a, b, reg, tol=tol, max_iter=max_iter, **kw)
        used = lib().orc_trace_used()
    finally:
        lib().orc_trace_set(None, 0)
    assert used <= buf.shape[0] and used % rec_len == 0
    recs = []
    for k in range(used // rec_len):
        r = buf[k * rec_len:(k + 1) * rec_len]
        v = r[TRACE_HDR:].reshape(5, N)
        recs.append(dict(hdr=r[:TRACE_HDR].copy(), gamma=v[0].copy(), tpoint=v[1].copy(),
                         direc=v[2].copy(), y=v[3].copy(), s=v[4].copy()))
    return res, recs


def bcd_iterations(M, a, b, reg, iters):
    """Wall seconds of ``iters`` BCD iterations (no convergence test)."""
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    g = np.empty(n + m - 1)
    return lib().orc_bcd_iterations(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), int(iters),
                                    _p(g))


# --------------------------------------------------------------------------
# Dense Newton, L-BFGS dual, APDAGD
# --------------------------------------------------------------------------
_LS_ERRORS = {
    -2: "the dense Hessian is not positive definite",
    -3: "the moving direction increases the objective function value",
    -4: "the line search routine failed, possibly due to insufficient numeric precision",
    -5: "the line search routine failed, unable to sufficiently decrease the function value",
}


def _check(rc):
    if rc < 0:
        raise RuntimeError(_LS_ERRORS.get(rc, "oracle error %d" % rc))
    return rc


def newton_direction(M, a, b, reg, gamma):
    """-H^{-1} g with the dense Hessian (src/sinkhorn_problem_obj.cpp:396-418)"""
    M, a, b, gamma = _d(M), _d(a), _d(b), _d(gamma)
    n, m = M.shape
    d = np.empty(n + m - 1)
    _check(lib().orc_newton_direction(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                                      _p(gamma), _p(d)))
    return d


def sinkhorn_newton(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0):
    """src/sinkhorn_newton.cpp:17-129"""
    out = _run(lib().orc_sinkhorn_newton, M, a, b, reg, tol, max_iter, verbose, None, False)
    _check(out["niter"])
    return out


def lbfgs_apply(S, Y, v, mm=6):
    """Two-loop product H v after the pairs (S[k], Y[k]) (BFGSMat.h:86-97,276-302)"""
    S, Y, v = _d(S), _d(Y), _d(v)
    res = np.empty_like(v)
    lib().orc_lbfgs_apply(ctypes.c_long(len(v)), int(mm), int(S.shape[0]), _p(S), _p(Y), _p(v),
                          _p(res))
    return res


def nw_line_search(M, a, b, reg, gamma, direc, init_step):
    """One LineSearchNocedalWright search -> (status, step, f, grad, last_f)"""
    M, a, b, gamma, direc = _d(M), _d(a), _d(b), _d(gamma), _d(direc)
    n, m = M.shape
    g = np.empty(n + m - 1)
    step, f, lf = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_double(0)
    rc = lib().orc_nw_line_search(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), _p(gamma),
                                  _p(direc), ctypes.c_double(init_step), ctypes.byref(step),
                                  ctypes.byref(f), _p(g), ctypes.byref(lf))
    return rc, step.value, f.value, g, lf.value


def sinkhorn_lbfgs_dual(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, hist_cap=None):
    """src/sinkhorn_lbfgs_dual.cpp:84-125 with LBFGSSolver::minimize (src/LBFGSpp/LBFGS.h)"""
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    cap = int(hist_cap if hist_cap else (max_iter + 1 if max_iter > 0 else 100001))
    dual = np.empty(n + m - 1)
    h = [np.empty(cap) for _ in range(3)]
    hl = ctypes.c_int(0)
    niter = lib().orc_sinkhorn_lbfgs_dual(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                                          ctypes.c_double(tol), int(max_iter), int(verbose), cap,
                                          _p(dual), _p(h[0]), _p(h[1]), _p(h[2]),
                                          ctypes.byref(hl))
    _check(niter)
    k = hl.value
    return dict(niter=niter, dual=dual, obj_vals=list(h[0][:k]), mar_errs=list(h[1][:k]),
                run_times=list(h[2][:k]), densities=[])


def sinkhorn_apdagd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0):
    """src/sinkhorn_apdagd.cpp:14-183 -> dict with the primal plan under "plan" """
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    dual = np.empty(n + m - 1)
    plan = np.empty((n, m))
    h = [np.empty(max_iter + 1) for _ in range(3)]
    hl = ctypes.c_int(0)
    niter = lib().orc_sinkhorn_apdagd(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                                      ctypes.c_double(tol), int(max_iter), int(verbose),
                                      _p(dual), _p(plan), _p(h[0]), _p(h[1]), _p(h[2]),
                                      ctypes.byref(hl))
    k = hl.value
    return dict(niter=niter, dual=dual, plan=plan, obj_vals=list(h[0][:k]),
                mar_errs=list(h[1][:k]), run_times=list(h[2][:k]), densities=[])


def approx_proj(plan, a, b):
    """Feasible rounding of a plan (src/approx_proj.h:6-26)"""
    plan, a, b = _d(plan), _d(a), _d(b)
    n, m = plan.shape
    out = np.empty((n, m))
    lib().orc_approx_proj(n, m, _p(plan), _p(a), _p(b), _p(out))
    return out


def primal_val(M, a, b, reg, gamma):
    """Primal objective of the rounded plan (src/sinkhorn_problem_obj.cpp:300-319)"""
    M, a, b, gamma = _d(M), _d(a), _d(b), _d(gamma)
    n, m = M.shape
    return lib().orc_primal_val(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), _p(gamma))



# --------------------------------------------------------------------------
# QROT (quadratically regularised OT): gamma = [alpha (n); beta (m)]
# --------------------------------------------------------------------------
def qrot_dual_obj_grad(M, a, b, reg, gamma, tau=0.0):
    """(f, grad): src/qrot_problem.cpp:37-96, 226-258"""
    M, a, b, gamma = _d(M), _d(a), _d(b), _d(gamma)
    n, m = M.shape
    g = np.empty(n + m)
    f = lib().orc_qrot_dual_obj_grad(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                                     ctypes.c_double(tau), _p(gamma), _p(g))
    return f, g


def qrot_pl_root(y, c):
    """Root of sum_k (x - y_k)_+ = c (src/qrot_problem.cpp:365-394)"""
    y = _d(y)
    return lib().orc_qrot_pl_root(y.shape[0], _p(y), ctypes.c_double(c))


def qrot_optimal_beta(M, b, reg, alpha):
    """src/qrot_problem.cpp:397-406"""
    M, b, alpha = _d(M), _d(b), _d(alpha)
    n, m = M.shape
    beta = np.empty(m)
    lib().orc_qrot_optimal_beta(n, m, _p(M), _p(b), ctypes.c_double(reg), _p(alpha), _p(beta))
    return beta


def qrot_optimal_alpha(M, a, reg, beta):
    """src/qrot_problem.cpp:409-418"""
    M, a, beta = _d(M), _d(a), _d(beta)
    n, m = M.shape
    alpha = np.empty(n)
    lib().orc_qrot_optimal_alpha(n, m, _p(M), _p(a), ctypes.c_double(reg), _p(beta), _p(alpha))
    return alpha


def qrot_get_plan(M, reg, gamma):
    """(alpha_i + beta_j - M_ij)_+ / reg (src/qrot_result.h:35-49)"""
    M, gamma = _d(M), _d(gamma)
    n, m = M.shape
    P = np.empty((n, m))
    lib().orc_qrot_get_plan(n, m, _p(M), ctypes.c_double(reg), _p(gamma), _p(P))
    return P


def qrot_primal_val(M, a, b, reg, gamma):
    """<P, M> + reg/2 ||P||^2 of the rounded plan (src/qrot_problem.cpp:10-26)"""
    M, a, b, gamma = _d(M), _d(a), _d(b), _d(gamma)
    n, m = M.shape
    return lib().orc_qrot_primal_val(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg), _p(gamma))


def _qrot_solve(solver, M, a, b, reg, tol, max_iter, verbose, tau=0.0, x0=None):
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    cap = max_iter + 2
    dual = np.empty(n + m)
    h = [np.empty(cap) for _ in range(4)]
    hl = ctypes.c_int(0)
    x0p = _p(_d(x0)) if x0 is not None else None
    plan = np.empty((n, m))
    niter = lib().orc_qrot_solve(int(solver), n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                                 ctypes.c_double(tau), ctypes.c_double(tol), int(max_iter),
                                 int(verbose), cap, x0p, _p(dual), _p(h[0]), _p(h[1]), _p(h[2]),
                                 _p(h[3]), ctypes.byref(hl), _p(plan))
    _check(niter)
    k = hl.value
    if solver not in (4, 8):
        plan = qrot_get_plan(M, reg, dual)
    return dict(niter=niter, dual=dual, plan=plan,
                obj_vals=list(h[0][:k]), prim_vals=list(h[1][:k]), mar_errs=list(h[2][:k]),
                run_times=list(h[3][:k]))


def qrot_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, x0=None):
    """src/qrot_bcd.cpp:15-101"""
    return _qrot_solve(0, M, a, b, reg, tol, max_iter, verbose, 0.0, x0)


def qrot_gd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, tau=0.0, x0=None):
    """src/qrot_gd.cpp:15-121"""
    return _qrot_solve(1, M, a, b, reg, tol, max_iter, verbose, tau, x0)


def qrot_lbfgs_dual(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, tau=0.0, x0=None):
    """src/qrot_lbfgs_dual.cpp:85-126 with LBFGSSolver::minimize"""
    return _qrot_solve(2, M, a, b, reg, tol, max_iter, verbose, tau, x0)


def qrot_lbfgs_semi_dual(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, x0=None):
    """src/qrot_lbfgs_semi_dual.cpp:78-118 with LBFGSSolver::minimize"""
    return _qrot_solve(3, M, a, b, reg, tol, max_iter, verbose, 0.0, x0)


def qrot_apdagd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, x0=None):
    """src/qrot_apdagd.cpp:14-181 -> dict with the averaged primal plan under "plan" """
    return _qrot_solve(4, M, a, b, reg, tol, max_iter, verbose, 0.0, x0)


def qrot_hess_solve(M, gamma, rhs, shift, tau=0.0, tol=1e-8, max_iter=None):
    """(H + shift I + tau vv')^-1 rhs, H the generalised Hessian at gamma
    (src/qrot_hess.cpp:42-96, 130-187; src/qrot_cg.cpp:9-51) -> (x, cg steps)"""
    M, gamma, rhs = _d(M), _d(gamma), _d(rhs)
    n, m = M.shape
    out = np.empty(n + m)
    it = lib().orc_qrot_hess_solve(n, m, _p(M), _p(gamma), _p(rhs), ctypes.c_double(shift),
                                   ctypes.c_double(tau), ctypes.c_double(tol),
                                   int(2 * m if max_iter is None else max_iter), _p(out))
    return out, it


def qrot_grssn(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, shift=1e-3, method=0,
               tau=0.0, x0=None):
    """src/qrot_grssn.cpp:16-133 -> dict (+ "cg_iters" per Newton step)"""
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    cap = max_iter + 2
    dual = np.empty(n + m)
    h = [np.empty(cap) for _ in range(4)]
    cg = np.zeros(cap, dtype=np.int32)
    hl = ctypes.c_int(0)
    x0p = _p(_d(x0)) if x0 is not None else None
    niter = lib().orc_qrot_grssn(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                                 ctypes.c_double(tau), ctypes.c_double(shift), int(method),
                                 ctypes.c_double(tol), int(max_iter), int(verbose), cap, x0p,
                                 _p(dual), _p(h[0]), _p(h[1]), _p(h[2]), _p(h[3]),
                                 ctypes.byref(hl), cg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    _check(niter)
    k = hl.value
    return dict(niter=niter, dual=dual, plan=qrot_get_plan(M, reg, dual),
                obj_vals=list(h[0][:k]), prim_vals=list(h[1][:k]), mar_errs=list(h[2][:k]),
                run_times=list(h[3][:k]), cg_iters=list(cg[:niter]))


def qrot_assn(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, method=0, tau=0.0, x0=None):
    """src/qrot_assn.cpp:16-170 -> dict (+ "cg_iters" per Newton step)"""
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    cap = max_iter + 2
    dual = np.empty(n + m)
    h = [np.empty(cap) for _ in range(4)]
    cg = np.zeros(cap, dtype=np.int32)
    hl = ctypes.c_int(0)
    x0p = _p(_d(x0)) if x0 is not None else None
    niter = lib().orc_qrot_assn(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                                ctypes.c_double(tau), int(method), ctypes.c_double(tol),
                                int(max_iter), int(verbose), cap, x0p, _p(dual), _p(h[0]),
                                _p(h[1]), _p(h[2]), _p(h[3]), ctypes.byref(hl),
                                cg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    _check(niter)
    k = hl.value
    return dict(niter=niter, dual=dual, plan=qrot_get_plan(M, reg, dual),
                obj_vals=list(h[0][:k]), prim_vals=list(h[1][:k]), mar_errs=list(h[2][:k]),
                run_times=list(h[3][:k]), cg_iters=list(cg[:niter]))


def qrot_s5n(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, method=0, mu0=1.0, tau=0.0,
             x0=None):
    """src/qrot_s5n.cpp:16-196 -> dict (+ "cg_iters" per Newton step)"""
    M, a, b = _d(M), _d(a), _d(b)
    n, m = M.shape
    cap = max_iter + 2
    dual = np.empty(n + m)
    h = [np.empty(cap) for _ in range(4)]
    cg = np.zeros(cap, dtype=np.int32)
    hl = ctypes.c_int(0)
    x0p = _p(_d(x0)) if x0 is not None else None
    niter = lib().orc_qrot_s5n(n, m, _p(M), _p(a), _p(b), ctypes.c_double(reg),
                               ctypes.c_double(tau), int(method), ctypes.c_double(mu0),
                               ctypes.c_double(tol), int(max_iter), int(verbose), cap, x0p,
                               _p(dual), _p(h[0]), _p(h[1]), _p(h[2]), _p(h[3]),
                               ctypes.byref(hl), cg.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    _check(niter)
    k = hl.value
    return dict(niter=niter, dual=dual, plan=qrot_get_plan(M, reg, dual),
                obj_vals=list(h[0][:k]), prim_vals=list(h[1][:k]), mar_errs=list(h[2][:k]),
                run_times=list(h[3][:k]), cg_iters=list(cg[:niter]))


def qrot_pdaam(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, x0=None):
    """src/qrot_pdaam.cpp:14-188 -> dict with the averaged primal plan under "plan" """
    return _qrot_solve(8, M, a, b, reg, tol, max_iter, verbose, 0.0, x0)
