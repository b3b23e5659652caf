"""regot -- MI355X (gfx950) drop-in for RegOT's entropic-OT solvers.

Drop-in functions with the reference's signature and result object
(yixuan/regot-python v0.0.3, src/solvers.cpp:507-528, :555-564)::

    sinkhorn_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs)
    sinkhorn_ssns(...), sinkhorn_sparse_newton(...), sinkhorn_splr(...)

Reference kwargs: ``method``, ``mu0``, ``shift``, ``density``
(src/solvers.cpp:314-341); unknown kwargs are ignored, as there.

Extensions (all optional):

* ``bulk``: ``"fp32"`` or ``"fp64"``.  Default: ``"fp32"`` when ``M`` is a
  float32 array, else ``"fp64"``.  FP32 bulk keeps M in FP32 and runs the
  exp/sum in FP32 with every potential, objective and test in FP64; FP64 bulk
  runs everything in FP64.
* ``x0``: warm start of length n+m-1 (the reference's ``SinkhornSolverOpts::x0``).
* ``cg_tol``, ``cg_max_iter``: CG controls of the Newton-type solvers.
* ``ls_max_iter``: strong-Wolfe trials of sparse Newton and SPLR (default 20,
  the reference's fixed value).
* ``device``: GPU ordinal; ``compute_plan``: form ``plan`` (default True).
* ``hess_storage``: how the Newton-type solvers hold the sparsified Hessian:
  -1 (default) matrix-free (products recomputed from M, FP32 bulk, m <= 16385)
  when the kept set is dense, else CSR + CSC; 0 always CSR + CSC; 1
  matrix-free wherever it applies.

Every computation runs in libregot_hip.so on the GPU; there is no CPU path.
"""
import numpy as np

from . import _lib
from ._lib import RegotError, check, dptr, lib

__all__ = [
    "sinkhorn_bcd", "sinkhorn_ssns", "sinkhorn_sparse_newton", "sinkhorn_splr",
    "sinkhorn_newton", "sinkhorn_lbfgs_dual", "sinkhorn_apdagd", "sinkhorn_result",
    "qrot_bcd", "qrot_gd", "qrot_lbfgs_dual", "qrot_lbfgs_semi_dual", "qrot_apdagd",
    "qrot_grssn", "qrot_assn", "qrot_s5n", "qrot_pdaam",
    "qrot_result",
    "Problem", "BatchProblem", "RegotError",
]

__version__ = "0.1.0"


class sinkhorn_result:
    """Result object with the fields of the reference's ``sinkhorn_result``.

    As in the reference (pybind11 ``def_readwrite``), each attribute read
    returns a fresh copy: ``plan`` a C-contiguous float64 (n, m) array,
    ``dual`` a float64 (n+m-1,) array, the histories Python lists of float.
    """

    __slots__ = ("_niter", "_dual", "_plan", "_obj_vals", "_mar_errs", "_run_times",
                 "_densities", "mar_err_fp64", "obj_fp64", "stats")

    def __init__(self):
        self._niter = 0
        self._dual = np.zeros(0)
        self._plan = np.zeros((0, 0))
        self._obj_vals = []
        self._mar_errs = []
        self._run_times = []
        self._densities = []
        self.mar_err_fp64 = None
        self.obj_fp64 = None
        self.stats = {}

    def _prop(name, kind):
        attr = "_" + name

        def get(self):
            v = getattr(self, attr)
            if kind == "arr":
                return np.array(v, dtype=np.float64, copy=True)
            if kind == "list":
                return list(v)
            return v

        def put(self, v):
            if kind == "arr":
                v = np.array(v, dtype=np.float64, copy=True)
            elif kind == "list":
                v = [float(x) for x in v]
            else:
                v = int(v)
            setattr(self, attr, v)

        return property(get, put)

    niter = _prop("niter", "int")
    dual = _prop("dual", "arr")
    plan = _prop("plan", "arr")
    obj_vals = _prop("obj_vals", "list")
    mar_errs = _prop("mar_errs", "list")
    run_times = _prop("run_times", "list")
    densities = _prop("densities", "list")
    del _prop

    def __repr__(self):
        return "<regot.sinkhorn_result niter=%d>" % self._niter


class qrot_result:
    """Result object with the fields of the reference's ``qrot_result``
    (src/solvers.cpp:531-538): niter, dual (n+m,), plan, obj_vals, prim_vals,
    mar_errs, run_times; each read returns a fresh copy."""

    __slots__ = ("_niter", "_dual", "_plan", "_obj_vals", "_prim_vals", "_mar_errs",
                 "_run_times")

    def __init__(self):
        self._niter = 0
        self._dual = np.zeros(0)
        self._plan = np.zeros((0, 0))
        self._obj_vals = []
        self._prim_vals = []
        self._mar_errs = []
        self._run_times = []

    niter = sinkhorn_result.niter
    dual = sinkhorn_result.dual
    plan = sinkhorn_result.plan
    obj_vals = sinkhorn_result.obj_vals
    mar_errs = sinkhorn_result.mar_errs
    run_times = sinkhorn_result.run_times

    @property
    def prim_vals(self):
        return list(self._prim_vals)

    @prim_vals.setter
    def prim_vals(self, v):
        self._prim_vals = [float(x) for x in v]

    def __repr__(self):
        return "<regot.qrot_result niter=%d>" % self._niter


def _as_bulk(M, bulk):
    M = np.asarray(M)
    if M.ndim != 2:
        raise ValueError("M must be a 2-D array (got shape %s)" % (M.shape,))
    if bulk is None:
        bulk = "fp32" if M.dtype == np.float32 else "fp64"
    if bulk not in ("fp32", "fp64"):
        raise ValueError("bulk must be 'fp32' or 'fp64'")
    dt = np.float32 if bulk == "fp32" else np.float64
    return np.ascontiguousarray(M, dtype=dt), bulk


def _vec(x, length, name):
    """FP64 contiguous copy of a 1-D input, checked against the length the C
    side reads (a short host buffer would be read past its end)."""
    v = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    if v.shape[0] != length:
        raise ValueError("%s must have %d entries (got %d)" % (name, length, v.shape[0]))
    return v


class Problem:
    """One EROT problem resident on a GPU (``regot_problem`` handle).

    Operators mirror ``Sinkhorn::Problem`` (src/sinkhorn_problem.h:9-147).
    """

    def __init__(self, M, a, b, reg, bulk=None, device=0):
        Mc, bulk = _as_bulk(M, bulk)
        n, m = Mc.shape
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
        if a.shape[0] != n or b.shape[0] != m:
            raise ValueError("a must have %d entries and b %d (got %d, %d)"
                             % (n, m, a.shape[0], b.shape[0]))
        self.n, self.m, self.reg, self.bulk = n, m, float(reg), bulk
        code = _lib.REGOT_F32 if bulk == "fp32" else _lib.REGOT_F64
        h = _lib._vp()
        check(lib().regot_problem_create(
            _lib.ctypes.byref(h), n, m, Mc.ctypes.data, code, code, 0, 0, dptr(a), dptr(b),
            float(reg), int(device)))
        self._h = h

    @classmethod
    def from_device(cls, ptr, n, m, dtype, a, b, reg, bulk=None, ldm=0, device=0):
        """Borrow (or convert) a device-resident M given by raw pointer."""
        self = cls.__new__(cls)
        a = _vec(a, n, "a")
        b = _vec(b, m, "b")
        dcode = _lib.REGOT_F32 if np.dtype(dtype) == np.float32 else _lib.REGOT_F64
        bulk = bulk or ("fp32" if dcode == _lib.REGOT_F32 else "fp64")
        bcode = _lib.REGOT_F32 if bulk == "fp32" else _lib.REGOT_F64
        self.n, self.m, self.reg, self.bulk = n, m, float(reg), bulk
        h = _lib._vp()
        check(lib().regot_problem_create(_lib.ctypes.byref(h), n, m, _lib._vp(ptr), dcode, bcode,
                                         1, int(ldm), dptr(a), dptr(b), float(reg), int(device)))
        self._h = h
        return self

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().regot_problem_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- operators -------------------------------------------------------
    def optimal_alpha(self, beta):
        """src/sinkhorn_problem_bcd.cpp:177-201"""
        beta = _vec(beta, self.m, "beta")
        out = np.empty(self.n)
        check(lib().regot_optimal_alpha(self._h, dptr(beta), dptr(out)))
        return out

    def optimal_beta(self, alpha):
        """src/sinkhorn_problem_bcd.cpp:149-174"""
        alpha = _vec(alpha, self.n, "alpha")
        out = np.empty(self.m)
        check(lib().regot_optimal_beta(self._h, dptr(alpha), dptr(out)))
        return out

    def dual_obj(self, gamma):
        """src/sinkhorn_problem_obj.cpp:324-342"""
        gamma = _vec(gamma, self.n + self.m - 1, "gamma")
        f = _lib.ctypes.c_double()
        check(lib().regot_dual_obj(self._h, dptr(gamma), _lib.ctypes.byref(f)))
        return f.value

    def dual_obj_grad(self, gamma):
        """src/sinkhorn_problem_obj.cpp:363-393 -> (f, grad)"""
        gamma = _vec(gamma, self.n + self.m - 1, "gamma")
        f = _lib.ctypes.c_double()
        g = np.empty(self.n + self.m - 1)
        check(lib().regot_dual_obj_grad(self._h, dptr(gamma), _lib.ctypes.byref(f), dptr(g)))
        return f.value, g

    def implicit_grad(self, gamma, G_ptr, G_dtype, ldg=0, loss="entropic", cg_tol=1e-10,
                      cg_max_iter=0):
        """dL/dM at the dual optimum ``gamma`` by the implicit function theorem,
        written to the device array at ``G_ptr`` (n x m, row pitch ``ldg``).
        ``loss`` is "entropic" (<T,M> - reg h(T)) or "sharp" (<T,M>).
        Returns (loss value, w, CG iterations, CG relative residual)."""
        gamma = _vec(gamma, self.n + self.m - 1, "gamma")
        code = {"entropic": 0, "sharp": 1}[loss]
        gcode = _lib.REGOT_F32 if np.dtype(G_dtype) == np.float32 else _lib.REGOT_F64
        lv = _lib.ctypes.c_double()
        w = np.empty(self.n + self.m - 1)
        it = _lib.ctypes.c_int()
        err = _lib.ctypes.c_double()
        check(lib().regot_implicit_grad(self._h, dptr(gamma), code, float(cg_tol),
                                        int(cg_max_iter), _lib._vp(int(G_ptr)), gcode,
                                        int(ldg or self.m), _lib.ctypes.byref(lv), dptr(w),
                                        _lib.ctypes.byref(it), _lib.ctypes.byref(err)))
        return lv.value, w, it.value, err.value

    def get_plan(self, gamma):
        """src/sinkhorn_result.h:32-50"""
        gamma = _vec(gamma, self.n + self.m - 1, "gamma")
        plan = np.empty((self.n, self.m))
        check(lib().regot_get_plan(self._h, dptr(gamma), dptr(plan)))
        return plan

    def round_plan(self, gamma):
        """approx_proj of plan(gamma): the nearest-feasible rounding onto the
        transport polytope (src/approx_proj.h:6-26)."""
        g = _vec(gamma, self.n + self.m - 1, "gamma")
        plan = np.empty((self.n, self.m))
        check(lib().regot_round_plan(self._h, dptr(g), dptr(plan)))
        return plan

    def primal_val(self, gamma):
        """Primal objective of the rounded plan (src/sinkhorn_problem_obj.cpp:300-319)."""
        g = _vec(gamma, self.n + self.m - 1, "gamma")
        v = _lib.ctypes.c_double()
        check(lib().regot_primal_val(self._h, dptr(g), _lib.ctypes.byref(v)))
        return v.value

    # --- QROT ---------------------------------------------------------------
    def qrot_obj_grad(self, gamma, tau=0.0):
        """QROT dual (f, grad) at gamma = [alpha; beta] (n+m)
        (src/qrot_problem.cpp:226-258)."""
        g = _vec(gamma, self.n + self.m, "gamma")
        f = _lib.ctypes.c_double()
        grad = np.empty(self.n + self.m)
        check(lib().regot_qrot_obj_grad(self._h, dptr(g), float(tau), _lib.ctypes.byref(f),
                                        dptr(grad)))
        return f.value, grad

    def qrot_optimal_beta(self, alpha):
        """QROT optimal_beta (src/qrot_problem.cpp:397-406)."""
        a = _vec(alpha, self.n, "alpha")
        beta = np.empty(self.m)
        check(lib().regot_qrot_optimal_beta(self._h, dptr(a), dptr(beta)))
        return beta

    def qrot_optimal_alpha(self, beta):
        """QROT optimal_alpha (src/qrot_problem.cpp:409-418)."""
        b = _vec(beta, self.m, "beta")
        alpha = np.empty(self.n)
        check(lib().regot_qrot_optimal_alpha(self._h, dptr(b), dptr(alpha)))
        return alpha

    def qrot_primal_val(self, gamma):
        """QROT primal value of the rounded plan (src/qrot_problem.cpp:10-26)."""
        g = _vec(gamma, self.n + self.m, "gamma")
        v = _lib.ctypes.c_double()
        check(lib().regot_qrot_primal_val(self._h, dptr(g), _lib.ctypes.byref(v)))
        return v.value

    def qrot_hess_solve(self, gamma, rhs, shift, tau=0.0, tol=1e-8, max_iter=0):
        """(H + shift I + tau vv')^-1 rhs with the QROT generalised Hessian at
        gamma (src/qrot_hess.cpp:42-96, src/qrot_cg.cpp:9-51) -> (x, CG steps)."""
        g = _vec(gamma, self.n + self.m, "gamma")
        r = _vec(rhs, self.n + self.m, "rhs")
        out = np.empty(self.n + self.m)
        it = _lib.ctypes.c_int(0)
        check(lib().regot_qrot_hess_solve(self._h, dptr(g), dptr(r), float(shift), float(tau),
                                          float(tol), int(max_iter), dptr(out),
                                          _lib.ctypes.byref(it)))
        return out, it.value

    def qrot_solve(self, solver, tol=1e-6, max_iter=1000, verbose=0, compute_plan=True,
                   **kwargs):
        """QROT solver (_lib.QROT_BCD / QROT_GD / QROT_LBFGS_DUAL /
        QROT_LBFGS_SEMI_DUAL / QROT_APDAGD / QROT_GRSSN / QROT_ASSN / QROT_S5N /
        QROT_PDAAM) -> qrot_result.
        Keywords: x0 (n+m warm start), tau (extra dual term; the reference's
        Python bindings never set it, src/solvers.cpp:33-51); GRSSN: method
        (default 0, CG), shift (default 1e-3), cg_tol, cg_max_iter; ASSN:
        method, cg_tol, cg_max_iter; S5N: those and mu0 (default 1)."""
        opts = _lib.Opts()
        lib().regot_opts_default(_lib.ctypes.byref(opts), 0)
        opts.method = int(kwargs.get("method", 0))
        opts.shift = float(kwargs.get("shift", 1e-3))
        opts.mu0 = float(kwargs.get("mu0", 1.0))
        if "cg_tol" in kwargs:
            opts.cg_tol = float(kwargs["cg_tol"])
        if "cg_max_iter" in kwargs:
            opts.cg_max_iter = int(kwargs["cg_max_iter"])
        if "tau" in kwargs:
            opts.tau = float(kwargs["tau"])
        x0 = kwargs.get("x0")
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
            if x0.shape[0] != self.n + self.m:
                raise ValueError("x0 must have n+m = %d entries" % (self.n + self.m))
            opts.x0 = dptr(x0)
        r = _lib._vp()
        check(lib().regot_qrot_solve(self._h, int(solver), float(tol), int(max_iter),
                                     int(verbose), _lib.ctypes.byref(opts), _lib.ctypes.byref(r)))
        try:
            L = lib()
            res = qrot_result()
            hl = L.regot_result_hist_len(r)
            dual = np.empty(L.regot_result_dual_len(r))
            h = [np.empty(max(hl, 1)) for _ in range(4)]
            check(L.regot_result_get(r, dptr(dual), dptr(h[0]), dptr(h[1]), dptr(h[2]), None))
            check(L.regot_result_prim_vals(r, dptr(h[3])))
            res._niter = int(L.regot_result_niter(r))
            res._dual = dual
            res._obj_vals = [float(x) for x in h[0][:hl]]
            res._mar_errs = [float(x) for x in h[1][:hl]]
            res._run_times = [float(x) for x in h[2][:hl]]
            res._prim_vals = [float(x) for x in h[3][:hl]]
            if compute_plan:
                plan = np.empty((self.n, self.m))
                check(L.regot_result_plan(self._h, r, dptr(plan)))
                res._plan = plan
        finally:
            lib().regot_result_free(r)
        return res

    # --- sparsified Hessian ------------------------------------------------
    def sparsified_hess(self, gamma, delta, density_hint=0.001):
        """Threshold rule at gamma (src/sinkhorn_sparsify3.cpp:276-405); returns nnz."""
        gamma = _vec(gamma, self.n + self.m - 1, "gamma")
        nnz = _lib.ctypes.c_longlong()
        check(lib().regot_sparsified_hess(self._h, dptr(gamma), float(delta), float(density_hint),
                                          _lib.ctypes.byref(nnz)))
        return nnz.value

    def sparsified_hess_with_density(self, gamma, density):
        """Top-k rule at gamma (src/sinkhorn_sparsify4.cpp:77-174); returns stored nnz."""
        gamma = _vec(gamma, self.n + self.m - 1, "gamma")
        nnz = _lib.ctypes.c_longlong()
        check(lib().regot_sparsified_hess_with_density(self._h, dptr(gamma), float(density),
                                                       _lib.ctypes.byref(nnz)))
        return nnz.value

    def set_hess_storage(self, mode):
        """-1: the solver drivers go matrix-free when the kept set is dense
        (FP32 bulk, m - 1 <= 16384), sparsified_hess* store CSR + CSC; 0: always
        stored; 1: matrix-free wherever it applies (hess_export then fails)."""
        check(lib().regot_set_hess_storage(self._h, int(mode)))

    def hess_info(self):
        """(stored nnz, FP64-zero kept entries, Hessian density)"""
        a = _lib.ctypes.c_longlong()
        z = _lib.ctypes.c_longlong()
        d = _lib.ctypes.c_double()
        check(lib().regot_hess_info(self._h, _lib.ctypes.byref(a), _lib.ctypes.byref(z),
                                    _lib.ctypes.byref(d)))
        return a.value, z.value, d.value

    def hess_export(self):
        """dict with h1, h2, CSC (colptr, rowind, cval) and CSR (rowptr, colind, rval)."""
        nnz = self.hess_info()[0]
        n, m = self.n, self.m
        out = dict(h1=np.empty(n), h2=np.empty(m - 1), colptr=np.empty(m, np.int32),
                   rowind=np.empty(max(nnz, 1), np.int32), cval=np.empty(max(nnz, 1)),
                   rowptr=np.empty(n + 1, np.int32), colind=np.empty(max(nnz, 1), np.int32),
                   rval=np.empty(max(nnz, 1)))
        check(lib().regot_hess_export(self._h, dptr(out["h1"]), dptr(out["h2"]),
                                      _lib.iptr(out["colptr"]), _lib.iptr(out["rowind"]),
                                      dptr(out["cval"]), _lib.iptr(out["rowptr"]),
                                      _lib.iptr(out["colind"]), dptr(out["rval"])))
        for k in ("rowind", "cval", "colind", "rval"):
            out[k] = out[k][:nnz]
        return out

    def hess_apply_delta(self, x, shift):
        x = _vec(x, self.m - 1, "x")
        out = np.empty(self.m - 1)
        check(lib().regot_hess_apply_delta(self._h, dptr(x), float(shift), dptr(out)))
        return out

    def hess_apply_hs(self, x):
        x = _vec(x, self.n + self.m - 1, "x")
        out = np.empty(self.n + self.m - 1)
        check(lib().regot_hess_apply_hs(self._h, dptr(x), dptr(out)))
        return out

    def hess_cg(self, rhs, shift, tol=1e-8, max_iter=0):
        """(Hs + shift I)^{-1} rhs -> (x, iterations, relative residual)"""
        rhs = _vec(rhs, self.n + self.m - 1, "rhs")
        out = np.empty(self.n + self.m - 1)
        it = _lib.ctypes.c_int()
        te = _lib.ctypes.c_double()
        check(lib().regot_hess_cg(self._h, dptr(rhs), float(shift), float(tol), int(max_iter),
                                  dptr(out), _lib.ctypes.byref(it), _lib.ctypes.byref(te)))
        return out, it.value, te.value

    def hess_cg2(self, rhs0, rhs1, shift, tol=1e-8, max_iter=0):
        """Two right-hand sides in the same CG launches -> ((x0, x1), (it0, it1),
        (err0, err1)); each as hess_cg would return it."""
        N = self.n + self.m - 1
        rhs0 = _vec(rhs0, N, "rhs0")
        rhs1 = _vec(rhs1, N, "rhs1")
        out0 = np.empty(N)
        out1 = np.empty(N)
        it = (_lib.ctypes.c_int * 2)()
        te = (_lib.ctypes.c_double * 2)()
        check(lib().regot_hess_cg2(self._h, dptr(rhs0), dptr(rhs1), float(shift), float(tol),
                                   int(max_iter), dptr(out0), dptr(out1), it, te))
        return (out0, out1), (it[0], it[1]), (te[0], te[1])

    def profile(self, on=True):
        check(lib().regot_profile_enable(self._h, int(bool(on))))

    def profile_stats(self, kind):
        """(launches, total ms) of kind 0 = row pass, 1 = column pass, 2 = single-read
        BCD pass, 3 = its column-sum combine."""
        cnt = _lib.ctypes.c_longlong()
        ms = _lib.ctypes.c_double()
        check(lib().regot_profile_get(self._h, int(kind), _lib.ctypes.byref(cnt),
                                      _lib.ctypes.byref(ms)))
        return cnt.value, ms.value

    def profile_reset(self):
        check(lib().regot_profile_reset(self._h))

    # --- solvers ---------------------------------------------------------
    def solve(self, solver, tol=1e-6, max_iter=1000, verbose=0, compute_plan=True,
              verify=False, **kwargs):
        opts = _lib.Opts()
        lib().regot_opts_default(_lib.ctypes.byref(opts), solver)
        if "method" in kwargs:
            opts.method = int(kwargs["method"])
        if "mu0" in kwargs:
            opts.mu0 = float(kwargs["mu0"])
        if "shift" in kwargs:
            opts.shift = float(kwargs["shift"])
        if "density" in kwargs:
            opts.density = float(kwargs["density"])
        if "cg_tol" in kwargs:
            opts.cg_tol = float(kwargs["cg_tol"])
        if "cg_max_iter" in kwargs:
            opts.cg_max_iter = int(kwargs["cg_max_iter"])
        if "bcd_batch" in kwargs:
            opts.bcd_batch = int(kwargs["bcd_batch"])
        if "bcd_fused" in kwargs:
            opts.bcd_fused = int(kwargs["bcd_fused"])
        if "ls_max_iter" in kwargs:
            opts.ls_max_iter = int(kwargs["ls_max_iter"])
        x0 = kwargs.get("x0")
        if x0 is not None:
            x0 = np.ascontiguousarray(x0, dtype=np.float64).reshape(-1)
            if x0.shape[0] != self.n + self.m - 1:
                raise ValueError("x0 must have n+m-1 = %d entries" % (self.n + self.m - 1))
            opts.x0 = dptr(x0)
        r = _lib._vp()
        check(lib().regot_solve(self._h, int(solver), float(tol), int(max_iter), int(verbose),
                                _lib.ctypes.byref(opts), _lib.ctypes.byref(r)))
        try:
            L = lib()
            res = sinkhorn_result()
            hl = L.regot_result_hist_len(r)
            dl = L.regot_result_dens_len(r)
            dual = np.empty(self.n + self.m - 1)
            h = [np.empty(max(hl, 1)) for _ in range(3)]
            dens = np.empty(max(dl, 1))
            check(L.regot_result_get(r, dptr(dual), dptr(h[0]), dptr(h[1]), dptr(h[2]),
                                     dptr(dens)))
            res._niter = int(L.regot_result_niter(r))
            res._dual = dual
            res._obj_vals = [float(x) for x in h[0][:hl]]
            res._mar_errs = [float(x) for x in h[1][:hl]]
            res._run_times = [float(x) for x in h[2][:hl]]
            res._densities = [float(x) for x in dens[:dl]]
            st = np.zeros(8)
            check(L.regot_result_stats(r, dptr(st)))
            res.stats = {"objgrad_ms": st[0], "objgrad_calls": int(st[1]),
                         "sparsify_ms": st[2], "sparsify_calls": int(st[3]),
                         "cg_ms": st[4], "cg_solves": int(st[5]), "cg_iters": int(st[6]),
                         "alg_bytes": float(st[7]),
                         "precise_from": int(L.regot_result_precise_from(r))}
            if compute_plan:
                plan = np.empty((self.n, self.m))
                check(L.regot_result_plan(self._h, r, dptr(plan)))
                res._plan = plan
            if verify:
                me = _lib.ctypes.c_double()
                ob = _lib.ctypes.c_double()
                check(L.regot_result_verify(self._h, r, _lib.ctypes.byref(me),
                                            _lib.ctypes.byref(ob)))
                res.mar_err_fp64, res.obj_fp64 = me.value, ob.value
        finally:
            lib().regot_result_free(r)
        return res


class BatchProblem:
    """Many independent EROT problems resident on one GPU, solved together by
    BCD (``regot_batch`` handle; no counterpart in the reference).

    ``Ms`` is a sequence of 2-D cost matrices (NumPy arrays, or
    ``(device_ptr, n, m, dtype)`` tuples of device-resident row-major
    matrices), ``as_``/``bs`` the marginals and ``regs`` one reg per problem
    (or a scalar).  ``bulk`` as for :class:`Problem` (default: FP32 when every
    M is float32).
    """

    def __init__(self, Ms, as_, bs, regs, bulk=None, device=0):
        Ms = list(Ms)
        P = len(Ms)
        if P == 0:
            raise ValueError("empty batch")
        regs = np.broadcast_to(np.asarray(regs, dtype=np.float64), (P,)).copy()
        on_dev = all(isinstance(x, tuple) for x in Ms)
        if not on_dev and any(isinstance(x, tuple) for x in Ms):
            raise ValueError("mix of host arrays and device pointers")
        keep = []
        if on_dev:
            dts = {np.dtype(x[3]) for x in Ms}
            if len(dts) != 1:
                raise ValueError("device matrices must share one dtype")
            mdt = dts.pop()
            ns = np.array([x[1] for x in Ms], dtype=np.int32)
            ms = np.array([x[2] for x in Ms], dtype=np.int32)
            ptrs = [int(x[0]) for x in Ms]
        else:
            if bulk is None:
                bulk = "fp32" if all(np.asarray(x).dtype == np.float32 for x in Ms) else "fp64"
            mdt = np.dtype(np.float32 if bulk == "fp32" else np.float64)
            arrs = [np.ascontiguousarray(x, dtype=mdt) for x in Ms]
            if any(x.ndim != 2 for x in arrs):
                raise ValueError("every M must be 2-D")
            keep += arrs
            ns = np.array([x.shape[0] for x in arrs], dtype=np.int32)
            ms = np.array([x.shape[1] for x in arrs], dtype=np.int32)
            ptrs = [x.ctypes.data for x in arrs]
        bulk = bulk or ("fp32" if mdt == np.float32 else "fp64")
        if bulk not in ("fp32", "fp64"):
            raise ValueError("bulk must be 'fp32' or 'fp64'")
        av = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in as_]
        bv = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in bs]
        if len(av) != P or len(bv) != P:
            raise ValueError("need one a and one b per problem")
        for p in range(P):
            if av[p].shape[0] != ns[p] or bv[p].shape[0] != ms[p]:
                raise ValueError("problem %d: marginal sizes do not match M" % p)
        keep += av + bv
        VP = _lib._vp * P
        mcode = _lib.REGOT_F32 if mdt == np.float32 else _lib.REGOT_F64
        bcode = _lib.REGOT_F32 if bulk == "fp32" else _lib.REGOT_F64
        h = _lib._vp()
        check(lib().regot_batch_create(
            _lib.ctypes.byref(h), P, _lib.iptr(ns), _lib.iptr(ms), VP(*ptrs), mcode, bcode,
            int(on_dev), VP(*[x.ctypes.data for x in av]), VP(*[x.ctypes.data for x in bv]),
            dptr(regs), int(device)))
        self._h = h
        self.nprob, self.bulk = P, bulk
        self.n, self.m = ns.copy(), ms.copy()
        tot = _lib.ctypes.c_longlong()
        dl = _lib.ctypes.c_longlong()
        check(lib().regot_batch_info(h, None, _lib.ctypes.byref(tot), _lib.ctypes.byref(dl)))
        self.total_elems, self.dual_len = tot.value, dl.value

    def solve_bcd(self, tol=1e-6, max_iter=1000, dual=True, graph=False, autotune=False,
                  single_read=True):
        """BCD on every problem.  Returns a dict of per-problem arrays:
        ``niter``, ``obj`` and ``mar_err`` (at the final iterate), ``dual``
        (a list of (n+m-1,) arrays, or None), ``rounds`` launched, the
        hipGraph instantiations / launches so far (``graph=True`` replays
        captured groups of 8 rounds instead of launching the rounds' kernels
        one by one; results are the same), the coarse
        tile configuration (``autotune=True`` times the candidates once per
        handle and keeps the fastest) and ``single_read``, per problem whether
        its rounds read its cost matrix once per iteration (FP32 problems
        with 1024 < m <= 4096 when ``single_read=True``) or twice."""
        check(lib().regot_batch_set_option(self._h, 0, 1 if graph else 0))
        check(lib().regot_batch_set_option(self._h, 1, 1 if autotune else 0))
        check(lib().regot_batch_set_option(self._h, 2, 1 if single_read else 0))
        P = self.nprob
        it = np.zeros(P, dtype=np.int32)
        ob = np.zeros(P)
        me = np.zeros(P)
        du = np.zeros(self.dual_len) if dual else None
        rounds = _lib.ctypes.c_longlong()
        check(lib().regot_batch_solve_bcd(self._h, float(tol), int(max_iter), _lib.iptr(it),
                                          dptr(ob), dptr(me), dptr(du) if dual else None,
                                          _lib.ctypes.byref(rounds)))
        st = np.zeros(12)
        check(lib().regot_batch_stats(self._h, dptr(st)))
        kinds = np.zeros(P, dtype=np.int32)
        check(lib().regot_batch_round_kinds(self._h, _lib.iptr(kinds)))
        out = {"niter": it, "obj": ob, "mar_err": me, "rounds": rounds.value, "dual": None,
               "graph_builds": int(st[3]), "graph_launches": int(st[4]),
               "rows_per_wave": int(st[0]), "block_rows": int(st[1]), "tuned": bool(st[2]),
               "tune_ms": [float(x) for x in st[5:11]], "single_read": kinds.astype(bool)}
        if dual:
            sizes = self.n.astype(np.int64) + self.m - 1
            out["dual"] = np.split(du, np.cumsum(sizes)[:-1])
        return out

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            lib().regot_batch_destroy(h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def sqeuclid_bf16(X_ptr, ldx, Y_ptr, ldy, n, m, d, M_ptr, ldm=0, device=0):
    """M = |x|^2 + |y|^2 - 2 X Y' from BF16 device arrays on the matrix cores
    (FP32 device output at ``M_ptr``); see :func:`regot.autograd.sqeuclid`."""
    check(lib().regot_sqeuclid_bf16(_lib._vp(int(X_ptr)), int(ldx), _lib._vp(int(Y_ptr)),
                                    int(ldy), int(n), int(m), int(d), _lib._vp(int(M_ptr)),
                                    int(ldm or m), int(device)))


def sinkhorn_bcd_batch(Ms, as_, bs, regs, tol=1e-6, max_iter=1000, bulk=None, device=0):
    """Batched ``sinkhorn_bcd`` over independent problems (one GPU).  Returns
    the per-problem dict of :meth:`BatchProblem.solve_bcd`."""
    with BatchProblem(Ms, as_, bs, regs, bulk=bulk, device=device) as B:
        return B.solve_bcd(tol=tol, max_iter=max_iter)


_HANDLE_KW = ("bulk", "device", "compute_plan", "verify", "hess_storage")


def _run(solver, M, a, b, reg, tol, max_iter, verbose, kwargs):
    rest = {k: v for k, v in kwargs.items() if k not in _HANDLE_KW}
    with Problem(M, a, b, reg, bulk=kwargs.get("bulk"), device=kwargs.get("device", 0)) as p:
        if "hess_storage" in kwargs:
            p.set_hess_storage(kwargs["hess_storage"])
        return p.solve(solver, tol=tol, max_iter=max_iter, verbose=verbose,
                       compute_plan=kwargs.get("compute_plan", True),
                       verify=kwargs.get("verify", False), **rest)


def sinkhorn_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Plain Sinkhorn (block coordinate descent on the dual).

    Drop-in for ``regot.sinkhorn_bcd`` (src/solvers.cpp:359-373,
    src/sinkhorn_bcd.cpp:14-124).
    """
    return _run(_lib.SOLVER_BCD, M, a, b, reg, tol, max_iter, verbose, kwargs)


def sinkhorn_ssns(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Safe and sparse Newton (src/solvers.cpp:425-441, src/sinkhorn_ssns.cpp:16-234)."""
    return _run(_lib.SOLVER_SSNS, M, a, b, reg, tol, max_iter, verbose, kwargs)


def sinkhorn_sparse_newton(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Sparse Newton (src/solvers.cpp:407-423, src/sinkhorn_sparse_newton.cpp:24-170)."""
    return _run(_lib.SOLVER_SPARSE_NEWTON, M, a, b, reg, tol, max_iter, verbose, kwargs)


def sinkhorn_splr(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Sparse-plus-low-rank quasi-Newton (src/solvers.cpp:443-459, src/sinkhorn_splr.cpp:24-199)."""
    return _run(_lib.SOLVER_SPLR, M, a, b, reg, tol, max_iter, verbose, kwargs)


def sinkhorn_newton(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Newton's method with the full dual Hessian and Armijo backtracking
    (src/solvers.cpp:391-405, src/sinkhorn_newton.cpp:17-129)."""
    return _run(_lib.SOLVER_NEWTON, M, a, b, reg, tol, max_iter, verbose, kwargs)


def sinkhorn_lbfgs_dual(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """L-BFGS on the dual (src/solvers.cpp:375-389, src/sinkhorn_lbfgs_dual.cpp:84-125).
    A failed line search raises RuntimeError with LBFGSpp's message."""
    return _run(_lib.SOLVER_LBFGS_DUAL, M, a, b, reg, tol, max_iter, verbose, kwargs)


def sinkhorn_apdagd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Accelerated primal-dual adaptive gradient descent; ``plan`` is the
    averaged primal plan (src/solvers.cpp:343-357, src/sinkhorn_apdagd.cpp:14-183)."""
    return _run(_lib.SOLVER_APDAGD, M, a, b, reg, tol, max_iter, verbose, kwargs)



def _run_qrot(solver, M, a, b, reg, tol, max_iter, verbose, kwargs):
    # the reference parses only method and shift for QROT (src/solvers.cpp:33-51);
    # other keywords are ignored, like there
    with Problem(M, a, b, reg, bulk=kwargs.get("bulk"), device=kwargs.get("device", 0)) as p:
        kw = {k: kwargs[k] for k in ("method", "shift") if k in kwargs}
        return p.qrot_solve(solver, tol=tol, max_iter=max_iter, verbose=verbose,
                            compute_plan=kwargs.get("compute_plan", True), **kw)


def qrot_bcd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Block coordinate descent on the QROT dual (src/solvers.cpp:85-99,
    src/qrot_bcd.cpp:15-101)."""
    return _run_qrot(_lib.QROT_BCD, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_gd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Gradient descent on the QROT dual (src/solvers.cpp:101-115, src/qrot_gd.cpp:15-121)."""
    return _run_qrot(_lib.QROT_GD, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_lbfgs_dual(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """L-BFGS on the QROT dual (src/solvers.cpp:134-148, src/qrot_lbfgs_dual.cpp:85-126)."""
    return _run_qrot(_lib.QROT_LBFGS_DUAL, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_lbfgs_semi_dual(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """L-BFGS on the QROT semi-dual (src/solvers.cpp:150-164,
    src/qrot_lbfgs_semi_dual.cpp:78-118)."""
    return _run_qrot(_lib.QROT_LBFGS_SEMI_DUAL, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_apdagd(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """APDAGD on the QROT dual (src/solvers.cpp:53-67, src/qrot_apdagd.cpp:14-181);
    the result's plan is the averaged primal plan."""
    return _run_qrot(_lib.QROT_APDAGD, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_grssn(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Globalised regularised semismooth Newton on the QROT dual
    (src/solvers.cpp:117-131 binding, src/qrot_grssn.cpp:16-133); keywords
    method (default 0) and shift (default 1e-3)."""
    return _run_qrot(_lib.QROT_GRSSN, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_assn(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Adaptive semismooth Newton on the QROT dual (src/solvers.cpp:69-83
    binding, src/qrot_assn.cpp:16-170); keyword method (default 0)."""
    return _run_qrot(_lib.QROT_ASSN, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_s5n(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Safeguarded semismooth Newton on the QROT dual (src/solvers.cpp:183-197
    binding, src/qrot_s5n.cpp:16-196); keyword method (default 0)."""
    return _run_qrot(_lib.QROT_S5N, M, a, b, reg, tol, max_iter, verbose, kwargs)


def qrot_pdaam(M, a, b, reg, tol=1e-6, max_iter=1000, verbose=0, **kwargs):
    """Primal-dual accelerated alternating minimisation on the QROT dual
    (src/solvers.cpp:166-180 binding, src/qrot_pdaam.cpp:14-188); the result's plan
    is the averaged primal plan."""
    return _run_qrot(_lib.QROT_PDAAM, M, a, b, reg, tol, max_iter, verbose, kwargs)
