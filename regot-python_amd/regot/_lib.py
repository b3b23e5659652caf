"""ctypes binding of libregot_hip.so (the C-ABI declared in include/regot_hip.h).

The shared library is built in-tree next to this file by
``regot-python_amd/build.py``.  There is no CPU fallback: if the library or a
GPU is missing, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libregot_hip.so")

REGOT_F32 = 0
REGOT_F64 = 1
SOLVER_BCD = 0
SOLVER_SSNS = 1
SOLVER_SPARSE_NEWTON = 2
SOLVER_SPLR = 3
SOLVER_NEWTON = 4
SOLVER_LBFGS_DUAL = 5
SOLVER_APDAGD = 6
QROT_BCD = 0
QROT_GD = 1
QROT_LBFGS_DUAL = 2
QROT_LBFGS_SEMI_DUAL = 3
QROT_APDAGD = 4
QROT_GRSSN = 5
QROT_ASSN = 6
QROT_S5N = 7
QROT_PDAAM = 8

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class RegotError(RuntimeError):
    """A C-ABI call returned a negative status."""


class Opts(ctypes.Structure):
    """Mirror of regot_opts (include/regot_hip.h)."""
    _fields_ = [
        ("method", ctypes.c_int),
        ("mu0", ctypes.c_double),
        ("shift", ctypes.c_double),
        ("density", ctypes.c_double),
        ("cg_tol", ctypes.c_double),
        ("cg_max_iter", ctypes.c_int),
        ("x0", _dp),
        ("bcd_batch", ctypes.c_int),
        ("bcd_fused", ctypes.c_int),
        ("tau", ctypes.c_double),
        ("ls_max_iter", ctypes.c_int),
    ]


_lib = None


def _torch_runtime_first():
    """PyTorch-ROCm bundles its own HIP runtime.  Two HIP runtimes can share a
    process and a device, but only if torch's initialises first (the other
    order leaves torch without GPUs), so when torch is installed its device
    runtime is brought up before this library is loaded."""
    if os.environ.get("REGOT_SKIP_TORCH_INIT") == "1":
        return
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


def lib():
    """Load the library once; raise if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    _torch_runtime_first()
    if not os.path.exists(LIB_PATH):
        raise RegotError(
            "libregot_hip.so is not built (expected at %s); run "
            "`python regot-python_amd/build.py` or __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    L.regot_last_error.restype = ctypes.c_char_p
    L.regot_problem_create.argtypes = [
        ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_longlong, _dp, _dp, ctypes.c_double, ctypes.c_int]
    L.regot_problem_destroy.argtypes = [_vp]
    L.regot_check_exp2.argtypes = [_dp, _dp, ctypes.c_longlong, ctypes.c_int]
    L.regot_check_log.argtypes = [_dp, _dp, ctypes.c_longlong, ctypes.c_int]
    L.regot_problem_dims.argtypes = [_vp, ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    for name in ("regot_optimal_alpha", "regot_optimal_beta"):
        getattr(L, name).argtypes = [_vp, _dp, _dp]
    L.regot_dual_obj.argtypes = [_vp, _dp, _dp]
    L.regot_dual_obj_grad.argtypes = [_vp, _dp, _dp, _dp]
    L.regot_dual_obj_grad_noalloc.argtypes = [
        _vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, _vp, _vp,
        ctypes.c_double, _vp, _vp, _vp, _vp, ctypes.c_int]
    L.regot_get_plan.argtypes = [_vp, _dp, _dp]
    L.regot_round_plan.argtypes = [_vp, _dp, _dp]
    L.regot_primal_val.argtypes = [_vp, _dp, _dp]
    L.regot_qrot_solve.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.POINTER(Opts), ctypes.POINTER(_vp)]
    L.regot_result_prim_vals.argtypes = [_vp, _dp]
    L.regot_result_dual_len.argtypes = [_vp]
    L.regot_qrot_obj_grad.argtypes = [_vp, _dp, ctypes.c_double, _dp, _dp]
    L.regot_qrot_optimal_beta.argtypes = [_vp, _dp, _dp]
    L.regot_qrot_optimal_alpha.argtypes = [_vp, _dp, _dp]
    L.regot_qrot_primal_val.argtypes = [_vp, _dp, _dp]
    L.regot_qrot_hess_solve.argtypes = [_vp, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, _dp,
                                        ctypes.POINTER(ctypes.c_int)]
    L.regot_opts_default.argtypes = [ctypes.POINTER(Opts), ctypes.c_int]
    L.regot_opts_default.restype = None
    L.regot_solve.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                              ctypes.POINTER(Opts), ctypes.POINTER(_vp)]
    for name in ("regot_result_niter", "regot_result_hist_len", "regot_result_dens_len"):
        getattr(L, name).argtypes = [_vp]
    L.regot_result_get.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp]
    L.regot_result_plan.argtypes = [_vp, _vp, _dp]
    L.regot_result_verify.argtypes = [_vp, _vp, _dp, _dp]
    L.regot_result_stats.argtypes = [_vp, _dp]
    L.regot_result_precise_from.argtypes = [_vp]
    L.regot_result_free.argtypes = [_vp]
    L.regot_result_free.restype = None
    _llp = ctypes.POINTER(ctypes.c_longlong)
    _ip = ctypes.POINTER(ctypes.c_int)
    L.regot_sparsified_hess.argtypes = [_vp, _dp, ctypes.c_double, ctypes.c_double, _llp]
    L.regot_sparsified_hess_with_density.argtypes = [_vp, _dp, ctypes.c_double, _llp]
    L.regot_hess_info.argtypes = [_vp, _llp, _llp, _dp]
    L.regot_set_hess_storage.argtypes = [_vp, ctypes.c_int]
    L.regot_hess_export.argtypes = [_vp, _dp, _dp, _ip, _ip, _dp, _ip, _ip, _dp]
    L.regot_hess_apply_delta.argtypes = [_vp, _dp, ctypes.c_double, _dp]
    L.regot_hess_apply_hs.argtypes = [_vp, _dp, _dp]
    L.regot_hess_cg2.argtypes = [_vp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                 _dp, _dp, _ip, _dp]
    L.regot_hess_cg.argtypes = [_vp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp,
                                _ip, _dp]
    _pp = ctypes.POINTER(_vp)
    L.regot_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, _ip, _ip, _pp,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, _pp, _pp, _dp,
                                     ctypes.c_int]
    L.regot_batch_destroy.argtypes = [_vp]
    L.regot_batch_info.argtypes = [_vp, _ip, _llp, _llp]
    L.regot_batch_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_longlong]
    L.regot_batch_stats.argtypes = [_vp, _dp]
    L.regot_batch_round_kinds.argtypes = [_vp, _ip]
    L.regot_batch_solve_bcd.argtypes = [_vp, ctypes.c_double, ctypes.c_int, _ip, _dp, _dp, _dp,
                                        _llp]
    L.regot_sqeuclid_bf16.argtypes = [_vp, ctypes.c_longlong, _vp, ctypes.c_longlong,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                      ctypes.c_longlong, ctypes.c_int]
    L.regot_implicit_grad.argtypes = [_vp, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                      _vp, ctypes.c_int, ctypes.c_longlong, _dp, _dp, _ip, _dp]
    L.regot_profile_enable.argtypes = [_vp, ctypes.c_int]
    L.regot_profile_get.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _dp]
    L.regot_profile_reset.argtypes = [_vp]
    _lib = L
    return L


def check(status):
    """Raise RegotError with the library's message on a negative status."""
    if status < 0:
        msg = lib().regot_last_error().decode("utf-8", "replace")
        raise RegotError("libregot_hip error %d: %s" % (status, msg))
    return status


def dptr(arr):
    return arr.ctypes.data_as(_dp)


def iptr(arr):
    return arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
