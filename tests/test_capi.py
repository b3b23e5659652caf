"""CPU checks of the drop-in boundary: the library loads, exports exactly the
declared C-ABI, and fails loudly (no silent CPU path) when no GPU is visible."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "regot_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(regot_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("regot_problem_create", "regot_optimal_alpha", "regot_optimal_beta",
                 "regot_dual_obj", "regot_dual_obj_grad", "regot_get_plan",
                 "regot_sparsified_hess", "regot_sparsified_hess_with_density",
                 "regot_hess_apply_delta", "regot_hess_apply_hs", "regot_hess_cg",
                 "regot_solve", "regot_result_plan"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from regot import _lib
    L = _lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert L.regot_version() >= 100


def test_no_gpu_fails_loudly():
    import regot
    from regot import _lib
    if _lib.lib().regot_device_count() > 0:
        pytest.skip("a GPU is visible; the no-GPU path is not reachable here")
    M = np.ones((3, 2))
    with pytest.raises(regot.RegotError, match="no GPU"):
        regot.sinkhorn_bcd(M, np.ones(3) / 3, np.ones(2) / 2, 0.1)


def test_input_validation_before_any_device_work():
    import regot
    with pytest.raises(ValueError):
        regot.sinkhorn_bcd(np.ones(3), np.ones(3), np.ones(3), 0.1)
    with pytest.raises(ValueError):
        regot.sinkhorn_bcd(np.ones((3, 2)), np.ones(4), np.ones(2), 0.1)
    with pytest.raises(ValueError):
        regot.Problem(np.ones((3, 2)), np.ones(3), np.ones(2), 0.1, bulk="fp16")


def test_result_object_copies_like_def_readwrite():
    import regot
    r = regot.sinkhorn_result()
    r.dual = np.arange(3.0)
    d = r.dual
    d[0] = 99.0
    assert r.dual[0] == 0.0
    r.obj_vals = [1, 2]
    assert r.obj_vals == [1.0, 2.0] and isinstance(r.obj_vals[0], float)
    r.niter = 5.0
    assert r.niter == 5 and isinstance(r.niter, int)


def test_operator_vectors_checked_against_the_handle():
    # every Problem operator hands a host pointer to C, which reads n, m,
    # n+m-1 or n+m doubles from it: a short vector must raise before the call
    import regot
    P = regot.Problem.__new__(regot.Problem)
    P.n, P.m, P.reg, P.bulk, P._h = 5, 4, 0.1, "fp64", None
    bad = np.zeros(3)
    for call in (lambda: P.optimal_alpha(bad), lambda: P.optimal_beta(bad),
                 lambda: P.dual_obj(bad), lambda: P.dual_obj_grad(bad),
                 lambda: P.get_plan(bad), lambda: P.round_plan(bad),
                 lambda: P.primal_val(bad), lambda: P.sparsified_hess(bad, 0.1),
                 lambda: P.sparsified_hess_with_density(bad, 0.1),
                 lambda: P.hess_apply_delta(np.zeros(4), 0.1), lambda: P.hess_apply_hs(bad),
                 lambda: P.hess_cg(bad, 0.1), lambda: P.qrot_obj_grad(np.zeros(8)),
                 lambda: P.qrot_optimal_alpha(bad), lambda: P.qrot_optimal_beta(bad),
                 lambda: P.qrot_primal_val(np.zeros(8)),
                 lambda: P.qrot_hess_solve(np.zeros(9), bad, 0.1),
                 lambda: P.implicit_grad(bad, 0, np.float32)):
        with pytest.raises(ValueError):
            call()
    with pytest.raises(ValueError):
        regot.Problem.from_device(0, 5, 4, np.float32, np.ones(4), np.ones(4), 0.1)
