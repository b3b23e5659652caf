"""The FP64 oracle's C restatement under AddressSanitizer and
UndefinedBehaviorSanitizer: oracle/sanitize_check.c runs every entry point
(the BCD, SSNS, sparse Newton, SPLR, Newton, L-BFGS and APDAGD drivers, the
dense passes, both sparsifiers, the Hessian operators and solves, rounding and
the primal value, the QROT solvers) on small problems; any memory error or
undefined behaviour aborts the run with the sanitizer's report."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("gcc") is None,
                    reason="needs make and gcc")
def test_oracle_runs_clean_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize", "CC=gcc"],
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ok (0 non-finite results)" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
