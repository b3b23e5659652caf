"""Shared pytest setup: the `gpu` marker, import paths and seeded inputs."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "regot-python_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def readme_example(n=100, m=80):
    """Input recipe of the reference README (README.md:111-122)."""
    from scipy.stats import expon, norm
    x1 = np.linspace(0.0, 5.0, num=n)
    x2 = np.linspace(0.0, 5.0, num=m)
    a = expon(scale=1.0).pdf(x1)
    a = a / np.sum(a)
    b = 0.2 * norm(loc=1.0, scale=0.2).pdf(x2) + 0.8 * norm(loc=3.0, scale=0.5).pdf(x2)
    b = b / np.sum(b)
    M = np.square(x1.reshape(n, 1) - x2.reshape(1, m))
    return M, a, b


def cloud_problem(n, m, d=16, seed=0, dtype=np.float64):
    """Gaussian point clouds, squared Euclidean cost / max, Dirichlet-mixed marginals."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, d))
    Y = rng.standard_normal((m, d))
    M = (X * X).sum(1)[:, None] + (Y * Y).sum(1)[None, :] - 2.0 * X @ Y.T
    M = np.maximum(M, 0.0)
    M /= M.max()
    M = M.astype(dtype)
    a = 0.9 * rng.dirichlet(np.ones(n)) + 0.1 / n
    b = 0.9 * rng.dirichlet(np.ones(m)) + 0.1 / m
    return M, a / a.sum(), b / b.sum()


@pytest.fixture
def gpu_lib():
    import regot
    from regot import _lib
    L = _lib.lib()
    if L.regot_device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests must run on an MI355X")
    return regot
