"""Writes tests/golden/*.npz: the entropic optimum of the reference README
recipe at config-1 size (n=1000, m=800, reg=0.1, README.md:111-122 scaled),
computed with an independent log-domain Sinkhorn in scipy (not the oracle),
iterated until the potentials stop changing in FP64.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
from scipy.special import logsumexp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from conftest import readme_example  # noqa: E402


def solve(M, a, b, reg, iters=200000):
    f = np.zeros(M.shape[0])
    g = np.zeros(M.shape[1])
    la, lb = np.log(a), np.log(b)
    for it in range(iters):
        f_new = reg * (la - logsumexp((g[None, :] - M) / reg, axis=1))
        g = reg * (lb - logsumexp((f_new[:, None] - M) / reg, axis=0))
        done = np.max(np.abs(f_new - f)) < 1e-14
        f = f_new
        if done:
            break
    return f, g, it


def main():
    n, m, reg = 1000, 800, 0.1
    M, a, b = readme_example(n, m)
    f, g, it = solve(M, a, b, reg)
    dual = np.concatenate([f + g[-1], g[:-1] - g[-1]])
    beta = np.append(dual[n:], 0.0)
    T = np.exp((dual[:n, None] + beta[None, :] - M) / reg)
    obj = reg * T.sum(1).sum() - dual[:n] @ a - dual[n:] @ b[:-1]
    marg = np.sqrt(np.sum((T.sum(1) - a) ** 2) + np.sum((T.sum(0)[:-1] - b[:-1]) ** 2))
    print("iterations", it, "marginal error", marg, "objective", obj)
    np.savez_compressed(os.path.join(HERE, "c1_readme_n1000_m800.npz"), a=a, b=b, reg=reg,
                        dual=dual, obj=obj, marg=marg)


if __name__ == "__main__":
    main()
