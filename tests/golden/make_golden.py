"""Writes the FP64 golden fixtures in this directory (run in the build container).

Every fixture is data computed here by numpy/scipy, independent of both the
oracle and the library:
  * exact pseudo-inverse solves Z = L^+ (Y - mean) of small graphs,
  * exact effective resistances R_eff(u,v) = (e_u - e_v)^T L^+ (e_u - e_v),
  * the reference sampling_test graph (src/tests.rs:265-317) and the
    probabilities that test assigns (0 on every second path edge per block,
    1 elsewhere), with the connected-component count it asserts (40).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from spec_spars_amd import graphs  # noqa: E402


def dense_laplacian(n, u, v, w):
    L = np.zeros((n, n))
    for a, b, x in zip(u, v, w):
        L[a, b] -= x
        L[b, a] -= x
        L[a, a] += x
        L[b, b] += x
    return L


def fixture(name, n, u, v, w, k, seed):
    L = dense_laplacian(n, u, v, w)
    Lp = np.linalg.pinv(L, rcond=1e-13, hermitian=True)
    rng = np.random.default_rng(seed)
    Y = rng.standard_normal((n, k))
    B = Y - Y.mean(0)
    Z = Lp @ B
    reff = np.array([Lp[a, a] + Lp[b, b] - 2 * Lp[a, b] for a, b in zip(u, v)])
    np.savez(os.path.join(HERE, name + ".npz"), n=n, u=u, v=v, w=w, Y=Y, Z=Z, reff=reff)


def main():
    n, u, v = graphs.grid2d(8, 8)
    w = graphs.loguniform_weights(len(u), seed=11, lo=1e-2, hi=1e2)
    fixture("grid8x8_loguniform", n, u, v, w, 6, 1)
    for seed in range(4, 100):                      # first connected instance
        n, u, v = graphs.rgg(120, seed=seed, mean_degree=7.0)
        A = np.zeros((n, n), bool)
        A[u, v] = A[v, u] = True
        reach = np.zeros(n, bool)
        reach[0] = True
        for _ in range(n):
            reach = reach | A[reach].any(0)
        if reach.all():
            break
    w = graphs.uniform_weights(len(u), seed=5)
    fixture("rgg120", n, u, v, w, 5, 2)
    # a tree (every leverage score is 1)
    rng = np.random.default_rng(9)
    par = np.array([rng.integers(0, i) for i in range(1, 60)])
    u, v, _ = graphs.canonical(par, np.arange(1, 60))
    w = graphs.uniform_weights(len(u), seed=3)
    fixture("tree60", 60, u, v, w, 4, 3)
    # reference sampling_test: path of 100 + 10 shortcuts, probabilities 0/1 by index
    n, v1, v2 = graphs.path_with_shortcuts()
    u, v, _ = graphs.canonical(v1, v2)
    m = len(u)
    probs = np.ones(m)
    for i in range(10):
        for j in range(5):
            idx = i * 11 + 2 + 2 * j
            if idx == m:
                break
            probs[idx] = 0.0
    np.savez(os.path.join(HERE, "sampling_test.npz"), n=n, v1=v1, v2=v2, u=u, v=v, probs=probs,
             components=40)


if __name__ == "__main__":
    main()
