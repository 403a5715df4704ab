"""The spectral checker's host side (spec_spars_amd/spectral.py), pinned on
small graphs against dense linear algebra: Lanczos extremes of L_H^+ L_G
equal the dense generalized eigenvalues, and the quotients of a graph with
itself are 1."""
import numpy as np
import scipy.linalg as sl


def _sub_graph():
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(20, 18)
    w = graphs.loguniform_weights(len(u), seed=2, lo=0.5, hi=2.0)
    rng = np.random.default_rng(3)
    p = rng.uniform(0.4, 1.0, len(u))
    keep = rng.random(len(u)) < p
    # a spanning tree stays, so H is connected: every edge of the first row
    # and of every column (grid2d ids are x + 20 y)
    keep |= (v - u == 1) & (u < 20)
    keep |= v - u == 20
    return n, u, v, w, keep, p


def test_lanczos_matches_dense_generalized_eigenvalues():
    from spec_spars_amd import spectral
    n, u, v, w, keep, p = _sub_graph()
    LG = spectral.laplacian(n, u, v, w)
    LH = spectral.laplacian(n, u[keep], v[keep], w[keep] / p[keep])
    FG = spectral.EdgeForm(n, u, v, w)
    FH = spectral.EdgeForm(n, u[keep], v[keep], w[keep] / p[keep])
    assert np.allclose(FG.matvec(np.arange(n) * 1.0), LG @ np.arange(n), rtol=1e-12)
    Q = sl.null_space(np.ones((1, n)))
    lam = sl.eigh(Q.T @ LG.toarray() @ Q, Q.T @ LH.toarray() @ Q, eigvals_only=True)
    LHp = np.linalg.pinv(LH.toarray())
    # the largest eigenvalue converges within 40 steps to rounding; the
    # smallest (a cluster) to 1e-3; both Ritz values lie inside the spectrum
    lo, hi, it, _, _ = spectral.lanczos_extremes(FG, FH, lambda b: (LHp @ b, 0.0), iters=40, seed=1)
    assert abs(hi - lam.max()) <= 1e-9 * lam.max()
    assert lam.min() <= lo <= (1 + 1e-2) * lam.min()
    lo, hi, it, _, _ = spectral.lanczos_extremes(FG, FH, lambda b: (LHp @ b, 0.0), iters=120, seed=1)
    assert abs(lo - lam.min()) <= 1e-4 * lam.min() and abs(hi - lam.max()) <= 1e-9 * lam.max()


def test_quotients_and_kept_ids():
    from spec_spars_amd import spectral
    n, u, v, w, keep, p = _sub_graph()
    FG = spectral.EdgeForm(n, u, v, w)
    rq = spectral.rayleigh_quotients(FG, FG, count=8)
    assert np.allclose(rq, 1.0, rtol=1e-12)
    ids = spectral.kept_edge_ids(n, u, v, u[keep], v[keep])
    assert np.array_equal(ids, np.nonzero(keep)[0])
    # w / sqrt(p) <= w / p edge by edge: the reference's H is below the textbook one
    FH1 = spectral.EdgeForm(n, u[keep], v[keep], w[keep] / np.sqrt(p[keep]))
    FH2 = spectral.EdgeForm(n, u[keep], v[keep], w[keep] / p[keep])
    assert np.all(spectral.rayleigh_quotients(FG, FH1, 16) <= spectral.rayleigh_quotients(FG, FH2, 16) + 1e-12)


def test_lanczos_inverse_pencil_on_g():
    """The (L_H, L_G) pencil, solves on G: its extremes are the reciprocals
    of the (L_G, L_H) pencil's."""
    from spec_spars_amd import spectral
    n, u, v, w, keep, p = _sub_graph()
    LG = spectral.laplacian(n, u, v, w)
    LH = spectral.laplacian(n, u[keep], v[keep], w[keep] / p[keep])
    Q = sl.null_space(np.ones((1, n)))
    lam = sl.eigh(Q.T @ LG.toarray() @ Q, Q.T @ LH.toarray() @ Q, eigvals_only=True)
    LGp = np.linalg.pinv(LG.toarray())
    FG = spectral.EdgeForm(n, u, v, w)
    FH = spectral.EdgeForm(n, u[keep], v[keep], w[keep] / p[keep])
    lo, hi, it, _, _ = spectral.lanczos_extremes(FH, FG, lambda b: (LGp @ b, 0.0), iters=120, seed=1)
    assert abs(lo - 1 / lam.max()) <= 1e-9 / lam.max() and abs(hi - 1 / lam.min()) <= 1e-4 / lam.min()


def test_lanczos_basis_stays_orthonormal_over_many_steps():
    """Weights over six decades and more steps than converge: the basis
    stays B-orthonormal (the constants, which the B product cannot see, are
    projected out every step) and every Ritz value lies in the spectrum."""
    from spec_spars_amd import graphs, spectral
    n, u, v = graphs.grid2d(20, 18)
    w = graphs.loguniform_weights(len(u), seed=4)
    rng = np.random.default_rng(5)
    p = rng.uniform(0.3, 1.0, len(u))
    keep = (rng.random(len(u)) < p) | ((v - u == 1) & (u < 20)) | (v - u == 20)
    LG = spectral.laplacian(n, u, v, w)
    LH = spectral.laplacian(n, u[keep], v[keep], w[keep] / p[keep])
    Q = sl.null_space(np.ones((1, n)))
    lam = sl.eigh(Q.T @ LG.toarray() @ Q, Q.T @ LH.toarray() @ Q, eigvals_only=True)
    LHp = np.linalg.pinv(LH.toarray())
    FG = spectral.EdgeForm(n, u, v, w)
    FH = spectral.EdgeForm(n, u[keep], v[keep], w[keep] / p[keep])
    lo, hi, it, ritz, d = spectral.lanczos_extremes(FG, FH, lambda b: (LHp @ b, 0.0), iters=200, seed=2)
    assert it == 200 and d["basis_departure"] < 1e-8
    assert lam.min() * (1 - 1e-6) <= lo and hi <= lam.max() * (1 + 1e-6)
    assert abs(hi - lam.max()) <= 1e-6 * lam.max()
