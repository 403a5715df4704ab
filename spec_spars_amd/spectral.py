"""Spectral quality of a sparsifier H of G: generalized Rayleigh quotients
x^T L_H x / x^T L_G x on random vectors and the extreme eigenvalues of the
pencils (L_H, L_G) and (L_G, L_H) by Lanczos, with L_G^+ and L_H^+ applied by
the device PCG.

H is an eps-spectral sparsifier when (1 - eps) L_G <= L_H <= (1 + eps) L_G,
i.e. the eigenvalues mu of L_G^+ L_H lie in [1 - eps, 1 + eps] (every
quotient lies in [mu_min, mu_max]) and those of L_H^+ L_G, lambda = 1 / mu,
in [1 / (1 + eps), 1 / (1 - eps)].  A disconnected H fails the lower bound
outright (mu_min = 0: an indicator of one of its components has
x^T L_H x = 0 < x^T L_G x); L_H^+ L_G is then not the inverse pencil and only
mu is reported.  The reference only reports quadratic-form errors on random
probes (src/experiments.rs:88-122).

Two weightings of the kept edges are checked: the reference's w / sqrt(p)
(src/sparsifier.rs:655-657), and the textbook unbiased w / p.  Where p < 1
the first gives E[L_H] below L_G, so it can fail the lower bound while
matching the reference bit for bit.
"""
import numpy as np
import scipy.linalg as sl
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components


def laplacian(n, u, v, w):
    """Graph Laplacian (CSR) of edges (u, v, w)."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.asarray(w, np.float64)
    A = sp.coo_matrix((np.concatenate([w, w]), (np.concatenate([u, v]), np.concatenate([v, u]))), shape=(n, n))
    A = A.tocsr()
    return (sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A).tocsr()


class EdgeForm:
    """The quadratic form of a Laplacian through its edges:
    x^T L y = sum_e w_e (x_u - x_v)(y_u - y_v).  With weights spanning many
    decades the vectors of a Krylov basis carry large smooth components, and
    x . (L x) summed over vertices cancels catastrophically; differences
    along edges are exact where x_u and x_v are close, and the edge sums
    keep the form's rounding relative to its value (a Gram matrix of edge
    vectors stays positive semidefinite)."""

    def __init__(self, n, u, v, w):
        m = len(u)
        rows = np.repeat(np.arange(m, dtype=np.int64), 2)
        cols = np.stack([np.asarray(u, np.int64), np.asarray(v, np.int64)], axis=1).ravel()
        vals = np.tile(np.array([1.0, -1.0]), m)
        self.n, self.m = n, m
        self.D = sp.csr_matrix((vals, (rows, cols)), shape=(m, n))
        self.w = np.asarray(w, np.float64)
        self.sw = np.sqrt(self.w)

    def matvec(self, x):
        """L x as D^T (w * D x)."""
        return self.D.T @ (self.w * (self.D @ x))

    def energy(self, x):
        """The edge vector sqrt(w) * D x, whose dot products give the form."""
        d = self.D @ x
        return self.sw * d if d.ndim == 1 else self.sw[:, None] * d

    def quad(self, X, chunk=8):
        """x^T L x for every column of the n x c matrix X."""
        return np.concatenate([np.sum(self.energy(X[:, j:j + chunk]) ** 2, axis=0)
                               for j in range(0, X.shape[1], chunk)])


def kept_edge_ids(n, u, v, ku, kv):
    """Indices into the canonical edge list (u, v) of the kept edges (a
    subsequence in the same order)."""
    key = np.asarray(v, np.int64) * n + np.asarray(u, np.int64)
    kk = np.asarray(kv, np.int64) * n + np.asarray(ku, np.int64)
    idx = np.searchsorted(key, kk)
    assert np.array_equal(key[idx], kk), "kept edges are not edges of G"
    return idx


def rayleigh_quotients(FG, FH, count=64, seed=0):
    """x^T L_H x / x^T L_G x for `count` Gaussian vectors projected off the
    constants (FG, FH: EdgeForm of G and H)."""
    x = np.random.default_rng(seed).standard_normal((FG.n, count))
    x -= x.mean(0)
    return FH.quad(x) / FG.quad(x)


def lanczos_extremes(A, B, solve_B, iters=40, seed=0):
    """Extreme eigenvalues of B^+ A on the complement of the constants, for
    the EdgeForms A, B of connected graphs: Lanczos on A x = lambda B x in
    the B inner product, fully reorthogonalized; solve_B(b) returns
    (B^+ b, relres) for a mean-zero b.  Returns (lambda_min, lambda_max,
    iterations, ritz, diagnostics: worst solve relres, the betas, the
    departure of V^T B V from the identity and the smallest eigenvalue of
    V^T A V)."""
    q = np.random.default_rng(seed).standard_normal(A.n)
    q -= q.mean()
    eb = B.energy(q)
    nrm = np.sqrt(eb @ eb)
    Q, EB, EA = [q / nrm], [eb / nrm], []
    alphas, betas, worst = [], [], 0.0
    for j in range(iters):
        ea = A.energy(Q[j])
        EA.append(ea)
        alphas.append(float(ea @ ea))
        Aq = A.matvec(Q[j])
        w, rr = solve_B(Aq - Aq.mean())
        worst = max(worst, rr)
        w = w - w.mean()
        ew = B.energy(w)
        for _ in range(2):                              # full reorthogonalization in the B product
            for qi, ebi in zip(Q, EB):
                c = ew @ ebi
                w -= c * qi
                ew -= c * ebi
        # the constants are invisible to the B product, so rounding in them
        # is never projected out and would grow by 1 / beta every step
        w -= w.mean()
        ew = B.energy(w)
        b = np.sqrt(ew @ ew)
        # an invariant subspace to the accuracy of the solves (H = G gives
        # w = q up to the solve's error)
        if j == iters - 1 or b < 1e-6 * abs(alphas[-1]):
            break
        betas.append(float(b))
        Q.append(w / b)
        EB.append(ew / b)
    # Rayleigh-Ritz on the basis: the pencil (V^T A V, V^T B V) as Gram
    # matrices of edge vectors, right when V is short of exact
    # B-orthonormality
    k = len(EA)
    GA = np.stack(EA, axis=1)
    GB = np.stack(EB[:k], axis=1)
    T = GA.T @ GA
    S = GB.T @ GB
    # drop directions S cannot resolve (the basis's rounding)
    sv, U = np.linalg.eigh(S)
    ok = sv > 1e-12 * sv[-1]
    W = U[:, ok] / np.sqrt(sv[ok])
    ritz = np.linalg.eigvalsh(W.T @ T @ W)
    diag = {"relres": worst, "betas": betas, "basis_departure": float(np.abs(S - np.eye(k)).max()),
            "basis_a_min_eig": float(np.linalg.eigvalsh(T)[0])}
    return float(ritz[0]), float(ritz[-1]), len(alphas), ritz, diag


def _device_solver(ctx, n, u, v, w):
    """(graph, solve(b) -> (L^+ b, relres)) by the device PCG at rel 1e-10
    on a rounds-order factor."""
    import spec_spars_amd as ss
    G = ss.Graph(ctx, n, u, v, w)
    G.set_factor(ss.Factor(*G.laplacian(), seed=0, ordering=ss.ORDER_ROUNDS))
    sv = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-10, max_iter=3000)

    F = EdgeForm(n, u, v, w)

    def solve(b):
        z, info = G.solve(b.reshape(n, 1), sv)
        z = np.ascontiguousarray(z[:, 0])
        # the true residual, on the host
        return z, float(np.linalg.norm(b - F.matvec(z)) / np.linalg.norm(b))
    return F, G, solve


def check(ctx, n, u, v, w, ku, kv, kw, p, eps, lanczos_iters=40, quotients=64, seed=0):
    """Both weightings of the kept edges of one sparsification pass of a
    connected G: w / sqrt(p) (the kernel's kw, the reference's) and w / p.
    Per weighting: the quotient range, the Lanczos extremes mu of
    L_G^+ L_H (solves on G) and, when H is connected, lambda of L_H^+ L_G
    (solves on H), and pass/fail against eps."""
    ids = kept_edge_ids(n, u, v, ku, kv)
    weights = {"w_over_sqrt_p": np.asarray(kw, np.float64),
               "w_over_p": np.asarray(w, np.float64)[ids] / np.asarray(p, np.float64)[ids]}
    out = {"epsilon": eps, "kept": int(len(ku)), "edges": int(len(u))}
    FG, GG, solve_G = _device_solver(ctx, n, u, v, w)
    try:
        for name, wh in weights.items():
            FH = EdgeForm(n, ku, kv, wh)
            comps, label = connected_components(laplacian(n, ku, kv, wh), directed=False)
            r = {"components": int(comps),
                 "isolated_vertices": int((np.bincount(label, minlength=comps) == 1).sum())}
            rq = rayleigh_quotients(FG, FH, count=quotients, seed=seed + 1)
            r.update(rayleigh_min=float(rq.min()), rayleigh_max=float(rq.max()), rayleigh_mean=float(rq.mean()))
            mlo, mhi, it, _, dg = lanczos_extremes(FH, FG, solve_G, iters=lanczos_iters, seed=seed)
            r.update(mu_min=0.0 if comps > 1 else mlo, mu_ritz_min=mlo, mu_max=mhi, mu_iters=it,
                     mu_solve_relres=dg["relres"], mu_diag=dg)
            if comps == 1:
                _, GH, solve_H = _device_solver(ctx, n, ku, kv, wh)
                try:
                    lo, hi, it, _, dg = lanczos_extremes(FG, FH, solve_H, iters=lanczos_iters, seed=seed)
                finally:
                    GH.close()
                r.update(lambda_min=lo, lambda_max=hi, lambda_iters=it, lambda_solve_relres=dg["relres"],
                         lambda_diag=dg)
            r["rayleigh_within_eps"] = bool(rq.min() >= 1 - eps and rq.max() <= 1 + eps)
            r["eigen_within_eps"] = bool(r["mu_min"] >= 1 - eps and mhi <= 1 + eps)
            out[name] = r
    finally:
        GG.close()
    return out
