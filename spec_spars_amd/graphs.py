"""Synthetic graph families with fixed seeds, in canonical edge order.

Canonical order is the CSC upper triangle of the Laplacian: edges (u, v) with
u < v, sorted by v then u -- the order InputStream::run_stream inserts edges
in (src/stream.rs:44-48) and the order probabilities are indexed by
(src/sparsifier.rs:547-548).
"""
import math

import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, count):
    """count outputs of splitmix64 started at `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        s = (np.arange(1, count + 1, dtype=np.uint64) * _GOLD + np.uint64(seed))
        z = s
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform01(seed, count):
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def canonical(u, v, w=None):
    """Orient u < v, sort by (v, u), drop self-loops and duplicates."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    a, b = np.minimum(u, v), np.maximum(u, v)
    keep = a != b
    a, b = a[keep], b[keep]
    ww = None if w is None else np.asarray(w, np.float64)[keep]
    order = np.lexsort((a, b))
    a, b = a[order], b[order]
    if ww is not None:
        ww = ww[order]
    if len(a):
        first = np.ones(len(a), bool)
        first[1:] = (a[1:] != a[:-1]) | (b[1:] != b[:-1])
        a, b = a[first], b[first]
        if ww is not None:
            ww = ww[first]
    return a.astype(np.int32), b.astype(np.int32), ww


def grid2d(nx, ny):
    """4-neighbour grid, id = x + nx*y."""
    idx = np.arange(nx * ny, dtype=np.int64).reshape(ny, nx)
    h = np.stack([idx[:, :-1].ravel(), idx[:, 1:].ravel()], 1)
    vt = np.stack([idx[:-1, :].ravel(), idx[1:, :].ravel()], 1)
    e = np.concatenate([h, vt])
    u, v, _ = canonical(e[:, 0], e[:, 1])
    return nx * ny, u, v


def grid3d(n1):
    """6-neighbour cube grid, id = x + n1*y + n1^2*z."""
    idx = np.arange(n1 ** 3, dtype=np.int64).reshape(n1, n1, n1)   # [z, y, x]
    parts = [np.stack([idx[:, :, :-1].ravel(), idx[:, :, 1:].ravel()], 1),
             np.stack([idx[:, :-1, :].ravel(), idx[:, 1:, :].ravel()], 1),
             np.stack([idx[:-1, :, :].ravel(), idx[1:, :, :].ravel()], 1)]
    e = np.concatenate(parts)
    u, v, _ = canonical(e[:, 0], e[:, 1])
    return n1 ** 3, u, v


def loguniform_weights(m, seed=0, lo=1e-3, hi=1e3):
    """w = 10^U(log10 lo, log10 hi) per edge in canonical order."""
    x = uniform01(seed, m)
    return np.power(10.0, math.log10(lo) + x * (math.log10(hi) - math.log10(lo)))


def uniform_weights(m, seed=0, lo=0.5, hi=2.0):
    return lo + uniform01(seed, m) * (hi - lo)


def chung_lu(n, m, gamma=2.2, seed=0, i0=None):
    """Chung-Lu graph with exactly m unique non-loop edges.

    Expected degrees d_i proportional to (i + i0)^(-1/(gamma-1)), capped at
    sqrt(2m) and rescaled so they sum to 2m; endpoints are drawn
    independently proportional to d until m distinct edges exist.
    """
    rng = np.random.default_rng(seed)
    if i0 is None:
        i0 = max(1.0, n / 1000.0)
    d = np.power(np.arange(n, dtype=np.float64) + i0, -1.0 / (gamma - 1.0))
    d *= 2.0 * m / d.sum()
    cap = math.sqrt(2.0 * m)
    for _ in range(20):
        d = np.minimum(d, cap)
        d *= 2.0 * m / d.sum()
    cdf = np.cumsum(d)
    cdf /= cdf[-1]
    keys = np.zeros(0, np.int64)
    while len(keys) < m:
        need = int((m - len(keys)) * 1.15) + 1024
        a = np.searchsorted(cdf, rng.random(need), side="right")
        b = np.searchsorted(cdf, rng.random(need), side="right")
        a = np.minimum(a, n - 1)
        b = np.minimum(b, n - 1)
        ok = a != b
        lo, hi = np.minimum(a[ok], b[ok]), np.maximum(a[ok], b[ok])
        keys = np.unique(np.concatenate([keys, hi.astype(np.int64) * n + lo]))
    if len(keys) > m:
        keys = np.sort(rng.choice(keys, m, replace=False))
    v = (keys // n).astype(np.int32)
    u = (keys % n).astype(np.int32)
    return n, u, v


def rgg(n, seed=0, mean_degree=8.0):
    """Random geometric graph in the unit square, radius for the given mean degree."""
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 2))
    r = math.sqrt(mean_degree / (math.pi * n))
    ncell = max(1, int(1.0 / r))
    cell = np.minimum((pts * ncell).astype(np.int64), ncell - 1)
    cid = cell[:, 0] * ncell + cell[:, 1]
    order = np.argsort(cid, kind="stable")
    cid_sorted = cid[order]
    starts = np.searchsorted(cid_sorted, np.arange(ncell * ncell))
    ends = np.searchsorted(cid_sorted, np.arange(ncell * ncell), side="right")
    us, vs = [], []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            nx_ = cell[:, 0] + dx
            ny_ = cell[:, 1] + dy
            ok = (nx_ >= 0) & (nx_ < ncell) & (ny_ >= 0) & (ny_ < ncell)
            src = np.nonzero(ok)[0]
            nc = nx_[ok] * ncell + ny_[ok]
            s, e = starts[nc], ends[nc]
            cnt = e - s
            rep_src = np.repeat(src, cnt)
            offs = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
            dst = order[np.repeat(s, cnt) + offs]
            sel = rep_src < dst
            rs, ds = rep_src[sel], dst[sel]
            dist2 = ((pts[rs] - pts[ds]) ** 2).sum(1)
            close = dist2 <= r * r
            us.append(rs[close])
            vs.append(ds[close])
    u, v, _ = canonical(np.concatenate(us), np.concatenate(vs))
    return n, u, v


def path_with_shortcuts():
    """The graph of the reference's sampling_test (src/tests.rs:265-317):
    a path over 100 nodes plus shortcuts (10i, 10i+9), in insertion order."""
    ins = [(i + 1, i) for i in range(99)] + [(10 * i + 9, 10 * i) for i in range(10)]
    v1 = np.array([a for a, _ in ins], np.int32)
    v2 = np.array([b for _, b in ins], np.int32)
    return 100, v1, v2


def beta_for(n, epsilon=0.3, beta_constant=4):
    """beta = eps^-2 * beta_constant * ln(n)  (src/sparsifier.rs:264)."""
    return math.pow(epsilon, -2.0) * float(beta_constant) * math.log(float(n))


def laplacian_csc(n, u, v, w):
    """CSC Laplacian with the diagonal (symmetric, so also CSR) of canonical
    edges: off-diagonals -w, diagonal = weighted degree summed in edge order
    (u's and v's contributions interleaved as the edges come), rows sorted
    within each column -- the same matrix the device graph holds."""
    u = np.asarray(u, np.int64)
    v = np.asarray(v, np.int64)
    w = np.asarray(w, np.float64)
    diag = np.zeros(n, np.float64)
    np.add.at(diag, np.stack([u, v], 1).ravel(), np.repeat(w, 2))
    rows = np.concatenate([v, u, np.arange(n)])
    cols = np.concatenate([u, v, np.arange(n)])
    vals = np.concatenate([-1.0 * w, -1.0 * w, diag])
    order = np.lexsort((rows, cols))
    colptr = np.zeros(n + 1, np.int32)
    colptr[1:] = np.cumsum(np.bincount(cols, minlength=n))
    return colptr, rows[order].astype(np.int32), vals[order]
