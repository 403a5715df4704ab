"""GPU parity of the HIP path against the CPU oracle (through the C ABI)."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import components, golden

pytestmark = pytest.mark.gpu


def _graph(kind, seed=0):
    from spec_spars_amd import graphs
    if kind == "grid":
        n, u, v = graphs.grid2d(24, 20)
        w = graphs.loguniform_weights(len(u), seed=seed)
    elif kind == "grid3d":
        n, u, v = graphs.grid3d(10)
        w = graphs.loguniform_weights(len(u), seed=seed)
    elif kind == "rgg":
        n, u, v = graphs.rgg(900, seed=seed, mean_degree=10.0)   # connected for seed 0
        w = np.ones(len(u))
    else:
        n, u, v = graphs.chung_lu(1500, 6000, seed=seed)
        w = graphs.uniform_weights(len(u), seed=seed)
    return n, u, v, w


# ---- (a) sketch --------------------------------------------------------------
@pytest.mark.parametrize("kind", ["grid", "rgg", "chunglu"])
@pytest.mark.parametrize("k", [1, 24, 48, 64, 80])
def test_sketch_bit_exact(ctx, oracle_mod, kind, k):
    n, u, v, w = _graph(kind)
    Y = ctx.sketch(u, v, w, n, k, seed=5, scale=1.0 / math.sqrt(k))
    Yo = oracle_mod.sketch(u, v, w, n, k, seed=5, scale=1.0 / math.sqrt(k))
    assert np.array_equal(Y, Yo)


def test_sketch_uniform_and_insertion_order(ctx, oracle_mod):
    g = golden("sampling_test")          # edges inserted with v1 > v2
    n = int(g["n"])
    w = np.linspace(0.5, 2.0, len(g["v1"]))
    for uniform in (False, True):
        Y = ctx.sketch(g["v1"], g["v2"], w, n, 16, seed=9, scale=math.sqrt(3.0), uniform=uniform)
        Yo = oracle_mod.sketch(g["v1"], g["v2"], w, n, 16, seed=9, scale=math.sqrt(3.0), uniform=uniform)
        assert np.array_equal(Y, Yo)
        assert np.allclose(Y.sum(0), 0.0, atol=1e-12)


def test_sketch_empty_and_isolated(ctx, oracle_mod):
    Y = ctx.sketch([], [], [], 5, 8, seed=1, scale=1.0)
    assert np.array_equal(Y, np.zeros((5, 8)))
    Y = ctx.sketch([0], [3], [4.0], 6, 8, seed=1, scale=1.0)
    assert np.array_equal(Y, oracle_mod.sketch([0], [3], [4.0], 6, 8, seed=1, scale=1.0))


# ---- (b)+(c) solve -----------------------------------------------------------
def _oracle_solution(oracle_mod, n, u, v, w, Y, seed, tol):
    import spec_spars_amd as ss
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    f = ss.Factor(cp, ri, va, seed=seed, ordering=ss.ORDER_NATURAL)
    urp, uci, uva, dg = f.export()
    Zo, it, rr = oracle_mod.solve_lap(Y, cp, ri, va, urp, uci, uva, dg, stop_rule=1, tol=tol, max_iter=2000)
    return cp, ri, va, Zo, it


def _rel_centered(Z, Zo):
    a, b = Z - Z.mean(0), Zo - Zo.mean(0)
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)


@pytest.mark.parametrize("kind", ["grid", "grid3d", "rgg"])
def test_solve_fp64_matches_oracle(ctx, oracle_mod, kind):
    import spec_spars_amd as ss
    n, u, v, w = _graph(kind)
    if components(n, u, v) != 1:
        pytest.skip("disconnected instance")
    k = 12
    Y = ctx.sketch(u, v, w, n, k, seed=2, scale=1.0 / math.sqrt(k))
    cp, ri, va, Zo, it = _oracle_solution(oracle_mod, n, u, v, w, Y, 3, 1e-12)
    Z, info = ctx.solve_lap(Y, cp, ri, va, ss.solve_opts(factor_seed=3, stop_rule=ss.STOP_REL, tol=1e-10,
                                                         max_iter=2000, ordering=ss.ORDER_NATURAL))
    assert info["max_relres"] < 1e-9
    assert _rel_centered(Z, Zo).max() < 1e-7
    # same factor, same method, same stop rule: iteration counts agree to within a couple
    _, _, _, _, it10 = _oracle_solution(oracle_mod, n, u, v, w, Y, 3, 1e-10)
    assert abs(int(info["iters"].max()) - int(it10.max())) <= 2


def test_solve_golden_exact(ctx, oracle_mod):
    import spec_spars_amd as ss
    for name in ("grid8x8_loguniform", "rgg120", "tree60"):
        g = golden(name)
        n = int(g["n"])
        cp, ri, va = oracle_mod.laplacian(n, g["u"], g["v"], g["w"])
        Z, info = ctx.solve_lap(g["Y"], cp, ri, va, ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-11))
        assert _rel_centered(Z, g["Z"]).max() < 1e-8


def test_solve_reference_stop_rule(ctx, oracle_mod):
    import spec_spars_amd as ss
    n, u, v, w = _graph("grid")
    Y = ctx.sketch(u, v, w, n, 8, seed=2, scale=math.sqrt(3.0))
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    Z, info = ctx.solve_lap(Y, cp, ri, va, ss.solve_opts(factor_seed=1, ordering=ss.ORDER_NATURAL))  # ||r||^2 <= 2e-7, 400 its
    assert info["max_relres"] < 1e-2
    f = ss.Factor(cp, ri, va, seed=1, ordering=ss.ORDER_NATURAL)
    urp, uci, uva, dg = f.export()
    Zo, it, rr = oracle_mod.solve_lap(Y, cp, ri, va, urp, uci, uva, dg, stop_rule=0, tol=2e-7)
    assert abs(int(info["iters"].max()) - int(it.max())) <= 2


def test_solve_mixed_precision(ctx, oracle_mod):
    import spec_spars_amd as ss
    n, u, v, w = _graph("grid3d")
    k = 16
    Y = ctx.sketch(u, v, w, n, k, seed=2, scale=1.0 / math.sqrt(k))
    cp, ri, va, Zo, it = _oracle_solution(oracle_mod, n, u, v, w, Y, 3, 1e-12)
    Z, info = ctx.solve_lap(Y, cp, ri, va, ss.solve_opts(factor_seed=3, stop_rule=ss.STOP_REL, tol=1e-8,
                                                         precision=ss.PREC_MIXED, max_iter=2000))
    assert info["max_relres"] <= 1e-8
    assert info["refinements"] >= 1
    assert _rel_centered(Z, Zo).max() < 1e-5


@pytest.mark.parametrize("kind", ["grid3d", "rgg"])
def test_mixed_reliable_and_restart_agree(ctx, oracle_mod, kind):
    """Every mixed-precision scheme (reliable updates, the default; the
    same with FP16 preconditioner rows; restarted FP32 sweeps) reaches the FP64 tolerance and agrees with the FP64 oracle;
    reliable updates keep the Krylov space, so they take no more iterations
    than the restarts, every inner_tol gives the same tolerance, and the
    FP16 rows cost at most two iterations over FP32 rows."""
    import spec_spars_amd as ss
    n, u, v, w = _graph(kind)
    k = 12
    Y = ctx.sketch(u, v, w, n, k, seed=4, scale=1.0 / math.sqrt(k))
    cp, ri, va, Zo, it = _oracle_solution(oracle_mod, n, u, v, w, Y, 3, 1e-12)
    out = {}
    for mode, inner in ((ss.MIXED_RELIABLE, 0.0), (ss.MIXED_RELIABLE, 0.1), (ss.MIXED_RELIABLE, 1e-4),
                        (ss.MIXED_RESTART, 0.0), (ss.MIXED_RELIABLE_F16, 0.0)):
        Z, info = ctx.solve_lap(Y, cp, ri, va, ss.solve_opts(factor_seed=3, stop_rule=ss.STOP_REL, tol=1e-8,
                                                             precision=ss.PREC_MIXED, max_iter=2000,
                                                             inner_tol=inner, mixed_mode=mode))
        L = sp.csc_matrix((va, ri, cp), shape=(n, n))
        Yp = Y - Y.mean(0)
        rr = np.linalg.norm(Yp - L @ Z, axis=0) / np.linalg.norm(Yp, axis=0)
        assert rr.max() <= 1.01e-8, (mode, inner, rr.max())
        assert info["max_relres"] <= 1e-8
        assert _rel_centered(Z, Zo).max() < 1e-5
        out[(mode, inner)] = info["max_iters"]
    assert out[(ss.MIXED_RELIABLE, 0.0)] <= out[(ss.MIXED_RESTART, 0.0)], out
    assert out[(ss.MIXED_RELIABLE_F16, 0.0)] <= out[(ss.MIXED_RELIABLE, 0.0)] + 2, out


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mixed_edge_cases(ctx, oracle_mod, mode):
    """Mixed precision (reliable updates with FP16 and FP32 preconditioner
    rows, restart sweeps) on the cases
    its per-column logic has to get right: a graph of three components and
    an isolated vertex (every component's residual at the tolerance, its
    own mean removed), a zero right-hand side (a zero solution, no
    iterations), the reference's absolute stop rule (|r|^2 <= 2e-7), and a
    budget too small to converge (SS_ERR_NOT_CONVERGED, not an abort)."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n1, u1, v1 = graphs.grid2d(20, 15)
    n2, u2, v2 = graphs.rgg(400, seed=0, mean_degree=12.0)
    n3, u3, v3 = graphs.grid2d(7, 9)
    u = np.concatenate([u1, u2 + n1, u3 + n1 + n2])
    v = np.concatenate([v1, v2 + n1, v3 + n1 + n2])
    n = n1 + n2 + n3 + 1   # the last vertex is isolated
    u, v, _ = graphs.canonical(u, v)
    w = graphs.loguniform_weights(len(u), seed=2, lo=0.1, hi=10.0)
    ncomp, lab = __import__("scipy.sparse.csgraph", fromlist=["x"]).connected_components(
        sp.coo_matrix((np.ones(len(u)), (u, v)), shape=(n, n)), directed=False)
    assert ncomp >= 4
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    L = sp.csc_matrix((va, ri, cp), shape=(n, n))
    rng = np.random.default_rng(5)
    Y = rng.standard_normal((n, 6))
    Y[:, 3] = 0.0
    Yp = Y.copy()
    for c in range(ncomp):
        Yp[lab == c] -= Y[lab == c].mean(0)
    o = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000, precision=ss.PREC_MIXED, mixed_mode=mode)
    Z, info = ctx.solve_lap(Y, cp, ri, va, o)
    R = Yp - L @ Z
    for c in range(ncomp):
        on = lab == c
        b = np.linalg.norm(Yp[on], axis=0)
        r = np.linalg.norm(R[on], axis=0)
        assert np.all(r <= 1.01e-8 * np.maximum(np.linalg.norm(Yp, axis=0), 1e-300)), (c, r / np.maximum(b, 1e-300))
    assert np.array_equal(Z[:, 3], np.zeros(n)) and info["iters"][3] == 0
    # the reference's absolute rule
    o = ss.solve_opts(stop_rule=ss.STOP_REF_ABS, precision=ss.PREC_MIXED, max_iter=3000, mixed_mode=mode)
    Z, info = ctx.solve_lap(Y, cp, ri, va, o)
    rr = (np.linalg.norm(Yp - L @ Z, axis=0) ** 2)
    assert np.all(rr <= 2e-7 * (1 + 1e-9)), rr
    # too few iterations: a status, not an abort (max_iter bounds the
    # reliable solve's iterations; restart sweeps take it per sweep)
    if mode != ss.MIXED_RESTART:
        o = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-12, max_iter=1, precision=ss.PREC_MIXED, mixed_mode=mode)
        with pytest.raises(ss.SSError) as ei:
            ctx.solve_lap(Y, cp, ri, va, o)
        assert ei.value.code == 3


def test_solve_not_converged_status(ctx, oracle_mod):
    import spec_spars_amd as ss
    n, u, v, w = _graph("grid")
    Y = ctx.sketch(u, v, w, n, 4, seed=2, scale=1.0)
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    with pytest.raises(ss.SSError) as ei:
        ctx.solve_lap(Y, cp, ri, va, ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-12, max_iter=1))
    assert ei.value.code == 3


# ---- (d) ER + sample + compact -----------------------------------------------
@pytest.mark.parametrize("kind", ["grid", "chunglu"])
def test_er_sample_bit_exact(ctx, oracle_mod, kind):
    from spec_spars_amd import graphs
    n, u, v, w = _graph(kind, seed=1)
    k = 24
    rng = np.random.default_rng(0)
    Z = rng.standard_normal((n, k)) * 0.05
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
    beta = graphs.beta_for(n, 0.3)
    d = oracle_mod.diff_norms(uu, vv, Z)
    p = oracle_mod.probs(lv, d, beta, k)
    assert 0.05 < np.mean(p < 1.0) < 0.95
    for seed in (0, 7):
        r = ctx.er_sample(uu, vv, lv, k, beta, seed, Z=Z)
        assert np.array_equal(r["d"], d)
        assert np.array_equal(r["p"], p)
        ref = oracle_mod.sample_reweight(n, uu, vv, lv, p, oracle_mod.coins(seed, len(uu)))
        assert np.array_equal(r["mask"], ref["mask"])
        assert r["kept"] == ref["kept"]
        assert np.array_equal(r["ku"], ref["ku"]) and np.array_equal(r["kv"], ref["kv"])
        assert np.array_equal(r["kw"], ref["kw"])


def test_sampling_test_kat_gpu(ctx, oracle_mod):
    g = golden("sampling_test")
    n = int(g["n"])
    cp, ri, va = oracle_mod.laplacian(n, g["v1"], g["v2"], np.ones(len(g["v1"])))
    u, v, lv = oracle_mod.upper_edges(n, cp, ri, va)
    r = ctx.er_sample(u, v, lv, 4, 1.0, 0, p_in=g["probs"], n=n)
    assert components(n, r["ku"], r["kv"]) == int(g["components"])


def test_sampling_verify_gpu(ctx, oracle_mod):
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(100, 100)
    w = graphs.uniform_weights(len(u), seed=8)
    for seed in range(3):
        r = ctx.er_sample(u, v, -w, 4, 1.0, seed, p_in=np.full(len(u), 0.5), n=n)
        ref = oracle_mod.sample_reweight(n, u, v, -w, np.full(len(u), 0.5), oracle_mod.coins(seed, len(u)))
        assert np.array_equal(r["mask"], ref["mask"])
        assert abs(len(u) / r["kept"] - 2.0) < 0.05
        assert np.allclose(r["kw"], math.sqrt(2.0) * w[r["mask"].astype(bool)], rtol=1e-15)


def test_er_sample_large_lookback(ctx, oracle_mod):
    # many tiles: the decoupled look-back must give every kept edge its place
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(600, 500)
    w = np.ones(len(u))
    rng = np.random.default_rng(3)
    p = rng.random(len(u))
    r = ctx.er_sample(u, v, -w, 4, 1.0, 11, p_in=p, n=n)
    ref = oracle_mod.sample_reweight(n, u, v, -w, p, oracle_mod.coins(11, len(u)))
    assert r["kept"] == ref["kept"]
    assert np.array_equal(r["mask"], ref["mask"])
    assert np.array_equal(r["ku"], ref["ku"]) and np.array_equal(r["kv"], ref["kv"])


# ---- whole device pipeline -------------------------------------------------------
def test_graph_sparsify_pipeline(ctx, oracle_mod):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v, w = _graph("grid", seed=4)
    k = 24
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    f = ss.Factor(cp, ri, va, seed=2, ordering=ss.ORDER_NATURAL)
    G.set_factor(f)
    beta = graphs.beta_for(n, 0.3)
    rep = G.sparsify(k, beta, scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0,
                     solve=ss.solve_opts(factor_seed=2, stop_rule=ss.STOP_REL, tol=1e-10, max_iter=2000))
    Y = G.read_block("Y")
    Yo = oracle_mod.sketch(u, v, w, n, k, seed=0, scale=1.0 / math.sqrt(k))
    assert np.array_equal(Y, Yo)
    Z = G.read_block("Z")
    urp, uci, uva, dg = f.export()
    Zo, it, rr = oracle_mod.solve_lap(Yo, cp, ri, va, urp, uci, uva, dg, stop_rule=1, tol=1e-12, max_iter=2000)
    assert _rel_centered(Z, Zo).max() < 1e-6
    # leverage scores within 1e-4 of the FP64 oracle using the same signs
    uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
    lev_o = -lv * oracle_mod.diff_norms(uu, vv, Zo) ** 2
    p = G.read_probs()
    po = oracle_mod.probs(lv, oracle_mod.diff_norms(uu, vv, Zo), beta, k)
    unclipped = (p < 1.0) & (po < 1.0)
    assert np.all(np.abs(p[unclipped] - po[unclipped]) <= 1e-4 * po[unclipped])
    assert lev_o.sum() > 0
    ku, kv, kw = G.read_kept()
    assert rep["kept"] == len(ku)
    ge = set(zip(u.tolist(), v.tolist()))
    assert all((a, b) in ge for a, b in zip(ku.tolist(), kv.tolist()))
    # sample mask given these probabilities is the oracle's
    ref = oracle_mod.sample_reweight(n, uu, vv, lv, p, oracle_mod.coins(0, len(uu)))
    assert np.array_equal(ref["ku"], ku) and np.array_equal(ref["kw"], kw)


@pytest.mark.parametrize("ordering", [1, 2])
def test_solve_ordered_factor(ctx, oracle_mod, ordering):
    import spec_spars_amd as ss
    n, u, v, w = _graph("grid3d")
    k = 16
    Y = ctx.sketch(u, v, w, n, k, seed=2, scale=1.0 / math.sqrt(k))
    cp, ri, va, Zo, it = _oracle_solution(oracle_mod, n, u, v, w, Y, 3, 1e-12)
    for prec in (ss.PREC_FP64, ss.PREC_MIXED):
        Z, info = ctx.solve_lap(Y, cp, ri, va, ss.solve_opts(factor_seed=3, stop_rule=ss.STOP_REL, tol=1e-9,
                                                             max_iter=2000, precision=prec, ordering=ordering))
        assert info["max_relres"] <= 1e-9
        assert _rel_centered(Z, Zo).max() < 1e-6
        assert info["depth_fwd"] < 100


def test_graph_sparsify_ordered_mixed(ctx, oracle_mod):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid3d(16)
    w = graphs.loguniform_weights(len(u), seed=1)
    k = 32
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    G.set_factor(ss.Factor(cp, ri, va, seed=2, ordering=ss.ORDER_ROUNDS))
    beta = graphs.beta_for(n, 0.3)
    G.profile(True)
    rep = G.sparsify(k, beta, sketch_seed=3, sampling_seed=4,
                     solve=ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=ss.PREC_MIXED))
    prof = G.profile_read()
    assert all(prof[x]["launches"] > 0 for x in ("sketch", "spmm", "tri_fwd", "tri_bwd", "er_sample"))
    assert rep["max_relres"] <= 1e-8
    Yo = oracle_mod.sketch(u, v, w, n, k, seed=3, scale=1.0 / math.sqrt(k))
    f = ss.Factor(cp, ri, va, seed=2, ordering=ss.ORDER_NATURAL)
    Zo, _, _ = oracle_mod.solve_lap(Yo, cp, ri, va, *f.export(), stop_rule=1, tol=1e-12, max_iter=2000)
    Z = G.read_block("Z")
    assert _rel_centered(Z, Zo).max() < 1e-6
    uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
    po = oracle_mod.probs(lv, oracle_mod.diff_norms(uu, vv, Zo), beta, k)
    p = G.read_probs()
    both = (p < 1) & (po < 1)
    assert np.all(np.abs(p[both] - po[both]) <= 1e-4 * po[both])
    ku, kv, kw = G.read_kept()
    ref = oracle_mod.sample_reweight(n, uu, vv, lv, p, oracle_mod.coins(4, len(uu)))
    assert np.array_equal(ref["ku"], ku) and np.array_equal(ref["kw"], kw)


def test_rounds_factor_solve_matches_permuted_oracle(ctx, oracle_mod):
    """The whole preconditioner on a power-law graph, checked against the
    oracle's PCG with the same rounds-order factor: the oracle runs in the
    factor's position space (L and Y permuted by the factor's order), so both
    apply the same M^-1, and the iteration counts and solutions agree.  The
    graph (one large component and isolated vertices, the right-hand side
    projected per component) has a dense tail of thousands of hub vertices,
    applied as the inverse of its block of L, and hub rows of more than
    1,024 entries.  A mixed-precision solve of the same system gives
    leverage scores within 1e-4 of the oracle's."""
    import scipy.sparse as sp
    import scipy.sparse.csgraph as cg
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.chung_lu(300000, 3000000, gamma=2.2, seed=0)
    w = graphs.uniform_weights(len(u), seed=0)
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    F = ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_ROUNDS)
    G.set_factor(F)
    bi = G.schedule_info(1)
    assert F.tail() > 1000 and bi["core_rows"] == F.tail(), (F.tail(), bi)
    nc, lab = cg.connected_components(sp.coo_matrix((np.ones(len(u)), (u, v)), shape=(n, n)), directed=False)
    cnt = np.bincount(lab, minlength=nc).astype(np.float64)

    def project(X):
        X = X.copy()
        for j in range(X.shape[1]):
            X[:, j] -= (np.bincount(lab, weights=X[:, j], minlength=nc) / cnt)[lab]
        return X

    Y = project(oracle_mod.sketch(u, v, w, n, 4, seed=1, scale=0.5))
    Z, info = G.solve(Y, ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-10, max_iter=300, precision=ss.PREC_FP64))
    Zm, _ = G.solve(Y, ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-10, max_iter=300, precision=ss.PREC_MIXED))
    G.close()
    order, _ = F.order()
    L = sp.csc_matrix((va, ri, cp), shape=(n, n))
    Lp = L[order][:, order].tocsc()
    Lp.sort_indices()
    Zp, it, rr = oracle_mod.solve_lap(Y[order], Lp.indptr.astype(np.int32), Lp.indices.astype(np.int32), Lp.data,
                                      *F.export(), stop_rule=1, tol=1e-10, max_iter=300)
    assert rr.max() <= 1e-10
    assert np.abs(info["iters"] - it).max() <= 1, (info["iters"], it)
    Zo = np.empty_like(Zp)
    Zo[order] = Zp
    a, b = project(Z), project(Zo)
    assert (np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)).max() < 1e-7
    # mixed precision (FP32 solves, FP64 reductions and refinement):
    # leverage scores within 1e-4 of the FP64 oracle's
    eu, ev, lv = oracle_mod.upper_edges(n, cp, ri, va)
    beta = graphs.beta_for(n, 0.3)
    po = oracle_mod.probs(lv, oracle_mod.diff_norms(eu, ev, Zo), beta, 4)
    pm = oracle_mod.probs(lv, oracle_mod.diff_norms(eu, ev, Zm), beta, 4)
    both = (pm < 1) & (po < 1)
    assert both.sum() > 100000
    assert np.all(np.abs(pm[both] - po[both]) <= 1e-4 * po[both])


def test_small_level_runs_match_oracle(ctx, oracle_mod):
    """Natural order on a 120 x 120 grid: the factor's solve DAG is nearly a
    chain, so outside the dense core (at most the last 4,096 positions) its levels
    are runs of small levels walked by one workgroup (more than 256 levels:
    several LDS stages), in both directions.  FP64 solution within 1e-7 of
    the oracle's with the same factor."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(120, 120)
    w = graphs.loguniform_weights(len(u), seed=5, lo=0.1, hi=10.0)
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    F = ss.Factor(cp, ri, va, seed=3, ordering=ss.ORDER_NATURAL)
    G.set_factor(F)
    for d in (0, 1):
        info = G.schedule_info(d)
        assert info["small_runs"] >= 1 and info["small_levels"] > 256 and 3000 < info["core_rows"] <= 4096, info
    Y = oracle_mod.sketch(u, v, w, n, 6, seed=2, scale=1.0)
    Z, rep = G.solve(Y, ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-11, max_iter=3000))
    G.close()
    Zo, it, rr = oracle_mod.solve_lap(Y, cp, ri, va, *F.export(), stop_rule=1, tol=1e-11, max_iter=3000)
    assert np.abs(rep["iters"] - it).max() <= 1
    a, b = Z - Z.mean(0), Zo - Zo.mean(0)
    assert (np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)).max() < 1e-7


@pytest.mark.parametrize("k", [130, 256])
def test_wide_blocks_match_oracle(ctx, oracle_mod, k):
    """k above 128 (one row per pass in the row-block kernels, several
    64-column chunks in the solves) up to the maximum, 256: the sketch is
    bit-exact and the FP64 solve matches the oracle; k = 257 is refused."""
    import spec_spars_amd as ss
    n, u, v, w = _graph("grid")
    Y = ctx.sketch(u, v, w, n, k, seed=2, scale=1.0 / math.sqrt(k))
    assert np.array_equal(Y, oracle_mod.sketch(u, v, w, n, k, seed=2, scale=1.0 / math.sqrt(k)))
    cp, ri, va, Zo, _ = _oracle_solution(oracle_mod, n, u, v, w, Y, 3, 1e-12)
    for ordering in (ss.ORDER_NATURAL, ss.ORDER_ROUNDS):
        Z, info = ctx.solve_lap(Y, cp, ri, va, ss.solve_opts(factor_seed=3, stop_rule=ss.STOP_REL, tol=1e-10,
                                                             max_iter=2000, ordering=ordering))
        assert info["max_relres"] < 1e-9
        assert _rel_centered(Z, Zo).max() < 1e-7
    with pytest.raises(ss.SSError):
        ctx.solve_lap(np.zeros((n, 257)), cp, ri, va, ss.solve_opts(factor_seed=3))
