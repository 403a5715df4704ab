"""Batches of uneven graphs on the GPU: each graph's results are bit-identical
alone and in a mix, and match the oracle graph by graph."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import components

pytestmark = pytest.mark.gpu


def _two_components():
    """Two weighted grids side by side plus two isolated vertices."""
    from spec_spars_amd import graphs
    n1, u1, v1 = graphs.grid2d(6, 5)
    n2, u2, v2 = graphs.grid2d(4, 4)
    u = np.concatenate([u1, u2 + n1])
    v = np.concatenate([v1, v2 + n1])
    u, v, _ = graphs.canonical(u, v)
    w = graphs.loguniform_weights(len(u), seed=5)
    return n1 + n2 + 2, u, v, w


def _mix():
    from spec_spars_amd import graphs
    out = []
    n, u, v = graphs.grid2d(12, 9)
    out.append((n, u, v, graphs.loguniform_weights(len(u), seed=1)))
    n, u, v = graphs.rgg(500, seed=3)
    out.append((n, u, v, np.ones(len(u))))
    n, u, v = graphs.chung_lu(700, 2500, seed=2)
    out.append((n, u, v, graphs.uniform_weights(len(u), seed=2)))
    out.append(_two_components())
    n = 50
    out.append((n, np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32), np.ones(n - 1)))
    n, u, v = graphs.grid3d(7)
    out.append((n, u, v, graphs.loguniform_weights(len(u), seed=4)))
    return out


def _factors(gs, seed=2, ordering=0):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    return [ss.Factor(*graphs.laplacian_csc(n, u, v, w), seed=seed + i, ordering=ordering)
            for i, (n, u, v, w) in enumerate(gs)]


def _run(ctx, gs, fs, idx, k, prec):
    """Sparsify graphs gs[idx] as one batch; per-graph seeds follow the index."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    B = ss.Batch(ctx, [gs[i] for i in idx])
    B.set_factors([fs[i] for i in idx])
    betas = [graphs.beta_for(gs[i][0], 0.3) for i in idx]
    rep = B.sparsify(k, betas, sketch_seeds=[10 + i for i in idx], sampling_seeds=[20 + i for i in idx],
                     solve=ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000, precision=prec))
    res = []
    for j in range(len(idx)):
        res.append({"Y": B.read_block(j, "Y"), "Z": B.read_block(j, "Z"), "p": B.read_probs(j),
                    "kept": B.read_kept(j), "rep": rep[j]})
    B.close()
    return res


def _laplacian(n, u, v, w):
    from spec_spars_amd import graphs
    cp, ri, va = graphs.laplacian_csc(n, u, v, w)
    return sp.csc_matrix((va, ri, cp), shape=(n, n))


def _component_projection(n, u, v, Y):
    import scipy.sparse.csgraph as cg
    A = sp.coo_matrix((np.ones(len(u)), (np.asarray(u), np.asarray(v))), shape=(n, n))
    nc, lab = cg.connected_components(A, directed=False)
    Yp = Y.copy()
    for c in range(nc):
        Yp[lab == c] -= Y[lab == c].mean(0)
    return Yp


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_alone_vs_mixed_bit_identical(ctx, prec):
    gs = _mix()
    fs = _factors(gs)
    k = 8
    mixed = _run(ctx, gs, fs, list(range(len(gs))), k, prec)
    for i in range(len(gs)):
        alone = _run(ctx, gs, fs, [i], k, prec)[0]
        for key in ("Y", "Z", "p"):
            assert np.array_equal(alone[key], mixed[i][key]), (i, key)
        for a, b in zip(alone["kept"], mixed[i]["kept"]):
            assert np.array_equal(a, b), i
        assert alone["rep"]["kept"] == mixed[i]["rep"]["kept"]
        assert alone["rep"]["max_iters"] == mixed[i]["rep"]["max_iters"]
        assert alone["rep"]["max_relres"] == mixed[i]["rep"]["max_relres"]
    # a reordered mix gives the same bits too
    order = list(range(len(gs)))[::-1]
    rev = _run(ctx, gs, fs, order, k, prec)
    for j, i in enumerate(order):
        assert np.array_equal(rev[j]["Z"], mixed[i]["Z"]) and np.array_equal(rev[j]["p"], mixed[i]["p"])


def _hubs(nh=100, seed=7, lo=35, hi=120):
    """Hubs on a ring, each with lo..hi-1 leaves (35..119 by default): the
    leaves go in the first elimination round, so each hub's forward row
    holds as many entries, plus one, in the second and third levels."""
    from spec_spars_amd import graphs
    rng = np.random.default_rng(seed)
    sizes = rng.integers(lo, hi, nh)
    u, v = [], []
    leaf = nh
    for h in range(nh):
        u.append(h)
        v.append((h + 1) % nh)
        for _ in range(sizes[h]):
            u.append(h)
            v.append(leaf)
            leaf += 1
    u, v, _ = graphs.canonical(np.array(u, np.int32), np.array(v, np.int32))
    return leaf, u, v, graphs.uniform_weights(len(u), seed=seed)


@pytest.mark.parametrize("prec", [0, 1])
def test_wave_rows_and_blocks_bit_identical(ctx, prec):
    """A hub row of 35..120 entries is finished by a block of waves when its
    level is small (the hub graph alone) and by one wave when a large grid
    shares the level (the mix); k = 72 takes two column chunks.  The bits
    agree, and every graph's residual meets the tolerance."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(200, 200)
    gs = [(n, u, v, graphs.loguniform_weights(len(u), seed=3)), _hubs()]
    fs = _factors(gs, seed=9, ordering=ss.ORDER_ROUNDS)
    k = 72
    mixed = _run(ctx, gs, fs, [0, 1], k, prec)
    for i in range(2):
        alone = _run(ctx, gs, fs, [i], k, prec)[0]
        for key in ("Y", "Z", "p"):
            assert np.array_equal(alone[key], mixed[i][key]), (i, key)
        n, u, v, w = gs[i]
        L = _laplacian(n, u, v, w)
        Yp = _component_projection(n, u, v, mixed[i]["Y"])
        rr = np.linalg.norm(Yp - L @ mixed[i]["Z"], axis=0) / np.linalg.norm(Yp, axis=0)
        assert rr.max() <= 1.01e-8, i


@pytest.mark.parametrize("prec", [0, 1])
def test_multi_workgroup_rows_bit_identical(ctx, prec):
    """Hub rows of 1100..9000 entries (more than 1024: 4 or 8 workgroups
    per row, folding through scratch and an arrival counter) alone and
    mixed with a grid, k = 72 (two column chunks): the same bits, every
    residual at the tolerance, a second solve gives the same bits again,
    and so do PCG iterations replayed from captured hipGraphs (the counters
    are reset by the kernel itself, inside the graph)."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(200, 200)
    gs = [(n, u, v, graphs.loguniform_weights(len(u), seed=3)), _hubs(nh=12, seed=5, lo=1100, hi=9000)]
    fs = _factors(gs, seed=9, ordering=ss.ORDER_ROUNDS)
    k = 72
    saved = ctx.tuning()
    try:
        ctx.set_tuning(tile_waves=saved["tile_waves"], level_rows_per_wave=saved["level_rows_per_wave"],
                       graph_iters=0)
        mixed = _run(ctx, gs, fs, [0, 1], k, prec)
        again = _run(ctx, gs, fs, [0, 1], k, prec)
        ctx.set_tuning(tile_waves=saved["tile_waves"], level_rows_per_wave=saved["level_rows_per_wave"],
                       graph_iters=2)
        replayed = _run(ctx, gs, fs, [0, 1], k, prec)
    finally:
        ctx.set_tuning(**{x: saved[x] for x in ("tile_waves", "level_rows_per_wave", "graph_iters", "xcd_slices")})
    for i in range(2):
        alone = _run(ctx, gs, fs, [i], k, prec)[0]
        for key in ("Y", "Z", "p"):
            assert np.array_equal(alone[key], mixed[i][key]), (i, key)
            assert np.array_equal(again[i][key], mixed[i][key]), (i, key)
            assert np.array_equal(replayed[i][key], mixed[i][key]), (i, key)
        n, u, v, w = gs[i]
        L = _laplacian(n, u, v, w)
        Yp = _component_projection(n, u, v, mixed[i]["Y"])
        rr = np.linalg.norm(Yp - L @ mixed[i]["Z"], axis=0) / np.linalg.norm(Yp, axis=0)
        assert rr.max() <= 1.01e-8, i


@pytest.mark.parametrize("prec", [0, 1])
def test_hub_tiles_bit_identical(ctx, oracle_mod, prec):
    """A power-law graph whose first row tile holds 58K matrix entries (its
    SpMM tile is split) and whose hub rows make the sketch's heaviest rows,
    alone and mixed with a grid of another tile size: the same bits, and the
    residual at the tolerance."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.chung_lu(20000, 200000, gamma=2.2, seed=3)
    # a path through every vertex makes it connected (the oracle solves
    # connected graphs)
    u, v, _ = graphs.canonical(np.concatenate([u, np.arange(n - 1)]), np.concatenate([v, np.arange(1, n)]))
    m, gu, gv = graphs.grid2d(40, 30)
    gs = [(n, u, v, graphs.uniform_weights(len(u), seed=3)), (m, gu, gv, graphs.loguniform_weights(len(gu), seed=6))]
    fs = _factors(gs, seed=11, ordering=ss.ORDER_ROUNDS)
    mixed = _run(ctx, gs, fs, [0, 1], 16, prec)
    for i in range(2):
        alone = _run(ctx, gs, fs, [i], 16, prec)[0]
        for key in ("Y", "Z", "p"):
            assert np.array_equal(alone[key], mixed[i][key]), (i, key)
        nn, uu, vv, ww = gs[i]
        Yo = oracle_mod.sketch(uu, vv, ww, nn, 16, seed=10 + i, scale=1.0 / math.sqrt(16))
        assert np.array_equal(mixed[i]["Y"], Yo), i
        L = _laplacian(nn, uu, vv, ww)
        Yp = _component_projection(nn, uu, vv, mixed[i]["Y"])
        rr = np.linalg.norm(Yp - L @ mixed[i]["Z"], axis=0) / np.linalg.norm(Yp, axis=0)
        assert rr.max() <= 1.01e-8, i
        if i == 0 and prec == ss.PREC_MIXED:
            # mixed precision: leverage scores within 1e-4 of the FP64 oracle
            assert components(nn, uu, vv) == 1
            cp, ri, va = graphs.laplacian_csc(nn, uu, vv, ww)
            Zo, _, _ = oracle_mod.solve_lap(Yo, cp, ri, va, *fs[0].export(), stop_rule=1, tol=1e-10, max_iter=3000)
            eu, ev, lv = oracle_mod.upper_edges(nn, cp, ri, va)
            po = oracle_mod.probs(lv, oracle_mod.diff_norms(eu, ev, Zo), graphs.beta_for(nn, 0.3), 16)
            both = (mixed[i]["p"] < 1) & (po < 1)
            assert both.sum() > 1000
            assert np.all(np.abs(mixed[i]["p"][both] - po[both]) <= 1e-4 * po[both])


def test_batch_matches_oracle_per_graph(ctx, oracle_mod):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    gs = _mix()
    fs = _factors(gs)
    k = 12
    res = _run(ctx, gs, fs, list(range(len(gs))), k, ss.PREC_FP64)
    for i, (n, u, v, w) in enumerate(gs):
        r = res[i]
        # (a) signs and incidence: bit-exact with the graph's own edge ids
        Yo = oracle_mod.sketch(u, v, w, n, k, seed=10 + i, scale=1.0 / math.sqrt(k))
        assert np.array_equal(r["Y"], Yo), i
        # (b)+(c) every component's residual at the requested tolerance
        L = _laplacian(n, u, v, w)
        Yp = _component_projection(n, u, v, r["Y"])
        res_norm = np.linalg.norm(Yp - L @ r["Z"], axis=0)
        assert np.all(res_norm <= 1.01e-8 * np.linalg.norm(Yp, axis=0) + 1e-300), i
        cp, ri, va = graphs.laplacian_csc(n, u, v, w)
        uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
        if components(n, u, v) == 1:
            urp, uci, uva, dg = fs[i].export()
            Zo, _, _ = oracle_mod.solve_lap(Yo, cp, ri, va, urp, uci, uva, dg, stop_rule=1, tol=1e-12,
                                            max_iter=3000)
            a, b = r["Z"] - r["Z"].mean(0), Zo - Zo.mean(0)
            assert (np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)).max() < 1e-6, i
            po = oracle_mod.probs(lv, oracle_mod.diff_norms(uu, vv, Zo), graphs.beta_for(n, 0.3), k)
            both = (r["p"] < 1) & (po < 1)
            assert np.all(np.abs(r["p"][both] - po[both]) <= 1e-4 * po[both]), i
        # (d) the mask and weights given these probabilities are the oracle's
        ref = oracle_mod.sample_reweight(n, uu, vv, lv, r["p"], oracle_mod.coins(20 + i, len(uu)))
        ku, kv, kw = r["kept"]
        assert np.array_equal(ref["ku"], ku) and np.array_equal(ref["kv"], kv) and np.array_equal(ref["kw"], kw), i


@pytest.mark.parametrize("prec", [0, 1])
def test_batch_solve_gaussian_rhs(ctx, prec):
    import spec_spars_amd as ss
    gs = _mix()
    fs = _factors(gs, seed=7, ordering=ss.ORDER_ROUNDS)
    B = ss.Batch(ctx, gs)
    B.set_factors(fs)
    rng = np.random.default_rng(0)
    k = 16
    rhs = []
    for n, u, v, w in gs:
        b = rng.standard_normal((n, k))
        rhs.append(b - b.mean(0))
    info = B.solve(rhs, ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000, precision=prec))
    assert info["relres"].max() <= 1e-8
    for i, (n, u, v, w) in enumerate(gs):
        Z = B.read_block(i, "Z")
        Yp = _component_projection(n, u, v, rhs[i])
        L = _laplacian(n, u, v, w)
        rr = np.linalg.norm(Yp - L @ Z, axis=0) / np.linalg.norm(Yp, axis=0)
        assert rr.max() <= 1.01e-8, i
        assert np.all(info["iters"][i] > 0)
    B.close()


def test_single_graph_api_is_batch_of_one(ctx):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    gs = _mix()
    n, u, v, w = gs[0]
    f = _factors(gs)[0]
    G = ss.Graph(ctx, n, u, v, w)
    G.set_factor(f)
    solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000)
    G.sparsify(8, graphs.beta_for(n, 0.3), sketch_seed=10, sampling_seed=20, solve=solve)
    one = _run(ctx, gs, [f], [0], 8, ss.PREC_FP64)[0]
    assert np.array_equal(G.read_block("Z"), one["Z"])
    assert np.array_equal(G.read_probs(), one["p"])
    # the device Laplacian is the host one
    cp, ri, va = G.laplacian()
    hc, hr, hv = graphs.laplacian_csc(n, u, v, w)
    assert np.array_equal(cp, hc) and np.array_equal(ri, hr) and np.array_equal(va, hv)


def test_batch_argument_errors(ctx):
    import spec_spars_amd as ss
    gs = _mix()[:2]
    fs = _factors(gs)
    B = ss.Batch(ctx, gs)
    with pytest.raises(ss.SSError):
        B.set_factors(fs[:1])
    B.set_factors(fs)
    from spec_spars_amd import _lib
    dummy = np.zeros(8)
    assert _lib.load().ss_batch_read_probs(B._h, 2, dummy.ctypes.data_as(_lib.c_f64p)) == _lib.SS_ERR_INVALID
    with pytest.raises(IndexError):
        B.read_probs(2)
    # per-graph k must agree: built by hand through the C ABI
    o = (_lib.SparsifyOpts * 2)()
    for i in range(2):
        o[i].k = 8 + i
        o[i].scale = 1.0
        o[i].beta = 10.0
        o[i].solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8)
    assert _lib.load().ss_batch_sparsify(B._h, o, None) == _lib.SS_ERR_INVALID
    B.close()


def test_c1_grid64_pipeline(ctx, oracle_mod):
    """2D 64x64 grid, unit weights, k=24 Rademacher columns scaled 1/sqrt(k),
    PCG to 1e-8 relative residual, eps=0.3, seed 0, against the oracle."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(64, 64)
    w = np.ones(len(u))
    assert (n, len(u)) == (4096, 8064)
    k = 24
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    f = ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_NATURAL)
    G.set_factor(f)
    beta = graphs.beta_for(n, 0.3)
    rep = G.sparsify(k, beta, scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0,
                     solve=ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000))
    assert rep["max_relres"] <= 1e-8
    Y = G.read_block("Y")
    Yo = oracle_mod.sketch(u, v, w, n, k, seed=0, scale=1.0 / math.sqrt(k))
    assert np.array_equal(Y, Yo)
    Zo, _, _ = oracle_mod.solve_lap(Yo, cp, ri, va, *f.export(), stop_rule=1, tol=1e-12, max_iter=3000)
    Z = G.read_block("Z")
    a, b = Z - Z.mean(0), Zo - Zo.mean(0)
    assert (np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)).max() < 1e-5
    uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
    dn = oracle_mod.diff_norms(uu, vv, Zo)
    lev_o = -lv * dn ** 2
    lev = -lv * oracle_mod.diff_norms(uu, vv, Z) ** 2
    assert np.all(np.abs(lev - lev_o) <= 1e-4 * lev_o)
    p = G.read_probs()
    ref = oracle_mod.sample_reweight(n, uu, vv, lv, p, oracle_mod.coins(0, len(uu)))
    ku, kv, kw = G.read_kept()
    assert np.array_equal(ref["ku"], ku) and np.array_equal(ref["kw"], kw)


def test_spectral_quality_dense_graph(ctx):
    """Complete graph on 700 vertices, eps = 0.3, unit-variance signs: about a
    quarter of the edges are dropped and the generalized eigenvalues of
    (L_H, L_G) and 64 Rayleigh quotients stay within 1 +- eps.  The kept
    weights are w / sqrt(p) as in src/sparsifier.rs:655-657, so E[L_H] is
    below L_G where p < 1: the bound holds while p stays near 1 (at eps = 0.5
    on this graph the smallest eigenvalue is about 0.42)."""
    import scipy.linalg as sl
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n = 700
    iu, iv = np.triu_indices(n, 1)
    u, v, _ = graphs.canonical(iu.astype(np.int32), iv.astype(np.int32))
    w = graphs.uniform_weights(len(u), seed=3)
    eps = 0.3
    k = 64
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    G.set_factor(ss.Factor(cp, ri, va, seed=0))
    rep = G.sparsify(k, graphs.beta_for(n, eps), scale=1.0, sketch_seed=1, sampling_seed=2,
                     solve=ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000))
    ku, kv, kw = G.read_kept()
    assert 0.5 * len(u) < rep["kept"] < 0.9 * len(u)
    LG = _laplacian(n, u, v, w).toarray()
    LH = np.zeros((n, n))
    np.add.at(LH, (ku, kv), -kw)
    np.add.at(LH, (kv, ku), -kw)
    LH[np.diag_indices(n)] = -LH.sum(1)
    # on the complement of the all-ones vector
    Q = sl.null_space(np.ones((1, n)))
    lam = sl.eigh(Q.T @ LH @ Q, Q.T @ LG @ Q, eigvals_only=True)
    assert lam.min() >= 1 - eps and lam.max() <= 1 + eps
    rng = np.random.default_rng(1)
    x = rng.standard_normal((n, 64))
    x -= x.mean(0)
    rq = np.einsum("ij,ij->j", x, LH @ x) / np.einsum("ij,ij->j", x, LG @ x)
    assert rq.min() >= 1 - eps and rq.max() <= 1 + eps


def test_autotune_and_tuning_shapes(ctx, oracle_mod):
    """The autotuner returns valid shapes; every shape gives solves within
    tolerance of the oracle, and batches created under one shape stay
    bit-identical alone and mixed."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    gs = _mix()
    fs = _factors(gs, ordering=ss.ORDER_ROUNDS)
    saved = ctx.tuning()
    try:
        B = ss.Batch(ctx, gs)
        B.set_factors(fs)
        t = ctx.autotune(B, 8, ss.PREC_FP64, reps=2)
        assert t["tile_waves"] in (1, 2, 4, 8, 16) and 1 <= t["level_rows_per_wave"] <= 64
        assert t["ms_level"] > 0 and t["ms_tile"] > 0
        B.close()
        for tw, rpw in ((1, 1), (16, 8), (8, 3)):
            ctx.set_tuning(tile_waves=tw, level_rows_per_wave=rpw)
            mixed = _run(ctx, gs, fs, list(range(len(gs))), 8, ss.PREC_FP64)
            alone = _run(ctx, gs, fs, [2], 8, ss.PREC_FP64)[0]
            assert np.array_equal(alone["Z"], mixed[2]["Z"]) and np.array_equal(alone["p"], mixed[2]["p"])
            n, u, v, w = gs[0]
            cp, ri, va = graphs.laplacian_csc(n, u, v, w)
            Yo = oracle_mod.sketch(u, v, w, n, 8, seed=10, scale=1.0 / math.sqrt(8))
            assert np.array_equal(mixed[0]["Y"], Yo)
            Zo, _, _ = oracle_mod.solve_lap(Yo, cp, ri, va, *fs[0].export(), stop_rule=1, tol=1e-12, max_iter=3000)
            a, b = mixed[0]["Z"] - mixed[0]["Z"].mean(0), Zo - Zo.mean(0)
            assert (np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)).max() < 1e-6
        with pytest.raises(ss.SSError):
            ctx.set_tuning(tile_waves=3)
    finally:
        ctx.set_tuning(**{k: saved[k] for k in ("tile_waves", "level_rows_per_wave", "graph_iters", "xcd_slices")})


@pytest.mark.parametrize("prec", [0, 1])
def test_graph_replay_equals_direct_launches(ctx, prec):
    """PCG iterations replayed from captured hipGraphs give the same bits as
    direct launches, over repeated solves with eager work between them and
    a reused capture."""
    import spec_spars_amd as ss
    gs = _mix()
    fs = _factors(gs, ordering=ss.ORDER_ROUNDS)
    saved = ctx.tuning()
    try:
        ctx.set_tuning(tile_waves=saved["tile_waves"], level_rows_per_wave=1, graph_iters=0)
        direct = _run(ctx, gs, fs, list(range(len(gs))), 8, prec)
        for gi in (1, 3):
            ctx.set_tuning(tile_waves=saved["tile_waves"], level_rows_per_wave=1, graph_iters=gi)
            B = ss.Batch(ctx, gs)
            B.set_factors(fs)
            for rep in range(3):   # the second and third solves replay the cached graph
                from spec_spars_amd import graphs
                betas = [graphs.beta_for(g[0], 0.3) for g in gs]
                reps = B.sparsify(8, betas, sketch_seeds=[10 + i for i in range(len(gs))],
                                  sampling_seeds=[20 + i for i in range(len(gs))],
                                  solve=ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000,
                                                      precision=prec))
                for i in range(len(gs)):
                    assert np.array_equal(B.read_block(i, "Z"), direct[i]["Z"]), (gi, rep, i)
                    assert np.array_equal(B.read_probs(i), direct[i]["p"])
                    assert reps[i]["max_iters"] == direct[i]["rep"]["max_iters"]
            B.close()
    finally:
        ctx.set_tuning(**{k: saved[k] for k in ("tile_waves", "level_rows_per_wave", "graph_iters", "xcd_slices")})


@pytest.mark.parametrize("prec", [0, 1])
def test_columns_independent_of_k(ctx, prec):
    """A column's result does not depend on how many other columns share the
    call: every per-column reduction (the tile partials, |R|^2, the SpMM's
    P.Q) keeps an order that is fixed by the graph alone, so solving 8
    right-hand sides, or the same 8 inside a call with 40, gives the same
    bits (what merged serving requests rely on)."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.rgg(3000, seed=5)
    w = graphs.loguniform_weights(len(u), seed=5)
    G = ss.Graph(ctx, n, u, v, w)
    try:
        cp, ri, va = G.laplacian()
        G.set_factor(ss.Factor(cp, ri, va, seed=3, ordering=ss.ORDER_ROUNDS))
        rng = np.random.default_rng(4)
        Y = rng.standard_normal((n, 40))
        opts = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=prec)
        Z8, i8 = G.solve(Y[:, :8].copy(), opts)
        Z40, i40 = G.solve(Y, opts)
        assert np.array_equal(i8["iters"], i40["iters"][:8])
        assert np.array_equal(Z8, Z40[:, :8])
    finally:
        G.close()


@pytest.mark.parametrize("prec", [0, 1])
def test_sparsify_many_matches_alone(ctx, prec):
    """Several sparsify requests merged into one pass on one graph (their
    sketch columns side by side, one multi-RHS solve, one sampling pass
    each) give every request the kept edges it gets alone, bit for bit,
    in FP64 and in mixed precision, on a graph with a dense tail."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.chung_lu(20000, 200000, gamma=2.2, seed=7)
    w = graphs.uniform_weights(len(u), seed=7)
    G = ss.Graph(ctx, n, u, v, w)
    F = ss.Factor(*G.laplacian(), seed=1, ordering=ss.ORDER_ROUNDS)
    assert F.tail() > 0
    G.set_factor(F)
    solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=prec)
    reqs = [dict(k=16, beta=graphs.beta_for(n, 0.3), sketch_seed=1, sampling_seed=2),
            dict(k=24, beta=graphs.beta_for(n, 0.5), sketch_seed=3, sampling_seed=4, scale=0.3),
            dict(k=8, beta=graphs.beta_for(n, 0.2), sketch_seed=5, sampling_seed=6)]
    reps = G.sparsify_many(reqs, solve=solve)
    many = [G.read_kept_many(r) for r in range(len(reqs))]
    for r, q in enumerate(reqs):
        rep = G.sparsify(q["k"], q["beta"], scale=q.get("scale", 1.0 / math.sqrt(q["k"])),
                         sketch_seed=q["sketch_seed"], sampling_seed=q["sampling_seed"], solve=solve)
        ku, kv, kw = G.read_kept()
        assert rep["kept"] == reps[r]["kept"] == len(ku) and rep["max_iters"] == reps[r]["max_iters"]
        assert np.array_equal(ku, many[r][0]) and np.array_equal(kv, many[r][1])
        assert np.array_equal(kw, many[r][2])
    G.close()


def test_sparsify_many_fp16_rows_match_alone(ctx):
    """FP16 preconditioner rows (MIXED_RELIABLE_F16), the served form's
    scheme: four merged requests of 64 columns (one 256-column pass, the
    solves and SpMM four columns per lane) give every request the kept
    edges, iterations and replacement rounds it gets alone (64 columns, two
    per lane): the lane width and the merged width change no column's
    arithmetic.  The residuals meet the FP64 tolerance."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.chung_lu(20000, 200000, gamma=2.2, seed=7)
    w = graphs.uniform_weights(len(u), seed=7)
    assert ss.mixed_mode_for(w) == ss.MIXED_RELIABLE_F16
    G = ss.Graph(ctx, n, u, v, w)
    F = ss.Factor(*G.laplacian(), seed=1, ordering=ss.ORDER_ROUNDS)
    G.set_factor(F)
    solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=ss.PREC_MIXED,
                          mixed_mode=ss.MIXED_RELIABLE_F16)
    reqs = [dict(k=64, beta=graphs.beta_for(n, e), sketch_seed=i, sampling_seed=i)
            for i, e in enumerate([0.3, 0.2, 0.5, 0.3])]
    reps = G.sparsify_many(reqs, solve=solve)
    many = [G.read_kept_many(r) for r in range(len(reqs))]
    for r, q in enumerate(reqs):
        rep = G.sparsify(q["k"], q["beta"], sketch_seed=q["sketch_seed"], sampling_seed=q["sampling_seed"],
                         solve=solve)
        ku, kv, kw = G.read_kept()
        assert rep["max_relres"] <= 1e-8 and reps[r]["max_relres"] <= 1e-8
        assert rep["kept"] == reps[r]["kept"] == len(ku) and rep["max_iters"] == reps[r]["max_iters"]
        assert np.array_equal(ku, many[r][0]) and np.array_equal(kv, many[r][1])
        assert np.array_equal(kw, many[r][2])
    G.close()
