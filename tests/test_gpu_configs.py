"""The benchmark configurations at their real sizes on the GPU, checked
against the oracle: the 128^3 grid in FP64 (k = 48), the 5M-vertex / 50M-edge
Chung-Lu graph in mixed precision (k = 64), the batch of 512 random
geometric graphs (k = 32 each, 16,384 right-hand sides) and a serving window
of Poisson arrivals over the three families.

Contracts (src/sparsifier.rs:490-691, include/custom_cg.hpp:95-274):
  * sketch Y bit-exact with the oracle's (every row, or every row of a
    random sample where the oracle would take minutes);
  * each column's FP64 true relative residual ||b - L z|| / ||b|| <= 1e-8;
  * the solution within 1e-7 of the oracle's PCG with the same factor (in
    the factor's position space), iteration counts within one;
  * probabilities bit-exact given the device's Z, and leverage scores within
    1e-4 relative of the FP64 oracle's;
  * mask, kept endpoints and kept weights bit-exact with the oracle's
    sample_reweight given the device's probabilities.
"""
import math
import os
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import scipy.sparse as sp

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 4))


def _say(*a):
    print("[configs]", *a, flush=True)


def _true_relres(cp, ri, va, B, Z):
    A = sp.csc_matrix((va, ri, cp), shape=(len(cp) - 1, len(cp) - 1)).tocsr()
    R = B - A @ Z
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


def _oracle_positions(oracle_mod, F, cp, ri, va, B, tol):
    """The oracle's PCG with the device's factor: in the factor's position
    space, every column on its own thread."""
    order, _ = F.order()
    pc, pr, pv = oracle_mod.permute_csc(cp, ri, va, order)
    urp, uci, uva, dg = F.export()
    Bp = np.ascontiguousarray(B[order])

    def one(j):
        return oracle_mod.solve_lap(Bp[:, j:j + 1], pc, pr, pv, urp, uci, uva, dg, stop_rule=1, tol=tol,
                                    max_iter=3000)

    with ThreadPoolExecutor(max_workers=min(THREADS, B.shape[1])) as ex:
        res = list(ex.map(one, range(B.shape[1])))
    Zo = np.empty_like(B)
    Zo[order] = np.concatenate([r[0] for r in res], axis=1)
    return Zo, np.array([int(r[1].max()) for r in res])


def _check_sampling(oracle_mod, n, cp, ri, va, p, seed, kept):
    """Mask and kept edges given the device's probabilities."""
    uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
    ref = oracle_mod.sample_reweight(n, uu, vv, lv, p, oracle_mod.coins(seed, len(uu)))
    ku, kv, kw = kept
    assert ref["kept"] == len(ku)
    assert np.array_equal(ref["ku"], ku) and np.array_equal(ref["kv"], kv) and np.array_equal(ref["kw"], kw)
    return uu, vv, lv


def _rel_diff(Z, Zo):
    a, b = Z - Z.mean(0), Zo - Zo.mean(0)
    return np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)


@pytest.mark.timeout(900)
def test_c2_grid128_fp64_at_size(ctx, oracle_mod):
    """3D 128^3 grid (n = 2,097,152, m = 6,242,304), weights 10^U(-3,3),
    k = 48, FP64 everywhere, PCG to rel 1e-8, eps = 0.3, seeds 0."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid3d(128)
    w = graphs.loguniform_weights(len(u), seed=0)
    k, m = 48, len(u)
    assert (n, m) == (2097152, 6242304)
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    F = ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_ROUNDS)
    G.set_factor(F)
    beta = graphs.beta_for(n, 0.3)
    sv = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=ss.PREC_FP64)
    rep = G.sparsify(k, beta, scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0, solve=sv)
    _say("c2 sparsify %.0f ms, %d iterations, relres %.2e, kept %d" %
         (rep["ms_total"], rep["max_iters"], rep["max_relres"], rep["kept"]))
    Y, Z, p = G.read_block("Y"), G.read_block("Z"), G.read_probs()
    kept = G.read_kept()
    # a second solve of two columns to rel 1e-10, against the oracle's
    B = Y - Y.mean(0)
    Z2, info2 = G.solve(B[:, :2], ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-10, max_iter=3000))
    G.close()
    assert np.array_equal(Y, oracle_mod.sketch(u, v, w, n, k, seed=0, scale=1.0 / math.sqrt(k)))
    _say("c2 sketch bit-exact")
    rr = _true_relres(cp, ri, va, B, Z)
    assert rr.max() <= 1.05e-8, rr.max()
    uu, vv, lv = _check_sampling(oracle_mod, n, cp, ri, va, p, 0, kept)
    assert np.array_equal(p, oracle_mod.probs(lv, oracle_mod.diff_norms(uu, vv, Z), beta, k))
    _say("c2 residuals, probabilities and sampling checked")
    Zo, it = _oracle_positions(oracle_mod, F, cp, ri, va, B[:, :2], 1e-10)
    assert np.abs(info2["iters"] - it).max() <= 1, (info2["iters"], it)
    assert _rel_diff(Z2, Zo).max() < 1e-7
    # leverage scores of the rel-1e-8 pass on these two columns vs the oracle's
    lo = -lv * oracle_mod.diff_norms(uu, vv, Zo) ** 2
    ld = -lv * oracle_mod.diff_norms(uu, vv, Z[:, :2]) ** 2
    big = lo > 1e-6 * lo.max()
    assert np.all(np.abs(ld[big] - lo[big]) <= 1e-4 * lo[big])


@pytest.mark.timeout(900)
def test_c3_chunglu5m_mixed_at_size(ctx, oracle_mod):
    """Chung-Lu n = 5,000,000, m = 50,000,000 (exponent 2.2), weights
    U[0.5, 2], k = 64, FP32 SpMM / triangular solves with FP64 reductions
    and refinement to rel 1e-8, eps = 0.3, the preconditioner's rows in FP16
    (the served form's scheme); the leverage scores against the FP64
    oracle's PCG on two columns."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.chung_lu(5_000_000, 50_000_000, gamma=2.2, seed=0)
    w = graphs.uniform_weights(len(u), seed=0)
    k, m = 64, len(u)
    assert m == 50_000_000
    _say("c3 graph generated")
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    F = ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_ROUNDS)
    assert F.tail() > 10000
    G.set_factor(F)
    _say("c3 factor uploaded (tail %d)" % F.tail())
    beta = graphs.beta_for(n, 0.3)
    sv = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=ss.PREC_MIXED,
                       mixed_mode=ss.mixed_mode_for(w))
    assert sv.mixed_mode == ss.MIXED_RELIABLE_F16
    rep = G.sparsify(k, beta, scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0, solve=sv)
    _say("c3 sparsify %.0f ms, %d iterations, %d sweeps, relres %.2e, kept %d" %
         (rep["ms_total"], rep["max_iters"], rep["refinements"], rep["max_relres"], rep["kept"]))
    assert rep["max_relres"] <= 1e-8
    Y, Z, p = G.read_block("Y"), G.read_block("Z"), G.read_probs()
    kept = G.read_kept()
    G.close()
    # sketch: 3,000 random rows and the 32 highest-degree hubs, bit-exact
    rng = np.random.default_rng(3)
    rows = np.unique(np.concatenate([rng.choice(n, 3000, replace=False), np.arange(32)]))
    inrow = np.zeros(n, bool)
    inrow[rows] = True
    sel = np.nonzero(inrow[u] | inrow[v])[0]
    Yo = oracle_mod.sketch_edges(sel, u[sel], v[sel], w[sel], n, k, seed=0, scale=1.0 / math.sqrt(k))
    assert np.array_equal(Y[rows], Yo[rows])
    del Yo
    _say("c3 sketch rows bit-exact (%d rows, %d edges)" % (len(rows), len(sel)))
    B = Y - Y.mean(0)
    del Y
    rr = _true_relres(cp, ri, va, B, Z)
    assert rr.max() <= 1.05e-8, rr.max()
    uu, vv, lv = _check_sampling(oracle_mod, n, cp, ri, va, p, 0, kept)
    assert np.array_equal(p, oracle_mod.probs(lv, oracle_mod.diff_norms(uu, vv, Z), beta, k))
    _say("c3 residuals, probabilities and sampling checked")
    Zo, it = _oracle_positions(oracle_mod, F, cp, ri, va, B[:, :2], 1e-10)
    lo = -lv * oracle_mod.diff_norms(uu, vv, Zo) ** 2
    ld = -lv * oracle_mod.diff_norms(uu, vv, Z[:, :2]) ** 2
    big = lo > 1e-6 * lo.max()
    assert big.sum() > 1_000_000
    err = np.abs(ld[big] - lo[big]) / lo[big]
    _say("c3 leverage vs FP64 oracle: max rel err %.2e over %d edges" % (err.max(), big.sum()))
    assert err.max() <= 1e-4


def _rgg512():
    from spec_spars_amd import graphs
    rng = np.random.default_rng(512)
    sizes = np.round(10.0 ** rng.uniform(3.0, math.log10(2e5), 512)).astype(np.int64)

    def one(i):
        _, u, v = graphs.rgg(int(sizes[i]), seed=1000 + i, mean_degree=8.0)
        return int(sizes[i]), u, v, np.ones(len(u))

    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        return list(ex.map(one, range(512)))


@pytest.mark.timeout(900)
def test_c4_rgg512_batch_at_size(ctx, oracle_mod):
    """512 random geometric graphs (n log-uniform in [1e3, 2e5], mean degree
    8, unit weights) in one batch, k = 32 each (16,384 right-hand sides),
    mixed precision: every graph's residual, sketch, probabilities and
    sampling against the oracle, and the oracle's FP64 leverage on four."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    gs = _rgg512()
    _say("c4 graphs generated: sum n %d, sum m %d" % (sum(g[0] for g in gs), sum(len(g[1]) for g in gs)))
    k = 32
    ss.set_factor_threads(1)
    try:
        with ThreadPoolExecutor(max_workers=THREADS) as ex:
            lap = list(ex.map(lambda g: graphs.laplacian_csc(*g), gs))
            fs = list(ex.map(lambda L: ss.Factor(*L, seed=0, ordering=ss.ORDER_ROUNDS), lap))
    finally:
        ss.set_factor_threads(0)
    Bt = ss.Batch(ctx, gs)
    Bt.set_factors(fs)
    betas = [graphs.beta_for(g[0], 0.3) for g in gs]
    sv = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000, precision=ss.PREC_MIXED)
    reps = Bt.sparsify(k, betas, sketch_seeds=list(range(512)), sampling_seeds=[1000 + i for i in range(512)],
                       solve=sv)
    _say("c4 batch sparsify %.0f ms, max relres %.2e" % (reps[0]["ms_total"], max(r["max_relres"] for r in reps)))
    assert max(r["max_relres"] for r in reps) <= 1e-8
    res = [(Bt.read_block(i, "Y"), Bt.read_block(i, "Z"), Bt.read_probs(i), Bt.read_kept(i)) for i in range(512)]
    Bt.close()
    lock = threading.Lock()
    done = [0]

    def check(i):
        n, u, v, w = gs[i]
        Y, Z, p, kept = res[i]
        assert np.array_equal(Y, oracle_mod.sketch(u, v, w, n, k, seed=i, scale=1.0 / math.sqrt(k))), i
        cp, ri, va = lap[i]
        uu, vv, lv = _check_sampling(oracle_mod, n, cp, ri, va, p, 1000 + i, kept)
        assert np.array_equal(p, oracle_mod.probs(lv, oracle_mod.diff_norms(uu, vv, Z), betas[i], k)), i
        with lock:
            done[0] += 1
        return True

    with ThreadPoolExecutor(max_workers=THREADS) as ex:
        assert all(ex.map(check, range(512)))
    assert done[0] == 512
    _say("c4 sketch, probabilities and sampling of all 512 graphs checked")
    # the oracle's FP64 PCG on four graphs: residuals and leverage
    for i in (0, 1, 100, 511):
        n, u, v, w = gs[i]
        Y, Z, p, kept = res[i]
        cp, ri, va = lap[i]
        B = Y - Y.mean(0)
        from conftest import components
        if components(n, u, v) == 1:
            assert _true_relres(cp, ri, va, B, Z).max() <= 1.05e-8
            Zo, _ = _oracle_positions(oracle_mod, fs[i], cp, ri, va, B[:, :4], 1e-11)
            uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
            lo = -lv * oracle_mod.diff_norms(uu, vv, Zo) ** 2
            ld = -lv * oracle_mod.diff_norms(uu, vv, Z[:, :4]) ** 2
            big = lo > 1e-6 * lo.max()
            assert np.all(np.abs(ld[big] - lo[big]) <= 1e-4 * lo[big]), i


@pytest.mark.timeout(900)
def test_c5_serving_alone_equals_mixed(ctx):
    """Poisson arrivals at 200 req/s for 4 s over resident graphs of the three
    families (n log-uniform in [1e4, 1e5]), half sparsify (eps in {0.2, 0.3,
    0.5}, k = 32, same-graph requests merged into one pass), half 16-RHS
    solves (same-graph requests merged into one solve): every completed
    request's result is bit-identical to the request run alone."""
    import spec_spars_amd as ss
    from spec_spars_amd.serve import Registry, Server, poisson_requests
    import bench
    pool, desc = bench.serve_pool(6, seed=5, n_max=1e5)
    reg = Registry(pool, threads=4)
    try:
        reqs = poisson_requests(rate=200, duration=4.0, graphs=len(pool), seed=11, k=32, rhs=16)
        out = Server(reg, slots=4, precision=ss.PREC_MIXED, reserve_bytes=1 << 30, keep_arrays=True).run(
            reqs, deadline=40.0)
        _say("c5 served: %d/%d completed, p50 %.1f ms, p99 %.1f ms" %
             (out["completed"], out["requests"], out["p50_ms"], out["p99_ms"]))
        assert out["completed"] == len(reqs) and out["failed"] == 0
        merged = sum(1 for r in reqs if r.result.get("group", 1) > 1)
        assert merged > 0
        for r in reqs:
            alone = reg.execute(r, ss.PREC_MIXED)
            if r.op == "sparsify":
                assert alone["kept"] == r.result["kept"]
                for a, b in zip(alone["arrays"], r.result["arrays"]):
                    assert np.array_equal(a, b), r.rid
            else:
                assert np.array_equal(alone["Z"], r.result["Z"]), r.rid
        _say("c5: %d requests (%d of them merged) equal to their runs alone" % (len(reqs), merged))
    finally:
        reg.close()
