"""Pins the CPU oracle against published vectors, the reference's own
known-answer tests and mathematical identities.  CPU only."""
import math

import networkx as nx
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import components, golden


def _hex(words):
    return bytes(b for x in words for b in int(x).to_bytes(4, "little")).hex()


# ---- CityHash --------------------------------------------------------------
def test_cityhash_table_entries(oracle_mod):
    # CityHash's published test table, first row (empty input):
    # CityHash64WithSeed(s, 0, 1234567) and CityHash64WithSeeds(s, 0, 1234567, k0).
    k0, k2 = 0xc3a5c85c97cb3127, 0x9ae16a3b2f90404f
    assert oracle_mod.hash128to64(0, 1234567) == 0x75106db890237a4a
    assert oracle_mod.hash128to64((k2 - 1234567) % 2 ** 64, k0) == 0x3feac5f636039766


def test_sign_statistics(oracle_mod):
    # parity of HashLen0to16 itself is unpinned (fasthash-sys source absent);
    # these catch a constant, biased or column-independent hash.
    s = np.array([[oracle_mod.sketch_value(e, j, 7) for j in range(32)] for e in range(400)])
    assert set(np.unique(s)) == {-1.0, 1.0}
    assert abs(s.mean()) < 4 / math.sqrt(s.size)
    assert abs((s[:, :-1] * s[:, 1:]).mean()) < 4 / math.sqrt(s.size)
    u = np.array([[oracle_mod.sketch_value(e, j, 7, 2.0, True) for j in range(32)] for e in range(400)])
    assert np.all(np.abs(u) < 2.0)
    assert abs(u.var() - 4.0 / 3.0) < 0.05
    assert oracle_mod.city64_seed16(3, 4, 1) != oracle_mod.city64_seed16(3, 4, 2)
    assert oracle_mod.city64_seed16(3, 4, 1, 0) != oracle_mod.city64_seed16(3, 4, 1, 1)


# ---- ChaCha / coins --------------------------------------------------------
def test_chacha20_rfc7539(oracle_mod):
    key = [int.from_bytes(bytes(range(4 * i, 4 * i + 4)), "little") for i in range(8)]
    # RFC 7539 2.3.2: counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00
    out = oracle_mod.chacha_block(key, 1 | (0x09000000 << 32), 0x4a000000, 20)
    assert " ".join("%08x" % x for x in out) == (
        "e4e7f110 15593bd1 1fdd0f50 c47120a3 c7f4d1c7 0368c033 9aaa2204 4e6cd4c3 "
        "466482d2 09aa9f07 05d7c214 a2028bd9 d19c12b5 b94e16de e883d0cb 4e3c50a2")
    z = np.zeros(8, np.uint32)
    # RFC 7539 A.1 test vectors 1 and 2 (zero key, counter 0 and 1)
    assert _hex(oracle_mod.chacha_block(z, 0, 0, 20)) == (
        "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
        "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert _hex(oracle_mod.chacha_block(z, 1, 0, 20)) == (
        "9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
        "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")


def test_coin_stream_layout(oracle_mod):
    # draw e = keystream words 2e (low), 2e+1 (high) of 12-round blocks
    seed = 12345
    key = oracle_mod.stdrng_key(seed)
    c = oracle_mod.coins(seed, 40)
    for e in (0, 1, 7, 8, 9, 39):
        blk = oracle_mod.chacha_block(key, e // 8, 0, 12)
        u = int(blk[2 * (e % 8)]) | (int(blk[2 * (e % 8) + 1]) << 32)
        assert c[e] == (u >> 12) * 2.0 ** -52
    # windowed generation agrees with the full stream
    assert np.array_equal(oracle_mod.coins(seed, 13, first=27), c[27:40])
    big = oracle_mod.coins(0, 200000)
    assert big.min() >= 0.0 and big.max() < 1.0
    assert abs(big.mean() - 0.5) < 0.005
    assert not np.array_equal(oracle_mod.coins(1, 16), oracle_mod.coins(2, 16))


# ---- Laplacian, incidence, edge order -------------------------------------
def test_laplacian_matches_dense_and_order(oracle_mod):
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(7, 5)
    w = graphs.loguniform_weights(len(u), seed=2)
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    A = sp.csc_matrix((va, ri, cp), shape=(n, n)).toarray()
    D = np.zeros((n, n))
    for a, b, x in zip(u, v, w):
        D[a, b] -= x; D[b, a] -= x; D[a, a] += x; D[b, b] += x
    assert np.allclose(A, D, rtol=0, atol=1e-12)
    assert np.allclose(A.sum(0), 0.0, atol=1e-10)       # lap_valid_random
    uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
    assert np.array_equal(uu, u) and np.array_equal(vv, v)
    assert np.array_equal(-lv, w)


def test_incidence_kat(oracle_mod):
    # evim_csc_equiv (src/tests.rs:363-369): +sqrt(w) at the second endpoint,
    # -sqrt(w) at the first, bit for bit.
    for (a, b, w) in [(0, 1, 2.0), (3, 1, 0.37), (2, 5, 1e-3)]:
        Y = oracle_mod.sketch([a], [b], [w], 6, 1, seed=0, scale=1.0)
        q = oracle_mod.sketch_value(0, 0, 0)
        assert Y[b, 0] == q * math.sqrt(w) and Y[a, 0] == -q * math.sqrt(w)


def test_sketch_brute_force(oracle_mod):
    from spec_spars_amd import graphs
    n, u, v = graphs.rgg(60, seed=1, mean_degree=5)
    w = graphs.uniform_weights(len(u), seed=1)
    k = 10
    for uniform in (False, True):
        Y = oracle_mod.sketch(u, v, w, n, k, seed=3, scale=0.5, uniform=uniform)
        B = np.zeros((n, len(u)))
        B[u, np.arange(len(u))] = -np.sqrt(w)
        B[v, np.arange(len(u))] = np.sqrt(w)
        Q = np.array([[oracle_mod.sketch_value(e, j, 3, 0.5, uniform) for j in range(k)] for e in range(len(u))])
        assert np.allclose(Y, B @ Q, rtol=1e-12, atol=1e-12)
        assert np.allclose(Y.sum(0), 0.0, atol=1e-12)     # jl_sketch_zero


# ---- level sets ------------------------------------------------------------
def test_levels_brute_force(oracle_mod):
    rng = np.random.default_rng(0)
    for trial in range(5):
        n = 40
        rows = [[i] for i in range(n)]
        for i in range(n):
            for c in rng.choice(np.arange(i + 1, n), size=min(3, n - i - 1), replace=False) if i < n - 1 else []:
                rows[i].append(int(c))
        urp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
        uci = np.array([c for r in rows for c in sorted(r)], np.int32)
        f, b = oracle_mod.levels(n, urp, uci)
        G = nx.DiGraph()
        G.add_nodes_from(range(n))
        for i, r in enumerate(rows):
            for c in r:
                if c != i:
                    G.add_edge(i, c)
        for x in range(n):
            anc = nx.ancestors(G, x) | {x}
            lf = max(nx.dag_longest_path_length(G.subgraph(nx.ancestors(G, x) | {x})), 0) + 1
            assert f[x] == lf
            desc = nx.descendants(G, x) | {x}
            lb = nx.dag_longest_path_length(G.subgraph(desc)) + 1
            assert b[x] == lb
            del anc


# ---- PCG -------------------------------------------------------------------
@pytest.mark.parametrize("name", ["grid8x8_loguniform", "rgg120", "tree60"])
def test_pcg_matches_exact_solve(oracle_mod, name):
    import spec_spars_amd as ss
    g = golden(name)
    n = int(g["n"])
    cp, ri, va = oracle_mod.laplacian(n, g["u"], g["v"], g["w"])
    f = ss.Factor(cp, ri, va, seed=4, ordering=ss.ORDER_NATURAL)
    urp, uci, uva, dg = f.export()
    Z, it, rr = oracle_mod.solve_lap(g["Y"], cp, ri, va, urp, uci, uva, dg, stop_rule=1, tol=1e-11)
    assert components(n, g["u"], g["v"]) == 1
    Zc = Z - Z.mean(0)
    rel = np.linalg.norm(Zc - g["Z"]) / np.linalg.norm(g["Z"])
    assert rel < 1e-8 and rr.max() < 1e-10


# ---- effective resistance / probabilities ----------------------------------
@pytest.mark.parametrize("name", ["grid8x8_loguniform", "rgg120", "tree60"])
def test_diff_norms_probs_foster(oracle_mod, name):
    g = golden(name)
    n, u, v, w = int(g["n"]), g["u"], g["v"], g["w"]
    m = len(u)
    L = np.zeros((n, n))
    for a, b, x in zip(u, v, w):
        L[a, b] -= x; L[b, a] -= x; L[a, a] += x; L[b, b] += x
    B = np.zeros((n, m))
    B[u, np.arange(m)] = -np.sqrt(w)
    B[v, np.arange(m)] = np.sqrt(w)
    Z = np.linalg.pinv(L, hermitian=True) @ B          # Q = I: d_e^2 = R_eff exactly
    d = oracle_mod.diff_norms(u, v, Z)
    assert np.allclose(d ** 2, g["reff"], rtol=1e-9)
    assert abs(np.sum(w * d ** 2) - (n - 1)) < 1e-8 * n   # Foster's theorem
    if name == "tree60":
        assert np.allclose(w * d ** 2, 1.0, rtol=1e-9)     # every tree edge has leverage 1
    beta = 0.5
    p = oracle_mod.probs(-w, d, beta, m)
    assert np.all(p <= 1.0) and np.all(p >= 0.0)
    assert abs(np.sum(p) * m / beta - (n - 1)) < 1e-8 * n  # no clipping at this beta
    p2 = oracle_mod.probs(-w, d, 1e9, m)
    assert np.all(p2 == 1.0)


# ---- sampling / reweighting ------------------------------------------------
def test_sampling_test_kat(oracle_mod):
    g = golden("sampling_test")
    n = int(g["n"])
    cp, ri, va = oracle_mod.laplacian(n, g["v1"], g["v2"], np.ones(len(g["v1"])))
    u, v, lv = oracle_mod.upper_edges(n, cp, ri, va)
    assert np.array_equal(u, g["u"]) and np.array_equal(v, g["v"])
    assert components(n, u, v) == 1
    for seed in (0, 1, 99):
        r = oracle_mod.sample_reweight(n, u, v, lv, g["probs"], oracle_mod.coins(seed, len(u)))
        assert components(n, r["ku"], r["kv"]) == int(g["components"])
        assert np.array_equal(r["mask"], (g["probs"] > 0.5).astype(np.uint8))


def test_sampling_verify_ratios(oracle_mod):
    # src/tests.rs:213-261 on a synthetic graph: p = 0.5 halves the edges and
    # divides the diagonal sum by sqrt(2) (w/sqrt(p) reweighting).
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(100, 100)
    w = graphs.uniform_weights(len(u), seed=8)
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    uu, vv, lv = oracle_mod.upper_edges(n, cp, ri, va)
    for seed in range(3):
        r = oracle_mod.sample_reweight(n, uu, vv, lv, np.full(len(uu), 0.5), oracle_mod.coins(seed, len(uu)))
        assert abs(len(uu) / r["kept"] - 2.0) < 0.05
        dg = np.zeros(n)
        np.add.at(dg, uu, w); np.add.at(dg, vv, w)
        dh = dg + r["ddiag"]
        assert abs(dg.sum() / dh.sum() - math.sqrt(2.0)) < 0.05
        # H is a valid Laplacian: rebuilt diagonal equals D_G + change
        dk = np.zeros(n)
        np.add.at(dk, r["ku"], r["kw"]); np.add.at(dk, r["kv"], r["kw"])
        assert np.allclose(dk, dh, rtol=1e-12, atol=1e-12)
        # H's edges are a subset of G's (check_for_additions)
        ge = set(zip(uu.tolist(), vv.tolist()))
        assert all((a, b) in ge for a, b in zip(r["ku"].tolist(), r["kv"].tolist()))
        assert np.allclose(r["kw"], r["kw"] * 0 + np.sqrt(2.0) * w[r["mask"].astype(bool)], rtol=1e-15)


def test_sketch_edges_and_permute_csc(oracle_mod):
    """The subset sketch gives the full sketch rows of every row whose edges
    are all in the subset (and the full sketch for all edges); the CSC
    permutation equals scipy's P A P^T."""
    import scipy.sparse as sp
    from spec_spars_amd import graphs
    n, u, v = graphs.chung_lu(3000, 20000, seed=1)
    w = graphs.uniform_weights(len(u), seed=1)
    Y = oracle_mod.sketch(u, v, w, n, 16, seed=3, scale=0.25)
    rows = np.array([5, 17, 1000, 2999])
    sel = np.nonzero(np.isin(u, rows) | np.isin(v, rows))[0]
    Ys = oracle_mod.sketch_edges(sel, u[sel], v[sel], w[sel], n, 16, seed=3, scale=0.25)
    assert np.array_equal(Ys[rows], Y[rows])
    assert np.array_equal(oracle_mod.sketch_edges(np.arange(len(u)), u, v, w, n, 16, seed=3, scale=0.25), Y)
    cp, ri, va = oracle_mod.laplacian(n, u, v, w)
    order = np.random.default_rng(0).permutation(n)
    pc, pr, pv = oracle_mod.permute_csc(cp, ri, va, order)
    Lp = sp.csc_matrix((va, ri, cp), shape=(n, n))[order][:, order].tocsc()
    Lp.sort_indices()
    assert np.array_equal(Lp.indptr, pc) and np.array_equal(Lp.indices, pr) and np.array_equal(Lp.data, pv)
