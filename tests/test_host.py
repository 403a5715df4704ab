"""Host-side logic of the library (no GPU): C-ABI exports, the approximate-
Cholesky factor, level sets, graph generators and error paths."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, components


def test_library_exports_every_declared_symbol():
    import spec_spars_amd._lib as L
    lib = ctypes.CDLL(L.LIB_PATH)
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(ssh?_[a-z0-9_]+)\s*\(", txt))
    assert len(names) >= 20
    for nm in sorted(names):
        assert hasattr(lib, nm), nm
    assert set(L.SIGNATURES) <= names


def _factor(n, u, v, w, seed=1):
    import oracle
    import spec_spars_amd as ss
    cp, ri, va = oracle.laplacian(n, u, v, w)
    return ss.Factor(cp, ri, va, seed=seed, ordering=ss.ORDER_NATURAL), (cp, ri, va)


def test_factor_invariants():
    from spec_spars_amd import graphs
    for n, u, v in (graphs.grid2d(20, 17), graphs.rgg(700, seed=2), graphs.chung_lu(800, 3000, seed=1)):
        w = graphs.loguniform_weights(len(u), seed=3)
        f, _ = _factor(n, u, v, w)
        urp, uci, uva, d = f.export()
        assert urp[-1] == f.info()["nnz"]
        for c in range(n):
            cols = uci[urp[c]:urp[c + 1]]
            vals = uva[urp[c]:urp[c + 1]]
            assert cols[0] == c and vals[0] == 1.0           # unit diagonal first
            assert np.all(np.diff(cols) > 0)                   # rows ascending, > c
            if len(cols) > 1:
                assert d[c] > 0
                assert abs(vals[1:].sum() + 1.0) < 1e-12        # -w/tot sums to -1
                assert np.all(vals[1:] < 0)
            else:
                assert d[c] == 0.0
        # exactly one grounded column per connected component
        assert int((d == 0).sum()) == components(n, u, v)


def test_factor_tree_is_exact():
    # leaf-first tree: every column has one lower entry, so no fill is sampled
    # and L D L^T reproduces the Laplacian (src/example.cc:262 loop is empty).
    rng = np.random.default_rng(4)
    n = 300
    parent = np.array([rng.integers(i + 1, n) for i in range(n - 1)])
    from spec_spars_amd import graphs
    u, v, _ = graphs.canonical(np.arange(n - 1), parent)
    w = graphs.uniform_weights(len(u), seed=1)
    f, (cp, ri, va) = _factor(n, u, v, w)
    urp, uci, uva, d = f.export()
    L = np.zeros((n, n))
    for c in range(n):
        L[uci[urp[c]:urp[c + 1]], c] = uva[urp[c]:urp[c + 1]]
    import scipy.sparse as sp
    A = sp.csc_matrix((va, ri, cp), shape=(n, n)).toarray()
    assert np.allclose(L @ np.diag(d) @ L.T, A, rtol=1e-12, atol=1e-12)


def test_factor_unbiased():
    # E[L D L^T] = L_G for the randomized elimination
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(5, 5)
    w = graphs.uniform_weights(len(u), seed=2)
    acc = np.zeros((n, n))
    trials = 300
    for s in range(trials):
        f, (cp, ri, va) = _factor(n, u, v, w, seed=s)
        urp, uci, uva, d = f.export()
        L = np.zeros((n, n))
        for c in range(n):
            L[uci[urp[c]:urp[c + 1]], c] = uva[urp[c]:urp[c + 1]]
        acc += L @ np.diag(d) @ L.T
    import scipy.sparse as sp
    A = sp.csc_matrix((va, ri, cp), shape=(n, n)).toarray()
    assert np.linalg.norm(acc / trials - A) / np.linalg.norm(A) < 0.05


def test_factor_deterministic_and_seeded():
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(30, 30)
    w = graphs.loguniform_weights(len(u), seed=0)
    a = _factor(n, u, v, w, seed=5)[0].export()
    b = _factor(n, u, v, w, seed=5)[0].export()
    c = _factor(n, u, v, w, seed=6)[0].export()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not all(len(x) == len(y) and np.array_equal(x, y) for x, y in zip(a, c))


def test_rounds_factor_with_dense_tail_same_for_every_thread_count():
    """A power-law graph whose rounds order ends in a dense tail (its
    eliminations spread over all threads, phases behind spin barriers):
    the same factor bits with 1, 2, 3 and the default number of threads."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.chung_lu(5000, 100000, gamma=2.2, seed=1)
    w = graphs.uniform_weights(len(u), seed=1)
    cp, ri, va = graphs.laplacian_csc(n, u, v, w)
    out = []
    try:
        for threads in (1, 2, 3, 0):
            ss.set_factor_threads(threads)
            f = ss.Factor(cp, ri, va, seed=4, ordering=ss.ORDER_ROUNDS)
            assert f.tail() > 100
            out.append((f.export(), f.order()[0], f.tail()))
    finally:
        ss.set_factor_threads(0)
    for exp, order, tail in out[1:]:
        assert tail == out[0][2]
        assert np.array_equal(order, out[0][1])
        assert all(np.array_equal(x, y) for x, y in zip(exp, out[0][0]))


def test_levels_match_oracle_dp():
    import oracle
    from spec_spars_amd import graphs
    for n, u, v in (graphs.grid2d(40, 30), graphs.grid3d(9), graphs.rgg(2000, seed=1)):
        f, _ = _factor(n, u, v, graphs.uniform_weights(len(u), seed=1))
        urp, uci, _, _ = f.export()
        lf, lb = f.levels()
        of, ob = oracle.levels(n, urp, uci)
        assert np.array_equal(lf, of) and np.array_equal(lb, ob)
        assert f.info()["depth_fwd"] == of.max() and f.info()["depth_bwd"] == ob.max()


def test_factor_rejects_bad_input():
    import spec_spars_amd as ss
    with pytest.raises(ss.SSError) as e:
        ss.Factor(np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32), np.ones(2))
    assert e.value.code == 1
    with pytest.raises(ss.SSError):
        ss.Factor(np.array([0, 1, 2], np.int32), np.array([0, 7], np.int32), np.ones(2))


def test_generators_canonical_and_sizes():
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(64, 64)
    assert (n, len(u)) == (4096, 8064)                 # c1
    assert 3 * 128 * 128 * 127 == 6242304             # c2 edge count
    n3, u3, v3 = graphs.grid3d(12)
    assert len(u3) == 3 * 12 * 12 * 11
    for (nn, a, b) in ((n, u, v), (n3, u3, v3), graphs.rgg(3000, seed=3), graphs.chung_lu(2000, 9000, seed=2)):
        assert np.all(a < b)
        key = b.astype(np.int64) * nn + a
        assert np.all(np.diff(key) > 0)
    n, u, v = graphs.chung_lu(5000, 20000, seed=1)
    assert len(u) == 20000
    w = graphs.loguniform_weights(100000, seed=0)
    assert w.min() >= 1e-3 and w.max() <= 1e3
    assert abs(np.log10(w).mean()) < 0.05


@pytest.mark.parametrize("ordering", [1, 2])
def test_ordered_factor_invariants_and_levels(ordering):
    import oracle
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    for n, u, v in (graphs.grid3d(8), graphs.rgg(900, seed=1), graphs.chung_lu(700, 2500, seed=3)):
        w = graphs.loguniform_weights(len(u), seed=1)
        cp, ri, va = oracle.laplacian(n, u, v, w)
        f = ss.Factor(cp, ri, va, seed=2, ordering=ordering)
        order, rounds = f.order()
        assert np.array_equal(np.sort(order), np.arange(n))          # a permutation
        urp, uci, uva, d = f.export()
        for c in range(n):
            cols = uci[urp[c]:urp[c + 1]]
            assert cols[0] == c and np.all(np.diff(cols) > 0)
            if len(cols) > 1:
                assert abs(uva[urp[c] + 1:urp[c + 1]].sum() + 1.0) < 1e-12
        assert int((d == 0).sum()) == components(n, u, v)
        lf, lb = f.levels()
        of, ob = oracle.levels(n, urp, uci)
        assert np.array_equal(lf, of) and np.array_equal(lb, ob)
        if ordering == 2:
            # a round is an independent set, so the depth never exceeds the rounds
            assert 1 <= f.info()["depth_fwd"] <= rounds and f.info()["depth_bwd"] <= rounds


def test_ordered_factor_preconditions_the_permuted_system():
    # the exported factor (position space) preconditions P L P^T: the oracle
    # PCG on the permuted system converges like the natural one
    import oracle
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid3d(12)
    w = graphs.loguniform_weights(len(u), seed=3)
    cp, ri, va = oracle.laplacian(n, u, v, w)
    Y = np.random.default_rng(0).standard_normal((n, 3))
    fn = ss.Factor(cp, ri, va, seed=1, ordering=ss.ORDER_NATURAL)
    Zn, itn, _ = oracle.solve_lap(Y, cp, ri, va, *fn.export(), stop_rule=1, tol=1e-10)
    f = ss.Factor(cp, ri, va, seed=1, ordering=2)
    order, _ = f.order()
    rank = np.empty(n, np.int64)
    rank[order] = np.arange(n)
    pu, pv, pw = graphs.canonical(rank[u], rank[v], w)
    pcp, pri, pva = oracle.laplacian(n, pu, pv, pw)
    Zp, itp, rr = oracle.solve_lap(Y[order], pcp, pri, pva, *f.export(), stop_rule=1, tol=1e-10)
    Z = np.empty_like(Zp)
    Z[order] = Zp
    a, b = Z - Z.mean(0), Zn - Zn.mean(0)
    assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-7
    assert itp.max() <= itn.max() + 5
    assert f.info()["depth_fwd"] * 20 < fn.info()["depth_fwd"]


def test_host_laplacian_matches_oracle():
    """graphs.laplacian_csc (the Laplacian batches factor) is the oracle's
    CSC Laplacian: same structure, same values (diagonal sums may differ in
    the last bit only through summation order)."""
    import oracle
    from spec_spars_amd import graphs
    oracle.load()
    for n, u, v in (graphs.grid2d(9, 7), graphs.rgg(300, seed=1), graphs.chung_lu(400, 1500, seed=3)):
        w = graphs.loguniform_weights(len(u), seed=2)
        cp, ri, va = graphs.laplacian_csc(n, u, v, w)
        # isolated vertices keep an explicit zero diagonal here (as on the
        # device); the oracle drops explicit zeros
        keep = va != 0.0
        col = np.repeat(np.arange(n), np.diff(cp))[keep]
        ri, va = ri[keep], va[keep]
        cp = np.concatenate([[0], np.cumsum(np.bincount(col, minlength=n))]).astype(np.int32)
        oc, orr, ov = oracle.laplacian(n, u, v, w)
        assert np.array_equal(cp, oc) and np.array_equal(ri, orr)
        assert np.allclose(va, ov, rtol=1e-14, atol=0)
        off = ri != np.repeat(np.arange(n), np.diff(cp))
        assert np.array_equal(va[off], ov[off])


def test_default_options_pick_the_shallow_order():
    """ss_solve_opts_default (what a run_solve_lap caller gets) selects the
    rounds order: on a 64^3 grid its solves are < 1,000 levels deep, where
    the natural order's are > 100,000."""
    import ctypes
    import spec_spars_amd as ss
    from spec_spars_amd import _lib, graphs
    o = _lib.SolveOpts()
    _lib.load().ss_solve_opts_default(ctypes.byref(o))
    assert o.ordering == ss.ORDER_ROUNDS
    assert (o.stop_rule, o.max_iter, o.tol, o.fail_relres) == (ss.STOP_REF_ABS, 400, 2e-7, 1e-2)
    assert ss.solve_opts().ordering == ss.ORDER_ROUNDS
    n, u, v = graphs.grid3d(64)
    w = graphs.loguniform_weights(len(u), seed=0)
    f = ss.Factor(*graphs.laplacian_csc(n, u, v, w), seed=0, ordering=o.ordering)
    info = f.info()
    assert info["depth_fwd"] < 1000 and info["depth_bwd"] < 1000


def test_mixed_mode_for_weight_spread():
    """FP16 preconditioner rows for weights spanning at most 2^10 (the
    Chung-Lu U[0.5, 2] and unit-weight families), FP32 rows beyond (the
    grids' 10^U(-3, 3))."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    assert ss.mixed_mode_for(graphs.uniform_weights(1000, seed=0)) == ss.MIXED_RELIABLE_F16
    assert ss.mixed_mode_for(np.ones(10)) == ss.MIXED_RELIABLE_F16
    assert ss.mixed_mode_for(graphs.loguniform_weights(1000, seed=0)) == ss.MIXED_RELIABLE
    assert ss.mixed_mode_for(np.array([1.0, 1024.0])) == ss.MIXED_RELIABLE_F16
    assert ss.mixed_mode_for(np.array([1.0, 1025.0])) == ss.MIXED_RELIABLE
