"""The reference's own tests (src/tests.rs), restated against the C++ host
layer (Sparsifier / SparsifierParameters) with the GPU path underneath."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import components, golden


def _params(**kw):
    import spec_spars_amd as ss
    return ss.SparsifierParameters(**kw)


def _graph_edges(L_cp, L_ri, n):
    A = sp.csc_matrix((np.ones(len(L_ri)), L_ri, L_cp), shape=(n, n)).tocoo()
    keep = A.row < A.col
    return A.row[keep], A.col[keep]


def test_lap_valid_random():
    # lap_valid_random (src/tests.rs:162-189): random updates keep a valid
    # Laplacian (host only, no GPU needed)
    import spec_spars_amd as ss
    for seed in range(3):
        rng = np.random.default_rng(seed)
        s = ss.Sparsifier(2000, _params(sketch_seed=seed, sampling_seed=seed))
        for i in range(0, 4000, 500):
            a = rng.integers(0, 2000, 500)
            b = rng.integers(0, 2000, 500)
            w = rng.uniform(-1.0, 1.0, 500)
            s.insert_many(a, b, w)
            s.form_laplacian(True)
            s.check_diagonal()


def test_insert_rejects_bad_vertex():
    import spec_spars_amd as ss
    s = ss.Sparsifier(10, _params())
    with pytest.raises(ss.SSError) as e:
        s.insert(3, 10, 1.0)
    assert e.value.code == 1


def test_parameters_match_reference():
    import spec_spars_amd as ss
    p = ss.SparsifierParameters.new_default()
    assert (p.epsilon, p.beta_constant, p.row_constant, p.jl_factor) == (0.5, 4, 2, 1.5)
    assert p.jl_scaling_factor == math.sqrt(3.0) and p.sketch_uniform == 1 and p.streaming == 0
    s = ss.Sparsifier(4096, _params(epsilon=0.3, jl_factor=2.0))
    assert s.jl_dim == 24                                          # ceil(12 * 2)
    assert s.beta == math.pow(0.3, -2.0) * 4.0 * math.log(4096.0)   # src/sparsifier.rs:264
    assert s.threshold == 4096 * round(s.beta) * 2


@pytest.mark.gpu
def test_sampling_verify(ctx):
    # sampling_verify (src/tests.rs:213-261) on a synthetic graph
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(120, 100)
    w = graphs.uniform_weights(len(u), seed=4)
    for seed in range(3):
        s = ss.Sparsifier(n, _params(sketch_seed=seed, sampling_seed=seed), ctx)
        s.insert_many(u, v, w)
        before_diag = 2.0 * w.sum()
        s.form_laplacian(True)
        m = s.num_edges()
        before, after = s.sample_and_apply(np.full(m, 0.5))
        assert before == len(u)
        cp, ri, va = s.current_laplacian()
        A = sp.csc_matrix((va, ri, cp), shape=(n, n))
        after_diag = A.diagonal().sum()
        assert abs(before / after - 2.0) < 0.05
        assert abs(before_diag / after_diag - math.sqrt(2.0)) < 0.05
        s.check_diagonal()


@pytest.mark.gpu
def test_sampling_test(ctx):
    # sampling_test (src/tests.rs:265-317): 40 components after deleting by index
    import spec_spars_amd as ss
    g = golden("sampling_test")
    s = ss.Sparsifier(100, _params(jl_factor=4.0), ctx)
    s.insert_many(g["v1"], g["v2"], np.ones(len(g["v1"])))
    s.form_laplacian(True)
    probs = np.ones(s.num_edges())
    for i in range(10):
        for j in range(5):
            idx = i * 11 + 2 + 2 * j
            if idx == s.num_edges():
                break
            probs[idx] = 0.0
    s.sample_and_apply(probs)
    cp, ri, va = s.current_laplacian()
    a, b = _graph_edges(cp, ri, 100)
    assert components(100, a, b) == 40


@pytest.mark.gpu
def test_jl_sketch_zero(ctx):
    # jl_sketch_zero (src/tests.rs:479-511): sketch columns sum to ~0
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.rgg(3000, seed=2)
    s = ss.Sparsifier(n, _params(), ctx)
    s.insert_many(u, v, graphs.uniform_weights(len(u), seed=1))
    Y = s.jl_sketch()
    assert Y.shape == (n, s.jl_dim)
    assert np.all(np.abs(Y.sum(0)) < 1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("ordering", [0, 2])
def test_check_for_additions_and_valid_output(ctx, ordering):
    # check_for_additions (src/tests.rs:636-667) + run_stream's check path:
    # the sparsifier's edges are a subset of the input's, the Laplacian stays
    # valid, and edges are removed at eps = 0.5.
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(48, 40)
    w = graphs.uniform_weights(len(u), seed=2)
    s = ss.Sparsifier(n, _params(epsilon=0.5, jl_factor=2.0, sketch_uniform=0, ordering=ordering,
                                 stop_rule=1, tol=1e-8, max_iter=2000), ctx)
    s.insert_many(u, v, w)
    before, after = s.sparsify(True)
    assert before == len(u) and 0 < after <= before
    s.check_diagonal()
    cp, ri, va = s.current_laplacian()
    a, b = _graph_edges(cp, ri, n)
    ge = set(zip(u.tolist(), v.tolist()))
    assert all((x, y) in ge for x, y in zip(a.tolist(), b.tolist()))


@pytest.mark.gpu
def test_input_stream_runs_like_inserting_by_hand(ctx, tmp_path):
    """InputStream::run_stream (src/stream.rs:41-64) on a MatrixMarket file
    written here -- a weighted grid stored as its lower triangle with
    diagonal entries and some weights split into two duplicate entries --
    inserts the same edges in the same order as inserting the graph's
    canonical edges by hand, so the sparsified Laplacians and the stats are
    identical; writeout writes that Laplacian, which reads back exactly."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(20, 15)
    w = graphs.loguniform_weights(len(u), seed=9, lo=0.1, hi=10.0)
    rng = np.random.default_rng(4)
    split = rng.random(len(u)) < 0.3
    lines = []
    for a, b, x, s in zip(u, v, w, split):
        if s:
            lines += ["%d %d %.17g" % (b + 1, a + 1, x / 2.0)] * 2       # halves sum exactly
        else:
            lines.append("%d %d %.17g" % (b + 1, a + 1, x))
    lines += ["%d %d 5.0" % (i + 1, i + 1) for i in range(0, n, 7)]
    path = tmp_path / "grid.mtx"
    path.write_text("%%%%MatrixMarket matrix coordinate real symmetric\n%% written by the test\n%d %d %d\n%s\n" %
                    (n, n, len(lines), "\n".join(lines)))
    params = _params(epsilon=4.0, jl_factor=2.0, sketch_uniform=0, sketch_seed=3, sampling_seed=4)
    stream = ss.InputStream(path, "grid")
    assert stream.num_nodes == n
    out = tmp_path / "h.mtx"
    s1, st1 = stream.run_stream(params, ctx=ctx, writeout=True, output_file=str(out))
    s2 = ss.Sparsifier(n, params, ctx)
    s2.insert_many(u, v, w)
    st2 = s2.sparsify()
    assert st1 == st2 and st1[1] < st1[0]
    L1, L2 = s1.current_laplacian(), s2.current_laplacian()
    assert all(np.array_equal(a, b) for a, b in zip(L1, L2))
    rows, cols, cp, ri, va = ss.read_mtx(out)
    assert rows == cols == n
    assert np.array_equal(cp, L1[0]) and np.array_equal(ri, L1[1]) and np.array_equal(va, L1[2])
    s1.check_diagonal()


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_streaming_trigger_resident_equals_host(ctx, precision):
    """Streaming (src/sparsifier.rs:357-376): once the formed Laplacian holds
    more than threshold = n * round(beta) * row_constant edges, every insert
    runs a pass.  Edges go in in batches, the trigger fires, and the
    device-resident pass (solution, probabilities, coins, sample and the
    reweighted Laplacian computed on the device, only the kept edges and the
    new diagonal read back) leaves the same Laplacian, bit for bit, as the
    host-driven pass after every batch."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(40, 40)
    w = graphs.loguniform_weights(len(u), seed=6, lo=0.1, hi=10.0)
    m = len(u)
    base = dict(epsilon=5.0, beta_constant=4, row_constant=1, sketch_uniform=0, sketch_seed=2, sampling_seed=3,
                streaming=1, precision=precision, stop_rule=ss.STOP_REL, tol=1e-10, max_iter=2000)
    runs = []
    for resident in (0, 1):
        s = ss.Sparsifier(n, _params(resident=resident, **base), ctx)
        assert s.threshold == n * 1 * 1 and s.threshold < m
        snaps = []
        s.insert_many(u[:2000], v[:2000], w[:2000])          # nothing formed yet: no trigger
        assert s.triggered() == 0
        s.form_laplacian()
        snaps.append(s.current_laplacian())
        for a in range(2000, m, 280):                       # the formed graph is above the threshold
            s.insert_many(u[a:a + 280], v[a:a + 280], w[a:a + 280])
            snaps.append(s.current_laplacian())
        stats = s.sparsify()
        snaps.append(s.current_laplacian())
        s.check_diagonal()
        runs.append((s.triggered(), stats, snaps))
        s.close()
    (t0, st0, sn0), (t1, st1, sn1) = runs
    assert t0 == t1 and t0 >= 1
    assert st0 == st1
    for a, b in zip(sn0, sn1):
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
