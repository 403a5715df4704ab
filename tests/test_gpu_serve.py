"""Serving on the GPU: requests that overlap on the device give the same
results as the same requests run one at a time."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pool():
    from spec_spars_amd import graphs
    out = []
    n, u, v = graphs.grid2d(20, 15)
    out.append((n, u, v, graphs.loguniform_weights(len(u), seed=1)))
    n, u, v = graphs.rgg(800, seed=2)
    out.append((n, u, v, np.ones(len(u))))
    n, u, v = graphs.chung_lu(1000, 4000, seed=3)
    out.append((n, u, v, graphs.uniform_weights(len(u), seed=3)))
    n, u, v = graphs.grid3d(8)
    out.append((n, u, v, graphs.loguniform_weights(len(u), seed=4)))
    n, u, v = graphs.rgg(2000, seed=5)
    out.append((n, u, v, np.ones(len(u))))
    return out


@pytest.mark.parametrize("prec", [0, 1])
def test_serving_results_equal_alone_and_mixed(prec):
    from spec_spars_amd.serve import Registry, Server, poisson_requests
    reg = Registry(_pool())
    try:
        reqs = poisson_requests(rate=400, duration=0.1, graphs=5, seed=2, k=16, rhs=8)
        assert len(reqs) > 20
        mixed = [dataclasses.replace(r) for r in reqs]
        out = Server(reg, slots=4, precision=prec, reserve_bytes=0).run(mixed)
        assert out["completed"] == len(reqs) and out["failed"] == 0 and out["dropped"] == 0
        assert out["p99_ms"] >= out["p50_ms"] > 0
        for r, m in zip(reqs, mixed):
            alone = reg.execute(r, prec)
            res = m.result
            assert res["relres"] <= 1e-8
            if r.op == "sparsify":
                assert res["kept"] == alone["kept"]
                for a, b in zip(res["arrays"], alone["arrays"]):
                    assert np.array_equal(a, b), r
            else:
                assert np.array_equal(res["Z"], alone["Z"]), r
    finally:
        reg.close()


def test_serving_solve_matches_direct_solve(oracle_mod):
    """A served solve request is the resident graph's PCG on the request's
    right-hand sides: residual at the tolerance, same as Graph.solve."""
    import scipy.sparse as sp
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    from spec_spars_amd.serve import Registry, Request
    pool = _pool()[:2]
    reg = Registry(pool)
    try:
        r = Request(0, 1, "solve", 0.0, k=8, seed=7)
        res = reg.execute(r, ss.PREC_FP64)
        n, u, v, w = pool[1]
        cp, ri, va = graphs.laplacian_csc(n, u, v, w)
        L = sp.csc_matrix((va, ri, cp), shape=(n, n))
        b = reg.rhs(1, 8, 7)
        import scipy.sparse.csgraph as cg
        nc, lab = cg.connected_components(sp.coo_matrix((np.ones(len(u)), (u, v)), shape=(n, n)), directed=False)
        for c in range(nc):
            b[lab == c] -= b[lab == c].mean(0)
        rr = np.linalg.norm(b - L @ res["Z"], axis=0) / np.linalg.norm(b, axis=0)
        assert rr.max() <= 1.01e-8
    finally:
        reg.close()
