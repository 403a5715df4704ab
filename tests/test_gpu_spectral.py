"""eps-spectral checks of the sparsifiers of the 64x64 grid (k = 24) and
the 128^3 grid with log-uniform weights (k = 48) at eps = 0.3: 64
generalized Rayleigh quotients and the Lanczos extremes of L_G^+ L_H
(device PCG on G) and of L_H^+ L_G (device PCG on H), for the reference's
kept weights w / sqrt(p) and for the textbook w / p.  Pass or fail against
eps is printed and written to gpurun_out/spectral_<case>.json; the asserts
hold what is true either way (the ordering of the two weightings, solve
accuracy, agreement of the two pencils)."""
import json
import math
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


def _run(ctx, n, u, v, w, k, ordering, lanczos_iters):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs, spectral
    G = ss.Graph(ctx, n, u, v, w)
    G.set_factor(ss.Factor(*G.laplacian(), seed=0, ordering=ordering))
    rep = G.sparsify(k, graphs.beta_for(n, 0.3), scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0,
                     solve=ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3000))
    ku, kv, kw = G.read_kept()
    p = G.read_probs()
    G.close()
    res = spectral.check(ctx, n, u, v, w, ku, kv, kw, p, 0.3, lanczos_iters=lanczos_iters)
    res["edges_with_p_below_1"] = int((p < 1).sum())
    res["kept_check"] = rep["kept"]
    return res


def _record(name, res):
    out = os.path.join(ROOT, "gpurun_out")
    os.makedirs(out, exist_ok=True)
    with open(os.path.join(out, "spectral_%s.json" % name), "w") as f:
        json.dump(res, f, indent=1)
    print("[spectral]", name, json.dumps(res), flush=True)


def _invariants(res):
    a, b = res["w_over_sqrt_p"], res["w_over_p"]
    # w / sqrt(p) <= w / p edge by edge: the same probes see the reference's
    # H at or below the textbook one, and so do the pencils' extremes
    assert a["rayleigh_max"] <= b["rayleigh_max"] + 1e-12 and a["rayleigh_min"] <= b["rayleigh_min"] + 1e-12
    assert a["components"] == b["components"]
    for r in (a, b):
        assert r["mu_solve_relres"] <= 1e-9
        assert 0 <= r["mu_min"] <= r["mu_max"]
        # Ritz values lie inside [mu_min, mu_max]; so do the quotients
        assert r["mu_ritz_min"] <= r["rayleigh_min"] * (1 + 1e-6) or r["components"] > 1
        if r["components"] == 1:
            assert r["lambda_solve_relres"] <= 1e-9
            # lambda = 1 / mu: the two pencils (solves on H, solves on G)
            # agree on both extremes once Lanczos has converged there
            assert abs(r["lambda_max"] * r["mu_min"] - 1) < 1e-2
            assert abs(r["lambda_min"] * r["mu_max"] - 1) < 1e-2


def test_spectral_c1_grid64(ctx):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid2d(64, 64)
    res = _run(ctx, n, u, v, np.ones(len(u)), 24, ss.ORDER_NATURAL, 60)
    _record("c1_grid64", res)
    _invariants(res)


@pytest.mark.timeout(900)
def test_spectral_c2_grid128(ctx):
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v = graphs.grid3d(128)
    w = graphs.loguniform_weights(len(u), seed=0)
    res = _run(ctx, n, u, v, w, 48, ss.ORDER_ROUNDS, 30)
    _record("c2_grid128", res)
    _invariants(res)
