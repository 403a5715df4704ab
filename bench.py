#!/usr/bin/env python3
"""Benchmark of the sparsification hot path on one MI355X.

Workload (default): the 3D 128^3 grid (n = 2,097,152, m = 6,242,304), weights
10^U(-3,3) in canonical edge order, k = 48 Rademacher sketch columns scaled
1/sqrt(k), FP64 SpMM / triangular solves, PCG to a 1e-8 relative residual,
eps = 0.3 sampling, sketch and sampling seed 0.  A step is one pass of the
hot path over the resident graph: sketch -> multi-RHS PCG -> fused
ER/coin/sample/compaction (the host factor is built and uploaded once, before
timing).  value = input edges / step time.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build runs on
one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sparsified edges/s at 1 GPU vs plain at ε=0.3; p50/p99 request latency"
PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(name):
    from spec_spars_amd import graphs
    if name == "grid3d128":
        n, u, v = graphs.grid3d(128)
        w = graphs.loguniform_weights(len(u), seed=0)
        return dict(name=name, desc="3D grid 128^3, weights 10^U(-3,3), k=48, FP64, PCG rel 1e-8, eps=0.3",
                    n=n, u=u, v=v, w=w, k=48, eps=0.3)
    if name == "grid3d64":
        n, u, v = graphs.grid3d(64)
        w = graphs.loguniform_weights(len(u), seed=0)
        return dict(name=name, desc="3D grid 64^3, weights 10^U(-3,3), k=48, FP64, PCG rel 1e-8, eps=0.3",
                    n=n, u=u, v=v, w=w, k=48, eps=0.3)
    if name == "chunglu5m":
        n, u, v = graphs.chung_lu(5_000_000, 50_000_000, gamma=2.2, seed=0)
        w = graphs.uniform_weights(len(u), seed=0)
        return dict(name=name, desc="Chung-Lu n=5e6 m=5e7 gamma=2.2, weights U[0.5,2], k=64, mixed precision, "
                    "PCG rel 1e-8, eps=0.3", n=n, u=u, v=v, w=w, k=64, eps=0.3, precision="mixed")
    if name == "rgg512":
        import numpy as np
        rng = np.random.default_rng(512)
        sizes = np.round(10.0 ** rng.uniform(3.0, math.log10(2e5), 512)).astype(np.int64)
        gs = []
        for i, ni in enumerate(sizes):
            _, u, v = graphs.rgg(int(ni), seed=1000 + i, mean_degree=8.0)
            gs.append((int(ni), u, v, np.ones(len(u))))
        return dict(name=name, desc="512 random geometric graphs, n log-uniform [1e3, 2e5], mean degree 8, "
                    "unit weights, k=32 each, mixed precision, PCG rel 1e-8, eps=0.3", batch=gs, k=32, eps=0.3,
                    precision="mixed", n=int(sizes.sum()), m=int(sum(len(g[1]) for g in gs)))
    if name == "grid2d64":
        n, u, v = graphs.grid2d(64, 64)
        import numpy as np
        return dict(name=name, desc="2D grid 64x64, unit weights, k=24, FP64, PCG rel 1e-8, eps=0.3",
                    n=n, u=u, v=v, w=np.ones(len(u)), k=24, eps=0.3)
    raise SystemExit("unknown workload " + name)


def kernel_bytes(n, m, G, k, kept, s=8):
    """Algorithmic (compulsory) HBM bytes per launch of each kernel kind:
    every array counted once per pass, gathers beyond first touch not
    counted.  n, m: total vertices / edges of the (batched) system."""
    nnzL = n + 2 * m
    fi, bi = G.schedule_info(0), G.schedule_info(1)

    def per_level(info, extra_per_row):
        # schedule row id + row pointer, (col, val) per entry, read in-row, write out-row
        if not info["level_launches"]:
            return None
        b = info["level_rows"] * (4 + 4 + 2 * k * s + extra_per_row) + info["level_nnz"] * (4 + s)
        return b / info["level_launches"]

    def small(info, extra_per_row, passes, extra_launches):
        # all runs of one application over their launches (runs, plus the
        # forward trailing run's head-segment launch)
        if not info["small_levels"]:
            return None
        total = info["small_rows"] * (4 + 4 + passes * 2 * k * s + extra_per_row) + info["small_nnz"] * (4 + s)
        return total / (info["small_runs"] + extra_launches)

    def core(info):
        # dense cores, per launch: rows' outside entries (forward), the
        # lower triangles once, every core row read and written once
        if not info["core_rows"]:
            return None
        total = info["core_outside_nnz"] * (4 + s) + info["core_rows"] * (4 + 2 * k * s) + \
            info["core_entries"] // 2 * s
        return total / (1 + info["core_steps"])

    return {
        # sketch: incidence lists (edge id + signed sqrt(w)) and Y
        "sketch": 4 * (n + 1) + 2 * m * (4 + 8) + n * k * s,
        # SpMM: rowptr, (col, val) per nonzero, read P, write Q
        "spmm": 4 * (n + 1) + nnzL * (4 + s) + 2 * n * k * s,
        "tri_fwd": per_level(fi, 0),
        "tri_bwd": per_level(bi, 8),
        "tri_fwd_small": small(fi, 0, 1, 0),
        "tri_bwd_small": small(bi, 8, 1, 0),
        "core_fwd": core(fi),
        "core_bwd": core(bi),
        # ER/sample: edges (u, v, w), Z once, p out, mask out, kept (u, v, w) out
        "er_sample": m * (4 + 4 + 8) + n * k * s + 8 * m + m + kept * 16,
        "vector": None,
    }


def cpu_baseline(W, ordering_seed=0, columns=1):
    """The oracle (CPU restatement of the reference path), one thread, on a
    bounded sample: full sketch, `columns` of the k PCG columns (scaled to k),
    full ER / probabilities / coins / sampling, natural-order factor."""
    import numpy as np
    import oracle
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v, w, k = W["n"], W["u"], W["v"], W["w"], W["k"]
    m = len(u)
    cp, ri, va = oracle.laplacian(n, u, v, w)
    f = ss.Factor(cp, ri, va, seed=ordering_seed, ordering=ss.ORDER_NATURAL)
    urp, uci, uva, dg = f.export()
    t0 = time.perf_counter()
    Y = oracle.sketch(u, v, w, n, k, seed=0, scale=1.0 / math.sqrt(k))
    t_sketch = time.perf_counter() - t0
    t0 = time.perf_counter()
    Zc, it, rr = oracle.solve_lap(Y[:, :columns], cp, ri, va, urp, uci, uva, dg, stop_rule=1, tol=1e-8,
                                  max_iter=1000)
    t_col = (time.perf_counter() - t0) / columns
    # ER over all edges needs a full Z: use the sampled columns tiled to k
    Z = np.tile(Zc, (1, (k + columns - 1) // columns))[:, :k]
    uu, vv, lv = oracle.upper_edges(n, cp, ri, va)
    t0 = time.perf_counter()
    d = oracle.diff_norms(uu, vv, Z)
    p = oracle.probs(lv, d, graphs.beta_for(n, W["eps"]), k)
    c = oracle.coins(0, m)
    r = oracle.sample_reweight(n, uu, vv, lv, p, c)
    t_er = time.perf_counter() - t0
    total = t_sketch + k * t_col + t_er
    return {
        "value": m / total, "unit": "edges/s", "cores": 1, "kind": "port",
        "sample": ("oracle on the full workload graph: full sketch (%.1fs), %d of %d PCG columns "
                   "(%.1fs each, %d iterations, natural-order factor, rel 1e-8) scaled to %d, full "
                   "ER/probabilities/coins/sampling (%.1fs); one thread" %
                   (t_sketch, columns, k, t_col, int(it.max()), k, t_er)),
        "seconds_measured": t_sketch + columns * t_col + t_er,
        "kept_sample": int(r["kept"]),
    }


def summarize(prof, kb, t_timed):
    """Per-kernel averages and the roofline of the dominant kernel kind."""
    kernels = {}
    for name, st in prof.items():
        if st["launches"]:
            avg = st["ms_total"] / st["launches"]
            kernels[name] = {"launches": st["launches"], "avg_ms": avg, "bytes_per_launch": kb[name],
                             "GBps": (kb[name] / (avg * 1e-3) / 1e9) if kb[name] else None,
                             "share": st["ms_total"] / (t_timed * 1e3)}
    roofline = None
    ranked = [x for x in kernels if kernels[x]["bytes_per_launch"]]
    if ranked:
        dom = max(ranked, key=lambda x: kernels[x]["avg_ms"] * kernels[x]["launches"])
        kd = kernels[dom]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(kd["GBps"], 1), "peak": PEAK_HBM_GBS,
                    "unit": "GB/s", "frac": round(kd["GBps"] / PEAK_HBM_GBS, 4), "traffic": None,
                    "bytes_per_launch": kd["bytes_per_launch"], "avg_launch_ms": kd["avg_ms"]}
    return kernels, roofline


def cpu_baseline_batch(W, count=4):
    """The oracle on the first `count` graphs of the batch (every column,
    natural-order factors, one thread): edges/s of that bounded sample."""
    import numpy as np
    import oracle
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    k = W["k"]
    t_total, m_total, iters = 0.0, 0, 0
    for n, u, v, w in W["batch"][:count]:
        cp, ri, va = oracle.laplacian(n, u, v, w)
        f = ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_NATURAL)
        urp, uci, uva, dg = f.export()
        t0 = time.perf_counter()
        Y = oracle.sketch(u, v, w, n, k, seed=0, scale=1.0 / math.sqrt(k))
        Z, it, rr = oracle.solve_lap(Y, cp, ri, va, urp, uci, uva, dg, stop_rule=1, tol=1e-8, max_iter=2000)
        uu, vv, lv = oracle.upper_edges(n, cp, ri, va)
        p = oracle.probs(lv, oracle.diff_norms(uu, vv, Z), graphs.beta_for(n, W["eps"]), k)
        oracle.sample_reweight(n, uu, vv, lv, p, oracle.coins(0, len(uu)))
        t_total += time.perf_counter() - t0
        m_total += len(u)
        iters = max(iters, int(it.max()))
    return {"value": m_total / t_total, "unit": "edges/s", "cores": 1, "kind": "port",
            "sample": "oracle on the first %d graphs of the batch (%d edges, all %d columns, natural-order "
                      "factors, up to %d PCG iterations), one thread, %.1fs" % (count, m_total, k, iters, t_total),
            "seconds_measured": t_total}


def bench_batch(W, args):
    """Many graphs as one resident batch: a step sparsifies all of them."""
    from concurrent.futures import ThreadPoolExecutor
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    gs, k = W["batch"], W["k"]
    n, m = W["n"], W["m"]
    prec = ss.PREC_FP64 if args.precision == "fp64" else ss.PREC_MIXED
    ordering = {"natural": ss.ORDER_NATURAL, "random": ss.ORDER_RANDOM, "rounds": ss.ORDER_ROUNDS}[args.ordering]
    log("workload %s: %d graphs, n=%d m=%d k=%d" % (W["name"], len(gs), n, m, k))
    ctx = ss.Context(0)
    t0 = time.perf_counter()
    B = ss.Batch(ctx, gs)
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()

    def build(g):
        return ss.Factor(*graphs.laplacian_csc(*g), seed=0, ordering=ordering)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        fs = list(ex.map(build, gs))
    t_factor = time.perf_counter() - t0
    fnnz = sum(f.info()["nnz"] for f in fs)
    t0 = time.perf_counter()
    B.set_factors(fs)
    t_upload = time.perf_counter() - t0
    log("batch %.1fs factors %.1fs (nnz %d) upload %.1fs" % (t_graph, t_factor, fnnz, t_upload))
    tune = ctx.tuning()
    if not args.no_tune:
        t0 = time.perf_counter()
        tune = ctx.autotune(B, k, prec, reps=3)
        log("autotune %.1fs: %s" % (time.perf_counter() - t0, tune))
    betas = [graphs.beta_for(g[0], W["eps"]) for g in gs]
    solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=prec)

    def step():
        return B.sparsify(k, betas, sketch_seeds=[0] * len(gs), sampling_seeds=[0] * len(gs), solve=solve)

    for i in range(args.warmup):
        reps = step()
        log("warmup %d: %.1f ms (sketch %.1f solve %.1f sample %.1f) iters %d kept %d" %
            (i, reps[0]["ms_total"], reps[0]["ms_sketch"], reps[0]["ms_solve"], reps[0]["ms_sample"],
             max(r["max_iters"] for r in reps), sum(r["kept"] for r in reps)))
    t0 = time.perf_counter()
    for i in range(args.steps):
        reps = step()
        log("step %d: %.1f ms" % (i, reps[0]["ms_total"]))
    t_timed = time.perf_counter() - t0
    prof, t_prof = {}, 1.0
    if not args.no_profile:
        B.profile(True)
        t0 = time.perf_counter()
        step()
        t_prof = time.perf_counter() - t0
        prof = B.profile_read()
        B.profile(False)
    ms_per_step = 1000.0 * t_timed / args.steps
    kept = sum(r["kept"] for r in reps)
    kb = kernel_bytes(n, m, B, k, kept, s=8 if prec == ss.PREC_FP64 else 4)
    kernels, roofline = summarize(prof, kb, t_prof)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline_batch(W)
        log("cpu baseline %.3g edges/s (%s)" % (cpu["value"], cpu["sample"]))
    r0 = reps[0]
    out = {
        "metric": METRIC, "value": m / (ms_per_step / 1000.0), "unit": "edges/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == ss.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": W["desc"], "graphs": len(gs), "n": n, "m": m, "k": k, "rhs": k * len(gs),
                   "epsilon": W["eps"], "precision": args.precision, "ordering": args.ordering,
                   "factor_nnz": fnnz, "pcg_iters": max(r["max_iters"] for r in reps),
                   "refinements": r0["refinements"], "max_relres": max(r["max_relres"] for r in reps),
                   "kept_edges": kept, "ms_sketch": r0["ms_sketch"], "ms_solve": r0["ms_solve"],
                   "ms_sample": r0["ms_sample"], "rhs_solves_per_s": k * len(gs) / (r0["ms_solve"] / 1000.0),
                   "host_factor_s": t_factor, "host_graph_s": t_graph, "upload_s": t_upload,
                   "parallelism": "1 GPU", "tuning": tune,
                   "schedule_fwd": B.schedule_info(0), "schedule_bwd": B.schedule_info(1)},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    B.close()
    ctx.close()
    return 0


def serve_pool(count, seed=5):
    """Resident graphs for the serving run: the three families mixed, n
    log-uniform in [1e4, 2e6], m <= 2e7."""
    import numpy as np
    from spec_spars_amd import graphs
    rng = np.random.default_rng(seed)
    pool, desc = [], []
    for i in range(count):
        n = int(round(10.0 ** rng.uniform(4.0, math.log10(2e6))))
        fam = ("grid3d", "chunglu", "rgg")[i % 3]
        if fam == "grid3d":
            side = max(3, int(round(n ** (1.0 / 3.0))))
            nn, u, v = graphs.grid3d(side)
            w = graphs.loguniform_weights(len(u), seed=100 + i)
        elif fam == "chunglu":
            nn, u, v = graphs.chung_lu(n, min(8 * n, 20_000_000), gamma=2.2, seed=100 + i)
            w = graphs.uniform_weights(len(u), seed=100 + i)
        else:
            nn, u, v = graphs.rgg(n, seed=100 + i, mean_degree=8.0)
            w = np.ones(len(u))
        pool.append((nn, u, v, w))
        desc.append("%s n=%d m=%d" % (fam, nn, len(u)))
    return pool, desc


def bench_serve(args):
    """Poisson arrivals of sparsify / 16-RHS solve requests over resident
    graphs, continuous batching by slots, admission by free device memory."""
    import spec_spars_amd as ss
    from spec_spars_amd.serve import Registry, Server, poisson_requests
    prec = ss.PREC_FP64 if args.precision == "fp64" else ss.PREC_MIXED
    t0 = time.perf_counter()
    pool, desc = serve_pool(args.pool)
    log("pool: " + "; ".join(desc))
    reg = Registry(pool, threads=min(16, os.cpu_count() or 1))
    t_build = time.perf_counter() - t0
    log("registry built in %.1fs (%.1f GB resident)" %
        (t_build, sum(G.resident_bytes() for G in reg.G) / 1e9))
    reqs = poisson_requests(rate=args.rate, duration=args.duration, graphs=len(pool), seed=11, k=32, rhs=16)
    log("%d requests over %.0fs" % (len(reqs), args.duration))
    srv = Server(reg, slots=args.slots, precision=prec, reserve_bytes=4 << 30, keep_arrays=True)
    out = srv.run(reqs, deadline=args.duration + 30.0)
    log("served: %s" % out)
    line = {
        "metric": METRIC, "value": out["sparsified_edges_per_s"], "unit": "edges/s", "n_gpus": 1,
        "steps": len(reqs), "warmup": 0, "ms_per_step": out["mean_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == ss.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson %.0f req/s for %.0fs over %d resident graphs (grid3d, Chung-Lu, "
                               "RGG; n log-uniform [1e4, 2e6], m <= 2e7); half sparsify (k=32, eps in "
                               "{0.2, 0.3, 0.5}), half 16-RHS solves; %d slots" %
                               (args.rate, args.duration, len(pool), args.slots),
                   "pool": desc, "precision": args.precision, "registry_build_s": t_build, **out},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    reg.close()
    return 0


def probe(args):
    """A short fixed program for hardware-counter runs: the workload's
    graph and factor, one sparsify pass capped at 3 PCG iterations with the
    shapes the autotuner picks on the 128^3 grid (16 waves per row tile, two
    tiny rows per wave), then one 1 GiB FP64 copy for calibration."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    W = workload(args.workload)
    n, u, v, w, k = W["n"], W["u"], W["v"], W["w"], W["k"]
    ctx = ss.Context(0)
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    G.set_factor(ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_ROUNDS))
    ctx.set_tuning(tile_waves=16, level_rows_per_wave=args.probe_rpw, graph_iters=0,
                   xcd_slices=0 if args.no_xcd else 1)
    prec = ss.PREC_FP64 if (args.precision or W.get("precision", "fp64")) == "fp64" else ss.PREC_MIXED
    solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3, precision=prec)
    solve.fail_relres = 1e300
    rep = G.sparsify(k, graphs.beta_for(n, W["eps"]), scale=1.0 / math.sqrt(k), solve=solve)
    log("probe pass %.1f ms" % rep["ms_total"])
    log("copy %.3f ms per GiB copy" % ctx.bench_copy(1 << 30, reps=1))
    G.close()
    ctx.close()
    return 0


def measured_traffic(workload, precision, kind):
    """HBM bytes per launch of a kernel kind from the committed counter
    summary (profiles/), when one exists for this workload."""
    path = os.path.join(ROOT, "profiles", "trisolve_%s_%s_pmc.md" % (workload, precision))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch %s:" % kind):
            return int(line.split(":")[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="grid3d128")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="default: the workload's (fp64 for the grids, mixed for chunglu5m and rgg512)")
    ap.add_argument("--ordering", choices=["natural", "random", "rounds"], default="rounds")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuner)")
    ap.add_argument("--probe", action="store_true",
                    help="one short sparsify pass (3 PCG iterations) and a 1 GiB copy, for counter runs")
    ap.add_argument("--no-xcd", action="store_true", help="probe: round-robin rows over the XCDs")
    ap.add_argument("--probe-rpw", type=int, default=2, help="probe: tiny rows per wave of the level launches")
    ap.add_argument("--rate", type=float, default=200.0, help="serve: Poisson arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve: arrival window in seconds")
    ap.add_argument("--pool", type=int, default=12, help="serve: resident graphs")
    ap.add_argument("--slots", type=int, default=4, help="serve: requests on the device at once")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU "
              "sharding design is documented, not built)" % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import spec_spars_amd as ss
    from spec_spars_amd import graphs

    if args.probe:
        return probe(args)
    if args.workload == "serve":
        if args.precision is None:
            args.precision = "mixed"
        return bench_serve(args)
    W = workload(args.workload)
    if args.precision is None:
        args.precision = W.get("precision", "fp64")
    if "batch" in W:
        return bench_batch(W, args)
    n, u, v, w, k = W["n"], W["u"], W["v"], W["w"], W["k"]
    m = len(u)
    prec = ss.PREC_FP64 if args.precision == "fp64" else ss.PREC_MIXED
    ordering = {"natural": ss.ORDER_NATURAL, "random": ss.ORDER_RANDOM, "rounds": ss.ORDER_ROUNDS}[args.ordering]
    log("workload %s n=%d m=%d k=%d" % (W["name"], n, m, k))

    ctx = ss.Context(0)
    t0 = time.perf_counter()
    G = ss.Graph(ctx, n, u, v, w)
    cp, ri, va = G.laplacian()
    t_graph = time.perf_counter() - t0
    t0 = time.perf_counter()
    F = ss.Factor(cp, ri, va, seed=0, ordering=ordering)
    t_factor = time.perf_counter() - t0
    finfo = F.info()
    _, rounds = F.order()
    t0 = time.perf_counter()
    G.set_factor(F)
    t_upload = time.perf_counter() - t0
    log("graph %.1fs factor %.1fs (nnz %d, depth %d/%d) upload %.1fs" %
        (t_graph, t_factor, finfo["nnz"], finfo["depth_fwd"], finfo["depth_bwd"], t_upload))
    tune = ctx.tuning()
    if not args.no_tune:
        t0 = time.perf_counter()
        tune = ctx.autotune(G, k, prec, reps=3)
        log("autotune %.1fs: %s" % (time.perf_counter() - t0, tune))
    beta = graphs.beta_for(n, W["eps"])
    solve = ss.solve_opts(factor_seed=0, stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=prec,
                          ordering=ordering)

    def step():
        return G.sparsify(k, beta, scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0, solve=solve)

    for i in range(args.warmup):
        rep = step()
        log("warmup %d: %.1f ms (sketch %.1f solve %.1f sample %.1f) iters %d relres %.2e kept %d" %
            (i, rep["ms_total"], rep["ms_sketch"], rep["ms_solve"], rep["ms_sample"], rep["max_iters"],
             rep["max_relres"], rep["kept"]))
    reps = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        reps.append(step())
        log("step %d: %.1f ms" % (i, reps[-1]["ms_total"]))
    t_timed = time.perf_counter() - t0
    # per-kernel times from one more step with HIP events around every
    # launch group (direct launches; the timed steps replay graphs)
    prof = {}
    t_prof = 1.0
    if not args.no_profile:
        G.profile(True)
        t0 = time.perf_counter()
        step()
        t_prof = time.perf_counter() - t0
        prof = G.profile_read()
        G.profile(False)
    ms_per_step = 1000.0 * t_timed / args.steps
    value = m / (ms_per_step / 1000.0)
    last = reps[-1]

    kb = kernel_bytes(n, m, G, k, last["kept"], s=8 if prec == ss.PREC_FP64 else 4)
    kernels, roofline = summarize(prof, kb, t_prof)
    if roofline:
        roofline["traffic"] = measured_traffic(W["name"], args.precision, roofline["kernel"])
    copy_ms = ctx.bench_copy(1 << 30, reps=10)
    copy_gbs = 2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(W)
        log("cpu baseline %.3g edges/s (%s)" % (cpu["value"], cpu["sample"]))

    out = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == ss.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": W["desc"], "n": n, "m": m, "k": k, "epsilon": W["eps"],
                   "precision": args.precision, "ordering": args.ordering, "factor_nnz": finfo["nnz"],
                   "depth_fwd": finfo["depth_fwd"], "depth_bwd": finfo["depth_bwd"], "rounds": rounds,
                   "pcg_iters": last["max_iters"], "refinements": last["refinements"],
                   "max_relres": last["max_relres"], "kept_edges": last["kept"],
                   "ms_sketch": last["ms_sketch"], "ms_solve": last["ms_solve"], "ms_sample": last["ms_sample"],
                   "rhs_solves_per_s": k / (last["ms_solve"] / 1000.0),
                   "host_factor_s": t_factor, "host_graph_s": t_graph, "upload_s": t_upload,
                   "copy_kernel_GBps": copy_gbs, "parallelism": "1 GPU", "tuning": tune,
                   "schedule_fwd": G.schedule_info(0), "schedule_bwd": G.schedule_info(1)},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    G.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
