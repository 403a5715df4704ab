#!/usr/bin/env python3
"""Benchmark of the sparsification hot path on one MI355X.

Default workload: the Chung-Lu power-law graph n = 5,000,000, m = 50,000,000
(degree exponent 2.2, weights U[0.5, 2]), k = 64 Rademacher sketch columns
scaled 1/sqrt(k), PCG to a 1e-8 relative residual, eps = 0.3.  A step is one
pass of the hot path over the resident graph (sketch -> multi-RHS PCG ->
fused ER/coin/sample/compaction); the host factor is built and uploaded once
before timing.  Legs of one run:

  plain   FP64 SpMM / triangular solves, one request per pass (k = 64).
  served  the mixed-precision served form: R requests (different eps and
          seeds) merged into one pass -- their sketch columns side by side,
          one multi-RHS solve with FP32 SpMM / triangular solves and FP64
          reductions and refinement, one sampling pass each.  Every request's
          kept edges equal what it gets alone (tested bit for bit).
  value = served edges/s = R * m / served step time.

Also reported: the plain method column by column (the reference's structure:
k single-RHS solves, FP64) on a timed sample of columns, leverage scores of
the mixed solve against the FP64 solve of the same sketch, a bounded serving
window (Poisson arrivals, p50/p99 latency), the dominant kernel's roofline
and a CPU baseline on the host cores.

Prints one JSON line.  --gpus above 1 is refused: this build runs on one GPU.
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
# a 4 KiB copy (k_copy_f64, 512 work-items) just before and just after the
# profiled served pass: the markers --trace-summary finds in a kernel trace
TRACE_MARK_BYTES = 4096


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def workload(name):
    from spec_spars_amd import graphs
    import numpy as np
    if name == "chunglu5m":
        n, u, v = graphs.chung_lu(5_000_000, 50_000_000, gamma=2.2, seed=0)
        w = graphs.uniform_weights(len(u), seed=0)
        return dict(name=name, desc="Chung-Lu n=5e6 m=5e7 gamma=2.2, weights U[0.5,2], k=64, PCG rel 1e-8, eps=0.3",
                    n=n, u=u, v=v, w=w, k=64, eps=0.3, precision="mixed", rows="fp16")
    if name == "grid3d128":
        n, u, v = graphs.grid3d(128)
        w = graphs.loguniform_weights(len(u), seed=0)
        return dict(name=name, desc="3D grid 128^3, weights 10^U(-3,3), k=48, FP64, PCG rel 1e-8, eps=0.3",
                    n=n, u=u, v=v, w=w, k=48, eps=0.3, precision="fp64")
    if name == "grid3d64":
        n, u, v = graphs.grid3d(64)
        w = graphs.loguniform_weights(len(u), seed=0)
        return dict(name=name, desc="3D grid 64^3, weights 10^U(-3,3), k=48, FP64, PCG rel 1e-8, eps=0.3",
                    n=n, u=u, v=v, w=w, k=48, eps=0.3, precision="fp64")
    if name == "rgg512":
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
        return dict(name=name, desc="2D grid 64x64, unit weights, k=24, FP64, PCG rel 1e-8, eps=0.3",
                    n=n, u=u, v=v, w=np.ones(len(u)), k=24, eps=0.3, precision="fp64")
    raise SystemExit("unknown workload " + name)


def kernel_bytes(n, m, G, k, kept, s=8, srow=None):
    """Algorithmic (compulsory) HBM bytes per launch of each kernel kind:
    every array counted once per pass, gathers beyond first touch not
    counted.  n, m: total vertices / edges of the (batched) system; k the
    columns of the pass; s the bytes of the solve's values; srow the bytes
    of a value of the preconditioner's intermediate rows (s, or 2 for FP16
    rows: the forward reads the residual at s and writes its rows at srow,
    the backward reads and writes rows at srow)."""
    srow = s if srow is None else srow
    nnzL = n + 2 * m
    fi, bi = G.schedule_info(0), G.schedule_info(1)
    row_io = {False: k * (s + srow), True: 2 * k * srow}   # forward / backward in-row + out-row

    def per_level(info, bwd):
        # schedule row id + row pointer, (col, val) per entry, read in-row, write out-row (+ D backward)
        if not info["level_launches"]:
            return None
        b = info["level_rows"] * (4 + 4 + row_io[bwd] + (8 if bwd else 0)) + info["level_nnz"] * (4 + s)
        return b / info["level_launches"]

    def small(info, bwd):
        if not info["small_levels"]:
            return None
        total = info["small_rows"] * (4 + 4 + row_io[bwd] + (8 if bwd else 0)) + info["small_nnz"] * (4 + s)
        return total / info["small_runs"]

    def core(info, launches):
        # dense cores, one direction: the inverse's triangle once, the core
        # rows read (gathered) and written once, the forward head's entries
        # outside the core; over the direction's launches
        if not info["core_rows"]:
            return None
        total = info["core_entries"] // 2 * s + info["core_rows"] * 2 * k * srow + \
            info["core_outside_nnz"] * (4 + s)
        return total / launches

    return {
        # sketch: incidence lists (edge id + signed sqrt(w)) and Y (FP64)
        "sketch": 4 * (n + 1) + 2 * m * (4 + 8) + n * k * 8,
        # SpMM: rowptr, (col, val) per nonzero, read P, write Q; with FP16
        # preconditioner rows Q = A Z + beta Q: read Z (FP16), P, Q, write Q
        "spmm": 4 * (n + 1) + nnzL * (4 + s) + (2 * n * k * s if srow == s else n * k * (srow + 3 * s)),
        "tri_fwd": per_level(fi, False),
        "tri_bwd": per_level(bi, True),
        "tri_fwd_small": small(fi, False),
        "tri_bwd_small": small(bi, True),
        "core_fwd": core(fi, 4),
        "core_bwd": core(bi, 3),
        # ER/sample: edges (u, v, w), Z (FP64) once, p out, mask out, kept (u, v, w) out
        "er_sample": m * (4 + 4 + 8) + n * k * 8 + 8 * m + m + kept * 16,
        "vector": None,
        "residual": None,
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


def add_traffic(roofline, traffic):
    """The counter-measured HBM/fabric bytes per launch of the roofline's
    kernel and the rate they move at: how close the kernel runs to the
    memory system when the gathers beyond first touch are counted too."""
    roofline["traffic"] = traffic
    if traffic and roofline.get("avg_launch_ms"):
        gbs = traffic / (roofline["avg_launch_ms"] * 1e-3) / 1e9
        roofline["traffic_GBps"] = round(gbs, 1)
        roofline["traffic_frac"] = round(gbs / PEAK_HBM_GBS, 4)
        roofline["traffic_over_algorithmic"] = round(traffic / roofline["bytes_per_launch"], 2)


def measured_traffic(workload, precision, kind, columns, rows=None):
    """HBM bytes per launch of a kernel kind, from the committed counter
    summary profiles/sparsify_kernels_<workload>_<precision>_pmc.md
    (FETCH_SIZE and WRITE_SIZE per dispatch, calibrated on a copy kernel),
    when one exists for a pass of the same number of columns (its
    `traffic_columns:` line) and, for mixed precision, the same
    preconditioner rows (its `traffic_rows:` line)."""
    path = os.path.join(ROOT, "profiles", "sparsify_kernels_%s_%s_pmc.md" % (workload, precision))
    if not os.path.exists(path):
        return None
    lines = open(path).read().splitlines()
    cols = [int(x.split(":")[1]) for x in lines if x.startswith("traffic_columns:")]
    if cols != [columns]:
        return None
    if rows is not None and [x.split(":")[1].strip() for x in lines if x.startswith("traffic_rows:")] != [rows]:
        return None
    for line in lines:
        if line.startswith("traffic_bytes_per_launch %s:" % kind):
            return int(line.split(":")[1])
    return None


def host_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def cpu_threads():
    """Host threads for the CPU legs: the process's CPU share, at most 16
    (the GPU box's share for one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


def cpu_baseline(W, F, cp, ri, va, t_factor, pcg_iters, columns=None, er_edges=5_000_000, sketch_edges=2_000_000,
                 pcg_iters_timed=2):
    """The CPU restatement of the reference path (the oracle, plain C) on the
    host cores, on a bounded sample of the same workload:
      sketch   sketch_edges edges timed on one thread, scaled to m and
               divided by the reference's ceil(m/n) sketch threads
               (src/sparsifier.rs:509-529);
      PCG      `columns` single-RHS solves (the product's factor in its
               position space), one per thread, pcg_iters_timed iterations
               each, scaled to pcg_iters iterations (the FP64 solve's count
               to rel 1e-8 on the GPU) and to k columns over those threads
               (the reference solves the k columns one after another,
               src/example.cc:552-569);
      ER, probabilities, coins, sampling: er_edges edges on one thread
               (the reference's loops are sequential), scaled to m.
    The factorization is the product's host factor (timed in the GPU leg)."""
    import numpy as np
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from spec_spars_amd import graphs
    n, u, v, w, k = W["n"], W["u"], W["v"], W["w"], W["k"]
    m = len(u)
    T = cpu_threads()
    columns = columns or T
    # sketch sample
    se = min(sketch_edges, m)
    t0 = time.perf_counter()
    oracle.sketch_edges(np.arange(se), u[:se], v[:se], w[:se], n, k, seed=0, scale=1.0 / math.sqrt(k))
    t_sk1 = time.perf_counter() - t0
    sk_threads = int(math.ceil(m / n))
    t_sketch = t_sk1 * (m / se) / sk_threads
    # PCG: columns single-RHS solves in parallel, position space
    order, _ = F.order()
    pc, pr, pv = oracle.permute_csc(cp, ri, va, order)
    urp, uci, uva, dg = F.export()
    rng = np.random.default_rng(0)
    Y = rng.standard_normal((n, columns))
    Y -= Y.mean(0)
    Yp = np.ascontiguousarray(Y[order])

    def one(j):
        return oracle.solve_lap(Yp[:, j:j + 1], pc, pr, pv, urp, uci, uva, dg, stop_rule=1, tol=1e-30,
                                max_iter=pcg_iters_timed)

    log("cpu baseline: sketch sample %.1fs; %d PCG columns x %d iterations on %d threads ..." %
        (t_sk1, columns, pcg_iters_timed, T))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=T) as ex:
        res = list(ex.map(one, range(columns)))
    t_cols = time.perf_counter() - t0
    # the solve takes pcg_iters iterations (the FP64 GPU solve's count on
    # the same system): scale the timed iterations to it and to k columns
    t_pcg = t_cols * (pcg_iters / pcg_iters_timed) * (k / columns)
    iters = pcg_iters
    # ER / probabilities / coins / sampling sample (canonical order: the
    # first er_edges edges, their endpoints' rows of a solution block)
    log("cpu baseline: PCG sample %.1fs; ER/sampling sample ..." % t_cols)
    ee = min(er_edges, m)
    Z = np.zeros((n, k))
    zr = rng.standard_normal((n, 1)) * 1e-3
    Z[:, :] = zr
    uu, vv, lv = u[:ee], v[:ee], -1.0 * w[:ee]
    t0 = time.perf_counter()
    d = oracle.diff_norms(uu, vv, Z)
    p = oracle.probs(lv, d, graphs.beta_for(n, W["eps"]), k)
    c = oracle.coins(0, ee)
    oracle.sample_reweight(n, uu, vv, lv, p, c)
    t_er = (time.perf_counter() - t0) * (m / ee)
    total = t_sketch + t_pcg + t_er
    return {
        "value": m / total, "unit": "edges/s", "cores": T, "kind": "port",
        "value_end_to_end": m / (total + t_factor),
        "sample": ("oracle (plain C restatement of the reference path) on the host: sketch %d of %d edges on one "
                   "thread (%.2fs) scaled to m over the reference's ceil(m/n)=%d sketch threads; %d single-RHS "
                   "FP64 PCG solves (this build's rounds-order factor) of %d iterations each on %d threads in %.1fs, "
                   "scaled to %d iterations (the FP64 solve to rel 1e-8) and k=%d columns; "
                   "ER/probabilities/coins/sampling on %d edges on one thread, scaled to m; factorization = this "
                   "build's host factor (%.1fs, shared with the GPU path)" %
                   (se, m, t_sk1, sk_threads, columns, pcg_iters_timed, T, t_cols, iters, k, ee, t_factor)),
        "seconds": {"sketch": t_sketch, "pcg": t_pcg, "er_sample": t_er, "factor": t_factor},
        "threads": {"pcg": T, "sketch": sk_threads, "factor": None},
        "host": host_info(),
    }


def serve_pool(count, seed=5, n_max=2e6, m_max=20_000_000):
    """Resident graphs for serving: the three families mixed, n log-uniform
    in [1e4, n_max], m <= m_max."""
    import numpy as np
    from spec_spars_amd import graphs
    rng = np.random.default_rng(seed)
    pool, desc = [], []
    for i in range(count):
        n = int(round(10.0 ** rng.uniform(4.0, math.log10(n_max))))
        fam = ("grid3d", "chunglu", "rgg")[i % 3]
        if fam == "grid3d":
            side = max(3, int(round(n ** (1.0 / 3.0))))
            nn, u, v = graphs.grid3d(side)
            w = graphs.loguniform_weights(len(u), seed=100 + i)
        elif fam == "chunglu":
            nn, u, v = graphs.chung_lu(n, min(8 * n, m_max), gamma=2.2, seed=100 + i)
            w = graphs.uniform_weights(len(u), seed=100 + i)
        else:
            nn, u, v = graphs.rgg(n, seed=100 + i, mean_degree=8.0)
            w = np.ones(len(u))
        pool.append((nn, u, v, w))
        desc.append("%s n=%d m=%d" % (fam, nn, len(u)))
    return pool, desc


def serve_window(rate, duration, pool_size, n_max, prec, slots=4, seed=11):
    """Poisson arrivals over a resident pool; returns the server summary."""
    from spec_spars_amd.serve import Registry, Server, poisson_requests
    t0 = time.perf_counter()
    pool, desc = serve_pool(pool_size, n_max=n_max)
    reg = Registry(pool, threads=4)
    t_build = time.perf_counter() - t0
    log("serve pool (%.1fs): %s" % (t_build, "; ".join(desc)))
    reqs = poisson_requests(rate=rate, duration=duration, graphs=len(pool), seed=seed, k=32, rhs=16)
    srv = Server(reg, slots=slots, precision=prec, reserve_bytes=4 << 30, keep_arrays=False)
    out = srv.run(reqs, deadline=duration + 30.0)
    reg.close()
    out.update({"rate": rate, "window_s": duration, "pool": desc, "registry_build_s": t_build, "slots": slots})
    return out


def step_loop(fn, warmup, steps, tag):
    for i in range(warmup):
        r = fn()
        log("%s warmup %d: %.1f ms" % (tag, i, r[0]["ms_total"] if isinstance(r, list) else r["ms_total"]))
    t0 = time.perf_counter()
    last = None
    for i in range(steps):
        last = fn()
        log("%s step %d: %.1f ms" % (tag, i, last[0]["ms_total"] if isinstance(last, list) else last["ms_total"]))
    return (time.perf_counter() - t0) / steps, last


# mixed-precision preconditioner rows -> ss_solve_opts.mixed_mode
MIXED_MODES = {"fp16": 2, "fp32": 0}


def bench_single(W, args):
    """One resident graph: plain FP64 leg, served mixed leg (merged requests),
    column-wise sample, leverage check, serving window, CPU baseline."""
    import numpy as np
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    n, u, v, w, k = W["n"], W["u"], W["v"], W["w"], W["k"]
    m = len(u)
    eps = W["eps"]
    beta = graphs.beta_for(n, eps)
    log("workload %s n=%d m=%d k=%d" % (W["name"], n, m, k))
    ctx = ss.Context(0)
    G = None
    try:
        t0 = time.perf_counter()
        G = ss.Graph(ctx, n, u, v, w)
        cp, ri, va = G.laplacian()
        t_graph = time.perf_counter() - t0
        t0 = time.perf_counter()
        F = ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_ROUNDS)
        t_factor = time.perf_counter() - t0
        finfo = F.info()
        _, rounds = F.order()
        t0 = time.perf_counter()
        G.set_factor(F)
        t_upload = time.perf_counter() - t0
        log("graph %.1fs factor %.1fs (nnz %d, tail %d, rounds %d) upload %.1fs" %
            (t_graph, t_factor, finfo["nnz"], F.tail(), rounds, t_upload))

        # ---- plain: FP64, one request per pass
        P64 = ss.PREC_FP64
        tune64 = ctx.autotune(G, k, P64, reps=3) if not args.no_tune else ctx.tuning()
        sv64 = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=P64)
        t_plain, rep_plain = step_loop(
            lambda: G.sparsify(k, beta, scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0, solve=sv64),
            args.warmup, args.steps, "plain fp64")
        p64 = G.read_probs()
        plain_eps = m / t_plain

        # ---- plain, column by column (the reference's structure): a sample
        # of single-RHS FP64 solves, scaled to k
        rhs = G.read_block("Y")[:, :args.col_sample]
        t0 = time.perf_counter()
        for j in range(rhs.shape[1]):
            _, info1 = G.solve(rhs[:, j:j + 1], sv64)
        t_col = (time.perf_counter() - t0) / rhs.shape[1]
        colwise_step = rep_plain["ms_sketch"] / 1e3 + k * t_col + rep_plain["ms_sample"] / 1e3
        log("column-wise: %.1f ms per single-RHS solve (%d iterations)" % (t_col * 1e3, int(info1["iters"].max())))

        # ---- served: R requests merged, mixed precision
        PM = ss.PREC_MIXED if args.precision != "fp64" else ss.PREC_FP64
        R = args.requests
        eps_list = [eps, 0.2, 0.5, eps, 0.2, 0.5][:R]
        reqs = [dict(k=k, beta=graphs.beta_for(n, e), sketch_seed=i, sampling_seed=i) for i, e in enumerate(eps_list)]
        K = R * k
        tuneM = ctx.autotune(G, K, PM, reps=3) if not args.no_tune else ctx.tuning()
        svM = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=PM,
                            mixed_mode=MIXED_MODES[args.mixed_rows])
        t_served, rep_served = step_loop(lambda: G.sparsify_many(reqs, solve=svM), args.warmup, args.steps,
                                         "served x%d" % R)
        value = R * m / t_served
        prof, t_prof = {}, 1.0
        if not args.no_profile:
            ctx.bench_copy(TRACE_MARK_BYTES, reps=1)
            G.profile(True)
            t0 = time.perf_counter()
            G.sparsify_many(reqs, solve=svM)
            t_prof = time.perf_counter() - t0
            prof = G.profile_read()
            G.profile(False)
            ctx.bench_copy(TRACE_MARK_BYTES, reps=1)
        kept = sum(r["kept"] for r in rep_served)
        kb = kernel_bytes(n, m, G, K, kept, s=8 if PM == ss.PREC_FP64 else 4,
                          srow=2 if PM == ss.PREC_MIXED and args.mixed_rows == "fp16" and K % 2 == 0 else None)
        kernels, roofline = summarize(prof, kb, t_prof)
        # er_sample runs once per request: its bytes are per request's k
        if "er_sample" in kernels:
            ke = kernels["er_sample"]
            ke["bytes_per_launch"] = kernel_bytes(n, m, G, k, kept // R, s=4)["er_sample"]
            ke["GBps"] = ke["bytes_per_launch"] / (ke["avg_ms"] * 1e-3) / 1e9
        if roofline:
            add_traffic(roofline, measured_traffic(W["name"], "mixed" if PM == ss.PREC_MIXED else "fp64",
                                                   roofline["kernel"], K,
                                                   rows=args.mixed_rows if PM == ss.PREC_MIXED else None))
        copy_ms = ctx.bench_copy(1 << 30, reps=10)
        copy_gbs = 2.0 * (1 << 30) / (copy_ms * 1e-3) / 1e9

        # ---- leverage scores of the mixed solve vs the FP64 solve (same
        # sketch: request 0's parameters are the plain leg's)
        rep_m = G.sparsify(k, beta, scale=1.0 / math.sqrt(k), sketch_seed=0, sampling_seed=0, solve=svM)
        pm = G.read_probs()
        both = (pm < 1.0) & (p64 < 1.0)
        lev_err = float(np.max(np.abs(pm[both] - p64[both]) / p64[both])) if both.any() else 0.0
        log("leverage error mixed vs FP64: %.3e over %d edges with p < 1" % (lev_err, int(both.sum())))
        sched = {"fwd": G.schedule_info(0), "bwd": G.schedule_info(1)}
    finally:
        if G is not None:
            G.close()
        ctx.close()

    serve = None
    if not args.no_serve:
        log("serving window: %.0f req/s for %.0fs" % (args.rate, args.serve_seconds))
        serve = serve_window(args.rate, args.serve_seconds, args.serve_pool, args.serve_nmax, PM)
        log("served: p50 %.1f ms p99 %.1f ms, %d/%d completed" %
            (serve["p50_ms"], serve["p99_ms"], serve["completed"], serve["requests"]))

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(W, F, cp, ri, va, t_factor, pcg_iters=rep_plain["max_iters"])
        log("cpu baseline %.3g edges/s (%s)" % (cpu["value"], cpu["sample"]))

    r0 = rep_served[0]
    out = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_served * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if PM == ss.PREC_MIXED else "f64", "data": "synthetic",
        "config": {"workload": W["desc"], "n": n, "m": m, "k": k, "epsilon": eps,
                   "served": {"requests_per_pass": R, "epsilons": eps_list, "columns": K,
                              "precision": "mixed" if PM == ss.PREC_MIXED else "fp64", "ms_per_pass": t_served * 1e3,
                              "edges_per_s": value, "kept_edges": [r["kept"] for r in rep_served],
                              "pcg_iters": r0["max_iters"], "refinements": r0["refinements"],
                              "max_relres": max(r["max_relres"] for r in rep_served),
                              "ms_sketch": r0["ms_sketch"], "ms_solve": r0["ms_solve"], "ms_sample": r0["ms_sample"],
                              "rhs_solves_per_s": K / (r0["ms_solve"] / 1000.0), "tuning": tuneM},
                   "plain": {"precision": "fp64", "ms_per_pass": t_plain * 1e3, "edges_per_s": plain_eps,
                             "pcg_iters": rep_plain["max_iters"], "max_relres": rep_plain["max_relres"],
                             "kept_edges": rep_plain["kept"], "ms_sketch": rep_plain["ms_sketch"],
                             "ms_solve": rep_plain["ms_solve"], "ms_sample": rep_plain["ms_sample"],
                             "rhs_solves_per_s": k / (rep_plain["ms_solve"] / 1000.0), "tuning": tune64},
                   "plain_columnwise": {"precision": "fp64", "columns_timed": int(rhs.shape[1]),
                                        "ms_per_column_solve": t_col * 1e3, "ms_per_pass_est": colwise_step * 1e3,
                                        "edges_per_s": m / colwise_step},
                   "served_vs_plain": value / plain_eps,
                   "served_vs_plain_columnwise": value / (m / colwise_step),
                   "leverage_vs_fp64": {"max_rel_err": lev_err, "edges_compared": int(both.sum()),
                                        "mixed_iters": rep_m["max_iters"], "mixed_relres": rep_m["max_relres"]},
                   "parity_unpinned": ["CityHash64WithSeed HashLen0to16 at len 16 (fasthash-sys version unknown)",
                                       "rand 0.8.5 StdRng seed_from_u64 + gen_range f64 coin stream"],
                   "ordering": "rounds", "factor_nnz": finfo["nnz"], "factor_tail_rows": F.tail(),
                   "depth_fwd": finfo["depth_fwd"], "depth_bwd": finfo["depth_bwd"], "rounds": rounds,
                   "host_factor_s": t_factor, "host_graph_s": t_graph, "upload_s": t_upload,
                   "factor_threads": int(os.environ.get("SPEC_SPARS_THREADS", "0")) or min(os.cpu_count() or 1, 32),
                   # the graph in host memory to H on the device: graph
                   # build and upload, host factor, factor upload, one pass
                   "edges_per_s_end_to_end": R * m / (t_served + t_graph + t_factor + t_upload),
                   "copy_kernel_GBps": copy_gbs, "parallelism": "1 GPU", "schedule": sched,
                   "serving": serve},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline_batch(W, count=4):
    """The oracle on the first `count` graphs of the batch (every column,
    natural-order factors, one thread): edges/s of that bounded sample."""
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
            "host": host_info()}


def bench_batch(W, args):
    """Many graphs as one resident batch: a step sparsifies all of them."""
    from concurrent.futures import ThreadPoolExecutor
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    gs, k = W["batch"], W["k"]
    n, m = W["n"], W["m"]
    prec = ss.PREC_FP64 if args.precision == "fp64" else ss.PREC_MIXED
    log("workload %s: %d graphs, n=%d m=%d k=%d" % (W["name"], len(gs), n, m, k))
    ctx = ss.Context(0)
    B = None
    try:
        t0 = time.perf_counter()
        B = ss.Batch(ctx, gs)
        t_graph = time.perf_counter() - t0
        t0 = time.perf_counter()
        ss.set_factor_threads(1)   # one graph per thread

        def build(g):
            return ss.Factor(*graphs.laplacian_csc(*g), seed=0, ordering=ss.ORDER_ROUNDS)

        with ThreadPoolExecutor(max_workers=cpu_threads()) as ex:
            fs = list(ex.map(build, gs))
        ss.set_factor_threads(0)
        t_factor = time.perf_counter() - t0
        fnnz = sum(f.info()["nnz"] for f in fs)
        t0 = time.perf_counter()
        B.set_factors(fs)
        t_upload = time.perf_counter() - t0
        log("batch %.1fs factors %.1fs (nnz %d) upload %.1fs" % (t_graph, t_factor, fnnz, t_upload))
        tune = ctx.autotune(B, k, prec, reps=3) if not args.no_tune else ctx.tuning()
        betas = [graphs.beta_for(g[0], W["eps"]) for g in gs]
        solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=2000, precision=prec,
                              mixed_mode=MIXED_MODES[args.mixed_rows])
        t_step, reps = step_loop(lambda: B.sparsify(k, betas, sketch_seeds=[0] * len(gs),
                                                    sampling_seeds=[0] * len(gs), solve=solve),
                                 args.warmup, args.steps, "batch")
        prof, t_prof = {}, 1.0
        if not args.no_profile:
            B.profile(True)
            t0 = time.perf_counter()
            B.sparsify(k, betas, sketch_seeds=[0] * len(gs), sampling_seeds=[0] * len(gs), solve=solve)
            t_prof = time.perf_counter() - t0
            prof = B.profile_read()
            B.profile(False)
        kept = sum(r["kept"] for r in reps)
        kb = kernel_bytes(n, m, B, k, kept, s=8 if prec == ss.PREC_FP64 else 4,
                          srow=2 if prec == ss.PREC_MIXED and args.mixed_rows == "fp16" and k % 2 == 0 else None)
        kernels, roofline = summarize(prof, kb, t_prof)
        if roofline:
            add_traffic(roofline, measured_traffic(W["name"], args.precision, roofline["kernel"], k))
        sched = {"fwd": B.schedule_info(0), "bwd": B.schedule_info(1)}
    finally:
        if B is not None:
            B.close()
        ctx.close()
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline_batch(W)
        log("cpu baseline %.3g edges/s (%s)" % (cpu["value"], cpu["sample"]))
    r0 = reps[0]
    out = {
        "metric": METRIC, "value": m / t_step, "unit": "edges/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_step * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == ss.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": W["desc"], "graphs": len(gs), "n": n, "m": m, "k": k, "rhs": k * len(gs),
                   "epsilon": W["eps"], "precision": args.precision, "ordering": "rounds",
                   "factor_nnz": fnnz, "pcg_iters": max(r["max_iters"] for r in reps),
                   "refinements": r0["refinements"], "max_relres": max(r["max_relres"] for r in reps),
                   "kept_edges": kept, "ms_sketch": r0["ms_sketch"], "ms_solve": r0["ms_solve"],
                   "ms_sample": r0["ms_sample"], "rhs_solves_per_s": k * len(gs) / (r0["ms_solve"] / 1000.0),
                   "host_factor_s": t_factor, "host_graph_s": t_graph, "upload_s": t_upload,
                   "parallelism": "1 GPU", "tuning": tune, "schedule": sched},
        "roofline": roofline,
        "kernels": kernels,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_serve(args):
    """Poisson arrivals of sparsify / 16-RHS solve requests over resident
    graphs, continuous batching by slots, admission by free device memory."""
    import spec_spars_amd as ss
    prec = ss.PREC_FP64 if args.precision == "fp64" else ss.PREC_MIXED
    out = serve_window(args.rate, args.duration, args.pool, 2e6, prec, slots=args.slots)
    log("served: %s" % {k: v for k, v in out.items() if k != "pool"})
    line = {
        "metric": METRIC, "value": out["sparsified_edges_per_s"], "unit": "edges/s", "n_gpus": 1,
        "steps": out["requests"], "warmup": 0, "ms_per_step": out["mean_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == ss.PREC_FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson %.0f req/s for %.0fs over %d resident graphs (grid3d, Chung-Lu, "
                               "RGG; n log-uniform [1e4, 2e6], m <= 2e7); half sparsify (k=32, eps in "
                               "{0.2, 0.3, 0.5}), half 16-RHS solves; %d slots" %
                               (args.rate, args.duration, args.pool, args.slots),
                   "precision": args.precision, **out},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


def probe(args):
    """A short fixed program for hardware-counter runs: the workload's graph
    and factor, the served form's merged pass (or one plain pass) capped at
    3 PCG iterations, then one 1 GiB FP64 copy for calibration."""
    import spec_spars_amd as ss
    from spec_spars_amd import graphs
    W = workload(args.workload)
    n, u, v, w, k = W["n"], W["u"], W["v"], W["w"], W["k"]
    ctx = ss.Context(0)
    G = ss.Graph(ctx, n, u, v, w)
    try:
        cp, ri, va = G.laplacian()
        G.set_factor(ss.Factor(cp, ri, va, seed=0, ordering=ss.ORDER_ROUNDS))
        log("probe: factor uploaded")
        prec = ss.PREC_FP64 if (args.precision or W["precision"]) == "fp64" else ss.PREC_MIXED
        rows = args.mixed_rows or W.get("rows", "fp32")
        solve = ss.solve_opts(stop_rule=ss.STOP_REL, tol=1e-8, max_iter=3, precision=prec,
                              mixed_mode=MIXED_MODES[rows])
        solve.fail_relres = 1e300
        if args.requests > 1:
            reqs = [dict(k=k, beta=graphs.beta_for(n, W["eps"]), sketch_seed=i, sampling_seed=i)
                    for i in range(args.requests)]
            rep = G.sparsify_many(reqs, solve=solve)[0]
        else:
            rep = G.sparsify(k, graphs.beta_for(n, W["eps"]), scale=1.0 / math.sqrt(k), solve=solve)
        log("probe pass %.1f ms" % rep["ms_total"])
        log("copy %.3f ms per GiB copy" % ctx.bench_copy(1 << 30, reps=1))
    finally:
        G.close()
        ctx.close()
    return 0


def trace_summary(path):
    """Per-kernel device time of the profiled served pass from a rocprofv3
    kernel trace of this script (--kernel-trace --output-format csv): the
    kernels between the two marker copies around that pass (the last
    marker launch before it, the first after it).  Prints markdown."""
    import csv
    rows = []
    with open(path) as f:
        rd = csv.DictReader(f)
        cols = rd.fieldnames or []
        if "Grid_Size" in cols:
            grid = lambda r: int(r["Grid_Size"])
        elif "Grid_Size_X" in cols:
            grid = lambda r: int(r["Grid_Size_X"]) * int(r.get("Grid_Size_Y") or 1) * int(r.get("Grid_Size_Z") or 1)
        else:
            raise SystemExit("trace-summary: no grid size column in %s (columns: %s)" % (path, ", ".join(cols)))
        for r in rd:
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"], grid(r)))
    rows.sort()
    marks = [i for i, r in enumerate(rows) if "k_copy_f64" in r[2] and r[3] <= TRACE_MARK_BYTES // 8]
    if len(marks) < 4:
        raise SystemExit("trace-summary: the marker copies are not in %s" % path)
    lo, hi = marks[1], marks[2]   # each marker is a warm launch plus one timed launch
    stats = {}
    for t0, t1, name, _ in rows[lo + 1:hi]:
        key = _kernel_key(name)
        s = stats.setdefault(key, [0, 0.0])
        s[0] += 1
        s[1] += (t1 - t0) / 1e6
    span = (rows[hi][0] - rows[lo][1]) / 1e6
    print("| kernel | launches | total ms | avg us |")
    print("|---|---:|---:|---:|")
    for k, (c, ms) in sorted(stats.items(), key=lambda x: -x[1][1]):
        print("| `%s` | %d | %.2f | %.1f |" % (k, c, ms, 1e3 * ms / c))
    print("\nKernel time in the window %.1f ms; the window spans %.1f ms." %
          (sum(v[1] for v in stats.values()), span))
    return 0


def _template_args(name):
    """Top-level template arguments of a demangled kernel name."""
    a = name.find("<")
    if a < 0:
        return []
    depth, cur, out = 0, "", []
    for ch in name[a + 1:]:
        if ch in "<(":
            depth += 1
        elif ch in ">)":
            if depth == 0:
                out.append(cur.strip())
                break
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    return out


def _kernel_base(name):
    """A kernel's name without template arguments or parameters, from its
    demangled or (for names rocprofv3 leaves mangled) Itanium-mangled form."""
    import re
    m = re.match(r"_ZN2ss(\d+)", name)
    if m:
        n = int(m.group(1))
        return "ss::" + name[m.end():m.end() + n]
    return name.split("(")[0].split("<")[0].replace("void ", "")


def _kernel_bwd(name, pos):
    """The bool template argument at position pos of a solve kernel (its
    direction: true = backward), demangled or mangled (Lb0E / Lb1E)."""
    if name.startswith("_Z"):
        i = name.find("Lb")
        return i >= 0 and name[i + 2] == "1"
    args = _template_args(name)
    return len(args) > pos and args[pos] == "true"


_DIR_ARG = {"ss::k_trisolve_level": 3, "ss::k_trisolve_huge": 3, "ss::k_level_copy": 3, "ss::k_tri_small": 3,
            "ss::k_core_gather": 2, "ss::k_core_apply": 1, "ss::k_core_apply_h": 0}


def _kernel_key(name):
    """A readable per-kernel key: the base name, with the direction of a
    solve kernel."""
    b = _kernel_base(name)
    return b + (" (%s)" % ("bwd" if _kernel_bwd(name, _DIR_ARG[b]) else "fwd") if b in _DIR_ARG else "")


def _counter_rows(path):
    """(dispatch id, kernel name, value) of one rocprofv3 counter CSV."""
    import csv
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return rows


def pmc_summary(fetch_csv, write_csv):
    """HBM traffic per dispatch from two rocprofv3 --pmc runs of the probe
    (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each calibrated on the
    probe's 1 GiB FP64 copy, and the traffic per launch of the kernel kinds
    bench.py times (a level launch with its rows of more than 1,024 entries;
    the forward core head, the level launch just before each forward core
    gather, counted with the dense cores).  Prints markdown."""
    fetch, write = _counter_rows(fetch_csv), _counter_rows(write_csv)
    if len(fetch) != len(write) or any(a[1] != b[1] for a, b in zip(fetch, write)):
        raise SystemExit("pmc-summary: the two runs dispatched different kernels")
    copy = [i for i, r in enumerate(fetch) if "k_copy_f64" in r[1]]
    if not copy:
        raise SystemExit("pmc-summary: no calibration copy in %s" % fetch_csv)
    cf = (1 << 20) / fetch[copy[-1]][2]    # the copy reads 1 GiB
    cw = (1 << 20) / write[copy[-1]][2]    # and writes 1 GiB
    print("Calibration: FETCH_SIZE reads %.0f KiB and WRITE_SIZE %.0f KiB for the copy's 1048576 KiB read and "
          "written,\nso traffic = %.2f x FETCH_SIZE + %.2f x WRITE_SIZE (KiB).\n" %
          (fetch[copy[-1]][2], write[copy[-1]][2], cf, cw))
    traffic = [(f[1], 1024.0 * (cf * f[2] + cw * w[2])) for f, w in zip(fetch, write)]
    short = [_kernel_base(t[0]) for t in traffic]
    per = {}
    for nm, b in zip((_kernel_key(t[0]) for t in traffic), (t[1] for t in traffic)):
        e = per.setdefault(nm, [0, 0.0])
        e[0] += 1
        e[1] += b
    print("| kernel | dispatches | traffic MB/dispatch |")
    print("|---|---:|---:|")
    for nm, (c, b) in sorted(per.items(), key=lambda x: -x[1][1]):
        print("| `%s` | %d | %.1f |" % (nm, c, b / c / 1e6))
    # kinds: walk the dispatches in order
    kinds = {}
    level = ("k_trisolve_level", "k_trisolve_huge", "k_level_copy")

    def add(kind, b, launch):
        e = kinds.setdefault(kind, [0, 0.0])
        e[0] += launch
        e[1] += b
    names = [t[0] for t in traffic]

    def fwd_level(i):
        return any(x in short[i] for x in level) and not _kernel_bwd(names[i], 3)
    head = set()
    for j, nm in enumerate(short):
        if "k_core_gather" in nm and not _kernel_bwd(names[j], 2) and j > 0 and fwd_level(j - 1):
            head.add(j - 1)
            if "k_trisolve_level" in short[j - 1] and j > 1 and "k_trisolve_huge" in short[j - 2] and fwd_level(j - 2):
                head.add(j - 2)
    for i, (nm, b) in enumerate(traffic):
        base = short[i]
        if any(x in base for x in level):
            bwd = _kernel_bwd(nm, 3)
            kind = "core_fwd" if i in head else ("tri_bwd" if bwd else "tri_fwd")
            add(kind, b, 0 if "k_trisolve_huge" in base else 1)
        elif "kb_spmm_dot" in base or "kb_spmm_z" in base:
            add("spmm", b, 1)
        elif "kb_er_sample" in base:
            add("er_sample", b, 1)
        elif "kb_sketch" in base:
            add("sketch", b, 1)
    print("\nPer launch of each kernel kind as bench.py times them (a level launch with its rows of more than\n"
          "1,024 entries; the forward core head counted with the dense cores), bytes:\n\n```")
    for kind in ("tri_bwd", "tri_fwd", "spmm", "er_sample", "sketch"):
        if kind in kinds and kinds[kind][0]:
            print("traffic_bytes_per_launch %s: %d" % (kind, kinds[kind][1] / kinds[kind][0]))
    print("```")
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="chunglu5m",
                    help="chunglu5m (default), grid3d128, grid3d64, grid2d64, rgg512 (batch), serve")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default=None,
                    help="served leg / batch precision (default: the workload's)")
    ap.add_argument("--requests", type=int, default=4, help="requests merged per served pass")
    ap.add_argument("--mixed-rows", choices=sorted(MIXED_MODES), default=None,
                    help="mixed precision: the preconditioner's intermediate rows in FP16 or FP32 (default: FP16 "
                         "for the merged Chung-Lu pass, FP32 elsewhere)")
    ap.add_argument("--col-sample", type=int, default=2, help="single-RHS solves timed for the column-wise plain leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-profile", action="store_true", help="skip per-kernel event timing")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes (no autotuner)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving window")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="serving window length")
    ap.add_argument("--serve-pool", type=int, default=6, help="graphs in the serving window's pool")
    ap.add_argument("--serve-nmax", type=float, default=2e5, help="largest graph of the serving window's pool")
    ap.add_argument("--probe", action="store_true", help="short pass for counter runs")
    ap.add_argument("--rate", type=float, default=200.0, help="serving: Poisson arrivals per second")
    ap.add_argument("--duration", type=float, default=60.0, help="serve workload: arrival window in seconds")
    ap.add_argument("--pool", type=int, default=12, help="serve workload: resident graphs")
    ap.add_argument("--slots", type=int, default=4, help="serve workload: requests on the device at once")
    ap.add_argument("--trace-summary", metavar="CSV", default=None,
                    help="summarize the profiled served pass of a rocprofv3 kernel trace of this script, and exit")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"), default=None,
                    help="traffic per dispatch and per kernel kind from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of "
                         "--probe, and exit")
    args = ap.parse_args()
    if args.trace_summary:
        return trace_summary(args.trace_summary)
    if args.pmc_summary:
        return pmc_summary(*args.pmc_summary)
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported (the multi-GPU "
              "sharding design is documented, not built)" % args.gpus, file=sys.stderr)
        return 2

    if args.probe:
        return probe(args)
    if args.workload == "serve":
        if args.precision is None:
            args.precision = "mixed"
        return bench_serve(args)
    W = workload(args.workload)
    if args.precision is None:
        args.precision = W["precision"]
    if args.mixed_rows is None:
        args.mixed_rows = W.get("rows", "fp32")
    if "batch" in W:
        return bench_batch(W, args)
    return bench_single(W, args)


if __name__ == "__main__":
    sys.exit(main())
