"""Serving sparsify and solve requests over resident graphs.

No reference counterpart: the reference sparsifies one graph per call
(src/sparsifier.rs:722-770) and solves through run_solve_lap
(src/lib.rs:103).  Here a registry keeps every graph's Laplacian, factor and
work blocks resident in device memory, each graph on its own context (HIP
stream).  Requests arrive over time; the dispatcher is continuous batching
by slots: whenever a slot is free it starts the oldest queued request whose
graph is idle and whose work blocks fit in free device memory, so requests
that arrive while others run join as soon as a slot and memory allow, and
requests on different graphs overlap on the device.  A request's result
depends only on its graph and parameters (each graph's kernels and
reductions run in a fixed order on its own stream), so it is the same alone
and in any mix.

Latency of a request = completion time - arrival time (queueing included);
the summary reports p50/p99 beside sparsified edges/s and RHS-solves/s.
"""
import dataclasses
import math
import threading
import time

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    graph: int             # registry index
    op: str                # "sparsify" or "solve"
    arrival: float         # seconds after the run starts
    eps: float = 0.3
    k: int = 32            # sketch columns (sparsify) or right-hand sides (solve)
    seed: int = 0
    start: float = None
    end: float = None
    result: object = None
    error: str = None


class Scheduler:
    """Admission: FIFO by arrival, skipping requests whose graph is busy or
    whose extra device bytes exceed what is free after the bytes reserved by
    running requests.  Pure host logic (no GPU)."""

    def __init__(self, slots, reserve_bytes=0):
        self.slots = slots
        self.reserve_bytes = reserve_bytes
        self.busy_graphs = set()
        self.reserved = 0
        self.running = 0

    def pick(self, queue, free_bytes, need):
        """Index in queue (arrival-ordered) of the request to start now, or
        None.  need(request) -> extra device bytes it would allocate."""
        if self.running >= self.slots:
            return None
        avail = free_bytes - self.reserve_bytes - self.reserved
        for idx, r in enumerate(queue):
            if r.graph in self.busy_graphs:
                continue
            if need(r) <= avail:
                return idx
        return None

    def started(self, r, nbytes):
        self.busy_graphs.add(r.graph)
        self.reserved += nbytes
        self.running += 1

    def finished(self, r, nbytes):
        self.busy_graphs.discard(r.graph)
        self.reserved -= nbytes
        self.running -= 1


def poisson_requests(rate, duration, graphs, seed=0, sparsify_frac=0.5, eps_choices=(0.2, 0.3, 0.5), k=32,
                     rhs=16):
    """Arrival-ordered requests: Poisson arrivals at `rate` per second over
    `duration` seconds, graph uniform over `graphs` registry entries, half
    sparsify (eps from eps_choices) and half solves with `rhs` right-hand
    sides."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        g = int(rng.integers(graphs))
        if rng.random() < sparsify_frac:
            out.append(Request(len(out), g, "sparsify", t, eps=float(rng.choice(eps_choices)), k=k,
                               seed=int(rng.integers(1 << 31))))
        else:
            out.append(Request(len(out), g, "solve", t, k=rhs, seed=int(rng.integers(1 << 31))))
    return out


class Registry:
    """Resident graphs: graphs[i] = (n, u, v, w) canonical edges.  Each gets
    its own context (stream), device graph and factor."""

    def __init__(self, graphs, ordering=None, factor_seed=0, threads=8):
        from concurrent.futures import ThreadPoolExecutor
        import spec_spars_amd as ss
        from . import graphs as gen
        self.ss = ss
        ordering = ss.ORDER_ROUNDS if ordering is None else ordering
        self.specs = graphs
        self.ctxs, self.G, self.F = [], [], []
        # the default mixed scheme on every graph (FP16 preconditioner rows
        # pay on wide merged requests; the served requests here are narrow)
        self.mixed_mode = [ss.MIXED_RELIABLE] * len(graphs)
        for n, u, v, w in graphs:
            ctx = ss.Context(0)
            self.ctxs.append(ctx)
            self.G.append(ss.Graph(ctx, n, u, v, w))

        def build(g):
            return ss.Factor(*gen.laplacian_csc(*g), seed=factor_seed, ordering=ordering)

        with ThreadPoolExecutor(max_workers=threads) as ex:
            self.F = list(ex.map(build, graphs))
        for G, F in zip(self.G, self.F):
            G.set_factor(F)
        # per graph one projected Gaussian block, drawn at registration; a
        # solve request takes seeded columns of it with seeded signs
        self.rhs_width = 64
        self.rhs_base = {}
        for gi, (n, _, _, _) in enumerate(graphs):
            base = np.random.default_rng(1000 + gi).standard_normal((n, self.rhs_width))
            base -= base.mean(0)
            self.rhs_base[gi] = base

    def close(self):
        for G in self.G:
            G.close()
        for c in self.ctxs:
            c.close()

    def free_bytes(self):
        return self.ctxs[0].mem_info()[0]

    def need(self, r, precision):
        return self.G[r.graph].work_bytes(r.k, precision)

    def _rhs_cols(self, k, seed):
        """A solve request's columns of the Gaussian block and their signs."""
        if not 1 <= k <= self.rhs_width:
            raise ValueError("a solve request takes 1..%d right-hand sides, not %d" % (self.rhs_width, k))
        rng = np.random.default_rng(seed)
        return rng.permutation(self.rhs_width)[:k], rng.choice([-1.0, 1.0], size=k)

    def rhs_block(self, gi, reqs):
        """Right-hand sides of solve requests reqs = [(k, seed), ...] on
        graph gi side by side: request (k, seed) takes k seeded columns of
        the graph's projected N(0, I) block, each sign-flipped by the seed.
        One gather over every column at once: gathering each request's
        columns apart made the host, not the device, the bottleneck of a
        merged solve on a 2M-vertex graph."""
        cols, signs = zip(*(self._rhs_cols(k, seed) for k, seed in reqs))
        b = np.take(self.rhs_base[gi], np.concatenate(cols), axis=1)
        b *= np.concatenate(signs)
        return b

    def rhs(self, gi, k, seed):
        """k right-hand sides for graph gi (one request's)."""
        return self.rhs_block(gi, [(k, seed)])

    def execute_group(self, group, precision, tol=1e-8, keep_arrays=True):
        """Run requests that share a graph (blocking).  Solve requests run as
        one multi-RHS solve over their concatenated right-hand sides (every
        column keeps its own PCG scalars, mask and reductions, so its result
        is the one it gets alone); sparsify requests as one merged pass (their
        sketch columns side by side, one solve, a sampling pass each, each
        result bit-identical to the request alone).  Returns one result per
        request."""
        if len(group) == 1 or len({r.op for r in group}) > 1:
            return [self.execute(r, precision, tol, keep_arrays) for r in group]
        G = self.G[group[0].graph]
        if group[0].op == "sparsify":
            from . import graphs as gen
            n = self.specs[group[0].graph][0]
            opts = self.ss.solve_opts(stop_rule=self.ss.STOP_REL, tol=tol, max_iter=2000, precision=precision,
                                      mixed_mode=self.mixed_mode[group[0].graph])
            reps = G.sparsify_many([dict(k=r.k, beta=gen.beta_for(n, r.eps), sketch_seed=r.seed,
                                         sampling_seed=r.seed) for r in group], solve=opts)
            return [{"kept": rep["kept"], "edges": G.m, "arrays": G.read_kept_many(i) if keep_arrays else None,
                     "iters": rep["max_iters"], "relres": rep["max_relres"], "group": len(group)}
                    for i, rep in enumerate(reps)]
        b = self.rhs_block(group[0].graph, [(r.k, r.seed) for r in group])
        opts = self.ss.solve_opts(stop_rule=self.ss.STOP_REL, tol=tol, max_iter=2000, precision=precision,
                                  mixed_mode=self.mixed_mode[group[0].graph])
        Z, info = G.solve(b, opts)
        out, c = [], 0
        for r in group:
            # each request's Z is a view of the merged block
            sl = slice(c, c + r.k)
            out.append({"Z": Z[:, sl], "rhs": r.k, "iters": int(info["iters"][sl].max()),
                        "relres": float(info["relres"][sl].max()), "group": len(group)})
            c += r.k
        return out

    def execute(self, r, precision, tol=1e-8, keep_arrays=True):
        """Run one request on its graph's stream (blocking)."""
        ss = self.ss
        from . import graphs as gen
        G = self.G[r.graph]
        n = self.specs[r.graph][0]
        opts = ss.solve_opts(stop_rule=ss.STOP_REL, tol=tol, max_iter=2000, precision=precision,
                             mixed_mode=self.mixed_mode[r.graph])
        if r.op == "sparsify":
            rep = G.sparsify(r.k, gen.beta_for(n, r.eps), sketch_seed=r.seed, sampling_seed=r.seed, solve=opts)
            kept = G.read_kept() if keep_arrays else None
            return {"kept": rep["kept"], "edges": G.m, "arrays": kept, "iters": rep["max_iters"],
                    "relres": rep["max_relres"]}
        b = self.rhs(r.graph, r.k, r.seed)
        Z, info = G.solve(b, opts)
        return {"Z": Z, "rhs": r.k, "iters": int(info["iters"].max()), "relres": float(info["relres"].max())}


class Server:
    """Slot dispatcher over a Registry (or any executor with free_bytes,
    need and execute_group).  A started request takes the queued requests of
    the same kind on the same graph with it: solves up to max_rhs
    right-hand sides, sparsifies up to max_cols sketch columns."""

    def __init__(self, registry, slots=4, precision=None, reserve_bytes=1 << 30, keep_arrays=True, max_rhs=64,
                 max_cols=256):
        self.reg = registry
        self.slots = slots
        self.precision = precision
        self.keep_arrays = keep_arrays
        self.max_rhs = max_rhs
        self.max_cols = max_cols
        self.sched = Scheduler(slots, reserve_bytes)

    def _group(self, r, queue):
        """r plus the queued requests of the same kind on its graph that fit
        (removed from the queue): solves up to max_rhs right-hand sides,
        sparsifies up to max_cols sketch columns."""
        group = [r]
        cap = self.max_rhs if r.op == "solve" else self.max_cols
        total, keep = r.k, []
        for q in queue:
            if q.op == r.op and q.graph == r.graph and total + q.k <= cap:
                group.append(q)
                total += q.k
            else:
                keep.append(q)
        queue[:] = keep
        return group

    def run(self, requests, deadline=None, time_scale=1.0):
        """Release requests at arrival * time_scale seconds, run them; stop
        starting new ones after `deadline` seconds (the rest are dropped).
        Returns the summary; request objects carry start/end/result."""
        pending = sorted(requests, key=lambda r: r.arrival)
        queue, done = [], []
        cv = threading.Condition()
        t0 = time.perf_counter()
        inflight = []

        def now():
            return time.perf_counter() - t0

        def work(group, nbytes):
            try:
                res = self.reg.execute_group(group, self.precision, keep_arrays=self.keep_arrays)
                for r, x in zip(group, res):
                    r.result = x
            except Exception as e:   # a failed request must not stop the server
                for r in group:
                    r.error = repr(e)
            t_end = now()
            with cv:
                for r in group:
                    r.end = t_end
                self.sched.finished(group[0], nbytes)
                done.extend(group)
                cv.notify_all()

        dropped = []
        with cv:
            while pending or queue or self.sched.running:
                t = now()
                while pending and pending[0].arrival * time_scale <= t:
                    queue.append(pending.pop(0))
                if deadline is not None and t > deadline:
                    dropped.extend(queue)
                    queue.clear()
                    pending_left = pending
                    dropped.extend(pending_left)
                    pending = []
                started = True
                while started and queue:
                    started = False
                    idx = self.sched.pick(queue, self.reg.free_bytes(),
                                          lambda r: self.reg.need(r, self.precision))
                    if idx is not None:
                        r = queue.pop(idx)
                        group = self._group(r, queue)
                        probe = dataclasses.replace(r, k=sum(q.k for q in group))
                        nbytes = self.reg.need(probe, self.precision)
                        self.sched.started(r, nbytes)
                        t_start = now()
                        for q in group:
                            q.start = t_start
                        th = threading.Thread(target=work, args=(group, nbytes), daemon=True)
                        inflight.append(th)
                        th.start()
                        started = True
                wait = None
                if pending:
                    wait = max(0.0, pending[0].arrival * time_scale - now())
                if queue or self.sched.running or pending:
                    cv.wait(timeout=wait if wait is not None else 0.05)
        for th in inflight:
            th.join()
        wall = now()
        return summarize(done, dropped, wall, time_scale)


def summarize(done, dropped, wall, time_scale=1.0):
    ok = [r for r in done if r.error is None]
    lat = np.array([r.end - r.arrival * time_scale for r in ok]) if ok else np.zeros(0)
    edges = sum(r.result["edges"] for r in ok if r.op == "sparsify")
    rhs = sum(r.result["rhs"] for r in ok if r.op == "solve")

    def pct(q):
        return float(np.percentile(lat, q)) * 1e3 if len(lat) else math.nan

    return {"requests": len(done) + len(dropped), "completed": len(ok), "failed": len(done) - len(ok),
            "dropped": len(dropped), "wall_s": wall,
            "p50_ms": pct(50), "p99_ms": pct(99), "mean_ms": float(lat.mean() * 1e3) if len(lat) else math.nan,
            "sparsified_edges_per_s": edges / wall if wall > 0 else 0.0,
            "rhs_solves_per_s": rhs / wall if wall > 0 else 0.0,
            "requests_per_s": len(ok) / wall if wall > 0 else 0.0}
