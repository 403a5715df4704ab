"""numpy-level wrappers over the C ABI.

Every call goes to libspecspars_hip.so; device work runs on the HIP device of
the Context.  Host-only pieces (Factor) need no GPU.
"""
import ctypes
import sys
import weakref

import numpy as np

from . import _lib
from ._lib import check, load

_i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
_f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)


def _p(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct)) if a is not None else None


def solve_opts(factor_seed=0, stop_rule=_lib.STOP_REF_ABS, max_iter=400, tol=None,
               fail_relres=1e-2, precision=_lib.PREC_FP64, ordering=_lib.ORDER_ROUNDS, inner_tol=0.0,
               mixed_mode=0):
    o = _lib.SolveOpts()
    load().ss_solve_opts_default(ctypes.byref(o))
    o.factor_seed = factor_seed
    o.stop_rule = stop_rule
    o.max_iter = max_iter
    o.tol = (2e-7 if stop_rule == _lib.STOP_REF_ABS else 1e-8) if tol is None else tol
    o.fail_relres = fail_relres
    o.precision = precision
    o.ordering = ordering
    o.inner_tol = inner_tol
    o.mixed_mode = mixed_mode
    return o


# The widest edge-weight spread (max / min) for which the mixed path keeps
# its preconditioner rows in FP16.
FP16_ROWS_MAX_SPREAD = 1024.0


def mixed_mode_for(w):
    """Mixed-precision scheme for wide merged requests on a graph with edge
    weights w: reliable updates with FP16 preconditioner rows
    (MIXED_RELIABLE_F16) when the weights span at most FP16_ROWS_MAX_SPREAD,
    else with FP32 rows (MIXED_RELIABLE).  A wider spread spreads the preconditioned
    vector's entries over more binary orders than FP16's 11-bit significand
    resolves, and the PCG then stalls between residual replacements (a
    random geometric graph with weights 10^U(-3, 3) left half its columns
    above 1e-4 after 60 iterations with FP16 rows, where FP32 rows reach
    1e-8).  Both schemes stop on the FP64 true residual."""
    w = np.asarray(w, dtype=np.float64)
    w = w[w > 0]
    if len(w) == 0:
        return _lib.MIXED_RELIABLE_F16
    return _lib.MIXED_RELIABLE_F16 if w.max() <= FP16_ROWS_MAX_SPREAD * w.min() else _lib.MIXED_RELIABLE


def set_factor_threads(threads=0):
    """Host threads of the factorization (0 = default); the factor does not
    depend on it."""
    check(load().ss_factor_set_threads(int(threads)))


class Factor:
    """Host approximate-Cholesky factor of a Laplacian given as CSC."""

    def __init__(self, colptr, rowidx, vals, seed=0, ordering=_lib.ORDER_ROUNDS):
        lib = load()
        self._cp, self._ri, self._v = _i32(colptr), _i32(rowidx), _f64(vals)
        self.n = len(self._cp) - 1
        h = ctypes.c_void_p()
        check(lib.ss_factor_build_ordered(self.n, _p(self._cp, ctypes.c_int32), _p(self._ri, ctypes.c_int32),
                                          _p(self._v, ctypes.c_double), seed, ordering, ctypes.byref(h)))
        self._h = h

    def order(self):
        o = np.zeros(self.n, np.int32)
        r = ctypes.c_int32()
        check(load().ss_factor_order(self._h, _p(o, ctypes.c_int32), ctypes.byref(r)))
        return o, r.value

    def tail(self):
        """Rows of the dense tail (rounds order): the last positions."""
        t = ctypes.c_int64()
        check(load().ss_factor_tail(self._h, ctypes.byref(t)))
        return t.value

    def info(self):
        n, nnz = ctypes.c_int64(), ctypes.c_int64()
        df, db = ctypes.c_int32(), ctypes.c_int32()
        check(load().ss_factor_info(self._h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(df),
                                    ctypes.byref(db)))
        return {"n": n.value, "nnz": nnz.value, "depth_fwd": df.value, "depth_bwd": db.value}

    def export(self):
        inf = self.info()
        rp = np.zeros(self.n + 1, np.int32)
        ci = np.zeros(inf["nnz"], np.int32)
        va = np.zeros(inf["nnz"], np.float64)
        d = np.zeros(self.n, np.float64)
        check(load().ss_factor_export(self._h, _p(rp, ctypes.c_int32), _p(ci, ctypes.c_int32),
                                      _p(va, ctypes.c_double), _p(d, ctypes.c_double)))
        return rp, ci, va, d

    def levels(self):
        f = np.zeros(self.n, np.int32)
        b = np.zeros(self.n, np.int32)
        check(load().ss_factor_levels(self._h, _p(f, ctypes.c_int32), _p(b, ctypes.c_int32)))
        return f, b

    def close(self):
        if getattr(self, "_h", None):
            load().ss_factor_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        if not sys.is_finalizing():
            self.close()


# blocks readable after a run: sketch, solution, and (FP64) the last
# preconditioner application's forward result and output
_BLOCKS = {"Y": 0, "Z": 1, "fwd": 2, "prec": 3}


class Context:
    """A device and its HIP stream."""

    def __init__(self, device=0):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.ss_context_create(device, ctypes.byref(h)))
        self._h = h
        self._children = weakref.WeakSet()   # graphs and batches on this stream

    def close(self):
        """Destroy the context; graphs still open on it are closed first."""
        if getattr(self, "_h", None):
            for child in list(self._children):
                child.close()
            load().ss_context_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        if not sys.is_finalizing():
            self.close()

    def mem_info(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        check(load().ss_context_mem_info(self._h, ctypes.byref(f), ctypes.byref(t)))
        return f.value, t.value

    def tuning(self):
        t = _lib.Tuning()
        check(load().ss_context_get_tuning(self._h, ctypes.byref(t)))
        return {"tile_waves": t.tile_waves, "level_rows_per_wave": t.level_rows_per_wave,
                "graph_iters": t.graph_iters, "xcd_slices": t.xcd_slices}

    def set_tuning(self, tile_waves=4, level_rows_per_wave=1, graph_iters=0, xcd_slices=1):
        t = _lib.Tuning(tile_waves, level_rows_per_wave, graph_iters, xcd_slices)
        check(load().ss_context_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, graph, k, precision=_lib.PREC_FP64, reps=5):
        """Time launch-shape candidates on graph (a Graph or Batch with its
        factor set); the fastest are kept in this context and the graph."""
        t = _lib.Tuning()
        ml, mt = ctypes.c_double(), ctypes.c_double()
        check(load().ss_context_autotune(self._h, graph._h, k, precision, reps, ctypes.byref(t),
                                         ctypes.byref(ml), ctypes.byref(mt)))
        return {"tile_waves": t.tile_waves, "level_rows_per_wave": t.level_rows_per_wave,
                "graph_iters": t.graph_iters, "xcd_slices": t.xcd_slices, "ms_level": ml.value,
                "ms_tile": mt.value}

    def bench_copy(self, nbytes, reps=20):
        """Average ms of a plain FP64 copy kernel over nbytes (read + write)."""
        ms = ctypes.c_double()
        check(load().ss_bench_copy(self._h, nbytes, reps, ctypes.byref(ms)))
        return ms.value

    def sketch(self, v1, v2, w, n, k, seed=0, scale=1.0, uniform=False):
        v1, v2, w = _i32(v1), _i32(v2), _f64(w)
        Y = np.zeros((n, k), np.float64)
        check(load().ss_sketch(self._h, _p(v1, ctypes.c_int32), _p(v2, ctypes.c_int32),
                               _p(w, ctypes.c_double), len(w), n, k, seed, scale, int(bool(uniform)),
                               _p(Y, ctypes.c_double)))
        return Y

    def solve_lap(self, Y, colptr, rowidx, vals, opts=None):
        Y = _f64(Y)
        n, k = Y.shape
        cp, ri, va = _i32(colptr), _i32(rowidx), _f64(vals)
        Z = np.zeros((n, k), np.float64)
        iters = np.zeros(k, np.int32)
        rel = np.zeros(k, np.float64)
        rep = _lib.SolveReport()
        rep.iters = _p(iters, ctypes.c_int32)
        rep.relres = _p(rel, ctypes.c_double)
        o = opts if opts is not None else solve_opts()
        check(load().ss_solve_lap(self._h, _p(Y, ctypes.c_double), n, k, _p(cp, ctypes.c_int32),
                                  _p(ri, ctypes.c_int32), _p(va, ctypes.c_double), len(va),
                                  ctypes.byref(o), _p(Z, ctypes.c_double), ctypes.byref(rep)))
        info = {f: getattr(rep, f) for f, _ in _lib.SolveReport._fields_ if f not in ("iters", "relres")}
        info["iters"] = iters
        info["relres"] = rel
        return Z, info

    def er_sample(self, u, v, lval, k, beta, seed, Z=None, p_in=None, n=None):
        u, v, lval = _i32(u), _i32(v), _f64(lval)
        m = len(u)
        Zc = _f64(Z) if Z is not None else None
        if n is None:
            n = Zc.shape[0]
        pin = _f64(p_in) if p_in is not None else None
        p = np.zeros(m, np.float64)
        d = np.zeros(m, np.float64)
        mask = np.zeros(m, np.uint8)
        ku = np.zeros(max(m, 1), np.int32)
        kv = np.zeros(max(m, 1), np.int32)
        kw = np.zeros(max(m, 1), np.float64)
        kept = ctypes.c_int64()
        check(load().ss_er_sample(self._h, _p(Zc, ctypes.c_double), n, k, _p(u, ctypes.c_int32),
                                  _p(v, ctypes.c_int32), _p(lval, ctypes.c_double), m, beta, seed,
                                  _p(pin, ctypes.c_double), _p(p, ctypes.c_double),
                                  _p(d, ctypes.c_double), _p(mask, ctypes.c_uint8), ctypes.byref(kept),
                                  _p(ku, ctypes.c_int32), _p(kv, ctypes.c_int32), _p(kw, ctypes.c_double)))
        c = kept.value
        return {"p": p, "d": d if p_in is None else None, "mask": mask, "kept": c,
                "ku": ku[:c].copy(), "kv": kv[:c].copy(), "kw": kw[:c].copy()}


class Graph:
    """Device-resident graph (canonical edges u < v sorted by (v, u))."""

    def __init__(self, ctx, n, u, v, w):
        self.ctx = ctx
        self.n = int(n)
        self._u, self._v, self._w = _i32(u), _i32(v), _f64(w)
        self.m = len(self._u)
        h = ctypes.c_void_p()
        check(load().ss_graph_create(ctx._h, self.n, self.m, _p(self._u, ctypes.c_int32),
                                     _p(self._v, ctypes.c_int32), _p(self._w, ctypes.c_double),
                                     ctypes.byref(h)))
        self._h = h
        self.k = None
        ctx._children.add(self)

    def close(self):
        if getattr(self, "_h", None):
            load().ss_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        if not sys.is_finalizing():
            self.close()

    @staticmethod
    def bytes_needed(n, m, factor_nnz, k, precision):
        b = ctypes.c_size_t()
        check(load().ss_graph_bytes(n, m, factor_nnz, k, precision, ctypes.byref(b)))
        return b.value

    def laplacian(self):
        nnz = self.n + 2 * self.m
        cp = np.zeros(self.n + 1, np.int32)
        ri = np.zeros(nnz, np.int32)
        va = np.zeros(nnz, np.float64)
        check(load().ss_graph_laplacian(self._h, _p(cp, ctypes.c_int32), _p(ri, ctypes.c_int32),
                                        _p(va, ctypes.c_double)))
        return cp, ri, va

    def set_factor(self, factor):
        check(load().ss_graph_set_factor(self._h, factor._h))

    def resident_bytes(self):
        b = ctypes.c_size_t()
        check(load().ss_graph_resident_bytes(self._h, ctypes.byref(b)))
        return b.value

    def work_bytes(self, k, precision=_lib.PREC_FP64):
        b = ctypes.c_size_t()
        check(load().ss_graph_work_bytes(self._h, k, precision, ctypes.byref(b)))
        return b.value

    def solve(self, rhs, solve=None):
        """Solve L Z = rhs (n x k, component means removed) on the resident
        graph and factor; returns (Z, info)."""
        sv = solve if solve is not None else solve_opts(stop_rule=_lib.STOP_REL, tol=1e-8)
        y = _f64(np.asarray(rhs).reshape(self.n, -1))
        k = y.shape[1]
        check(load().ss_batch_write_rhs(self._h, 0, k, sv.precision, _p(y, ctypes.c_double)))
        iters = np.zeros(k, np.int32)
        relres = np.zeros(k, np.float64)
        refin = ctypes.c_int32()
        check(load().ss_batch_solve(self._h, k, ctypes.byref(sv), _p(iters, ctypes.c_int32),
                                    _p(relres, ctypes.c_double), ctypes.byref(refin)))
        self.k = k
        return self.read_block("Z"), {"iters": iters, "relres": relres, "refinements": refin.value}

    def sparsify(self, k, beta, scale=None, sketch_seed=0, sampling_seed=0, uniform=False, solve=None):
        o = _lib.SparsifyOpts()
        o.k = k
        o.uniform = int(bool(uniform))
        o.scale = 1.0 / np.sqrt(k) if scale is None else scale
        o.sketch_seed = sketch_seed
        o.sampling_seed = sampling_seed
        o.beta = beta
        o.solve = solve if solve is not None else solve_opts(stop_rule=_lib.STOP_REL, tol=1e-8)
        rep = _lib.SparsifyReport()
        check(load().ss_graph_sparsify(self._h, ctypes.byref(o), ctypes.byref(rep)))
        self.k = k
        return {f: getattr(rep, f) for f, _ in _lib.SparsifyReport._fields_}

    def sparsify_many(self, requests, solve=None, uniform=False):
        """Several sparsify requests in one pass: requests is a list of dicts
        with k, beta and optionally scale (1/sqrt(k)), sketch_seed,
        sampling_seed.  Returns one report per request; read_kept_many(r)
        gives request r's kept edges."""
        R = len(requests)
        arr = (_lib.SparsifyOpts * R)()
        sv = solve if solve is not None else solve_opts(stop_rule=_lib.STOP_REL, tol=1e-8)
        for i, q in enumerate(requests):
            o = arr[i]
            o.k = int(q["k"])
            o.uniform = int(bool(uniform))
            o.scale = float(q.get("scale", 1.0 / np.sqrt(q["k"])))
            o.sketch_seed = int(q.get("sketch_seed", 0))
            o.sampling_seed = int(q.get("sampling_seed", 0))
            o.beta = float(q["beta"])
            o.solve = sv
        reps = (_lib.SparsifyReport * R)()
        check(load().ss_graph_sparsify_many(self._h, R, arr, reps))
        self.k = sum(int(q["k"]) for q in requests)
        return [{f: getattr(r, f) for f, _ in _lib.SparsifyReport._fields_} for r in reps]

    def read_kept_many(self, r):
        kept = ctypes.c_int64()
        cap = self.m
        ku = np.zeros(max(cap, 1), np.int32)
        kv = np.zeros(max(cap, 1), np.int32)
        kw = np.zeros(max(cap, 1), np.float64)
        check(load().ss_graph_read_kept_many(self._h, r, cap, _p(ku, ctypes.c_int32), _p(kv, ctypes.c_int32),
                                             _p(kw, ctypes.c_double), ctypes.byref(kept)))
        c = kept.value
        return ku[:c].copy(), kv[:c].copy(), kw[:c].copy()

    def read_kept(self):
        kept = ctypes.c_int64()
        cap = self.m
        ku = np.zeros(max(cap, 1), np.int32)
        kv = np.zeros(max(cap, 1), np.int32)
        kw = np.zeros(max(cap, 1), np.float64)
        check(load().ss_graph_read_kept(self._h, cap, _p(ku, ctypes.c_int32), _p(kv, ctypes.c_int32),
                                        _p(kw, ctypes.c_double), ctypes.byref(kept)))
        c = kept.value
        return ku[:c].copy(), kv[:c].copy(), kw[:c].copy()

    def read_block(self, which):
        out = np.zeros((self.n, self.k), np.float64)
        check(load().ss_graph_read_block(self._h, _BLOCKS[which], _p(out, ctypes.c_double)))
        return out

    def profile(self, enable=True):
        check(load().ss_graph_profile(self._h, int(bool(enable))))

    def profile_read(self):
        arr = (_lib.KernelStat * len(_lib.KERNELS))()
        check(load().ss_graph_profile_read(self._h, arr))
        return {name: {"ms_total": arr[i].ms_total, "launches": arr[i].launches}
                for i, name in enumerate(_lib.KERNELS)}

    def schedule_info(self, direction):
        si = _lib.ScheduleInfo()
        check(load().ss_graph_schedule_info(self._h, direction, ctypes.byref(si)))
        return {f: getattr(si, f) for f, _ in _lib.ScheduleInfo._fields_}

    def read_probs(self):
        p = np.zeros(self.m, np.float64)
        check(load().ss_graph_read_probs(self._h, _p(p, ctypes.c_double)))
        return p


class Batch:
    """Many graphs resident as one block-diagonal system.

    ``graphs`` is a list of (n, u, v, w) with canonical edges (u < v sorted by
    (v, u)) and vertex ids local to each graph.  Every stage runs once over
    the whole batch; each graph's results equal those of running it alone.
    """

    def __init__(self, ctx, graphs):
        self.ctx = ctx
        self.count = len(graphs)
        self.n = np.array([int(g[0]) for g in graphs], np.int64)
        self.m = np.array([len(g[1]) for g in graphs], np.int64)
        self.edge_ptr = np.zeros(self.count + 1, np.int64)
        self.edge_ptr[1:] = np.cumsum(self.m)
        self._u = _i32(np.concatenate([np.asarray(g[1]) for g in graphs]) if self.count else [])
        self._v = _i32(np.concatenate([np.asarray(g[2]) for g in graphs]) if self.count else [])
        self._w = _f64(np.concatenate([np.asarray(g[3]) for g in graphs]) if self.count else [])
        h = ctypes.c_void_p()
        check(load().ss_batch_create(ctx._h, self.count, _p(self.n, ctypes.c_int64),
                                     _p(self.edge_ptr, ctypes.c_int64), _p(self._u, ctypes.c_int32),
                                     _p(self._v, ctypes.c_int32), _p(self._w, ctypes.c_double),
                                     ctypes.byref(h)))
        self._h = h
        self.k = None
        ctx._children.add(self)

    def close(self):
        if getattr(self, "_h", None):
            load().ss_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        if not sys.is_finalizing():
            self.close()

    def _index(self, i):
        if not 0 <= i < self.count:
            raise IndexError("graph %d outside the batch of %d" % (i, self.count))
        return i

    def laplacian(self, i):
        """Host CSC Laplacian of graph i (built on the host from its edges)."""
        from .graphs import laplacian_csc
        e0, e1 = self.edge_ptr[i], self.edge_ptr[i + 1]
        return laplacian_csc(int(self.n[i]), self._u[e0:e1], self._v[e0:e1], self._w[e0:e1])

    def set_factors(self, factors):
        arr = (ctypes.c_void_p * self.count)(*[f._h for f in factors])
        check(load().ss_batch_set_factors(self._h, self.count, arr))

    def sparsify(self, k, betas, scales=None, sketch_seeds=None, sampling_seeds=None, uniform=False,
                 solve=None):
        """Sparsify every graph; per-graph lists of beta, scale and seeds."""
        o = (_lib.SparsifyOpts * self.count)()
        sv = solve if solve is not None else solve_opts(stop_rule=_lib.STOP_REL, tol=1e-8)
        for i in range(self.count):
            o[i].k = k
            o[i].uniform = int(bool(uniform))
            o[i].scale = 1.0 / np.sqrt(k) if scales is None else float(scales[i])
            o[i].sketch_seed = 0 if sketch_seeds is None else int(sketch_seeds[i])
            o[i].sampling_seed = 0 if sampling_seeds is None else int(sampling_seeds[i])
            o[i].beta = float(betas[i])
            o[i].solve = sv
        rep = (_lib.SparsifyReport * self.count)()
        check(load().ss_batch_sparsify(self._h, o, rep))
        self.k = k
        return [{f: getattr(rep[i], f) for f, _ in _lib.SparsifyReport._fields_}
                for i in range(self.count)]

    def solve(self, rhs, solve=None):
        """Solve L_i Z_i = rhs[i] for every graph (rhs[i] is n_i x k)."""
        sv = solve if solve is not None else solve_opts(stop_rule=_lib.STOP_REL, tol=1e-8)
        k = int(np.asarray(rhs[0]).shape[1])
        for i in range(self.count):
            y = _f64(np.asarray(rhs[i]).reshape(int(self.n[i]), k))
            check(load().ss_batch_write_rhs(self._h, i, k, sv.precision, _p(y, ctypes.c_double)))
        iters = np.zeros(self.count * k, np.int32)
        relres = np.zeros(self.count * k, np.float64)
        refin = ctypes.c_int32()
        check(load().ss_batch_solve(self._h, k, ctypes.byref(sv), _p(iters, ctypes.c_int32),
                                    _p(relres, ctypes.c_double), ctypes.byref(refin)))
        self.k = k
        return {"iters": iters.reshape(self.count, k), "relres": relres.reshape(self.count, k),
                "refinements": refin.value}

    def read_kept(self, i):
        self._index(i)
        kept = ctypes.c_int64()
        cap = int(self.m[i])
        ku = np.zeros(max(cap, 1), np.int32)
        kv = np.zeros(max(cap, 1), np.int32)
        kw = np.zeros(max(cap, 1), np.float64)
        check(load().ss_batch_read_kept(self._h, i, cap, _p(ku, ctypes.c_int32), _p(kv, ctypes.c_int32),
                                        _p(kw, ctypes.c_double), ctypes.byref(kept)))
        c = kept.value
        return ku[:c].copy(), kv[:c].copy(), kw[:c].copy()

    def read_block(self, i, which):
        self._index(i)
        out = np.zeros((int(self.n[i]), self.k), np.float64)
        check(load().ss_batch_read_block(self._h, i, _BLOCKS[which], _p(out, ctypes.c_double)))
        return out

    def read_probs(self, i):
        self._index(i)
        p = np.zeros(int(self.m[i]), np.float64)
        check(load().ss_batch_read_probs(self._h, i, _p(p, ctypes.c_double)))
        return p

    def profile(self, enable=True):
        check(load().ss_graph_profile(self._h, int(bool(enable))))

    def profile_read(self):
        arr = (_lib.KernelStat * len(_lib.KERNELS))()
        check(load().ss_graph_profile_read(self._h, arr))
        return {name: {"ms_total": arr[i].ms_total, "launches": arr[i].launches}
                for i, name in enumerate(_lib.KERNELS)}

    def schedule_info(self, direction):
        si = _lib.ScheduleInfo()
        check(load().ss_graph_schedule_info(self._h, direction, ctypes.byref(si)))
        return {f: getattr(si, f) for f, _ in _lib.ScheduleInfo._fields_}
