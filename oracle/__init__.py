"""Test-only binding of the CPU oracle (oracle.c).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker / CPU baseline.  The shipped library never loads it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        import subprocess
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c99", "-o", LIB,
                        os.path.join(_HERE, "oracle.c"), "-lm"], check=True)
    L = ctypes.CDLL(LIB)
    sig = {
        "or_hash128to64": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
        "or_city64_seed16": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]),
        "or_sketch_value": (ctypes.c_double, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double,
                                              ctypes.c_int, ctypes.c_int]),
        "or_sketch": (None, [ctypes.c_int64, _i32p, _i32p, _f64p, ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64,
                             ctypes.c_double, ctypes.c_int, ctypes.c_int, _f64p]),
        "or_sketch_edges": (None, [ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), _i32p, _i32p, _f64p,
                                   ctypes.c_int64, ctypes.c_int32, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, _f64p]),
        "or_permute_csc": (None, [ctypes.c_int64, _i32p, _i32p, _f64p, _i32p, _i32p, _i32p, _f64p]),
        "or_laplacian": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p]),
        "or_upper_edges": (ctypes.c_int64, [ctypes.c_int64, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p]),
        "or_levels_fwd": (None, [ctypes.c_int64, _i32p, _i32p, _i32p]),
        "or_levels_bwd": (None, [ctypes.c_int64, _i32p, _i32p, _i32p]),
        "or_pcg": (ctypes.c_int, [ctypes.c_int64, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p]),
        "or_precond_apply": (None, [ctypes.c_int64, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]),
        "or_solve_lap": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _f64p, _i32p, _i32p, _f64p, _i32p, _i32p,
                                        _f64p, _f64p, ctypes.c_int, ctypes.c_double, ctypes.c_int, _f64p, _i32p,
                                        _f64p]),
        "or_diff_norms": (None, [ctypes.c_int64, _i32p, _i32p, ctypes.c_int32, _f64p, _f64p]),
        "or_probs": (None, [ctypes.c_int64, _f64p, _f64p, ctypes.c_double, ctypes.c_int32, _f64p]),
        "or_chacha_block": (None, [_u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _u32p]),
        "or_stdrng_seed": (None, [ctypes.c_uint64, _u32p]),
        "or_coins": (None, [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, _f64p]),
        "or_sample_reweight": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, _i32p, _i32p, _f64p, _f64p, _f64p,
                                                _u8p, _i32p, _i32p, _f64p, _f64p]),
    }
    for k, (r, a) in sig.items():
        f = getattr(L, k)
        f.restype = r
        f.argtypes = a
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t)) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def city64_seed16(e, j, seed, variant=0):
    return load().or_city64_seed16(e, j, seed, variant)


def hash128to64(lo, hi):
    return load().or_hash128to64(lo, hi)


def sketch_value(e, j, seed, scale=1.0, uniform=False, variant=0):
    return load().or_sketch_value(e, j, seed, scale, int(uniform), variant)


def sketch(v1, v2, w, n, k, seed=0, scale=1.0, uniform=False, variant=0):
    v1, v2, w = _c(v1, np.int32), _c(v2, np.int32), _c(w, np.float64)
    Y = np.zeros((n, k), np.float64)
    load().or_sketch(len(w), _p(v1, ctypes.c_int32), _p(v2, ctypes.c_int32), _p(w, ctypes.c_double), n, k, seed,
                     scale, int(uniform), variant, _p(Y, ctypes.c_double))
    return Y


def sketch_edges(eid, v1, v2, w, n, k, seed=0, scale=1.0, uniform=False, variant=0, Y=None):
    """Sketch of the edges with ids eid (ascending; v1, v2, w are theirs),
    added into Y (n x k, zeros if None).  Rows whose incident edges are all
    in the subset get their full sketch rows."""
    eid = _c(eid, np.int64)
    v1, v2, w = _c(v1, np.int32), _c(v2, np.int32), _c(w, np.float64)
    if Y is None:
        Y = np.zeros((n, k), np.float64)
    load().or_sketch_edges(len(eid), eid.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _p(v1, ctypes.c_int32),
                           _p(v2, ctypes.c_int32), _p(w, ctypes.c_double), n, k, seed, scale, int(uniform),
                           variant, _p(Y, ctypes.c_double))
    return Y


def permute_csc(cp, ri, va, order):
    """The CSC matrix in the position space of an elimination order
    (position p holds vertex order[p]): P A P^T, rows sorted."""
    n = len(cp) - 1
    cp, ri, va = _c(cp, np.int32), _c(ri, np.int32), _c(va, np.float64)
    rank = np.empty(n, np.int32)
    rank[np.asarray(order)] = np.arange(n, dtype=np.int32)
    pc = np.zeros(n + 1, np.int32)
    pr = np.zeros(len(ri), np.int32)
    pv = np.zeros(len(va), np.float64)
    load().or_permute_csc(n, _p(cp, ctypes.c_int32), _p(ri, ctypes.c_int32), _p(va, ctypes.c_double),
                          _p(rank, ctypes.c_int32), _p(pc, ctypes.c_int32), _p(pr, ctypes.c_int32),
                          _p(pv, ctypes.c_double))
    return pc, pr, pv


def laplacian(n, v1, v2, w):
    v1, v2, w = _c(v1, np.int32), _c(v2, np.int32), _c(w, np.float64)
    m = len(w)
    cp = np.zeros(n + 1, np.int32)
    ri = np.zeros(n + 2 * m, np.int32)
    va = np.zeros(n + 2 * m, np.float64)
    nnz = load().or_laplacian(n, m, _p(v1, ctypes.c_int32), _p(v2, ctypes.c_int32), _p(w, ctypes.c_double),
                              _p(cp, ctypes.c_int32), _p(ri, ctypes.c_int32), _p(va, ctypes.c_double))
    if nnz < 0:
        raise ValueError("bad vertex id")
    return cp, ri[:nnz].copy(), va[:nnz].copy()


def upper_edges(n, cp, ri, va):
    cp, ri, va = _c(cp, np.int32), _c(ri, np.int32), _c(va, np.float64)
    u = np.zeros(len(ri), np.int32)
    v = np.zeros(len(ri), np.int32)
    lv = np.zeros(len(ri), np.float64)
    m = load().or_upper_edges(n, _p(cp, ctypes.c_int32), _p(ri, ctypes.c_int32), _p(va, ctypes.c_double),
                              _p(u, ctypes.c_int32), _p(v, ctypes.c_int32), _p(lv, ctypes.c_double))
    return u[:m].copy(), v[:m].copy(), lv[:m].copy()


def levels(n, urp, uci):
    urp, uci = _c(urp, np.int32), _c(uci, np.int32)
    f = np.zeros(n, np.int32)
    b = np.zeros(n, np.int32)
    load().or_levels_fwd(n, _p(urp, ctypes.c_int32), _p(uci, ctypes.c_int32), _p(f, ctypes.c_int32))
    load().or_levels_bwd(n, _p(urp, ctypes.c_int32), _p(uci, ctypes.c_int32), _p(b, ctypes.c_int32))
    return f, b


def precond_apply(n, urp, uci, uva, diag, r):
    urp, uci, uva, diag, r = (_c(urp, np.int32), _c(uci, np.int32), _c(uva, np.float64), _c(diag, np.float64),
                              _c(r, np.float64))
    z = np.zeros(n, np.float64)
    load().or_precond_apply(n, _p(urp, ctypes.c_int32), _p(uci, ctypes.c_int32), _p(uva, ctypes.c_double),
                            _p(diag, ctypes.c_double), _p(r, ctypes.c_double), _p(z, ctypes.c_double))
    return z


def solve_lap(Y, cp, ri, va, urp, uci, uva, diag, stop_rule=0, tol=2e-7, max_iter=400):
    Y = _c(Y, np.float64)
    n, k = Y.shape
    cp, ri, va = _c(cp, np.int32), _c(ri, np.int32), _c(va, np.float64)
    urp, uci, uva, diag = _c(urp, np.int32), _c(uci, np.int32), _c(uva, np.float64), _c(diag, np.float64)
    Z = np.zeros((n, k), np.float64)
    it = np.zeros(k, np.int32)
    rr = np.zeros(k, np.float64)
    load().or_solve_lap(n, k, _p(Y, ctypes.c_double), _p(cp, ctypes.c_int32), _p(ri, ctypes.c_int32),
                        _p(va, ctypes.c_double), _p(urp, ctypes.c_int32), _p(uci, ctypes.c_int32),
                        _p(uva, ctypes.c_double), _p(diag, ctypes.c_double), stop_rule, tol, max_iter,
                        _p(Z, ctypes.c_double), _p(it, ctypes.c_int32), _p(rr, ctypes.c_double))
    return Z, it, rr


def diff_norms(u, v, Z):
    u, v, Z = _c(u, np.int32), _c(v, np.int32), _c(Z, np.float64)
    d = np.zeros(len(u), np.float64)
    load().or_diff_norms(len(u), _p(u, ctypes.c_int32), _p(v, ctypes.c_int32), Z.shape[1], _p(Z, ctypes.c_double),
                         _p(d, ctypes.c_double))
    return d


def probs(lval, d, beta, k):
    lval, d = _c(lval, np.float64), _c(d, np.float64)
    p = np.zeros(len(d), np.float64)
    load().or_probs(len(d), _p(lval, ctypes.c_double), _p(d, ctypes.c_double), beta, k, _p(p, ctypes.c_double))
    return p


def chacha_block(key, counter, stream, rounds):
    key = _c(key, np.uint32)
    out = np.zeros(16, np.uint32)
    load().or_chacha_block(_p(key, ctypes.c_uint32), counter, stream, rounds, _p(out, ctypes.c_uint32))
    return out


def stdrng_key(seed):
    key = np.zeros(8, np.uint32)
    load().or_stdrng_seed(seed, _p(key, ctypes.c_uint32))
    return key


def coins(seed, count, first=0):
    c = np.zeros(count, np.float64)
    load().or_coins(seed, first, count, _p(c, ctypes.c_double))
    return c


def sample_reweight(n, u, v, lval, p, c):
    u, v, lval, p, c = _c(u, np.int32), _c(v, np.int32), _c(lval, np.float64), _c(p, np.float64), _c(c, np.float64)
    m = len(u)
    mask = np.zeros(m, np.uint8)
    ku = np.zeros(max(m, 1), np.int32)
    kv = np.zeros(max(m, 1), np.int32)
    kw = np.zeros(max(m, 1), np.float64)
    dd = np.zeros(n, np.float64)
    kept = load().or_sample_reweight(m, n, _p(u, ctypes.c_int32), _p(v, ctypes.c_int32), _p(lval, ctypes.c_double),
                                     _p(p, ctypes.c_double), _p(c, ctypes.c_double), _p(mask, ctypes.c_uint8),
                                     _p(ku, ctypes.c_int32), _p(kv, ctypes.c_int32), _p(kw, ctypes.c_double),
                                     _p(dd, ctypes.c_double))
    return {"mask": mask, "kept": kept, "ku": ku[:kept].copy(), "kv": kv[:kept].copy(), "kw": kw[:kept].copy(),
            "ddiag": dd}
