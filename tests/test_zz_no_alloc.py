"""ss_er_sample_noalloc (device arrays, no device allocation inside the
call, look-back words in a fixed per-device array) against ss_er_sample
(host arrays) through the C ABI: bit-identical probabilities, distances,
mask and kept edges on two calls in a row, the second smaller, so the
fixed array's leftover contents from the first call must not leak into the
second; p_in mode; the size limit and the device-side edge checks."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(kind, seed):
    from spec_spars_amd import graphs
    if kind == "grid3d":
        n, u, v = graphs.grid3d(48)
    else:
        n, u, v = graphs.grid2d(100, 80)
    w = graphs.loguniform_weights(len(u), seed=seed)
    return n, np.asarray(u, np.int32), np.asarray(v, np.int32), -w


class _Hip:
    """Device buffers for the test through the HIP runtime the library links
    (the library and the test share one runtime in the process)."""

    def __init__(self):
        self.rt = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
        self.bufs = []

    def _ok(self, rc, what):
        assert rc == 0, "%s failed: %d" % (what, rc)

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self._ok(self.rt.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 8))), "hipMalloc")
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        self._ok(self.rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1), "hipMemcpy")
        return p

    def filled(self, count, dtype, value):
        return self.put(np.full(max(count, 1), value, dtype))

    def get(self, p, count, dtype):
        a = np.empty(count, dtype)
        if count:
            self._ok(self.rt.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(a.nbytes), 2),
                     "hipMemcpy")
        return a

    def free(self):
        self._ok(self.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        for p in self.bufs:
            self._ok(self.rt.hipFree(p), "hipFree")
        self.bufs = []


def _noalloc(ctx, n, u, v, lval, k, beta, seed, Z=None, p_in=None):
    from spec_spars_amd import _lib
    from spec_spars_amd.api import check
    m = len(u)
    h = _Hip()
    try:
        du, dv, dl = h.put(u), h.put(v), h.put(lval)
        dZ = h.put(Z) if Z is not None else None
        dpin = h.put(p_in) if p_in is not None else None
        # outputs start as sentinels, so an entry left unwritten shows
        p = h.filled(m, np.float64, -7.0)
        d = h.filled(m, np.float64, -7.0)
        mask = h.filled(m, np.uint8, 9)
        ku = h.filled(m, np.int32, -1)
        kv = h.filled(m, np.int32, -1)
        kw = h.filled(m, np.float64, 0.0)
        h._ok(h.rt.hipDeviceSynchronize(), "hipDeviceSynchronize")
        kept = ctypes.c_int64(-1)
        check(_lib.load().ss_er_sample_noalloc(ctx._h, dZ, n, k, du, dv, dl, m, beta, seed, dpin, p, d, mask,
                                               ctypes.byref(kept), ku, kv, kw))
        c = kept.value
        return {"p": h.get(p, m, np.float64), "d": h.get(d, m, np.float64) if p_in is None else None,
                "mask": h.get(mask, m, np.uint8), "kept": c, "ku": h.get(ku, c, np.int32),
                "kv": h.get(kv, c, np.int32), "kw": h.get(kw, c, np.float64)}
    finally:
        h.free()


def _same(a, b):
    assert a["kept"] == b["kept"] and 0 < a["kept"]
    for f in ("p", "d", "kw"):
        if a[f] is not None:
            assert np.array_equal(a[f].view(np.uint64), b[f].view(np.uint64)), f
    for f in ("mask", "ku", "kv"):
        assert np.array_equal(a[f], b[f]), f


def test_noalloc_matches_er_sample_twice_then_smaller(ctx):
    runs = [("grid3d", 24, 3, 0.8), ("grid2d", 8, 11, 2.5)]      # second call: fewer edges and tiles
    for kind, k, seed, beta in runs:
        n, u, v, lval = _case(kind, seed)
        Z = np.random.default_rng(seed).standard_normal((n, k))
        print("[no_alloc] %s n=%d m=%d k=%d seed=%d" % (kind, n, len(u), k, seed), flush=True)
        ref = ctx.er_sample(u, v, lval, k, beta, seed, Z=Z)
        got = _noalloc(ctx, n, u, v, lval, k, beta, seed, Z=Z)
        assert 0 < (ref["p"] < 1).sum() < len(u)        # both branches of min(1, .)
        _same(got, ref)


def test_noalloc_probabilities_in(ctx):
    n, u, v, lval = _case("grid2d", 5)
    p_in = np.random.default_rng(5).uniform(0.0, 1.0, len(u))
    p_in[::7] = 1.0
    p_in[3::11] = 0.0
    print("[no_alloc] p_in n=%d m=%d" % (n, len(u)), flush=True)
    ref = ctx.er_sample(u, v, lval, 1, 1.0, 21, p_in=p_in, n=n)
    got = _noalloc(ctx, n, u, v, lval, 1, 1.0, 21, p_in=p_in)
    _same(got, ref)


def test_noalloc_rejects_oversize_and_bad_edges(ctx):
    from spec_spars_amd import _lib
    lib = _lib.load()
    h = _Hip()
    try:
        q = h.alloc(128)
        kept = ctypes.c_int64(-1)
        # more edges than the fixed array covers: refused before any device work
        rc = lib.ss_er_sample_noalloc(ctx._h, q, 4, 1, q, q, q, (1 << 26) + 1, 1.0, 0, None, None, None, None,
                                      ctypes.byref(kept), q, q, q)
    finally:
        h.free()
    assert rc == _lib.SS_ERR_INVALID and kept.value == -1
    buf = ctypes.create_string_buffer(512)
    lib.ss_last_error(buf, 512)
    assert b"allocation-free maximum" in buf.value
    # an edge with v >= n is flagged on the device and Z is never read for it
    n, u, v, lval = _case("grid2d", 1)
    v = v.copy()
    v[17] = n + 5
    Z = np.zeros((n, 4))
    with pytest.raises(Exception, match="u < v < n"):
        _noalloc(ctx, n, u, v, lval, 4, 1.0, 0, Z=Z)
    # and the next call is clean again
    n, u, v, lval = _case("grid2d", 2)
    Z = np.random.default_rng(2).standard_normal((n, 4))
    _same(_noalloc(ctx, n, u, v, lval, 4, 2.0, 4, Z=Z), ctx.er_sample(u, v, lval, 4, 2.0, 4, Z=Z))
