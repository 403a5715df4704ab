"""Python view of the C++ host layer (include/spec_spars.hpp): the
reference's SparsifierParameters / Sparsifier interface (src/sparsifier.rs),
MatrixMarket files (src/utils.rs:149-191) and InputStream (src/stream.rs)
with the GPU path underneath.  Method names and argument meanings follow the
Rust code; its panics surface as SSError."""
import ctypes
import sys

import numpy as np

from . import _lib
from ._lib import check, load


class SparsifierParameters:
    """SparsifierParameters (src/sparsifier.rs:146-210)."""

    FIELDS = [f for f, _ in _lib.HostParams._fields_]

    def __init__(self, **kw):
        p = _lib.HostParams()
        load().ssh_params_default(ctypes.byref(p), 0)
        for f in self.FIELDS:
            setattr(self, f, getattr(p, f))
        for k, v in kw.items():
            if k not in self.FIELDS:
                raise TypeError("unknown parameter " + k)
            setattr(self, k, v)

    @classmethod
    def new_default(cls, bench=False):
        return cls()

    def _c(self):
        p = _lib.HostParams()
        for f in self.FIELDS:
            setattr(p, f, getattr(self, f))
        return p


class Sparsifier:
    """Sparsifier (src/sparsifier.rs:240-826); ctx=None allows the host-only
    methods without a GPU."""

    def __init__(self, num_nodes, parameters, ctx=None, _handle=None):
        if _handle is None:
            h = ctypes.c_void_p()
            self._params = parameters._c()
            check(load().ssh_sparsifier_create(ctx._h if ctx is not None else None, int(num_nodes),
                                               ctypes.byref(self._params), ctypes.byref(h)))
        else:
            h = _handle
        self._h = h
        self.num_nodes = int(num_nodes)
        b, k, t = ctypes.c_double(), ctypes.c_int32(), ctypes.c_int64()
        check(load().ssh_sparsifier_info(self._h, ctypes.byref(b), ctypes.byref(k), ctypes.byref(t)))
        self.beta, self.jl_dim, self.threshold = b.value, k.value, t.value

    def close(self):
        if getattr(self, "_h", None):
            load().ssh_sparsifier_destroy(self._h)
            self._h = None

    def __del__(self):
        # at interpreter exit the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        if not sys.is_finalizing():
            self.close()

    def insert(self, v1, v2, value):
        self.insert_many([v1], [v2], [value])

    def insert_many(self, v1, v2, values):
        a = np.ascontiguousarray(v1, np.int32)
        b = np.ascontiguousarray(v2, np.int32)
        c = np.ascontiguousarray(values, np.float64)
        check(load().ssh_sparsifier_insert_many(self._h, len(a), a.ctypes.data_as(_lib.c_i32p),
                                                b.ctypes.data_as(_lib.c_i32p), c.ctypes.data_as(_lib.c_f64p)))

    def form_laplacian(self, check_=False):
        check(load().ssh_sparsifier_form_laplacian(self._h, int(check_)))

    def check_diagonal(self):
        check(load().ssh_sparsifier_check_diagonal(self._h))

    def num_edges(self):
        x = ctypes.c_int64()
        check(load().ssh_sparsifier_num_edges(self._h, ctypes.byref(x)))
        return x.value

    def sparsify(self, check_=False):
        """Returns SparsificationStats as (original_num_edges, final_num_edges)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(load().ssh_sparsifier_sparsify(self._h, int(check_), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def triggered(self):
        """Passes the streaming trigger ran (src/sparsifier.rs:370-374)."""
        x = ctypes.c_int64()
        check(load().ssh_sparsifier_triggered(self._h, ctypes.byref(x)))
        return x.value

    def sample_and_apply(self, probs):
        """sample_and_reweight(probs) then apply_reweightings(report=false)."""
        p = np.ascontiguousarray(probs, np.float64)
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(load().ssh_sparsifier_sample_apply(self._h, p.ctypes.data_as(_lib.c_f64p), len(p),
                                                 ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def current_laplacian(self):
        nnz = ctypes.c_int64()
        check(load().ssh_sparsifier_laplacian(self._h, ctypes.byref(nnz), None, None, None))
        cp = np.zeros(self.num_nodes + 1, np.int32)
        ri = np.zeros(nnz.value, np.int32)
        va = np.zeros(nnz.value, np.float64)
        check(load().ssh_sparsifier_laplacian(self._h, None, cp.ctypes.data_as(_lib.c_i32p),
                                              ri.ctypes.data_as(_lib.c_i32p), va.ctypes.data_as(_lib.c_f64p)))
        return cp, ri, va

    def jl_sketch(self):
        Y = np.zeros((self.num_nodes, self.jl_dim), np.float64)
        check(load().ssh_sparsifier_sketch(self._h, Y.ctypes.data_as(_lib.c_f64p)))
        return Y


def read_mtx(path):
    """read_mtx (src/utils.rs:149-154): (rows, cols, colptr, rowidx, vals),
    CSC with rows sorted in each column; symmetric storage mirrored,
    duplicates summed, explicit zeros kept."""
    h = ctypes.c_void_p()
    check(load().ssh_mtx_read(str(path).encode(), ctypes.byref(h)))
    try:
        r, c, nnz = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        check(load().ssh_matrix_info(h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(nnz)))
        cp = np.zeros(c.value + 1, np.int32)
        ri = np.zeros(nnz.value, np.int32)
        va = np.zeros(nnz.value, np.float64)
        check(load().ssh_matrix_csc(h, cp.ctypes.data_as(_lib.c_i32p), ri.ctypes.data_as(_lib.c_i32p),
                                    va.ctypes.data_as(_lib.c_f64p)))
        return r.value, c.value, cp, ri, va
    finally:
        load().ssh_matrix_destroy(h)


def write_mtx(path, colptr, rowidx, vals, rows=None):
    """write_mtx (src/utils.rs:189-191): coordinate real general, CSC order."""
    cp = np.ascontiguousarray(colptr, np.int32)
    ri = np.ascontiguousarray(rowidx, np.int32)
    va = np.ascontiguousarray(vals, np.float64)
    cols = len(cp) - 1
    check(load().ssh_mtx_write(str(path).encode(), cols if rows is None else int(rows), cols,
                               cp.ctypes.data_as(_lib.c_i32p), ri.ctypes.data_as(_lib.c_i32p),
                               va.ctypes.data_as(_lib.c_f64p)))


class InputStream:
    """InputStream (src/stream.rs:8-64): a square MatrixMarket file, its
    stored diagonal zeroed; run_stream feeds its upper triangle in CSC order
    to a new Sparsifier and sparsifies."""

    def __init__(self, filename, dataset_name=""):
        h = ctypes.c_void_p()
        check(load().ssh_stream_open(str(filename).encode(), str(dataset_name).encode(), ctypes.byref(h)))
        self._h = h
        self.dataset_name = dataset_name
        n, st, dz = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        check(load().ssh_stream_info(h, ctypes.byref(n), ctypes.byref(st), ctypes.byref(dz)))
        self.num_nodes, self.stored, self.diag_zeros = n.value, st.value, dz.value

    def run_stream(self, parameters, test=False, writeout=False, ctx=None, output_file=None):
        """Returns (Sparsifier, (original_num_edges, final_num_edges))."""
        p = parameters._c()
        h = ctypes.c_void_p()
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(load().ssh_stream_run(self._h, ctx._h if ctx is not None else None, ctypes.byref(p), int(test),
                                    int(writeout), output_file.encode() if output_file else None, ctypes.byref(h),
                                    ctypes.byref(a), ctypes.byref(b)))
        return Sparsifier(self.num_nodes, parameters, _handle=h), (a.value, b.value)

    def close(self):
        if getattr(self, "_h", None):
            load().ssh_stream_destroy(self._h)
            self._h = None

    def __del__(self):
        if not sys.is_finalizing():
            self.close()
