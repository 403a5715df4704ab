"""ctypes binding of libspecspars_hip.so (the C ABI in include/spec_spars_hip.h).

The library is loaded from the package directory only; if it is missing the
import fails loudly -- there is no fallback implementation.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libspecspars_hip.so")

SS_OK = 0
SS_ERR_INVALID = 1
SS_ERR_HIP = 2
SS_ERR_NOT_CONVERGED = 3
SS_ERR_NOMEM = 4
SS_ERR_TIMEOUT = 5
SS_ERR_CAPACITY = 6
PREC_FP64 = 0
PREC_MIXED = 1
MIXED_RELIABLE = 0
MIXED_RESTART = 1
MIXED_RELIABLE_F16 = 2   # reliable updates with FP16 preconditioner rows
STOP_REF_ABS = 0
STOP_REL = 1
ORDER_NATURAL = 0
ORDER_RANDOM = 1
ORDER_ROUNDS = 2

c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_f64p = ctypes.POINTER(ctypes.c_double)
c_u8p = ctypes.POINTER(ctypes.c_uint8)
c_vp = ctypes.c_void_p


class SolveOpts(ctypes.Structure):
    _fields_ = [("factor_seed", ctypes.c_uint64), ("stop_rule", ctypes.c_int32),
                ("max_iter", ctypes.c_int32), ("tol", ctypes.c_double),
                ("fail_relres", ctypes.c_double), ("precision", ctypes.c_int32),
                ("ordering", ctypes.c_int32), ("inner_tol", ctypes.c_double),
                ("mixed_mode", ctypes.c_int32)]


class SolveReport(ctypes.Structure):
    _fields_ = [("iters", c_i32p), ("relres", c_f64p), ("factor_nnz", ctypes.c_int64),
                ("depth_fwd", ctypes.c_int32), ("depth_bwd", ctypes.c_int32),
                ("max_iters", ctypes.c_int32), ("refinements", ctypes.c_int32),
                ("max_relres", ctypes.c_double), ("ms_factor", ctypes.c_double),
                ("ms_solve", ctypes.c_double)]


class SparsifyOpts(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int32), ("uniform", ctypes.c_int32), ("scale", ctypes.c_double),
                ("sketch_seed", ctypes.c_uint64), ("sampling_seed", ctypes.c_uint64),
                ("beta", ctypes.c_double), ("solve", SolveOpts)]


class Tuning(ctypes.Structure):
    _fields_ = [("tile_waves", ctypes.c_int32), ("level_rows_per_wave", ctypes.c_int32),
                ("graph_iters", ctypes.c_int32), ("xcd_slices", ctypes.c_int32)]


class SparsifyReport(ctypes.Structure):
    _fields_ = [("kept", ctypes.c_int64), ("max_iters", ctypes.c_int32),
                ("refinements", ctypes.c_int32), ("max_relres", ctypes.c_double),
                ("ms_sketch", ctypes.c_double), ("ms_solve", ctypes.c_double),
                ("ms_sample", ctypes.c_double), ("ms_total", ctypes.c_double)]


class HostParams(ctypes.Structure):
    _fields_ = [("epsilon", ctypes.c_double), ("beta_constant", ctypes.c_int32),
                ("row_constant", ctypes.c_int32), ("verbose", ctypes.c_int32),
                ("sketch_uniform", ctypes.c_int32), ("jl_factor", ctypes.c_double),
                ("jl_scaling_factor", ctypes.c_double), ("sketch_seed", ctypes.c_uint64),
                ("sampling_seed", ctypes.c_uint64), ("streaming", ctypes.c_int32),
                ("ordering", ctypes.c_int32), ("factor_seed", ctypes.c_uint64),
                ("precision", ctypes.c_int32), ("stop_rule", ctypes.c_int32),
                ("max_iter", ctypes.c_int32), ("resident", ctypes.c_int32), ("tol", ctypes.c_double)]


KERNELS = ["sketch", "spmm", "tri_fwd", "tri_bwd", "er_sample", "tri_fwd_small", "tri_bwd_small", "vector",
           "core_fwd", "core_bwd", "residual"]


class ScheduleInfo(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int32), ("level_launches", ctypes.c_int32),
                ("small_levels", ctypes.c_int32), ("small_runs", ctypes.c_int32),
                ("level_rows", ctypes.c_int64), ("level_nnz", ctypes.c_int64),
                ("small_rows", ctypes.c_int64), ("small_nnz", ctypes.c_int64),
                ("core_rows", ctypes.c_int64), ("core_entries", ctypes.c_int64),
                ("core_outside_nnz", ctypes.c_int64), ("core_steps", ctypes.c_int32),
                ("core_graphs", ctypes.c_int32)]


class KernelStat(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("launches", ctypes.c_int64)]


# name -> (restype, argtypes); every symbol include/spec_spars_hip.h declares
SIGNATURES = {
    "ss_last_error": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "ss_version": (ctypes.c_char_p, []),
    "ss_context_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_vp)]),
    "ss_context_destroy": (ctypes.c_int, [c_vp]),
    "ss_context_mem_info": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_size_t),
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "ss_factor_build": (ctypes.c_int, [ctypes.c_int64, c_i32p, c_i32p, c_f64p, ctypes.c_uint64,
                                       ctypes.POINTER(c_vp)]),
    "ss_factor_build_ordered": (ctypes.c_int, [ctypes.c_int64, c_i32p, c_i32p, c_f64p, ctypes.c_uint64,
                                               ctypes.c_int32, ctypes.POINTER(c_vp)]),
    "ss_factor_order": (ctypes.c_int, [c_vp, c_i32p, c_i32p]),
    "ss_factor_set_threads": (ctypes.c_int, [ctypes.c_int32]),
    "ss_factor_tail": (ctypes.c_int, [c_vp, c_i64p]),
    "ss_factor_info": (ctypes.c_int, [c_vp, c_i64p, c_i64p, c_i32p, c_i32p]),
    "ss_factor_export": (ctypes.c_int, [c_vp, c_i32p, c_i32p, c_f64p, c_f64p]),
    "ss_factor_levels": (ctypes.c_int, [c_vp, c_i32p, c_i32p]),
    "ss_factor_destroy": (ctypes.c_int, [c_vp]),
    "ss_sketch": (ctypes.c_int, [c_vp, c_i32p, c_i32p, c_f64p, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_int32, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                                 c_f64p]),
    "ss_solve_opts_default": (None, [ctypes.POINTER(SolveOpts)]),
    "ss_solve_lap": (ctypes.c_int, [c_vp, c_f64p, ctypes.c_int64, ctypes.c_int32, c_i32p, c_i32p,
                                    c_f64p, ctypes.c_int64, ctypes.POINTER(SolveOpts), c_f64p,
                                    ctypes.POINTER(SolveReport)]),
    "ss_er_sample": (ctypes.c_int, [c_vp, c_f64p, ctypes.c_int64, ctypes.c_int32, c_i32p, c_i32p,
                                    c_f64p, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64,
                                    c_f64p, c_f64p, c_f64p, c_u8p, c_i64p, c_i32p, c_i32p,
                                    c_f64p]),
    "ss_er_sample_noalloc": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int64, ctypes.c_int32, c_vp, c_vp, c_vp,
                                            ctypes.c_int64, ctypes.c_double, ctypes.c_uint64, c_vp, c_vp,
                                            c_vp, c_vp, c_i64p, c_vp, c_vp, c_vp]),
    "ss_graph_create": (ctypes.c_int, [c_vp, ctypes.c_int64, ctypes.c_int64, c_i32p, c_i32p,
                                       c_f64p, ctypes.POINTER(c_vp)]),
    "ss_graph_destroy": (ctypes.c_int, [c_vp]),
    "ss_graph_bytes": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_int32, ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_size_t)]),
    "ss_graph_set_factor": (ctypes.c_int, [c_vp, c_vp]),
    "ss_graph_create_csc": (ctypes.c_int, [c_vp, ctypes.c_int64, c_i32p, c_i32p, c_f64p, ctypes.POINTER(c_vp)]),
    "ss_graph_sample": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, c_i64p]),
    "ss_graph_apply_reweight": (ctypes.c_int, [c_vp, c_f64p]),
    "ss_graph_resident_bytes": (ctypes.c_int, [c_vp, ctypes.POINTER(ctypes.c_size_t)]),
    "ss_graph_work_bytes": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_size_t)]),
    "ss_graph_laplacian": (ctypes.c_int, [c_vp, c_i32p, c_i32p, c_f64p]),
    "ss_graph_sparsify": (ctypes.c_int, [c_vp, ctypes.POINTER(SparsifyOpts),
                                         ctypes.POINTER(SparsifyReport)]),
    "ss_graph_read_kept": (ctypes.c_int, [c_vp, ctypes.c_int64, c_i32p, c_i32p, c_f64p, c_i64p]),
    "ss_graph_sparsify_many": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.POINTER(SparsifyOpts),
                                              ctypes.POINTER(SparsifyReport)]),
    "ss_graph_read_kept_many": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_int64, c_i32p, c_i32p, c_f64p,
                                               c_i64p]),
    "ss_graph_read_block": (ctypes.c_int, [c_vp, ctypes.c_int, c_f64p]),
    "ss_graph_read_probs": (ctypes.c_int, [c_vp, c_f64p]),
    "ss_graph_profile": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "ss_graph_profile_read": (ctypes.c_int, [c_vp, ctypes.POINTER(KernelStat)]),
    "ss_graph_schedule_info": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.POINTER(ScheduleInfo)]),
    "ss_batch_create": (ctypes.c_int, [c_vp, ctypes.c_int32, c_i64p, c_i64p, c_i32p, c_i32p,
                                       c_f64p, ctypes.POINTER(c_vp)]),
    "ss_batch_count": (ctypes.c_int, [c_vp, c_i32p]),
    "ss_batch_set_factors": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.POINTER(c_vp)]),
    "ss_batch_sparsify": (ctypes.c_int, [c_vp, ctypes.POINTER(SparsifyOpts),
                                         ctypes.POINTER(SparsifyReport)]),
    "ss_batch_write_rhs": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          c_f64p]),
    "ss_batch_solve": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.POINTER(SolveOpts), c_i32p,
                                      c_f64p, c_i32p]),
    "ss_batch_read_kept": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_int64, c_i32p, c_i32p,
                                          c_f64p, c_i64p]),
    "ss_batch_read_block": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.c_int, c_f64p]),
    "ss_batch_read_probs": (ctypes.c_int, [c_vp, ctypes.c_int32, c_f64p]),
    "ss_context_get_tuning": (ctypes.c_int, [c_vp, ctypes.POINTER(Tuning)]),
    "ss_context_set_tuning": (ctypes.c_int, [c_vp, ctypes.POINTER(Tuning)]),
    "ss_context_autotune": (ctypes.c_int, [c_vp, c_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                           ctypes.POINTER(Tuning), c_f64p, c_f64p]),
    "ss_bench_copy": (ctypes.c_int, [c_vp, ctypes.c_size_t, ctypes.c_int, c_f64p]),
    # host layer (include/spec_spars_host.h)
    "ssh_params_default": (None, [ctypes.POINTER(HostParams), ctypes.c_int]),
    "ssh_sparsifier_create": (ctypes.c_int, [c_vp, ctypes.c_int32, ctypes.POINTER(HostParams),
                                             ctypes.POINTER(c_vp)]),
    "ssh_sparsifier_destroy": (ctypes.c_int, [c_vp]),
    "ssh_sparsifier_insert_many": (ctypes.c_int, [c_vp, ctypes.c_int64, c_i32p, c_i32p, c_f64p]),
    "ssh_sparsifier_form_laplacian": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "ssh_sparsifier_check_diagonal": (ctypes.c_int, [c_vp]),
    "ssh_sparsifier_num_edges": (ctypes.c_int, [c_vp, c_i64p]),
    "ssh_sparsifier_sparsify": (ctypes.c_int, [c_vp, ctypes.c_int, c_i64p, c_i64p]),
    "ssh_sparsifier_sample_apply": (ctypes.c_int, [c_vp, c_f64p, ctypes.c_int64, c_i64p, c_i64p]),
    "ssh_sparsifier_laplacian": (ctypes.c_int, [c_vp, c_i64p, c_i32p, c_i32p, c_f64p]),
    "ssh_sparsifier_info": (ctypes.c_int, [c_vp, c_f64p, c_i32p, c_i64p]),
    "ssh_sparsifier_sketch": (ctypes.c_int, [c_vp, c_f64p]),
    "ssh_sparsifier_triggered": (ctypes.c_int, [c_vp, c_i64p]),
    "ssh_mtx_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(c_vp)]),
    "ssh_matrix_info": (ctypes.c_int, [c_vp, c_i32p, c_i32p, c_i64p]),
    "ssh_matrix_csc": (ctypes.c_int, [c_vp, c_i32p, c_i32p, c_f64p]),
    "ssh_matrix_destroy": (ctypes.c_int, [c_vp]),
    "ssh_mtx_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, c_i32p, c_i32p, c_f64p]),
    "ssh_stream_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(c_vp)]),
    "ssh_stream_info": (ctypes.c_int, [c_vp, c_i32p, c_i64p, c_i32p]),
    "ssh_stream_run": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(HostParams), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_char_p, ctypes.POINTER(c_vp), c_i64p, c_i64p]),
    "ssh_stream_destroy": (ctypes.c_int, [c_vp]),
}

_lib = None


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libspecspars_hip.so is not built; run __graft_entry__.build() "
            "(or python -m spec_spars_amd.build) first")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SSError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("status %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != SS_OK:
        buf = ctypes.create_string_buffer(1024)
        load().ss_last_error(buf, 1024)
        raise SSError(rc, buf.value.decode(errors="replace"))
