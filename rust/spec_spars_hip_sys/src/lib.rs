//! Raw bindings of the C ABI in `include/spec_spars_hip.h` (struct layouts
//! and signatures field for field; every function returns an `SS_*` status,
//! the message of a failure is `ss_last_error`).
#![allow(non_camel_case_types)]
use std::os::raw::{c_char, c_int};

pub const SS_OK: c_int = 0;
pub const SS_ERR_INVALID: c_int = 1;
pub const SS_ERR_HIP: c_int = 2;
pub const SS_ERR_NOT_CONVERGED: c_int = 3;
pub const SS_ERR_NOMEM: c_int = 4;
pub const SS_ERR_TIMEOUT: c_int = 5;
pub const SS_ERR_CAPACITY: c_int = 6;
pub const SS_PREC_FP64: i32 = 0;
pub const SS_PREC_MIXED: i32 = 1;
pub const SS_STOP_REF_ABS: i32 = 0;
pub const SS_STOP_REL: i32 = 1;
pub const SS_ORDER_NATURAL: i32 = 0;
pub const SS_ORDER_RANDOM: i32 = 1;
pub const SS_ORDER_ROUNDS: i32 = 2;
pub const SS_MIXED_RELIABLE: i32 = 0;
pub const SS_MIXED_RESTART: i32 = 1;
pub const SS_MIXED_RELIABLE_F16: i32 = 2;

#[repr(C)]
pub struct ss_context {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ss_factor {
    _p: [u8; 0],
}
#[repr(C)]
pub struct ss_graph {
    _p: [u8; 0],
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ss_solve_opts {
    pub factor_seed: u64,
    pub stop_rule: i32,
    pub max_iter: i32,
    pub tol: f64,
    pub fail_relres: f64,
    pub precision: i32,
    pub ordering: i32,
    pub inner_tol: f64,
    pub mixed_mode: i32,
}

#[repr(C)]
pub struct ss_solve_report {
    pub iters: *mut i32,
    pub relres: *mut f64,
    pub factor_nnz: i64,
    pub depth_fwd: i32,
    pub depth_bwd: i32,
    pub max_iters: i32,
    pub refinements: i32,
    pub max_relres: f64,
    pub ms_factor: f64,
    pub ms_solve: f64,
}

#[repr(C)]
#[derive(Clone, Copy, Debug)]
pub struct ss_sparsify_opts {
    pub k: i32,
    pub uniform: i32,
    pub scale: f64,
    pub sketch_seed: u64,
    pub sampling_seed: u64,
    pub beta: f64,
    pub solve: ss_solve_opts,
}

#[repr(C)]
#[derive(Clone, Copy, Debug, Default)]
pub struct ss_sparsify_report {
    pub kept: i64,
    pub max_iters: i32,
    pub refinements: i32,
    pub max_relres: f64,
    pub ms_sketch: f64,
    pub ms_solve: f64,
    pub ms_sample: f64,
    pub ms_total: f64,
}

#[link(name = "specspars_hip")]
extern "C" {
    pub fn ss_last_error(buf: *mut c_char, len: usize) -> c_int;
    pub fn ss_version() -> *const c_char;

    pub fn ss_context_create(device: c_int, out: *mut *mut ss_context) -> c_int;
    pub fn ss_context_destroy(ctx: *mut ss_context) -> c_int;

    pub fn ss_factor_set_threads(threads: i32) -> c_int;
    pub fn ss_factor_build_ordered(n: i64, colptr: *const i32, rowidx: *const i32, vals: *const f64,
                                   seed: u64, ordering: i32, out: *mut *mut ss_factor) -> c_int;
    pub fn ss_factor_destroy(f: *mut ss_factor) -> c_int;

    // src/sparsifier.rs:490-539
    pub fn ss_sketch(ctx: *mut ss_context, v1: *const i32, v2: *const i32, w: *const f64, m: i64, n: i64,
                     k: i32, sketch_seed: u64, scale: f64, uniform: c_int, y: *mut f64) -> c_int;
    // src/lib.rs:103, src/example.cc:668-703
    pub fn ss_solve_opts_default(o: *mut ss_solve_opts);
    pub fn ss_solve_lap(ctx: *mut ss_context, y: *const f64, n: i64, k: i32, colptr: *const i32,
                        rowidx: *const i32, vals: *const f64, nnz: i64, opts: *const ss_solve_opts, z: *mut f64,
                        rep: *mut ss_solve_report) -> c_int;
    // src/sparsifier.rs:542-680
    pub fn ss_er_sample(ctx: *mut ss_context, z: *const f64, n: i64, k: i32, u: *const i32, v: *const i32,
                        lval: *const f64, m: i64, beta: f64, sampling_seed: u64, p_in: *const f64,
                        p_out: *mut f64, d_out: *mut f64, mask_out: *mut u8, kept: *mut i64, ku: *mut i32,
                        kv: *mut i32, kw: *mut f64) -> c_int;
    // no counterpart in the reference: ss_er_sample over device pointers,
    // allocating nothing (at most SS_NOALLOC_MAX_EDGES = 2^26 edges)
    pub fn ss_er_sample_noalloc(ctx: *mut ss_context, z: *const f64, n: i64, k: i32, u: *const i32,
                                v: *const i32, lval: *const f64, m: i64, beta: f64, sampling_seed: u64,
                                p_in: *const f64, p_out: *mut f64, d_out: *mut f64, mask_out: *mut u8,
                                kept: *mut i64, ku: *mut i32, kv: *mut i32, kw: *mut f64) -> c_int;

    // resident graphs: src/sparsifier.rs:722-770 on the device
    pub fn ss_graph_create(ctx: *mut ss_context, n: i64, m: i64, u: *const i32, v: *const i32, w: *const f64,
                           out: *mut *mut ss_graph) -> c_int;
    pub fn ss_graph_create_csc(ctx: *mut ss_context, n: i64, colptr: *const i32, rowidx: *const i32,
                               vals: *const f64, out: *mut *mut ss_graph) -> c_int;
    pub fn ss_graph_destroy(g: *mut ss_graph) -> c_int;
    pub fn ss_graph_set_factor(g: *mut ss_graph, f: *const ss_factor) -> c_int;
    pub fn ss_graph_sparsify(g: *mut ss_graph, o: *const ss_sparsify_opts, rep: *mut ss_sparsify_report) -> c_int;
    pub fn ss_graph_sparsify_many(g: *mut ss_graph, count: i32, opts: *const ss_sparsify_opts,
                                  rep: *mut ss_sparsify_report) -> c_int;
    pub fn ss_graph_read_kept(g: *const ss_graph, cap: i64, ku: *mut i32, kv: *mut i32, kw: *mut f64,
                              kept: *mut i64) -> c_int;
    pub fn ss_graph_read_kept_many(g: *const ss_graph, r: i32, cap: i64, ku: *mut i32, kv: *mut i32,
                                   kw: *mut f64, kept: *mut i64) -> c_int;
    pub fn ss_batch_write_rhs(g: *mut ss_graph, i: i32, k: i32, precision: i32, y: *const f64) -> c_int;
    pub fn ss_batch_solve(g: *mut ss_graph, k: i32, opts: *const ss_solve_opts, iters: *mut i32,
                          relres: *mut f64, refinements: *mut i32) -> c_int;
    pub fn ss_graph_sample(g: *mut ss_graph, k: i32, beta: f64, sampling_seed: u64, kept: *mut i64) -> c_int;
    pub fn ss_graph_apply_reweight(g: *mut ss_graph, diag_out: *mut f64) -> c_int;
}

/// The last error message of the calling thread.
pub fn last_error() -> String {
    let mut buf = vec![0u8; 1024];
    unsafe { ss_last_error(buf.as_mut_ptr() as *mut c_char, buf.len()) };
    let end = buf.iter().position(|&b| b == 0).unwrap_or(buf.len());
    String::from_utf8_lossy(&buf[..end]).into_owned()
}
