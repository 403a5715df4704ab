// A module for the reference crate (copied in as src/hip.rs, declared with
// `mod hip;` in src/lib.rs): run_solve_lap with the exact signature of the
// cxx bridge function it replaces (src/lib.rs:103), on the MI355X path.
// The one edit to src/sparsifier.rs is the call at :593, which names
// hip::run_solve_lap instead of ffi::run_solve_lap.  The crate's Cargo.toml
// gains `spec_spars_hip_sys = { path = "<checkout>/rust/spec_spars_hip_sys" }`.
use crate::ffi::FlattenedVec;
use spec_spars_hip_sys as sys;

thread_local! {
    // one device context (HIP stream) per calling thread, as the reference
    // calls run_solve_lap from the thread running sparsify
    static CTX: *mut sys::ss_context = {
        let mut c = std::ptr::null_mut();
        let rc = unsafe { sys::ss_context_create(0, &mut c) };
        assert_eq!(rc, sys::SS_OK, "ss_context_create: {}", sys::last_error());
        c
    };
}

pub fn run_solve_lap(shared_jl_cols: FlattenedVec, rust_col_ptrs: Vec<i32>, rust_row_indices: Vec<i32>,
                     rust_values: Vec<f64>, _solver_output_filename: &str, num_nodes: i32,
                     _verbose: bool) -> FlattenedVec {
    let (n, k) = (shared_jl_cols.num_rows, shared_jl_cols.num_cols);
    assert_eq!(n, num_nodes as usize);
    let mut z = vec![0.0f64; n * k];
    let mut o: sys::ss_solve_opts = unsafe { std::mem::zeroed() };
    // the reference's stop rule (||r||^2 <= 2e-7, 400 iterations, fail at
    // a relative residual of 1e-2) with the rounds-order factor
    unsafe { sys::ss_solve_opts_default(&mut o) };
    let rc = CTX.with(|c| unsafe {
        sys::ss_solve_lap(*c, shared_jl_cols.vec.as_ptr(), n as i64, k as i32, rust_col_ptrs.as_ptr(),
                          rust_row_indices.as_ptr(), rust_values.as_ptr(), rust_values.len() as i64, &o,
                          z.as_mut_ptr(), std::ptr::null_mut())
    });
    // the reference aborts on these (src/example.cc:564-567): so does this
    assert_eq!(rc, sys::SS_OK, "ss_solve_lap: {}", sys::last_error());
    // row-major n x k, the layout flatten_vector returns (src/example.cc:606-618)
    FlattenedVec { vec: z, num_cols: k, num_rows: n }
}
