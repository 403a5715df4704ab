/*
 * spec_spars_host.h -- C wrapper of the C++ host layer (spec_spars.hpp),
 * the reference's Sparsifier interface (src/sparsifier.rs) over the GPU
 * path.  Bindings in other languages (and the Python tests) use it; the
 * drop-in boundary for a Rust caller that keeps its own sparsifier is
 * spec_spars_hip.h.
 */
#ifndef SPEC_SPARS_HOST_H
#define SPEC_SPARS_HOST_H
#include <stdint.h>

#include "spec_spars_hip.h"
#ifdef __cplusplus
extern "C" {
#endif

typedef struct ssh_sparsifier ssh_sparsifier;

/* SparsifierParameters (src/sparsifier.rs:146-210) plus the solve settings
 * the reference hard-codes in C++. */
typedef struct ssh_params {
    double epsilon;
    int32_t beta_constant;
    int32_t row_constant;
    int32_t verbose;
    int32_t sketch_uniform;
    double jl_factor;
    double jl_scaling_factor;
    uint64_t sketch_seed;
    uint64_t sampling_seed;
    int32_t streaming;
    int32_t ordering;
    uint64_t factor_seed;
    int32_t precision;
    int32_t stop_rule;
    int32_t max_iter;
    int32_t resident;       /* device-resident pass (see spec_spars.hpp)   */
    double tol;
} ssh_params;

/* SparsifierParameters::new_default (src/sparsifier.rs:182-209) */
void ssh_params_default(ssh_params *p, int bench);
/* Sparsifier::new (src/sparsifier.rs:262-297) */
int ssh_sparsifier_create(ss_context *ctx, int32_t num_nodes, const ssh_params *p, ssh_sparsifier **out);
int ssh_sparsifier_destroy(ssh_sparsifier *s);
/* Sparsifier::insert for each (v1[i], v2[i], value[i]) (src/sparsifier.rs:357-376) */
int ssh_sparsifier_insert_many(ssh_sparsifier *s, int64_t count, const int32_t *v1, const int32_t *v2,
                               const double *value);
/* Sparsifier::form_laplacian (src/sparsifier.rs:693-709) */
int ssh_sparsifier_form_laplacian(ssh_sparsifier *s, int check);
/* Sparsifier::check_diagonal (src/sparsifier.rs:806-817): SS_OK or SS_ERR_INVALID */
int ssh_sparsifier_check_diagonal(const ssh_sparsifier *s);
/* Sparsifier::num_edges (src/sparsifier.rs:346-353) */
int ssh_sparsifier_num_edges(const ssh_sparsifier *s, int64_t *out);
/* Sparsifier::sparsify (src/sparsifier.rs:722-770), returning SparsificationStats */
int ssh_sparsifier_sparsify(ssh_sparsifier *s, int check, int64_t *before, int64_t *after);
/* sample_and_reweight(probs) + apply_reweightings(report = false)
 * (src/sparsifier.rs:635-691) */
int ssh_sparsifier_sample_apply(ssh_sparsifier *s, const double *probs, int64_t len, int64_t *before,
                                int64_t *after);
/* current_laplacian as CSC (pass NULL arrays to query nnz first) */
int ssh_sparsifier_laplacian(const ssh_sparsifier *s, int64_t *nnz, int32_t *colptr, int32_t *rowidx,
                             double *vals);
int ssh_sparsifier_info(const ssh_sparsifier *s, double *beta, int32_t *jl_dim, int64_t *threshold);
/* passes the streaming trigger ran (src/sparsifier.rs:370-374) */
int ssh_sparsifier_triggered(const ssh_sparsifier *s, int64_t *count);
/* jl_sketch_colwise_flat of new_entries (src/sparsifier.rs:532-539), n x jl_dim row-major */
int ssh_sparsifier_sketch(const ssh_sparsifier *s, double *Y);

/* ---- MatrixMarket files and the stream driver (src/utils.rs:149-191,
 * src/stream.rs:18-64) ---------------------------------------------------- */
typedef struct ssh_matrix ssh_matrix;
typedef struct ssh_stream ssh_stream;
/* read_mtx: a coordinate file (real / integer / pattern values; general,
 * symmetric or skew-symmetric storage, mirrored) as CSC, rows sorted in each
 * column, duplicates summed, explicit zeros kept. */
int ssh_mtx_read(const char *path, ssh_matrix **out);
int ssh_matrix_info(const ssh_matrix *m, int32_t *rows, int32_t *cols, int64_t *nnz);
/* colptr (cols + 1), rowidx and vals (nnz); NULL arrays are skipped */
int ssh_matrix_csc(const ssh_matrix *m, int32_t *colptr, int32_t *rowidx, double *vals);
int ssh_matrix_destroy(ssh_matrix *m);
/* write_mtx: coordinate real general, entries in CSC storage order */
int ssh_mtx_write(const char *path, int32_t rows, int32_t cols, const int32_t *colptr,
                  const int32_t *rowidx, const double *vals);
/* InputStream::new: read the file (square), zero its stored diagonal */
int ssh_stream_open(const char *path, const char *dataset_name, ssh_stream **out);
/* vertices, stored entries, diagonal positions absent from the file */
int ssh_stream_info(const ssh_stream *s, int32_t *num_nodes, int64_t *stored, int32_t *diag_zeros);
/* run_stream: a sparsifier of num_nodes vertices with parameters p gets
 * every stored (row, col, value), row < col, in CSC order, then sparsify;
 * with writeout (and not test) the sparsified Laplacian goes to
 * output_file.  *out receives the sparsifier (destroy it with
 * ssh_sparsifier_destroy); before / after are SparsificationStats. */
int ssh_stream_run(const ssh_stream *s, ss_context *ctx, const ssh_params *p, int test, int writeout,
                   const char *output_file, ssh_sparsifier **out, int64_t *before, int64_t *after);
int ssh_stream_destroy(ssh_stream *s);

#ifdef __cplusplus
}
#endif
#endif
