/*
 * spec_spars_hip.h -- C ABI of the MI355X (gfx950) sparsification hot path.
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point returns
 * an int status (SS_OK == 0); the message of the last failure on the calling
 * thread is available from ss_last_error().  Nothing throws or aborts.
 *
 * Threading: one HIP stream per context; calls on different contexts may run
 * concurrently from different threads; calls on one context are serialized
 * by the caller.  Host-only entry points (ss_factor_*, ss_graph_check_*) need
 * no GPU.
 *
 * Layouts: dense n x k blocks (sketch Y, solution Z) are row-major, node i's
 * k values contiguous -- the FlattenedVec layout of the reference
 * (src/lib.rs:116-142).  Laplacians are CSC with int32 indices and the
 * diagonal stored, as handed over by Sparsifier::get_probs
 * (src/sparsifier.rs:589-593); being symmetric, CSC and CSR coincide.
 * Canonical edge order is the Laplacian's CSC upper triangle: column
 * ascending, row ascending, row < column (src/sparsifier.rs:547-548).
 */
#ifndef SPEC_SPARS_HIP_H
#define SPEC_SPARS_HIP_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define SS_OK 0
#define SS_ERR_INVALID 1        /* bad argument / malformed input           */
#define SS_ERR_HIP 2            /* HIP runtime error                        */
#define SS_ERR_NOT_CONVERGED 3  /* solve ended with rel. residual >= limit  */
#define SS_ERR_NOMEM 4          /* host or device allocation failed         */
#define SS_ERR_TIMEOUT 5        /* a device wait exceeded its spin limit    */
#define SS_ERR_CAPACITY 6       /* size beyond a documented maximum         */

#define SS_PREC_FP64 0          /* FP64 everywhere                           */
#define SS_PREC_MIXED 1         /* FP32 SpMM/tri-solves, FP64 reductions and
                                   FP64 true-residual refinement             */
#define SS_STOP_REF_ABS 0       /* ||r||^2 <= tol (include/custom_cg.hpp:159,
                                   :213-222), tol = 2e-7                      */
#define SS_STOP_REL 1           /* ||r|| <= tol * ||b||                      */

#define SS_ORDER_NATURAL 0      /* the reference's natural elimination order */
#define SS_ORDER_RANDOM 1       /* seeded random permutation                 */
#define SS_ORDER_ROUNDS 2       /* rounds of independent low-degree vertices
                                   of the fill-augmented graph: every round
                                   is one level of both triangular solves   */

typedef struct ss_context ss_context;
typedef struct ss_factor ss_factor;
typedef struct ss_graph ss_graph;

/* Last error message of the calling thread (truncated to len-1 chars).
 * Returns the full message length. */
int ss_last_error(char *buf, size_t len);
const char *ss_version(void);

/* ---------------------------------------------------------------------- */
/* Context: device + one HIP stream.                                       */
int ss_context_create(int device, ss_context **out);
/* Fails (SS_ERR_INVALID) while graphs created on the context still exist. */
int ss_context_destroy(ss_context *ctx);
/* Free / total device memory seen by this context (bytes). */
int ss_context_mem_info(ss_context *ctx, size_t *free_b, size_t *total_b);

/* ---------------------------------------------------------------------- */
/* Host approximate-Cholesky factor (no GPU).  Same randomized algorithm as
 * factorization_driver / perform_factorization_device
 * (src/example.cc:90-471, include/auxilliary.hpp:90-173), natural order,
 * made deterministic: one random stream per column keyed by (seed, column),
 * duplicate merges and value sorts with a total order.                    */
int ss_factor_build(int64_t n, const int32_t *colptr, const int32_t *rowidx,
                    const double *vals, uint64_t seed, ss_factor **out);
/* Same elimination with a chosen order (SS_ORDER_*).  The natural order's
 * triangular-solve DAG is nearly a chain (depth ~ n/2 on grids, fill edges
 * link consecutive columns); the other orders keep the algorithm and give
 * the level-scheduled solves a shallow DAG.  No reference counterpart. */
int ss_factor_build_ordered(int64_t n, const int32_t *colptr, const int32_t *rowidx,
                            const double *vals, uint64_t seed, int32_t ordering,
                            ss_factor **out);
/* Host threads the factorization uses (0 = default: SPEC_SPARS_THREADS, else
 * the hardware threads capped at 32, the reference's OpenMP team size,
 * src/example.cc:672).  The factor does not depend on it.  Process-wide. */
int ss_factor_set_threads(int32_t threads);
/* Rows of the dense tail (SS_ORDER_ROUNDS): positions [n - tail, n) were
 * eliminated one at a time once the remaining graph was nearly dense; the
 * device applies that block of L as a dense inverse.  0 for other orders. */
int ss_factor_tail(const ss_factor *f, int64_t *tail);
/* Elimination order: order[p] = vertex eliminated at position p (identity
 * for the natural order); rounds = number of rounds (SS_ORDER_ROUNDS). */
int ss_factor_order(const ss_factor *f, int32_t *order, int32_t *rounds);
/* nnz counts the unit diagonal, as "factor nnz" at src/example.cc:448-450;
 * depth_* are the longest dependency paths (src/example.cc:531-545). */
int ss_factor_info(const ss_factor *f, int64_t *n, int64_t *nnz,
                   int32_t *depth_fwd, int32_t *depth_bwd);
/* The factor in the reference's precond_M layout (src/example.cc:422-470):
 * CSR, row c = column c of the unit-lower factor L (diagonal first, rows
 * ascending), plus D.  Arrays sized n+1, nnz, nnz, n.  Indices are
 * elimination positions (vertex = order[position]). */
int ss_factor_export(const ss_factor *f, int32_t *urowptr, int32_t *ucol,
                     double *uval, double *diag);
/* 1-based level of every row: forward = the DP at src/example.cc:531-542,
 * backward = the same DP run on L^T. */
int ss_factor_levels(const ss_factor *f, int32_t *lvl_fwd, int32_t *lvl_bwd);
int ss_factor_destroy(ss_factor *f);

/* ---------------------------------------------------------------------- */
/* (a) JL sketch Y = B Q, replaces jl_sketch_colwise_flat
 * (src/sparsifier.rs:490-539).  Edge e = (v1[e], v2[e], w[e]) in insertion
 * order; B[:,e] = -sqrt(w) at v1, +sqrt(w) at v2 (src/sparsifier.rs:90-114);
 * Q[e,j] = scale * sign(CityHash64WithSeed(LE64(e)||LE64(j), seed))
 * (Rademacher, src/sparsifier.rs:379-407) or scale * uniform(-1,1)
 * (src/sparsifier.rs:385-420) when uniform != 0.  Y: n x k row-major FP64.
 * Edges must have v1 != v2, w > 0.                                         */
int ss_sketch(ss_context *ctx, const int32_t *v1, const int32_t *v2,
              const double *w, int64_t m, int64_t n, int32_t k,
              uint64_t sketch_seed, double scale, int uniform, double *Y);

/* ---------------------------------------------------------------------- */
/* (b)+(c) Laplacian solve L Z = Y - mean, replaces ffi::run_solve_lap
 * (src/lib.rs:103, include/example.h:18, src/example.cc:668-703) and the
 * MKL PCG it drives (include/custom_cg.hpp:95-274).                        */
/* Mixed precision (SS_PREC_MIXED) around an FP32 PCG with FP64 dots:
 * RELIABLE keeps one PCG running and replaces a column's recursive residual
 * by the FP64 true residual b - A x whenever it has fallen by inner_tol
 * (reliable updates; the search directions are kept; max_iter bounds the
 * iterations); RESTART runs FP32 PCG sweeps, each on the FP64 residual of
 * the last, until the FP64 rule holds (iterative refinement; at most 16
 * sweeps of at most min(max_iter, 300) iterations each).  RELIABLE_F16 is
 * RELIABLE with the preconditioner's intermediate rows (the triangular
 * solves' outputs, which every factor entry gathers) stored in FP16, each
 * column scaled to unit mean square before the solve (FP32 arithmetic; k
 * even, else FP32 rows): fewer bytes per gather, which pays on wide merged
 * requests over weights spanning little, at the cost of iterations
 * elsewhere.  The preconditioner's precision changes the iteration count
 * only: every scheme stops on the FP64 true residual. */
#define SS_MIXED_RELIABLE 0
#define SS_MIXED_RESTART 1
#define SS_MIXED_RELIABLE_F16 2

typedef struct ss_solve_opts {
    uint64_t factor_seed;   /* seed of the randomized factor                 */
    int32_t stop_rule;      /* SS_STOP_REF_ABS or SS_STOP_REL                */
    int32_t max_iter;       /* 400 in the reference                          */
    double tol;             /* 2e-7 (REF_ABS) or e.g. 1e-8 (REL)             */
    double fail_relres;     /* final rel. residual >= this -> NOT_CONVERGED;
                               1e-2 in the reference (custom_cg.hpp:254)     */
    int32_t precision;      /* SS_PREC_FP64 or SS_PREC_MIXED                 */
    int32_t ordering;       /* SS_ORDER_* of the factor ss_solve_lap builds  */
    double inner_tol;       /* SS_PREC_MIXED: relative residual reduction
                               between two FP64 residual replacements
                               (reliable; 0 = 3e-3) or asked of each FP32
                               PCG sweep (restart; 0 = 1e-3)                 */
    int32_t mixed_mode;     /* SS_PREC_MIXED: SS_MIXED_RELIABLE (default),
                               SS_MIXED_RESTART or SS_MIXED_RELIABLE_F16     */
} ss_solve_opts;

typedef struct ss_solve_report {
    int32_t *iters;         /* [k] or NULL: PCG iterations per column        */
    double *relres;         /* [k] or NULL: final true relative residual     */
    int64_t factor_nnz;     /* out */
    int32_t depth_fwd;      /* out */
    int32_t depth_bwd;      /* out */
    int32_t max_iters;      /* out */
    int32_t refinements;    /* out: refinement sweeps (mixed precision)      */
    double max_relres;      /* out */
    double ms_factor;       /* out: host factorization time                  */
    double ms_solve;        /* out: device solve time                        */
} ss_solve_report;

/* The reference's stop rule (REF_ABS, tol 2e-7, 400 iterations, fail at a
 * relative residual of 1e-2), FP64, factor seed 0 and SS_ORDER_ROUNDS: the
 * rounds order keeps the level-scheduled solves shallow (the natural order
 * gives a DAG of depth ~n/2 on grids, one launch per level). */
void ss_solve_opts_default(ss_solve_opts *o);
int ss_solve_lap(ss_context *ctx, const double *Y, int64_t n, int32_t k,
                 const int32_t *colptr, const int32_t *rowidx,
                 const double *vals, int64_t nnz, const ss_solve_opts *opts,
                 double *Z, ss_solve_report *rep);

/* ---------------------------------------------------------------------- */
/* (d) effective resistance + probabilities + coins + sampling + compaction,
 * replaces compute_diff_norms / compute_probs / flip_coins /
 * sample_and_reweight (src/sparsifier.rs:542-680).  Edges (u<v, lval=L_uv)
 * in canonical order; Z n x k row-major FP64.
 *   p_e = min(1, beta * -1 * L_uv * d_e^2 / k),  d_e = ||Z_u - Z_v||
 *   c_e = e-th StdRng(ChaCha12)::seed_from_u64(sampling_seed) draw in [0,1)
 *   keep iff c_e < p_e; kept weight = -(L_uv + (w - w/sqrt(p))).
 * p_in != NULL replaces the computed probabilities (Z may then be NULL).
 * p_out, d_out, mask_out (one byte per edge) may be NULL.  Kept edges are
 * written in canonical order to ku/kv/kw (capacity m).                    */
int ss_er_sample(ss_context *ctx, const double *Z, int64_t n, int32_t k,
                 const int32_t *u, const int32_t *v, const double *lval,
                 int64_t m, double beta, uint64_t sampling_seed,
                 const double *p_in, double *p_out, double *d_out,
                 uint8_t *mask_out, int64_t *kept, int32_t *ku, int32_t *kv,
                 double *kw);

/* No counterpart in the reference: ss_er_sample for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  Same arguments and results, but
 * Z, u, v, lval, p_in and every output array except `kept` are DEVICE
 * pointers (ku, kv, kw of capacity m; p_in, p_out, d_out, mask_out may be
 * NULL); `kept` is host memory (may be NULL).  The call allocates nothing:
 * its counters and look-back words live in a fixed per-device array of
 * 8 * (8 + SS_NOALLOC_MAX_EDGES / 256) bytes (2 MiB), held until the process
 * ends and set up by ss_context_create; m > SS_NOALLOC_MAX_EDGES is
 * rejected.  Edges are checked on the device (u < v < n, L_uv < 0, p_in in
 * [0,1]; SS_ERR_INVALID).  Runs on the context's stream and returns after
 * it drains.  Every call on one GPU shares that array: calls that overlap
 * on one device race. */
#define SS_NOALLOC_MAX_EDGES ((int64_t)1 << 26)
int ss_er_sample_noalloc(ss_context *ctx, const double *Z, int64_t n, int32_t k,
                         const int32_t *u, const int32_t *v, const double *lval,
                         int64_t m, double beta, uint64_t sampling_seed,
                         const double *p_in, double *p_out, double *d_out,
                         uint8_t *mask_out, int64_t *kept, int32_t *ku, int32_t *kv,
                         double *kw);

/* ---------------------------------------------------------------------- */
/* Device-resident graph: canonical edge list, Laplacian CSR, incidence ids,
 * factor, sketch/solution blocks and sampling output stay in HBM.          */
int ss_graph_create(ss_context *ctx, int64_t n, int64_t m, const int32_t *u,
                    const int32_t *v, const double *w, ss_graph **out);
int ss_graph_destroy(ss_graph *g);
/* Device bytes the graph would need for k columns at a precision. */
int ss_graph_bytes(int64_t n, int64_t m, int64_t factor_nnz, int32_t k,
                   int32_t precision, size_t *bytes);
/* A device graph from a CSC Laplacian with its diagonal (as ss_solve_lap
 * reads it): the edges are its strictly upper entries (weight -L_uv, canonical
 * order), the diagonal is kept as given. */
int ss_graph_create_csc(ss_context *ctx, int64_t n, const int32_t *colptr, const int32_t *rowidx,
                        const double *vals, ss_graph **out);
int ss_graph_set_factor(ss_graph *g, const ss_factor *f);
/* Device bytes the graph holds now (structure, factor, work blocks), and
 * the bytes a run with k columns at a precision would add (0 when the work
 * blocks already allocated cover it): the admission test of a server.  No
 * reference counterpart. */
int ss_graph_resident_bytes(const ss_graph *g, size_t *bytes);
int ss_graph_work_bytes(const ss_graph *g, int32_t k, int32_t precision, size_t *bytes);
/* Laplacian of the graph as CSC with diagonal (host copy, sizes n+1 and
 * n+2m) -- the input ss_factor_build expects. */
int ss_graph_laplacian(const ss_graph *g, int32_t *colptr, int32_t *rowidx,
                       double *vals);

typedef struct ss_sparsify_opts {
    int32_t k;              /* sketch columns                                */
    int32_t uniform;        /* 0 = Rademacher signs                          */
    double scale;           /* jl_scaling_factor                             */
    uint64_t sketch_seed;
    uint64_t sampling_seed;
    double beta;            /* eps^-2 * beta_constant * ln n                 */
    ss_solve_opts solve;
} ss_sparsify_opts;

typedef struct ss_sparsify_report {
    int64_t kept;
    int32_t max_iters;
    int32_t refinements;
    double max_relres;
    double ms_sketch;       /* device time per phase (HIP events)           */
    double ms_solve;
    double ms_sample;
    double ms_total;
} ss_sparsify_report;

/* Sketch -> solve -> ER/sample/compact, all on the device; the factor must
 * have been set.  Results stay resident until read back.  With
 * SS_PREC_MIXED and a reliable scheme, a request's columns replace their
 * residuals together, and a column at its stop target keeps iterating
 * until every column of its request is there (one FP64 verification for
 * the request); each still stops on its FP64 true residual. */
int ss_graph_sparsify(ss_graph *g, const ss_sparsify_opts *o,
                      ss_sparsify_report *rep);
int ss_graph_read_kept(const ss_graph *g, int64_t cap, int32_t *ku,
                       int32_t *kv, double *kw, int64_t *kept);
/* Several sparsify requests on one graph in one pass (no reference
 * counterpart: a served form merging requests that arrive together).  The
 * requests' sketch columns sit side by side (request r's column j hashes
 * index j with r's sketch seed and scale), one multi-RHS solve takes every
 * column (each keeps its own PCG scalars and stop test), then one sampling
 * pass per request reads its own columns with its own beta and sampling
 * seed.  Each request's kept edges are bit-identical to ss_graph_sparsify
 * with its options alone.  uniform and the solve options must agree; the
 * requests' k add up to at most 256.  rep[r] gets request r's kept count,
 * iterations and residual; the phase times are the pass's. */
int ss_graph_sparsify_many(ss_graph *g, int32_t count, const ss_sparsify_opts *opts,
                           ss_sparsify_report *rep);
/* Request r's kept edges of the last ss_graph_sparsify_many call. */
int ss_graph_read_kept_many(const ss_graph *g, int32_t r, int64_t cap, int32_t *ku,
                            int32_t *kv, double *kw, int64_t *kept);
/* Sketch Y / solution Z of the last run as FP64 n x k (host).  After an
 * FP64 solve, which = 2 / 3 give the last preconditioner application's
 * forward result L^{-1} r and its output L^{-T} D^+ L^{-1} r. */
int ss_graph_read_block(const ss_graph *g, int which /*0=Y,1=Z,2,3*/, double *out);
int ss_graph_read_probs(const ss_graph *g, double *p);
/* The fused ER/coin/sample/compaction pass alone over the solution of the
 * last solve (k columns; ss_batch_write_rhs + ss_batch_solve), with beta and
 * sampling seed: compute_probs + flip_coins + sample_and_reweight
 * (src/sparsifier.rs:542-680).  Kept edges readable with ss_graph_read_kept. */
int ss_graph_sample(ss_graph *g, int32_t k, double beta, uint64_t sampling_seed, int64_t *kept);
/* apply_reweightings of the last sample on the device (src/sparsifier.rs:
 * 682-691): the new Laplacian is the kept edges (ss_graph_read_kept: L_uv =
 * -kw) and diag_out (host, n values) = the old diagonal plus each vertex's
 * changes summed in edge order -- the values Triplet + sparse add produce.
 * Exact zeros (diagonals of isolated vertices) are for the caller to drop. */
int ss_graph_apply_reweight(ss_graph *g, double *diag_out);

/* ---------------------------------------------------------------------- */
/* Batches.  No reference counterpart: the reference sparsifies one graph per
 * Sparsifier (src/sparsifier.rs:722-770).  G independent graphs are held as
 * one block-diagonal system and every stage runs once over all of them;
 * all reductions are segmented per (graph, column) in a fixed order, so each
 * graph's results are bit-identical to running it alone.  An ss_graph from
 * ss_graph_create is a batch of one; the ss_graph_* calls above operate on
 * batches of one only.
 *
 * Graph i has n[i] vertices and the canonical edges (u < v, sorted by
 * (v, u)) at positions [edge_ptr[i], edge_ptr[i+1]) of u/v/w, with vertex
 * ids local to the graph. */
int ss_batch_create(ss_context *ctx, int32_t count, const int64_t *n,
                    const int64_t *edge_ptr, const int32_t *u, const int32_t *v,
                    const double *w, ss_graph **out);
int ss_batch_count(const ss_graph *g, int32_t *count);
/* One factor per graph (factors[i] built from graph i's Laplacian). */
int ss_batch_set_factors(ss_graph *g, int32_t count,
                         const ss_factor *const *factors);
/* Sparsify every graph of the batch: opts[i] for graph i.  k, uniform and
 * the solve options must agree across the batch; seeds, scale and beta are
 * per graph.  rep[i] (may be NULL) gets graph i's kept count, iterations
 * and residual; the phase times are the batch's. */
int ss_batch_sparsify(ss_graph *g, const ss_sparsify_opts *opts,
                      ss_sparsify_report *rep);
/* Solve L_i Z_i = Y_i for every graph with the right-hand sides uploaded by
 * ss_batch_write_rhs; iters/relres (may be NULL) have count*k entries. */
int ss_batch_write_rhs(ss_graph *g, int32_t i, int32_t k, int32_t precision,
                       const double *Y /* n_i x k */);
int ss_batch_solve(ss_graph *g, int32_t k, const ss_solve_opts *opts,
                   int32_t *iters, double *relres, int32_t *refinements);
/* Graph i's part of the last run's results (local vertex ids). */
int ss_batch_read_kept(const ss_graph *g, int32_t i, int64_t cap, int32_t *ku,
                       int32_t *kv, double *kw, int64_t *kept);
int ss_batch_read_block(const ss_graph *g, int32_t i, int which /*0=Y,1=Z*/,
                        double *out);
int ss_batch_read_probs(const ss_graph *g, int32_t i, double *p);

/* Per-kernel device time, measured with HIP events on the graph's stream
 * around every launch while enabled (enabling clears the counters). */
#define SS_KERNEL_SKETCH 0      /* (a) sketch                                */
#define SS_KERNEL_SPMM 1        /* (b) Laplacian SpMM (+ p.q)                */
#define SS_KERNEL_TRI_FWD 2     /* (c) one level of the forward solve        */
#define SS_KERNEL_TRI_BWD 3     /* (c) one level of the backward solve       */
#define SS_KERNEL_ER_SAMPLE 4   /* (d) ER + coin + sample + compaction       */
#define SS_KERNEL_TRI_FWD_SMALL 5 /* forward run of small levels (1 workgroup) */
#define SS_KERNEL_TRI_BWD_SMALL 6 /* backward run of small levels            */
#define SS_KERNEL_VECTOR 7      /* PCG vector updates and reductions         */
#define SS_KERNEL_CORE_FWD 8    /* dense cores, forward: outside entries, gather,
                                   inverse product, reduction               */
#define SS_KERNEL_CORE_BWD 9    /* dense cores, backward: gather of y/D,
                                   transposed inverse product, reduction    */
#define SS_KERNEL_RESIDUAL 10   /* FP64 true residuals b - A x (refinement)  */
#define SS_KERNEL_COUNT 11
typedef struct ss_kernel_stat {
    double ms_total;
    int64_t launches;
} ss_kernel_stat;
int ss_graph_profile(ss_graph *g, int enable);
/* Shape of the uploaded solve schedules: per direction (0 = forward,
 * 1 = backward) the level count, the rows and entries handled by per-level
 * launches and by the single-workgroup runs of small levels (small_runs of
 * them, small_levels levels in all), and the dense cores -- per graph the
 * last rows of its elimination order (the factor's dense tail, else its
 * trailing small levels), applied as the inverse of their block of L:
 * core_rows rows, core_entries = sum of padded core_rows_g^2 stored values
 * of the inverses, core_steps = product pieces per direction (64 rows times
 * at most 2,048 inner rows each), core_outside_nnz entries of core rows
 * outside the cores (forward only). */
typedef struct ss_schedule_info {
    int32_t levels, level_launches, small_levels, small_runs;
    int64_t level_rows, level_nnz, small_rows, small_nnz;
    int64_t core_rows, core_entries, core_outside_nnz;
    int32_t core_steps, core_graphs;
} ss_schedule_info;
int ss_graph_schedule_info(const ss_graph *g, int direction, ss_schedule_info *out);
int ss_graph_profile_read(const ss_graph *g, ss_kernel_stat *out /* [SS_KERNEL_COUNT] */);

/* ---------------------------------------------------------------------- */
/* Launch shapes.  No reference counterpart.  tile_waves = wavefronts per
 * block of the row-tile kernels (SpMM, vector updates, reductions);
 * level_rows_per_wave = consecutive rows each wavefront of a per-level
 * triangular-solve launch finishes.  A graph keeps the tile_waves in effect
 * when it was created (its reductions are shaped by it, so compare results
 * bit for bit only between graphs created under the same setting);
 * level_rows_per_wave = consecutive short rows (at most 32 entries) each
 * wavefront of a per-level triangular-solve launch finishes; graph_iters =
 * PCG iterations captured in one hipGraph and replayed between two reads of
 * the active-column count (0 = launch every kernel directly; per-kernel
 * profiling always launches directly).  Both change timing only. */
typedef struct ss_tuning {
    int32_t tile_waves;            /* 1, 2, 4, 8 or 16 (default 4)           */
    int32_t level_rows_per_wave;   /* 1 .. 64 (default 1)                    */
    int32_t graph_iters;           /* 0 .. 16 (default 0, which the autotuner
                                      keeps: replay never beat direct launches);
                                      iterations whose capture would exceed
                                      1,024 kernel nodes launch directly      */
    int32_t xcd_slices;            /* 1: each of the 8 XCDs takes one contiguous
                                      slice of a level's rows; 0: round robin
                                      (default; the autotuner picks)          */
} ss_tuning;
int ss_context_get_tuning(const ss_context *ctx, ss_tuning *out);
int ss_context_set_tuning(ss_context *ctx, const ss_tuning *t);
/* Autotuner: times the candidate shapes on graph g (factor set) for k
 * columns at a precision -- one preconditioner application per
 * rows-per-wave candidate, one SpMM + vector-update pass per tile-waves
 * candidate, median of reps runs each -- and keeps the fastest in the
 * context and in g.  ms_level / ms_tile (may be NULL) receive the best
 * times. */
int ss_context_autotune(ss_context *ctx, ss_graph *g, int32_t k, int32_t precision,
                        int32_t reps, ss_tuning *best, double *ms_level,
                        double *ms_tile);

/* Achievable-bandwidth probe: a plain FP64 copy kernel over `bytes` bytes
 * (read + write), repeated reps times; *ms = average duration per copy. */
int ss_bench_copy(ss_context *ctx, size_t bytes, int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif
