/*
 * oracle.h -- CPU restatement (FP64, plain C) of the spec_spars sparsification
 * hot path.  TEST INFRASTRUCTURE ONLY: nothing in the shipped library links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / CPU baseline.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 *
 * Pinning status (what checks each function beyond retyping its formula):
 *   or_hash128to64 / seeded wrapper  -- CityHash test-table entries (len 0)
 *   or_city64_seed16 (HashLen0to16)  -- parity unpinned against fasthash-sys
 *                                       (crate source absent; version v1.1
 *                                       assumed, v1.0.3 selectable)
 *   or_chacha_block                  -- RFC 7539 2.3.2 / A.1 vectors (20 rounds)
 *   or_stdrng_seed / or_coin         -- parity unpinned against rand 0.8.5
 *                                       (crate source absent); invariants only
 *   or_sketch                        -- brute-force dense B*Q, zero column sums
 *   or_laplacian / edge order        -- reference sampling_test KAT (40 CCs),
 *                                       evim_csc_equiv KAT
 *   or_levels_fwd / or_levels_bwd    -- brute-force longest path on tiny DAGs
 *   or_pcg                           -- scipy direct solve fixtures
 *   or_diff_norms / or_probs         -- Foster's theorem with exact R_eff,
 *                                       tree leverage == 1
 *   or_sample_reweight               -- sampling_test (40 CCs), sampling_verify
 *                                       ratios (2 and sqrt 2), H subset of G
 */
#ifndef SPEC_SPARS_ORACLE_H
#define SPEC_SPARS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- CityHash64WithSeed on the 16-byte key LE64(e) || LE64(j) ---------- */
uint64_t or_hash128to64(uint64_t lo, uint64_t hi);
uint64_t or_city64_seed16(uint64_t e, uint64_t j, uint64_t seed, int variant);
double   or_sketch_value(uint64_t e, uint64_t j, uint64_t seed, double scale,
                         int uniform, int variant);

/* ---- sketch Y = B Q (n x k row-major) ---------------------------------- */
void or_sketch(int64_t m, const int32_t *v1, const int32_t *v2, const double *w,
               int64_t n, int32_t k, uint64_t seed, double scale, int uniform,
               int variant, double *Y);

void or_sketch_edges(int64_t count, const int64_t *eid, const int32_t *v1, const int32_t *v2,
                     const double *w, int64_t n, int32_t k, uint64_t seed, double scale,
                     int uniform, int variant, double *Y);
void or_permute_csc(int64_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals,
                    const int32_t *rank, int32_t *pcolptr, int32_t *prowidx, double *pvals);

/* ---- Laplacian assembly (CSC with diagonal; symmetric so CSC == CSR) --- */
/* Returns nnz written (<= n + 2m) or -1 on a bad vertex. */
int64_t or_laplacian(int64_t n, int64_t m, const int32_t *v1, const int32_t *v2,
                     const double *w, int32_t *colptr, int32_t *rowidx,
                     double *vals);
/* Upper-triangle edge list of a CSC Laplacian, in CSC order. */
int64_t or_upper_edges(int64_t n, const int32_t *colptr, const int32_t *rowidx,
                       const double *vals, int32_t *u, int32_t *v, double *lval);

/* ---- level sets of the factor (U = L^T in CSR, diagonal first) --------- */
void or_levels_fwd(int64_t n, const int32_t *urowptr, const int32_t *ucol,
                   int32_t *level);
void or_levels_bwd(int64_t n, const int32_t *urowptr, const int32_t *ucol,
                   int32_t *level);

/* ---- preconditioned CG, one RHS ---------------------------------------- */
/* stop_rule 0: reference rule ||r||^2 <= tol (absolute, tol = 2e-7)
 * stop_rule 1: relative ||r|| <= tol * ||b||
 * Returns iterations; writes final true relative residual to *relres.      */
int or_pcg(int64_t n, const int32_t *arowptr, const int32_t *acol,
           const double *aval, const int32_t *urowptr, const int32_t *ucol,
           const double *uval, const double *diag, const double *b_in,
           double *x, int stop_rule, double tol, int max_iter, double *relres);
void or_precond_apply(int64_t n, const int32_t *urowptr, const int32_t *ucol,
                      const double *uval, const double *diag, const double *r,
                      double *z);
/* All k columns (Y row-major n x k) -> Z row-major, sequential per column. */
int or_solve_lap(int64_t n, int32_t k, const double *Y, const int32_t *arowptr,
                 const int32_t *acol, const double *aval, const int32_t *urowptr,
                 const int32_t *ucol, const double *uval, const double *diag,
                 int stop_rule, double tol, int max_iter, double *Z,
                 int32_t *iters, double *relres);

/* ---- effective resistance, probabilities ------------------------------- */
void or_diff_norms(int64_t m, const int32_t *u, const int32_t *v, int32_t k,
                   const double *Z, double *d);
void or_probs(int64_t m, const double *lval, const double *d, double beta,
              int32_t k, double *p);

/* ---- coins: rand 0.8.5 StdRng (ChaCha12) gen_range(0.0..1.0) ----------- */
void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                     int rounds, uint32_t out[16]);
void or_stdrng_seed(uint64_t seed, uint32_t key[8]);
void or_coins(uint64_t seed, int64_t first, int64_t count, double *c);

/* ---- sampling and reweighting ------------------------------------------ */
/* mask[e] = c_e < p_e; kept edges compacted in edge order with the
 * reweighted Laplacian weight; ddiag[v] = diagonal change (sequential in
 * edge order).  Returns number kept.                                       */
int64_t or_sample_reweight(int64_t m, int64_t n, const int32_t *u,
                           const int32_t *v, const double *lval,
                           const double *p, const double *c, uint8_t *mask,
                           int32_t *ku, int32_t *kv, double *kw, double *ddiag);

#ifdef __cplusplus
}
#endif
#endif
