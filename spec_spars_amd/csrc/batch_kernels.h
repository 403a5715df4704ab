// Launchers of the tile-based batched kernels (batch_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hash_dev.h"
#include "kernels.h"

namespace ss {

static const int kRowTile = 2048;    // rows per row tile, at most
static const int kMinRowTile = 64;   // rows per row tile, at least
// Rows per row tile of a graph of n rows: about n / 256 (a power of two in
// [kMinRowTile, kRowTile]), so a small graph still spreads its streaming and
// reduction kernels over many workgroups.  A function of n alone: a graph's
// tiles, and so its reductions' shape, are the same alone and in a batch.
inline int64_t row_tile_rows(int64_t n) {
    int64_t r = kMinRowTile;
    while (r < kRowTile && r * 256 < n) r *= 2;
    return r;
}
static const int kSpmmTileNnz = 8192;   // SpMM tiles: matrix entries per tile, about
static const int kSpmmMaxSplit = 256;   // SpMM tiles per row tile, at most
static const int kEdgeTile = 256;    // edges per edge tile (one block)

static const int kMaxTileWaves = 16;

// Tiles of one kind: tile t covers [start[t], start[t+1]) of graph graph[t].
// Row tiles are processed by a block of `waves` wavefronts (1..16) striding
// the tile's rows; per-tile partials combine the waves in order.
struct Tiles {
    const int32_t *start;
    const int32_t *graph;
    int count;
    int waves;
};

struct BatchSampleArgs {
    int64_t m;                 // total edges (bound for kept positions)
    Tiles et;                  // edge tiles
    const int32_t *edge_off;   // per graph: first global edge
    const int32_t *eu, *ev;
    const double *ew;
    const double *Z;           // the first column sampled
    int k;                     // columns sampled
    int ldz;                   // row stride of Z
    const double *beta;        // per graph
    const uint32_t *keys;      // per graph: 8 ChaCha key words
    const double *p_in;
    double *p_out, *d_out;
    uint8_t *mask;
    int32_t *ku, *kv;
    double *kw;
    uint32_t *ticket;
    uint64_t *tile_state;
    int64_t *tile_incl;        // inclusive kept count after each tile
    uint32_t epoch;
    int *err;
    int spin_limit;
};

// Per (graph, column) sketch parameters, indexed g * k + j: the sign
// hash's seed and column index, and the scale.
struct SketchCols {
    const uint64_t *seed;
    const double *scale;
    const int32_t *j;
};
// S (may be null; Rademacher only): scratch of ceil(k / 64) words per edge
// of the edge tiles et (edge_off: per graph, its first global edge) for the
// signs, hashed once per edge instead of once per incidence.
void bt_sketch(hipStream_t s, Tiles rt, const int32_t *iptr, const int32_t *ieid, const double *ib, int k,
               SketchCols cols, int uniform, double *Y, Tiles et = Tiles{}, const int32_t *edge_off = nullptr,
               uint64_t *S = nullptr);
template <typename TA, typename TB>
void bt_tile_partial(hipStream_t s, Tiles rt, int k, const TA *A, const TB *B, const double *diag, int op,
                     double *partial);
// psc (may be null): the FP16 preconditioner rows' per-(graph, column)
// scale, which the r . z partials (p_rz) carry squared.
void bt_pcg_init(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p_bb, const double *p_rz,
                 PcgScalars sc, const float *psc = nullptr);
// FP16 preconditioner rows: sc[gj] = sqrt(n_g / |r|^2) and inv[gj] = 1 /
// sc[gj] (0 for r = 0), |r|^2 from the tile partials p_rr or, with p_rr
// null, from rr[gj]; n_g from row_off.
void bt_prec_scale(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p_rr, const double *rr,
                   const int32_t *row_off, float *sc, float *inv);
void bt_pcg_alpha(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, PcgScalars sc, int *nactive,
                  int *flag = nullptr);
void bt_pcg_check(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, int rule, double tol,
                  PcgScalars sc, int *nactive);
void bt_pcg_beta(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, PcgScalars sc,
                 const float *psc = nullptr);
void bt_graph_sums(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, double *out);
void bt_small_comp_sums(hipStream_t s, int nsmall, const int32_t *small_id, const int32_t *cptr, const int32_t *cvert,
                        int k, const double *Y, double *csum);
void bt_comp_means(hipStream_t s, int ncomp, int k, const int32_t *comp_graph, const int32_t *comp_off,
                   const int32_t *cptr, const double *gsum, double *csum, double *cmean);
template <typename TS, typename T>
void bt_project(hipStream_t s, Tiles rt, int k, const TS *Y, const int32_t *comp_of_row, const double *cmean, T *R,
                T *B, T *X);
// P = Z + beta P, with the previous iteration's X += alpha P folded in.
// zinv (may be null): Z holds FP16 rows scaled per (graph, column), and z
// = Z zinv[gj].
template <typename T, typename TZ>
void bt_update_px(hipStream_t s, Tiles rt, int k, const double *alpha, const double *beta, const TZ *Z,
                  const float *zinv, T *P, T *X);
// The same for FP16 preconditioner rows (k even): Z scaled by 1 / zinv per
// column, two or four columns per thread.
void bt_update_px_h(hipStream_t s, Tiles rt, int k, const double *alpha, const double *beta, const _Float16 *Z,
                    const float *zinv, float *P, float *X);
// r . z = y^T D^+ y per (tile, column) of FP16 preconditioner rows (k even).
void bt_tile_rz_h(hipStream_t s, Tiles rt, int k, const _Float16 *W, const double *diag, double *partial);
// X += alpha P (the last iteration's solution update).
template <typename T>
void bt_update_x(hipStream_t s, Tiles rt, int k, const double *alpha, const T *P, T *X);
template <typename T>
void bt_spmm_dot(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const T *val, int k, const T *P,
                 T *Q, double *partial);
// Q = A Z zinv + beta Q (FP16 preconditioner rows, k even): the search
// directions' recurrence with the gathers on the FP16 Z rows; partials of
// P . Q from the FP32 P rows.
void bt_spmm_z(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const float *val, int k,
               const _Float16 *Z, const float *zinv, const double *beta, const float *P, float *Q, double *partial);
// R -= alpha Q, per-tile partials of |R|^2.
template <typename T>
void bt_update_r_rr(hipStream_t s, Tiles rt, int k, const double *alpha, const T *Q, T *R, double *partial);
// out = B - A X (FP64) with per-tile partials of |out|^2; mask (per
// (graph, column), or NULL for all) selects the columns computed.
void bt_residual64(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const double *val, int k,
                   const double *X, const double *B, double *out, double *partial, const int *mask = nullptr);
void bt_convert_masked(hipStream_t s, Tiles rt, int k, const double *src, const int *keep, float *dst);
void bt_accumulate(hipStream_t s, Tiles rt, int k, const float *d, double *x);
// Mixed precision with reliable updates (per (graph, column)): bnorm from
// rr_ref = |B64|^2; the recursive residual check flagging columns for
// replacement; the fold of flagged columns' FP32 iterate x into X64 (XF =
// x, X32 = 0); R64 -= A XF and R32 = (float) R64 on the flagged columns;
// the FP64 stop test and new reference.  status = {active, err, flagged}.
void bt_reliable_init(hipStream_t s, int gk, PcgScalars sc, double *rr_ref, int *flag, int *status);
// defer: a column at its stop target waits for its replacement group
// (bt_reliable_group must follow).
void bt_reliable_check(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, int rule, double tol,
                       double delta2, bool defer, PcgScalars sc, const double *rr_ref, int *flag, int *status);
// Flag every active column of a group (columns [gs[j], ge[j]) of its
// graph) that has a flagged column or whose active columns are all at
// their stop target; unflag the other stop candidates.
void bt_reliable_group(hipStream_t s, int gk, int k, const int32_t *gs, const int32_t *ge, PcgScalars sc, int *flag,
                       int *status);
void bt_reliable_fold(hipStream_t s, Tiles rt, int k, const int *flag, const double *alpha, const float *P,
                      float *X32, double *X64, float *XF);
void bt_reliable_apply(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, int rule, double tol,
                       PcgScalars sc, double *rr_ref, int *flag, int *status);
void bt_reliable_residual(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const double *val,
                          int k, const int *flag, const float *X, double *R, float *R32, double *partial);
void bt_apply_diag(hipStream_t s, int32_t n, const int32_t *rowptr, const int32_t *col, const double *val,
                   const int32_t *iptr, const int32_t *ieid, const double *ew, const double *p, const uint8_t *mask,
                   double *diag_new);
void bt_er_sample(hipStream_t s, const BatchSampleArgs &a, int ntiles);

}  // namespace ss
