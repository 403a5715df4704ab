// Launchers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hash_dev.h"
#include "solve_shape.h"

namespace ss {

// Per-column PCG state, FP64, resident on the device.
struct PcgScalars {
    double *rz;      // r . z
    double *alpha;
    double *beta;
    double *rr;      // ||r||^2 after the last update
    double *bnorm;   // ||b|| (projected)
    int *active;     // 1 while the column iterates
    int *iters;      // updates applied
};

// Two adjacent columns of a row: FP32 pairs and FP16 pairs (one lane of a
// triangular solve moves both with one access).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Scaling of FP16 solve rows (mixed precision): the forward solve's start
// values are r times s[g * k + j] for a row of graph g (rg[row], graph 0
// without rg); s == nullptr: no scaling (rows stored in the arithmetic's
// own type).
struct RowScale {
    const float *s;
    const int32_t *rg;
};

template <typename TZ>
struct ErSampleArgs {
    int64_t m;
    int64_t ntiles;
    int64_t n;              // vertex count: edges not u < v < n are flagged (err 2), never read
    const int32_t *eu;
    const int32_t *ev;
    const double *ew;       // w = -L_uv, or L_uv itself when ew_is_lval
    int ew_is_lval;
    const TZ *Z;
    int k;
    double beta;
    CoinKey ck;
    const double *p_in;
    double *p_out;
    double *d_out;
    uint8_t *mask;
    int32_t *ku;
    int32_t *kv;
    double *kw;
    uint32_t *ticket;
    uint64_t *tile_state;
    uint32_t epoch;
    int64_t *kept_total;
    int *err;               // 1: a look-back wait hit spin_limit; 2: an invalid edge or probability
    int spin_limit;
};

template <typename OutT>
void launch_sketch(hipStream_t s, int32_t n, const int32_t *iptr, const int32_t *ieid, const double *ib, int k,
                   uint64_t seed, double scale, int uniform, int variant, OutT *Y);
void launch_fill_u32(hipStream_t s, int64_t count, uint32_t *p, uint32_t v);
void launch_copy_f64(hipStream_t s, int64_t count, const double *a, double *b);
// Rows of one level, positions [t0, t1): the first nhuge of them have more
// than kWgRowMax entries and take hm workgroups each, folding their parts
// across workgroups through scratch (nhuge x kMaxWg x k values) and one
// zeroed counter per row; the next nlong take a block of `waves`
// wavefronts each (4, 8 or 16), the next nmed one wave each, the rest (at
// most kTinyRow entries each) go rpw to a wave.
// waves == 0: no row of the level has entries (a flat copy of start values).
// xcd != 0: each of the 8 XCDs takes one contiguous slice of the wave rows.
// Arithmetic in T; start values read as TI (the solve's input), rows
// written and gathered as V: T, or with FP32 arithmetic pairs of columns
// (f32x2, or f16x2 scaled by rs; k even), TI the same width.  scratch holds
// the huge rows' subtrees in the lane type (FP32 / FP64 values).
template <typename T, typename TI, typename V, bool BWD>
void launch_trisolve_level(hipStream_t s, int32_t t0, int32_t nhuge, int32_t hm, int32_t nlong, int32_t nmed,
                           int32_t t1, int waves, int rpw, int xcd, const int32_t *rows, const int32_t *ptr,
                           const int32_t *col, const T *val, const double *diag, int k, const TI *in, V *out,
                           void *scratch, int *counter, RowScale rs);
// Small levels [l0, l1) (each of at most kSmallRows rows) walked by one
// workgroup; level l holds positions [level_ptr[l], level_ptr[l+1]).
template <typename T, typename TI, typename V, bool BWD>
void launch_tri_small(hipStream_t s, int32_t l0, int32_t l1, const int32_t *level_ptr, const int32_t *rows,
                      const int32_t *ptr, const int32_t *col, const T *val, const double *diag, int k, const TI *in,
                      V *out, RowScale rs);
// Dense cores (per graph, the last rows of its elimination order; see
// core_kernels.hip).  Forward: launch_core_head (the core rows' entries
// outside the core, schedule positions [t0, t1), the first nhuge of more
// than kWgRowMax entries, writing r_c - L_co y_o to the core rows), then
// launch_core_apply<T, false>.  Backward: launch_core_apply<T, true> first.
template <typename T, typename TI, typename V>
void launch_core_head(hipStream_t s, int32_t t0, int32_t nhuge, int32_t t1, const int32_t *rows, const int32_t *ptr,
                      const int32_t *col, const T *val, const double *diag, int k, const TI *in, V *out,
                      RowScale rs);
// A piece of a core product: rows [r0, r0 + kCoreTile) of graph `graph`'s
// core times inner rows [m0, m1) (at most kCoreSeg of them).
struct CorePiece {
    int32_t graph, r0, m0, m1;
};
// Device arrays of the dense cores of a batch.
struct CoreApply {
    int64_t rows = 0;                  // sum of padded core sizes (rows of the gathered block)
    int32_t pieces = 0, tiles = 0;     // pieces per direction; row tiles
    const int32_t *vert = nullptr;     // [rows]: vertex of each padded core row, -1 on padding
    const int32_t *cpad = nullptr;     // [G]: padded core size
    const int64_t *dense_off = nullptr;// [G]: offset of the graph's cp x cp inverse
    const int64_t *tc_off = nullptr;   // [G]: first gathered row of the graph
    const CorePiece *piece_fwd = nullptr, *piece_bwd = nullptr;   // [pieces], grouped by tile
    const int32_t *tile_row = nullptr;          // [tiles]: first gathered row of the tile
    const int32_t *tile_piece_fwd = nullptr, *tile_piece_bwd = nullptr;   // [tiles + 1]
};
void launch_core_scatter(hipStream_t s, int64_t count, const int64_t *pos, const double *val, double *dense);
// M = L^{-1} of one unit-lower cp x cp block (strictly lower part in L).
void launch_core_invert(hipStream_t s, int cp, const double *L, double *M);
void launch_to_f32(hipStream_t s, int64_t count, const double *a, float *b);
void launch_to_f16(hipStream_t s, int64_t count, const double *a, _Float16 *b);
// Forward: out[core rows] = M (out[core rows]); backward: out[core rows] =
// M^T (D^+ src[core rows]).  tc: rows x k scratch; part: pieces x kCoreTile x k.
// src and out rows are stored as V (T, or scaled FP16).
template <typename T, typename V, bool BWD>
void launch_core_apply(hipStream_t s, const CoreApply &c, const T *M, int k, const double *diag, const V *src,
                       T *tc, T *part, V *out);
// The same for FP16 rows with M in FP16 (matrix-core FP16 products, FP32
// accumulation): tc holds FP16 gathered rows, part FP32 partials.
template <bool BWD>
void launch_core_apply_h(hipStream_t s, const CoreApply &c, const _Float16 *M, int k, const double *diag,
                         const _Float16 *src, _Float16 *tc, float *part, _Float16 *out);
template <typename TZ>
void launch_er_sample(hipStream_t s, const ErSampleArgs<TZ> &a);

}  // namespace ss
