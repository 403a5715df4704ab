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

template <typename TZ>
struct ErSampleArgs {
    int64_t m;
    int64_t ntiles;
    const int32_t *eu;
    const int32_t *ev;
    const double *ew;
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
    int *err;
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
template <typename T, bool BWD>
void launch_trisolve_level(hipStream_t s, int32_t t0, int32_t nhuge, int32_t hm, int32_t nlong, int32_t nmed,
                           int32_t t1, int waves, int rpw, int xcd, const int32_t *rows, const int32_t *ptr,
                           const int32_t *col, const T *val, const double *diag, int k, const T *in, T *out,
                           T *scratch, int *counter);
// Small levels [l0, l1) (each of at most kSmallRows rows) walked by one
// workgroup; level l holds positions [level_ptr[l], level_ptr[l+1]).
template <typename T, bool BWD>
void launch_tri_small(hipStream_t s, int32_t l0, int32_t l1, const int32_t *level_ptr, const int32_t *rows,
                      const int32_t *ptr, const int32_t *col, const T *val, const double *diag, int k, const T *in,
                      T *out);
// Dense cores (per graph: core_n[g] rows, vertex ids vert[core_off[g] ..],
// dense matrix at D + dense_off[g], row-major core_n x core_n).
// Forward: launch_core_head (the entries outside the core, schedule
// positions [t0, t1), the first nhuge of more than kWgRowMax entries), then launch_core_step<T, false> for steps 0 .. max
// blocks - 1 with D = L.  Backward: launch_core_bwd_init, then
// launch_core_step<T, true> for the same steps with D = L^T.
template <typename T>
void launch_core_head(hipStream_t s, int32_t t0, int32_t nhuge, int32_t t1, const int32_t *rows, const int32_t *ptr,
                      const int32_t *col, const T *val, const double *diag, int k, const T *in, T *out);
template <typename T>
void launch_core_bwd_init(hipStream_t s, int32_t count, const int32_t *vert, const double *diag, int k, const T *in,
                          T *out);
template <typename T, bool BWD>
void launch_core_step(hipStream_t s, int step, int graphs, int max_rows, const int32_t *core_n,
                      const int32_t *core_off, const int64_t *dense_off, const int32_t *vert, const T *D, int k,
                      T *out);
template <typename TZ>
void launch_er_sample(hipStream_t s, const ErSampleArgs<TZ> &a);

}  // namespace ss
