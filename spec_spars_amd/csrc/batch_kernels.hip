// Tile-based kernels of the batched device path.
//
// A batch packs G graphs into one block-diagonal system: graph g owns rows
// [row_off[g], row_off[g+1]) and edges [edge_off[g], edge_off[g+1]).  Rows
// are cut into tiles of row_tile_rows(n) rows and edges into tiles of kEdgeTile
// edges, never straddling two graphs; tile t knows its first row/edge and
// its graph.  Every reduction is a per-tile partial followed by a per
// (graph, column) sum over that graph's tiles in a fixed shape, so a graph's
// results are bit-identical alone and inside any batch.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch_kernels.h"
#include "hash_dev.h"

namespace ss {

#define BT_LANE (threadIdx.x & 63)
#define BT_WAVE (threadIdx.x >> 6)
#define BT_NW ((int)(blockDim.x >> 6))   // waves per tile block

__device__ __forceinline__ int bt_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Rows a wave of a streaming tile kernel handles per pass (rows w, w + NW,
// ... of the tile, their loads in flight together; reductions still take
// the rows in order).
constexpr int kBtRows = 4;

// Explicit fused multiply-adds: one rounding per term whatever the loop
// layout (contraction only joins a multiply and an add in one basic block).
__device__ __forceinline__ double bt_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float bt_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
__device__ __forceinline__ T bt_dot_sub(T acc, int q0, int q1, const int32_t *__restrict__ col,
                                        const T *__restrict__ val, const T *X, int k, int j) {
    int q = q0;
    for (; q + 7 < q1; q += 8) {
        T x[8], v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            x[u] = X[(size_t)col[q + u] * k + j];
            v[u] = val[q + u];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc = bt_fma(-v[u], x[u], acc);
    }
    for (; q < q1; q++) acc = bt_fma(-val[q], X[(size_t)col[q] * k + j], acc);
    return acc;
}

// Per-tile column partial of f over the tile's rows; the block's waves
// stride the rows, lanes take columns, waves combine in order.  Writes partial[t*k+j].
__device__ __forceinline__ void bt_tile_store(double acc, int k, int jb, double *partial, double (*sm)[64]) {
    const int w = BT_WAVE, lane = BT_LANE, j = jb + lane;
    sm[w][lane] = acc;
    __syncthreads();
    if (w == 0 && j < k) {
        double s = sm[0][lane];
        for (int u = 1; u < BT_NW; u++) s += sm[u][lane];
        partial[(size_t)blockIdx.x * k + j] = s;
    }
    __syncthreads();
}

// The same for a 32-column chunk held by both half-waves (lanes l and l + 32
// hold column jb + l of two row sets): the column's value is the sum over
// waves, in order, of the two halves' sums.
__device__ __forceinline__ void bt_tile_store32(double acc, int k, int jb, double *partial, double (*sm)[64]) {
    const int w = BT_WAVE, lane = BT_LANE, j = jb + (lane & 31);
    sm[w][lane] = acc;
    __syncthreads();
    if (w == 0 && lane < 32 && j < k) {
        double s = 0.0;
        for (int u = 0; u < BT_NW; u++) {
            s += sm[u][lane];
            s += sm[u][lane + 32];
        }
        partial[(size_t)blockIdx.x * k + j] = s;
    }
    __syncthreads();
}

// The partial of one column (lane value acc of column j) over the block's
// waves, in wave order, into partial[blockIdx.x * k + j].
__device__ __forceinline__ void bt_tile_store_col(double acc, int j, int k, double *partial, double (*sm)[64]) {
    const int w = BT_WAVE, lane = BT_LANE;
    sm[w][lane] = acc;
    __syncthreads();
    if (w == 0 && j < k) {
        double s = sm[0][lane];
        for (int u = 1; u < BT_NW; u++) s += sm[u][lane];
        partial[(size_t)blockIdx.x * k + j] = s;
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// (a) sketch: block per row tile, wave per row, seed/scale of the tile's graph
// ---------------------------------------------------------------------------
constexpr int kSketchSplit = 128;   // blocks per row tile in the sketch

// The Rademacher signs once per edge (the sketch visits every edge from both
// ends): bit l of S[E * nw + c] (nw = ceil(k / 64)) is 1 where column 64 c +
// l of global edge E (graph-local index E - edge_off[g], the hash input) is
// +1.  Block per edge tile, wave per edge, lane per column.
__global__ __launch_bounds__(256) void kb_sketch_signs(Tiles et, const int32_t *__restrict__ edge_off, int k,
                                                       SketchCols cols, int variant, uint64_t *__restrict__ S) {
    const int t = blockIdx.x, g = et.graph[t];
    const int e0 = et.start[t], e1 = et.start[t + 1];
    const int nw = (k + 63) / 64;
    const int eo = edge_off[g];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE, jc = g * k + (j < k ? j : k - 1);
        const uint64_t seed = cols.seed[jc];
        const uint64_t jh = (uint64_t)cols.j[jc];
        for (int e = e0 + BT_WAVE; e < e1; e += BT_NW) {
            const uint32_t x = (uint32_t)sign_hash((uint64_t)(e - eo), jh, seed, variant);
            const uint64_t b = __ballot((x & 0x80000000u) != 0u);
            if (BT_LANE == 0) S[(size_t)e * nw + jb / 64] = b;
        }
    }
}

// A lane's double from lane t (t wave-uniform).
__device__ __forceinline__ double bt_readlane_f64(double v, int t) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, t);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), t);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// With SIGNS, the Rademacher signs come from kb_sketch_signs' bits (the
// same values: +-1 times the scale, exactly) instead of a hash per incidence.
template <bool SIGNS>
__global__ __launch_bounds__(1024) void kb_sketch(Tiles rt, const int32_t *__restrict__ iptr,
                                                 const int32_t *__restrict__ ieid, const double *__restrict__ ib,
                                                 int k, SketchCols cols, int uniform, int variant,
                                                 const int32_t *__restrict__ edge_off,
                                                 const uint64_t *__restrict__ S, double *__restrict__ Y) {
#pragma clang fp contract(off)
    // kSketchSplit blocks per row tile, one row per wave: a row's cost is
    // its degree, and the rows of a power-law graph's hubs often sit
    // together, so rows spread over as many waves as there are rows
    const int t = blockIdx.x / kSketchSplit, sb = blockIdx.x % kSketchSplit;
    const int g = rt.graph[t];
    const int per = (rt.start[t + 1] - rt.start[t] + kSketchSplit - 1) / kSketchSplit;
    const int r0 = rt.start[t] + sb * per;
    const int r1 = r0 + per < rt.start[t + 1] ? r0 + per : rt.start[t + 1];
    const int nw = (k + 63) / 64;
    for (int row = r0 + BT_WAVE; row < r1; row += BT_NW) {
        const int r = bt_uniform(row);
        const int q0 = iptr[r], q1 = iptr[r + 1];
        for (int jb = 0; jb < k; jb += 64) {
            const int j = jb + BT_LANE, jc = g * k + (j < k ? j : k - 1);
            // column j's seed, scale and the column index its signs hash
            const uint64_t seed = cols.seed[jc];
            const double scale = cols.scale[jc];
            const uint64_t jh = (uint64_t)cols.j[jc];
            double acc = 0.0;
            if constexpr (SIGNS) {
                // 64 incidences at a time, one per lane (their sign words
                // loaded together), then summed in incidence order with
                // each one's word and value broadcast from its lane
                const uint64_t *Sg = S + (size_t)edge_off[g] * nw + jb / 64;
                for (int qb = q0; qb < q1; qb += 64) {
                    const int qq = qb + BT_LANE, cnt = q1 - qb < 64 ? q1 - qb : 64;
                    uint64_t wl = 0;
                    double bl = 0.0;
                    if (qq < q1) {
                        wl = Sg[(size_t)ieid[qq] * nw];
                        bl = ib[qq];
                    }
                    const uint32_t wlo = (uint32_t)wl, whi = (uint32_t)(wl >> 32);
                    for (int t = 0; t < cnt; t++) {
                        const uint32_t half = BT_LANE < 32 ? (uint32_t)__builtin_amdgcn_readlane((int)wlo, t)
                                                           : (uint32_t)__builtin_amdgcn_readlane((int)whi, t);
                        const double b = bt_readlane_f64(bl, t);
                        const double sv = ((half >> (BT_LANE & 31)) & 1u) ? scale : -scale;
                        const double term = sv * b;
                        acc = acc + term;
                    }
                }
            } else {
                for (int q = q0; q < q1; q++) {
                    const double sv = sign_value((uint64_t)ieid[q], jh, seed, uniform, variant) * scale;
                    const double term = sv * ib[q];
                    acc = acc + term;
                }
            }
            if (j < k) Y[(size_t)r * k + j] = acc;
        }
    }
}

// ---------------------------------------------------------------------------
// per-tile column partials: op 0 sum A*B, op 1 sum A, op 2 sum y (y / D)
// (y^T D^+ y with the reference's skip of D == 0), in the row split of
// update_r_rr (32-column chunks over both half-waves)
// ---------------------------------------------------------------------------
template <typename TA, typename TB>
__global__ __launch_bounds__(1024) void kb_tile_partial(Tiles rt, int k, const TA *__restrict__ A,
                                                       const TB *__restrict__ B, const double *__restrict__ diag,
                                                       int op, double *__restrict__ partial) {
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x;
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    // 32-column chunks; half-wave h of wave w takes rows 2 (w + NW m) + h
    const int half = BT_LANE >> 5, NW2 = 2 * BT_NW;
    for (int jb = 0; jb < k; jb += 32) {
        const int j = jb + (BT_LANE & 31);
        double acc = 0.0;
        if (j < k) {
            // kBtRows rows per pass (their loads in flight together), added in row order
            for (int rb = r0 + 2 * BT_WAVE + half; rb < r1; rb += kBtRows * NW2) {
                double a[kBtRows], b[kBtRows];
#pragma unroll
                for (int u = 0; u < kBtRows; u++) {
                    const int r = rb + u * NW2;
                    const bool on = r < r1;
                    a[u] = on ? (double)A[(size_t)r * k + j] : 0.0;
                    b[u] = on && op == 0 ? (double)B[(size_t)r * k + j] : (on && op == 2 ? diag[r] : 0.0);
                }
#pragma unroll
                for (int u = 0; u < kBtRows; u++) {
                    if (rb + u * NW2 >= r1) break;
                    if (op == 0) acc = bt_fma(a[u], b[u], acc);
                    else if (op == 1) acc += a[u];
                    else acc = b[u] != 0.0 ? bt_fma(a[u], a[u] / b[u], acc) : bt_fma(a[u], a[u], acc);
                }
            }
        }
        bt_tile_store32(acc, k, jb, partial, sm);
    }
}

// Per (graph, column) sum of the graph's tile partials by one 64-lane block:
// lane l adds tiles l, l+64, ... in order, then a fixed butterfly.
__device__ __forceinline__ double bt_graph_sum(const double *partial, const int32_t *gtile, int g, int k, int j) {
    // lane l adds tiles l, l + 64, ... in order; 8 of its loads in flight
    double s = 0.0;
    const int t1 = gtile[g + 1];
    for (int t = gtile[g] + BT_LANE; t < t1; t += 64 * 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = t + 64 * u < t1 ? partial[(size_t)(t + 64 * u) * k + j] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (t + 64 * u < t1) s += v[u];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    return s;
}

// ---------------------------------------------------------------------------
// PCG scalars, one 64-lane block per (graph, column), index gj = g*k + j
// ---------------------------------------------------------------------------
// psc (FP16 preconditioner rows): r . z was summed over rows scaled by
// psc[gj], so it is divided by psc[gj]^2.
__device__ __forceinline__ double bt_unscale_rz(double rz, const float *psc, int gj) {
    if (!psc) return rz;
    const double sc = (double)psc[gj];
    return sc > 0.0 ? rz / (sc * sc) : 0.0;
}

__global__ __launch_bounds__(64) void kb_pcg_init(int k, const int32_t *gtile, const double *p_bb,
                                                  const double *p_rz, PcgScalars s, const float *psc) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double bb = bt_graph_sum(p_bb, gtile, g, k, j);
    const double rz = bt_unscale_rz(bt_graph_sum(p_rz, gtile, g, k, j), psc, gj);
    if (BT_LANE == 0) {
        s.bnorm[gj] = sqrt(bb);
        s.rz[gj] = rz;
        s.active[gj] = bb > 0.0 ? 1 : 0;
        s.iters[gj] = 0;
        s.alpha[gj] = 0.0;
        s.beta[gj] = 0.0;
    }
}

// A column whose p . A p is not positive stops; with flag (reliable
// updates) it is flagged for a residual replacement (2: with a new search
// direction) instead.
__global__ __launch_bounds__(64) void kb_pcg_alpha(int k, const int32_t *gtile, const double *p_pq, PcgScalars s,
                                                   int *nactive, int *flag) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double pq = bt_graph_sum(p_pq, gtile, g, k, j);
    if (BT_LANE == 0) {
        double a = 0.0;
        if (s.active[gj]) {
            if (pq > 0.0) a = s.rz[gj] / pq;
            else if (flag) flag[gj] = 2;
            else s.active[gj] = 0;
        }
        s.alpha[gj] = a;
        if (gj == 0) *nactive = 0;
    }
}

__global__ __launch_bounds__(64) void kb_pcg_check(int k, const int32_t *gtile, const double *p_rr, int rule,
                                                   double tol, PcgScalars s, int *nactive) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double rr = bt_graph_sum(p_rr, gtile, g, k, j);
    if (BT_LANE == 0) {
        s.rr[gj] = rr;
        if (s.active[gj]) {
            s.iters[gj] += 1;
            const double rn = sqrt(rr);
            const bool stop = rule == 0 ? (rn <= sqrt(tol)) : (rn <= tol * s.bnorm[gj]);
            if (stop) s.active[gj] = 0;
        }
        if (s.active[gj]) atomicAdd(nactive, 1);
    }
}

__global__ __launch_bounds__(64) void kb_pcg_beta(int k, const int32_t *gtile, const double *p_rz, PcgScalars s,
                                                  const float *psc) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double rzn = bt_unscale_rz(bt_graph_sum(p_rz, gtile, g, k, j), psc, gj);
    if (BT_LANE == 0) {
        double b = 0.0;
        if (s.active[gj] && s.rz[gj] != 0.0) b = rzn / s.rz[gj];
        s.beta[gj] = b;
        s.rz[gj] = rzn;
    }
}

// Scale of the FP16 preconditioner rows of (graph, column) gj: s =
// sqrt(n_g / |r|^2), so the scaled residual has unit mean square and its
// preconditioned rows stay far inside FP16's range; inv = 1 / s (both 0
// for r = 0).  |r|^2 from the tile partials p_rr, else from rr[gj].
__global__ __launch_bounds__(64) void kb_prec_scale(int k, const int32_t *gtile, const double *p_rr, const double *rr,
                                                    const int32_t *row_off, float *sc, float *inv) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double v = p_rr ? bt_graph_sum(p_rr, gtile, g, k, j) : rr[gj];
    if (BT_LANE == 0) {
        const double n = (double)(row_off[g + 1] - row_off[g]);
        const bool on = v > 0.0 && n > 0.0;
        sc[gj] = on ? (float)sqrt(n / v) : 0.0f;
        inv[gj] = on ? (float)sqrt(v / n) : 0.0f;
    }
}

// out[gj] = graph sum (plain)
__global__ __launch_bounds__(64) void kb_graph_sums(int k, const int32_t *gtile, const double *partial,
                                                    double *out) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double v = bt_graph_sum(partial, gtile, g, k, j);
    if (BT_LANE == 0) out[gj] = v;
}

// ---------------------------------------------------------------------------
// component means: small components summed by one wave each (vertices in
// order); the graph's main component gets (graph sum - small sums) / size
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void kb_small_comp_sums(int nsmall, const int32_t *__restrict__ small_id,
                                                          const int32_t *__restrict__ cptr,
                                                          const int32_t *__restrict__ cvert, int k,
                                                          const double *__restrict__ Y, double *__restrict__ csum) {
    const int w = blockIdx.x * 4 + BT_WAVE;
    if (w >= nsmall) return;
    const int c = bt_uniform(small_id[w]);
    const int v0 = cptr[c], v1 = cptr[c + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE;
        if (j >= k) break;
        double s = 0.0;
        for (int q = v0; q < v1; q++) s += Y[(size_t)cvert[q] * k + j];
        csum[(size_t)c * k + j] = s;
    }
}

__global__ __launch_bounds__(256) void kb_comp_means(int ncomp, int k, const int32_t *__restrict__ comp_graph,
                                                     const int32_t *__restrict__ comp_off,
                                                     const int32_t *__restrict__ cptr, const double *__restrict__ gsum,
                                                     double *__restrict__ csum, double *__restrict__ cmean) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)ncomp * k) return;
    const int c = (int)(idx / k), j = (int)(idx - (int64_t)c * k);
    const int g = comp_graph[c];
    const double size = (double)(cptr[c + 1] - cptr[c]);
    double s;
    if (c == comp_off[g]) {          // the graph's main component
        s = gsum[(size_t)g * k + j];
        for (int c2 = comp_off[g] + 1; c2 < comp_off[g + 1]; c2++) s -= csum[(size_t)c2 * k + j];
    } else {
        s = csum[(size_t)c * k + j];
    }
    cmean[(size_t)c * k + j] = s / size;
}

template <typename TS, typename T>
__global__ __launch_bounds__(1024) void kb_project(Tiles rt, int k, const TS *__restrict__ Y,
                                                  const int32_t *__restrict__ comp_of_row,
                                                  const double *__restrict__ cmean, T *__restrict__ R,
                                                  T *__restrict__ B, T *__restrict__ X) {
    const int t = blockIdx.x;
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    for (int r = r0 + BT_WAVE; r < r1; r += BT_NW) {
        const int c = comp_of_row[r];
        for (int jb = 0; jb < k; jb += 64) {
            const int j = jb + BT_LANE;
            if (j >= k) break;
            const size_t o = (size_t)r * k + j;
            const T v = (T)((double)Y[o] - cmean[(size_t)c * k + j]);
            R[o] = v;
            B[o] = v;
            X[o] = 0;
        }
    }
}

// ---------------------------------------------------------------------------
// PCG vector kernels (block per row tile, per-(graph, column) scalars)
// ---------------------------------------------------------------------------
// P = Z + beta P, and the previous iteration's X += alpha P with the old P
// (the solution update rides along with the search-direction update, so
// P is read once for both).  A column with alpha == 0 keeps X untouched
// (on the first iteration P holds no direction yet).  Thread i of the
// tile's block works on column j = i % k of rows i / k, i / k + RP, ...
// (RP = blockDim / k rows per pass), so a pass moves RP whole rows as one
// contiguous run and no lane idles when k is below 64.
// With zinv (FP16 preconditioner rows, Z scaled by its column's s), z =
// Z zinv[gj]; a value outside FP16's range (never seen; it would only
// slow the column's convergence) counts as zero rather than spreading
// through P.
template <typename T, typename TZ>
__global__ __launch_bounds__(1024) void kb_update_px(Tiles rt, int k, const double *__restrict__ alpha,
                                                    const double *__restrict__ beta, const TZ *__restrict__ Z,
                                                    const float *__restrict__ zinv, T *__restrict__ P,
                                                    T *__restrict__ X) {
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    const int RP = (int)blockDim.x / k, i = (int)threadIdx.x;
    if (i >= RP * k) return;
    const int j = i % k, ro = i / k;
    const T a = (T)alpha[(size_t)g * k + j];
    const T b = (T)beta[(size_t)g * k + j];
    const T zs = zinv ? (T)zinv[(size_t)g * k + j] : (T)1;
    for (int rb = r0 + ro; rb < r1; rb += kBtRows * RP) {
        T p[kBtRows], z[kBtRows], x[kBtRows];
#pragma unroll
        for (int u = 0; u < kBtRows; u++) {
            const int r = rb + u * RP;
            const bool on = r < r1;
            const size_t o = (size_t)r * k + j;
            p[u] = on && (a != (T)0 || b != (T)0) ? P[o] : (T)0;
            z[u] = on ? (T)Z[o] : (T)0;
            x[u] = on && a != (T)0 ? X[o] : (T)0;
        }
        if (zinv) {
#pragma unroll
            for (int u = 0; u < kBtRows; u++) z[u] = __builtin_isfinite(z[u]) ? z[u] * zs : (T)0;
        }
#pragma unroll
        for (int u = 0; u < kBtRows; u++) {
            const int r = rb + u * RP;
            if (r >= r1) break;
            const size_t o = (size_t)r * k + j;
            if (a != (T)0) X[o] = bt_fma(a, p[u], x[u]);
            P[o] = b == (T)0 ? z[u] : bt_fma(b, p[u], z[u]);
        }
    }
}

// The same for FP16 preconditioner rows (Z scaled by 1 / zinv per column),
// NC = 2 or 4 columns per thread (one 4- or 8-byte load of Z, 8 or 16 bytes
// of P and X): thread i of the tile's block takes element i % ke of rows
// i / ke, i / ke + RP, ... (ke = k / NC elements per row).  Every column's
// arithmetic is kb_update_px's.
template <typename VZ, typename LV>
__global__ __launch_bounds__(1024) void kb_update_px_h(Tiles rt, int k, const double *__restrict__ alpha,
                                                      const double *__restrict__ beta, const VZ *__restrict__ Z,
                                                      const float *__restrict__ zinv, LV *__restrict__ P,
                                                      LV *__restrict__ X) {
    constexpr int NC = sizeof(LV) / 4;
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    const int ke = k / NC;
    const int RP = (int)blockDim.x / ke, i = (int)threadIdx.x;
    if (i >= RP * ke) return;
    const int e = i % ke, ro = i / ke;
    const size_t gc = (size_t)g * k + NC * e;
    LV a, b, zs;
    bool any_a = false, any_ab = false;
#pragma unroll
    for (int h = 0; h < NC; h++) {
        a[h] = (float)alpha[gc + h];
        b[h] = (float)beta[gc + h];
        zs[h] = zinv[gc + h];
        any_a = any_a || a[h] != 0.0f;
        any_ab = any_ab || a[h] != 0.0f || b[h] != 0.0f;
    }
    for (int rb = r0 + ro; rb < r1; rb += kBtRows * RP) {
        LV p[kBtRows], x[kBtRows], z[kBtRows];
#pragma unroll
        for (int u = 0; u < kBtRows; u++) {
            const int r = rb + u * RP;
            const bool on = r < r1;
            const size_t o = (size_t)r * ke + e;
            p[u] = on && any_ab ? P[o] : LV();
            z[u] = on ? __builtin_convertvector(Z[o], LV) : LV();
            x[u] = on && any_a ? X[o] : LV();
        }
#pragma unroll
        for (int u = 0; u < kBtRows; u++) {
            const int r = rb + u * RP;
            if (r >= r1) break;
            const size_t o = (size_t)r * ke + e;
            LV xn = x[u], pn;
#pragma unroll
            for (int h = 0; h < NC; h++) {
                const float zh = __builtin_isfinite(z[u][h]) ? z[u][h] * zs[h] : 0.0f;
                if (a[h] != 0.0f) xn[h] = __builtin_fmaf(a[h], p[u][h], x[u][h]);
                pn[h] = b[h] == 0.0f ? zh : __builtin_fmaf(b[h], p[u][h], zh);
            }
            if (any_a) X[o] = xn;
            P[o] = pn;
        }
    }
}

// y^T D^+ y per (tile, column) for FP16 preconditioner rows y (the r . z
// partials of the FP16-row path; D == 0 skipped as in kb_tile_partial's op
// 2): NC = 2 or 4 columns per lane (one 4- or 8-byte load), wave w of the
// block taking rows r0 + w, r0 + w + NW, ... (kBtRows of them in flight),
// each column's terms added in row order within a wave and the waves added
// in order: an order fixed by the tile alone, whatever k and NC are.
template <typename VZ, typename LV>
__global__ __launch_bounds__(1024) void kb_tile_rz_h(Tiles rt, int k, const VZ *__restrict__ W,
                                                    const double *__restrict__ diag, double *__restrict__ partial) {
    constexpr int NC = sizeof(LV) / 4;
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x, NW = BT_NW;
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    const int ke = k / NC;
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + BT_LANE, ee = e < ke ? e : ke - 1;
        double acc[NC];
#pragma unroll
        for (int h = 0; h < NC; h++) acc[h] = 0.0;
        for (int rb = r0 + BT_WAVE; rb < r1; rb += kBtRows * NW) {
            LV y[kBtRows];
            double d[kBtRows];
#pragma unroll
            for (int u = 0; u < kBtRows; u++) {
                const int r = rb + u * NW;
                const bool on = r < r1;
                y[u] = on ? __builtin_convertvector(W[(size_t)r * ke + ee], LV) : LV();
                d[u] = on ? diag[r] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kBtRows; u++) {
                if (rb + u * NW >= r1) break;
#pragma unroll
                for (int h = 0; h < NC; h++) {
                    const double a = (double)y[u][h];
                    acc[h] = d[u] != 0.0 ? bt_fma(a, a / d[u], acc[h]) : bt_fma(a, a, acc[h]);
                }
            }
        }
#pragma unroll
        for (int h = 0; h < NC; h++) bt_tile_store_col(acc[h], NC * e + h, k, partial, sm);
    }
}

// X += alpha P: the last iteration's solution update.
template <typename T>
__global__ __launch_bounds__(1024) void kb_update_x(Tiles rt, int k, const double *__restrict__ alpha,
                                                   const T *__restrict__ P, T *__restrict__ X) {
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE;
        if (j >= k) break;
        const T a = (T)alpha[(size_t)g * k + j];
        if (a == (T)0) continue;
        for (int r = r0 + BT_WAVE; r < r1; r += BT_NW) {
            const size_t o = (size_t)r * k + j;
            X[o] = bt_fma(a, P[o], X[o]);
        }
    }
}

// Lane l's value of v (l uniform; the lane need not be active).
__device__ __forceinline__ int bt_lane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float bt_lane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double bt_lane(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Q = A P with per-tile partials of P . Q.  Wave w of the tile takes rows
// r0 + w, r0 + w + NW, ... in order, kSpmmGroup of them at a time: when no
// row of the group has more than kSpmmShort entries, each row's entries load
// one per lane and every gather of the group is in flight together; longer
// rows go one at a time.  Every row is one running subtraction from zero
// over its entries in order and the wave's dot takes its rows in order, so
// the grouping does not change a bit.
constexpr int kSpmmGroup = 4;
constexpr int kSpmmShort = 8;

// Lane values of the SpMM: one column (T), or two adjacent FP32 columns
// (f32x2, k even) moved by one 8-byte access.
__device__ __forceinline__ f32x2 bt_fma(float a, f32x2 b, f32x2 c) {
    f32x2 r;
    r.x = __builtin_fmaf(a, b.x, c.x);
    r.y = __builtin_fmaf(a, b.y, c.y);
    return r;
}
template <typename V> __device__ __forceinline__ double bt_col(V v, int h) { return (double)v; }
template <> __device__ __forceinline__ double bt_col<f32x2>(f32x2 v, int h) { return (double)(h ? v.y : v.x); }
__device__ __forceinline__ f32x2 bt_widen(f16x2 x) { return __builtin_convertvector(x, f32x2); }
__device__ __forceinline__ f32x4 bt_widen(f16x4 x) { return __builtin_convertvector(x, f32x4); }
__device__ __forceinline__ f32x4 bt_fma(float a, f32x4 b, f32x4 c) {
    f32x4 r;
    r.x = __builtin_fmaf(a, b.x, c.x);
    r.y = __builtin_fmaf(a, b.y, c.y);
    r.z = __builtin_fmaf(a, b.z, c.z);
    r.w = __builtin_fmaf(a, b.w, c.w);
    return r;
}

// V: T, or f32x2 (two columns per lane, rows of k / 2 elements).  Every
// column's arithmetic is the same for either.
template <typename T, typename V>
__global__ __launch_bounds__(1024) void kb_spmm_dot(Tiles rt, const int32_t *__restrict__ rowptr,
                                                   const int32_t *__restrict__ col, const T *__restrict__ val, int k,
                                                   const V *__restrict__ P, V *__restrict__ Q,
                                                   double *__restrict__ partial) {
    // rows of up to 16 entries (4-byte lanes; 8 for 8-byte lanes, which
    // would not fit the registers at 16) take the all-gathers-in-flight path
    constexpr int NC = sizeof(V) / sizeof(T);   // columns per lane
    constexpr int G = kSpmmGroup, E = sizeof(V) == 4 ? 2 * kSpmmShort : kSpmmShort;
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x, NW = BT_NW, lane = BT_LANE;
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    const int ke = k / NC;
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + lane, ee = e < ke ? e : ke - 1;
        double acc[NC];
#pragma unroll
        for (int h = 0; h < NC; h++) acc[h] = 0.0;
        for (int rb = bt_uniform(r0 + BT_WAVE); rb < r1; rb += G * NW) {
            int rr[G], q0[G], len[G];
            int maxlen = 0;
            // the group's row records: lane u < G loads row u's (rows past
            // the tile read row rb's) in one vector load, read back per lane
            const int lu = BT_LANE < G ? BT_LANE : 0;
            const int ru = rb + lu * NW < r1 ? rb + lu * NW : rb;
            const int va = rowptr[ru], vb = rowptr[ru + 1];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const bool on = rb + u * NW < r1;
                rr[u] = on ? rb + u * NW : rb;
                const int a = bt_lane(va, u), b = bt_lane(vb, u);
                q0[u] = on ? a : 0;
                len[u] = on ? b - a : 0;
                maxlen = len[u] > maxlen ? len[u] : maxlen;
            }
            if (maxlen <= E) {
                int c[G];
                T v[G];
                V own[G], x[G][E];
#pragma unroll
                for (int u = 0; u < G; u++) {
                    c[u] = lane < len[u] ? col[q0[u] + lane] : 0;
                    v[u] = lane < len[u] ? val[q0[u] + lane] : (T)0;
                }
#pragma unroll
                for (int u = 0; u < G; u++) {
                    own[u] = P[(size_t)rr[u] * ke + ee];
#pragma unroll
                    for (int q = 0; q < E; q++) x[u][q] = q < len[u] ? P[(size_t)bt_lane(c[u], q) * ke + ee] : V();
                }
#pragma unroll
                for (int u = 0; u < G; u++) {
                    V a = V();
#pragma unroll
                    for (int q = 0; q < E; q++) a = bt_fma(-bt_lane(v[u], q), x[u][q], a);
                    if (rb + u * NW < r1 && e < ke) {
                        Q[(size_t)rr[u] * ke + e] = -a;
#pragma unroll
                        for (int h = 0; h < NC; h++) acc[h] = bt_fma(bt_col(own[u], h), -bt_col(a, h), acc[h]);
                    }
                }
            } else {
                // one row at a time: 64 entries at a time one per lane, 16
                // gathers in flight (zero-valued slots past the row's end add
                // exact zeros)
#pragma unroll
                for (int u = 0; u < G; u++) {
                    if (rb + u * NW >= r1) break;
                    V a = V();
                    for (int base = 0; base < len[u]; base += 64) {
                        const int n = len[u] - base;
                        const int cc = lane < n ? col[q0[u] + base + lane] : 0;
                        const T vv = lane < n ? val[q0[u] + base + lane] : (T)0;
                        for (int e0 = 0; e0 < 64 && e0 < n; e0 += 16) {
                            V x[16];
#pragma unroll
                            for (int q = 0; q < 16; q++)
                                x[q] = e0 + q < n ? P[(size_t)bt_lane(cc, e0 + q) * ke + ee] : V();
#pragma unroll
                            for (int q = 0; q < 16; q++) a = bt_fma(-bt_lane(vv, e0 + q), x[q], a);
                        }
                    }
                    if (e < ke) {
                        const size_t o = (size_t)rr[u] * ke + e;
                        Q[o] = -a;
                        const V pv = P[o];
#pragma unroll
                        for (int h = 0; h < NC; h++) acc[h] = bt_fma(bt_col(pv, h), -bt_col(a, h), acc[h]);
                    }
                }
            }
        }
#pragma unroll
        for (int h = 0; h < NC; h++) bt_tile_store_col(acc[h], NC * e + h, k, partial, sm);
    }
}

// Q = A P through the search directions' recurrence, for FP16
// preconditioner rows: P = Z + beta P_old gives A P = A Z + beta Q_old, so
// the gathers read the FP16 Z rows (scaled by 1 / zinv per column: half the
// bytes of the FP32 P rows) and each row then takes q = (A Z_s) zinv +
// beta q_old, with the per-tile partials of P . Q from the row's own FP32
// P.  Q_old is read only where beta is not zero (a first or restarted
// direction).  Two FP32 columns per lane (k even); rows and groups as in
// kb_spmm_dot, FP16 elements never loaded under a branch (a slot past a
// row's entries loads the row's own element and its term is selected
// away).  Q drifts from A P by FP32 rounding across iterations; the
// reliable updates' FP64 residuals are computed afresh, so the drift only
// shows in the recursive residual between replacements.  VZ: f16x2 (two
// columns per lane, LV = f32x2) or f16x4 (four, LV = f32x4; k a multiple of
// 4; 8 gathers in flight per row of a short group, for the registers).
template <typename VZ, typename LV>
__global__ __launch_bounds__(1024) void kb_spmm_z(Tiles rt, const int32_t *__restrict__ rowptr,
                                                 const int32_t *__restrict__ col, const float *__restrict__ val,
                                                 int k, const VZ *__restrict__ Z, const float *__restrict__ zinv,
                                                 const double *__restrict__ beta, const LV *__restrict__ P,
                                                 LV *__restrict__ Q, double *__restrict__ partial) {
    constexpr int NC = sizeof(LV) / 4;   // columns per lane
    constexpr int G = kSpmmGroup, E = NC == 2 ? 2 * kSpmmShort : kSpmmShort;
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x, NW = BT_NW, lane = BT_LANE, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    const int ke = k / NC;
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + lane, ee = e < ke ? e : ke - 1;
        const size_t gc = (size_t)g * k + NC * ee;
        LV zs, bq;
        bool use_old = false;
#pragma unroll
        for (int h = 0; h < NC; h++) {
            zs[h] = zinv[gc + h];
            bq[h] = (float)beta[gc + h];
            use_old = use_old || bq[h] != 0.0f;
        }
        double acc[NC];
#pragma unroll
        for (int h = 0; h < NC; h++) acc[h] = 0.0;
        for (int rb = bt_uniform(r0 + BT_WAVE); rb < r1; rb += G * NW) {
            int rr[G], q0[G], len[G];
            int maxlen = 0;
            const int lu = BT_LANE < G ? BT_LANE : 0;
            const int ru = rb + lu * NW < r1 ? rb + lu * NW : rb;
            const int va = rowptr[ru], vb = rowptr[ru + 1];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const bool on = rb + u * NW < r1;
                rr[u] = on ? rb + u * NW : rb;
                const int a = bt_lane(va, u), b = bt_lane(vb, u);
                q0[u] = on ? a : 0;
                len[u] = on ? b - a : 0;
                maxlen = len[u] > maxlen ? len[u] : maxlen;
            }
            LV az[G];
            if (maxlen <= E) {
                int c[G];
                float v[G];
                VZ x[G][E];
#pragma unroll
                for (int u = 0; u < G; u++) {
                    c[u] = lane < len[u] ? col[q0[u] + lane] : 0;
                    v[u] = lane < len[u] ? val[q0[u] + lane] : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < G; u++)
#pragma unroll
                    for (int q = 0; q < E; q++) x[u][q] = Z[(size_t)(q < len[u] ? bt_lane(c[u], q) : rr[u]) * ke + ee];
#pragma unroll
                for (int u = 0; u < G; u++) {
                    LV a = LV();
#pragma unroll
                    for (int q = 0; q < E; q++) a = bt_fma(-bt_lane(v[u], q), q < len[u] ? bt_widen(x[u][q]) : LV(), a);
                    az[u] = -a;
                }
            } else {
#pragma unroll
                for (int u = 0; u < G; u++) {
                    LV a = LV();
                    for (int base = 0; base < len[u]; base += 64) {
                        const int n = len[u] - base;
                        const int cc = lane < n ? col[q0[u] + base + lane] : 0;
                        const float vv = lane < n ? val[q0[u] + base + lane] : 0.0f;
                        for (int e0 = 0; e0 < 64 && e0 < n; e0 += 16) {
                            VZ x[16];
#pragma unroll
                            for (int q = 0; q < 16; q++)
                                x[q] = Z[(size_t)(e0 + q < n ? bt_lane(cc, e0 + q) : rr[u]) * ke + ee];
#pragma unroll
                            for (int q = 0; q < 16; q++)
                                a = bt_fma(-bt_lane(vv, e0 + q), e0 + q < n ? bt_widen(x[q]) : LV(), a);
                        }
                    }
                    az[u] = -a;
                }
            }
#pragma unroll
            for (int u = 0; u < G; u++) {
                if (rb + u * NW < r1 && e < ke) {
                    const size_t o = (size_t)rr[u] * ke + e;
                    LV qv = az[u] * zs;
                    if (use_old) {
                        const LV qo = Q[o];
#pragma unroll
                        for (int h = 0; h < NC; h++) qv[h] = __builtin_fmaf(bq[h], qo[h], az[u][h] * zs[h]);
                    }
                    Q[o] = qv;
                    const LV pv = P[o];
#pragma unroll
                    for (int h = 0; h < NC; h++) acc[h] = bt_fma((double)pv[h], (double)qv[h], acc[h]);
                }
            }
        }
#pragma unroll
        for (int h = 0; h < NC; h++) bt_tile_store_col(acc[h], NC * e + h, k, partial, sm);
    }
}

// R -= alpha Q with per-tile partials of |R|^2.  Columns go in chunks of
// 32, each held by both half-waves, and the half-waves and waves split the
// rows, so every lane is busy for k = 32 and a column's sum order does not
// depend on k (a request's columns give the same bits alone or merged with
// others').
template <typename T>
__global__ __launch_bounds__(1024) void kb_update_r_rr(Tiles rt, int k, const double *__restrict__ alpha,
                                                      const T *__restrict__ Q, T *__restrict__ R,
                                                      double *__restrict__ partial) {
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    // 32-column chunks; half-wave h of wave w takes rows 2 (w + NW m) + h
    const int half = BT_LANE >> 5, NW2 = 2 * BT_NW;
    for (int jb = 0; jb < k; jb += 32) {
        const int j = jb + (BT_LANE & 31);
        double acc = 0.0;
        if (j < k) {
            const T a = (T)alpha[(size_t)g * k + j];
            for (int rb = r0 + 2 * BT_WAVE + half; rb < r1; rb += kBtRows * NW2) {
                T q[kBtRows], rv[kBtRows];
#pragma unroll
                for (int u = 0; u < kBtRows; u++) {
                    const int r = rb + u * NW2;
                    const bool on = r < r1;
                    q[u] = on ? Q[(size_t)r * k + j] : (T)0;
                    rv[u] = on ? R[(size_t)r * k + j] : (T)0;
                }
#pragma unroll
                for (int u = 0; u < kBtRows; u++) {
                    const int r = rb + u * NW2;
                    if (r >= r1) break;
                    const T v = bt_fma(-a, q[u], rv[u]);
                    R[(size_t)r * k + j] = v;
                    acc = bt_fma((double)v, (double)v, acc);
                }
            }
        }
        bt_tile_store32(acc, k, jb, partial, sm);
    }
}

// out = B - A X in FP64, with per-tile partials of |out|^2; with a mask,
// only the (graph, column) pairs it marks are computed (the others keep
// out and get zero partials)
__global__ __launch_bounds__(1024) void kb_residual64(Tiles rt, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col, const double *__restrict__ val,
                                                     int k, const double *__restrict__ X, const double *__restrict__ B,
                                                     const int *__restrict__ mask, double *__restrict__ out,
                                                     double *__restrict__ partial) {
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE;
        const bool on = j < k && (!mask || mask[(size_t)g * k + j] != 0);
        double acc = 0.0;
        for (int row = r0 + BT_WAVE; row < r1; row += BT_NW) {
            const int r = bt_uniform(row);
            const int q0 = rowptr[r], q1 = rowptr[r + 1];
            if (on) {
                const size_t o = (size_t)r * k + j;
                const double v = bt_dot_sub<double>(B[o], q0, q1, col, val, X, k, j);
                out[o] = v;
                acc = bt_fma(v, v, acc);
            }
        }
        bt_tile_store(acc, k, jb, partial, sm);
    }
}

// FP64 -> FP32 with columns of converged (graph, column) pairs zeroed
// (keep == NULL converts every column)
__global__ __launch_bounds__(1024) void kb_convert_masked(Tiles rt, int k, const double *__restrict__ src,
                                                         const int *__restrict__ keep, float *__restrict__ dst) {
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE;
        if (j >= k) break;
        const bool on = !keep || keep[(size_t)g * k + j] != 0;
        for (int r = r0 + BT_WAVE; r < r1; r += BT_NW) {
            const size_t o = (size_t)r * k + j;
            dst[o] = on ? (float)src[o] : 0.0f;
        }
    }
}

__global__ __launch_bounds__(1024) void kb_accumulate(Tiles rt, int k, const float *__restrict__ d,
                                                     double *__restrict__ x) {
    const int t = blockIdx.x;
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE;
        if (j >= k) break;
        for (int r = r0 + BT_WAVE; r < r1; r += BT_NW) x[(size_t)r * k + j] += (double)d[(size_t)r * k + j];
    }
}

// ---------------------------------------------------------------------------
// Mixed precision with reliable updates.  One FP32 PCG runs on; a column
// whose recursive FP32 residual has fallen by delta since its last
// replacement (or to the stop target) folds its FP32 iterate into the FP64
// solution and takes the FP64 true residual b - A x as its residual, its
// search direction kept.  Every step is per (graph, column): a column's
// arithmetic never depends on when the others replace.
// ---------------------------------------------------------------------------

// bnorm = |B64| and rr_ref = |B64|^2 from rr_ref (uploaded |B64|^2); flags
// and the flag count cleared.
__global__ __launch_bounds__(256) void kb_reliable_init(int gk, PcgScalars s, double *rr_ref, int *flag, int *status) {
    const int gj = blockIdx.x * blockDim.x + threadIdx.x;
    if (gj == 0) status[2] = 0;
    if (gj >= gk) return;
    s.bnorm[gj] = sqrt(rr_ref[gj]);
    flag[gj] = 0;
}

// A column that reached its stop target while others of its replacement
// group have not; it waits for them (kb_reliable_group).
constexpr int kStopCandidate = 4;

// The recursive residual's |R|^2 from the tiles; an active column is
// flagged when it reached rr_ref * delta2 or the stop target.  With defer
// (replacement groups), a column at its stop target alone is only marked
// kStopCandidate: its group verifies its stop when every active column of
// it is there, or at its next replacement.  Counts the active (status[0],
// cleared by kb_pcg_alpha) and flagged (status[2]) columns.
__global__ __launch_bounds__(64) void kb_reliable_check(int k, const int32_t *gtile, const double *p_rr, int rule,
                                                        double tol, double delta2, int defer, PcgScalars s,
                                                        const double *rr_ref, int *flag, int *status) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double rr = bt_graph_sum(p_rr, gtile, g, k, j);
    if (BT_LANE == 0) {
        s.rr[gj] = rr;
        if (s.active[gj]) {
            s.iters[gj] += 1;
            const bool target = rule == 0 ? rr <= tol : sqrt(rr) <= tol * s.bnorm[gj];
            if (flag[gj] || rr <= delta2 * rr_ref[gj] || (target && !defer)) {
                if (!flag[gj]) flag[gj] = 1;
                atomicAdd(status + 2, 1);
            } else if (target) {
                flag[gj] = kStopCandidate;
            }
            atomicAdd(status, 1);
        }
    }
}

// Replacement groups (columns [gs[j], ge[j]) of a graph, the columns of
// one request): a group replaces when one of its active columns is flagged
// (drift, or a non-positive p . A p) or when every active column of it is
// at its stop target; its other active columns are then flagged too, so a
// request's columns replace together.  A stop candidate of a group that
// does not replace goes back to unflagged and iterates on.  The decision is
// a function of the flags kb_reliable_check left, and a flag rewritten here
// is rewritten the way that decision goes (1 when the group replaces, 0 when
// not), so it does not depend on the order the threads run in.
__global__ __launch_bounds__(256) void kb_reliable_group(int gk, int k, const int32_t *__restrict__ gs,
                                                        const int32_t *__restrict__ ge, PcgScalars s, int *flag,
                                                        int *status) {
    const int gj = blockIdx.x * blockDim.x + threadIdx.x;
    if (gj >= gk) return;
    const int g = gj / k, j = gj - g * k;
    const int own = flag[gj];
    if (!s.active[gj] || (own & 3)) return;
    bool any = false, all = true;
    for (int c = gs[j]; c < ge[j]; c++) {
        if (!s.active[g * k + c]) continue;
        const int f = __hip_atomic_load(&flag[g * k + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (f & 3) {
            any = true;
            break;
        }
        if (f == 0) all = false;
    }
    if (any || all) {
        __hip_atomic_store(&flag[gj], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        atomicAdd(status + 2, 1);
    } else if (own) {
        __hip_atomic_store(&flag[gj], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Replacement, step 1 (flagged columns): x = X32 + alpha P (the FP32
// iterate with the iteration's pending update), X64 += x, XF = x (what the
// FP64 residual gathers) and X32 = 0 (every flagged column's FP32 iterate
// starts again: its residual is replaced, or it stops).
__global__ __launch_bounds__(1024) void kb_reliable_fold(Tiles rt, int k, const int *__restrict__ flag,
                                                        const double *__restrict__ alpha, const float *__restrict__ P,
                                                        float *__restrict__ X32, double *__restrict__ X64,
                                                        float *__restrict__ XF) {
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE;
        if (j >= k) break;
        if (!flag[(size_t)g * k + j]) continue;
        const float a = (float)alpha[(size_t)g * k + j];
        for (int r = r0 + BT_WAVE; r < r1; r += BT_NW) {
            const size_t o = (size_t)r * k + j;
            const float x = a != 0.0f ? bt_fma(a, P[o], X32[o]) : X32[o];
            X64[o] += (double)x;
            XF[o] = x;
            X32[o] = 0.0f;
        }
    }
}

// Replacement, step 2 (flagged columns): R64 -= A x with x the folded
// FP32 iterate (R64 held b - A X64 before the fold, so it becomes the FP64
// true residual of the new X64; the gathers read FP32 rows), with per-tile
// partials of |R64|^2 (zero for the other columns); R32 = (float) R64 on
// the same columns (a column the next step finds done keeps it unused:
// it no longer iterates).
__global__ __launch_bounds__(1024) void kb_reliable_residual(Tiles rt, const int32_t *__restrict__ rowptr,
                                                            const int32_t *__restrict__ col,
                                                            const double *__restrict__ val, int k,
                                                            const int *__restrict__ flag,
                                                            const float *__restrict__ X, double *__restrict__ R,
                                                            float *__restrict__ R32, double *__restrict__ partial) {
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + BT_LANE;
        const bool on = j < k && flag[(size_t)g * k + j] != 0;
        double acc = 0.0;
        // a chunk with no replaced column skips its rows (the waves still
        // meet at the partial's barrier)
        for (int row = r0 + BT_WAVE; __ballot(on) != 0 && row < r1; row += BT_NW) {
            const int r = bt_uniform(row);
            const int q0 = rowptr[r], q1 = rowptr[r + 1];
            if (on) {
                const size_t o = (size_t)r * k + j;
                double v = R[o];
                int q = q0;
                for (; q + 7 < q1; q += 8) {
                    float x[8];
                    double w[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        x[u] = X[(size_t)col[q + u] * k + j];
                        w[u] = val[q + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) v = bt_fma(-w[u], (double)x[u], v);
                }
                for (; q < q1; q++) v = bt_fma(-val[q], (double)X[(size_t)col[q] * k + j], v);
                R[o] = v;
                R32[o] = (float)v;
                acc = bt_fma(v, v, acc);
            }
        }
        bt_tile_store(acc, k, jb, partial, sm);
    }
}

// The same with four columns per lane (k a multiple of 4): FP32 rows of X
// gathered 16 bytes at a time, FP64 rows of R 32; the lane's columns that
// are not flagged keep their R and add nothing to the partials.  Every
// column's arithmetic is kb_reliable_residual's.
__global__ __launch_bounds__(1024) void kb_reliable_residual4(Tiles rt, const int32_t *__restrict__ rowptr,
                                                             const int32_t *__restrict__ col,
                                                             const double *__restrict__ val, int k,
                                                             const int *__restrict__ flag,
                                                             const f32x4 *__restrict__ X, f64x4 *__restrict__ R,
                                                             float *__restrict__ R32, double *__restrict__ partial) {
    __shared__ double sm[kMaxTileWaves][64];
    const int t = blockIdx.x, g = rt.graph[t];
    const int r0 = rt.start[t], r1 = rt.start[t + 1];
    const int ke = k / 4;
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + BT_LANE;
        bool onh[4], on = false;
#pragma unroll
        for (int h = 0; h < 4; h++) {
            onh[h] = e < ke && flag[(size_t)g * k + 4 * e + h] != 0;
            on = on || onh[h];
        }
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        // a chunk with no replaced column skips its rows (the waves still
        // meet at the partials' barriers)
        for (int row = r0 + BT_WAVE; __ballot(on) != 0 && row < r1; row += BT_NW) {
            const int r = bt_uniform(row);
            const int q0 = rowptr[r], q1 = rowptr[r + 1];
            if (on) {
                const size_t o = (size_t)r * ke + e;
                const f64x4 v0 = R[o];
                f64x4 v = v0;
                int q = q0;
                for (; q + 7 < q1; q += 8) {
                    f32x4 x[8];
                    double w[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        x[u] = X[(size_t)col[q + u] * ke + e];
                        w[u] = val[q + u];
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++)
#pragma unroll
                        for (int h = 0; h < 4; h++) v[h] = bt_fma(-w[u], (double)x[u][h], v[h]);
                }
                for (; q < q1; q++) {
                    const f32x4 x = X[(size_t)col[q] * ke + e];
                    const double w = val[q];
#pragma unroll
                    for (int h = 0; h < 4; h++) v[h] = bt_fma(-w, (double)x[h], v[h]);
                }
#pragma unroll
                for (int h = 0; h < 4; h++) {
                    if (!onh[h]) {
                        v[h] = v0[h];
                    } else {
                        acc[h] = bt_fma(v[h], v[h], acc[h]);
                        R32[4 * o + h] = (float)v[h];
                    }
                }
                R[o] = v;
            }
        }
#pragma unroll
        for (int h = 0; h < 4; h++) bt_tile_store_col(acc[h], 4 * e + h, k, partial, sm);
    }
}

// Replacement, step 3, with |B64 - A X64|^2 of the flagged columns from the
// SpMM tiles: a flagged column that meets the FP64 stop rule is done; the
// others take it as their reference (their FP32 residual was replaced in
// step 2).  Where the true residual exceeds the recursive one by more
// than kReliableJump (the FP32 recursion ran below what its arithmetic
// resolves) or p . A p broke down, the column's next search direction is
// its preconditioned residual alone (rz = 0 makes beta zero).  Flagged
// columns' alpha is cleared (their update is folded); flags and the flag
// count are cleared.
constexpr double kReliableJump = 4.0;   // |r|^2 ratio, true over recursive
__global__ __launch_bounds__(64) void kb_reliable_apply(int k, const int32_t *gtile, const double *p_rr, int rule,
                                                        double tol, PcgScalars s, double *rr_ref, int *flag,
                                                        int *status) {
    const int gj = blockIdx.x, g = gj / k, j = gj - g * k;
    const double rr = bt_graph_sum(p_rr, gtile, g, k, j);
    if (BT_LANE == 0) {
        if (flag[gj]) {
            s.alpha[gj] = 0.0;
            const bool done = rr == 0.0 || (rule == 0 ? rr <= tol : sqrt(rr) <= tol * s.bnorm[gj]);
            if (done) {
                s.active[gj] = 0;
                atomicSub(status, 1);
            } else {
                if (flag[gj] == 2 || rr > kReliableJump * s.rr[gj]) s.rz[gj] = 0.0;
                s.rr[gj] = rr;   // the replaced FP32 residual's |r|^2
            }
            rr_ref[gj] = rr;   // a done column's final |b - A x|^2
            flag[gj] = 0;
        }
        if (gj == 0) status[2] = 0;
    }
}

// ---------------------------------------------------------------------------
// (d) ER + probability + coin + sample + compaction over edge tiles.
// Coins use the graph's key and the edge's index within its graph.
// ---------------------------------------------------------------------------
#define BT_EPOCH_SHIFT 40
#define BT_FLAG_SHIFT 38
#define BT_VALUE_MASK ((1ull << 38) - 1)

constexpr int kErCols = 16;   // Z columns staged per pass of the ER kernel

__global__ __launch_bounds__(256) void kb_er_sample(BatchSampleArgs a) {
#pragma clang fp contract(off)
    __shared__ uint32_t s_tile;
    __shared__ int s_wcnt[4];
    __shared__ int64_t s_excl;
    __shared__ double s_t[4][64][kErCols + 1];
    if (threadIdx.x == 0) s_tile = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const int g = a.et.graph[tile];
    const int64_t e0 = a.et.start[tile], e1 = a.et.start[tile + 1];
    const int64_t e = e0 + threadIdx.x;
    const int lane = BT_LANE, wv = BT_WAVE;
    bool keep = false;
    int32_t eu = 0, ev = 0;
    double kwv = 0.0;
    if (e < e1) {
        eu = a.eu[e];
        ev = a.ev[e];
    }
    // s = ||Z_u - Z_v||^2 summed in column order.  The rows are gathered
    // coalesced: in each pass a wave stages (Z_u - Z_v) of kErCols columns
    // for its 64 edges, four edges per load (16 lanes each, one 128-byte
    // run per row), then every lane sums its own edge's terms in order.
    double s = 0.0;
    if (!a.p_in) {
        const int sub = lane >> 4, cl = lane & (kErCols - 1);
        for (int jb = 0; jb < a.k; jb += kErCols) {
            const int j = jb + cl;
#pragma unroll
            for (int i0 = 0; i0 < 64; i0 += 4) {
                const int i = i0 + sub;
                const int u = __shfl(eu, i, 64), v = __shfl(ev, i, 64);
                double t = 0.0;
                if (j < a.k) t = a.Z[(size_t)u * a.ldz + j] - a.Z[(size_t)v * a.ldz + j];
                s_t[wv][i][cl] = t;
            }
            __syncthreads();
            const int nj = a.k - jb < kErCols ? a.k - jb : kErCols;
            for (int c = 0; c < nj; c++) {
                const double t = s_t[wv][lane][c];
                s = s + t * t;
            }
            __syncthreads();
        }
    }
    if (e < e1) {
        const double w = a.ew[e];
        const double lval = -1.0 * w;
        double p;
        if (a.p_in) {
            p = a.p_in[e];
        } else {
            const double d = sqrt(s);
            if (a.d_out) a.d_out[e] = d;
            const double d2 = d * d;
            const double x = a.beta[g] * -1.0 * lval * d2 / (double)a.k;
            p = 1.0 * (x < 1.0 ? x : 1.0);
        }
        if (a.p_out) a.p_out[e] = p;
        CoinKey ck;
#pragma unroll
        for (int i = 0; i < 8; i++) ck.k[i] = a.keys[g * 8 + i];
        const double c = coin_value(ck, (uint64_t)(e - a.edge_off[g]));
        keep = c < p;
        if (a.mask) a.mask[e] = keep ? 1 : 0;
        if (keep) {
            const double target = w / sqrt(p);
            const double change = w - target;
            kwv = -1.0 * (lval + change);
        }
    }
    const uint64_t bal = __ballot(keep);
    const int wpre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_wcnt[wv] = __popcll(bal);
    __syncthreads();
    int woff = 0;
    for (int w2 = 0; w2 < wv; w2++) woff += s_wcnt[w2];
    if (threadIdx.x == 0) {
        const uint64_t total = (uint64_t)(s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3]);
        const uint64_t ep = (uint64_t)a.epoch << BT_EPOCH_SHIFT;
        int64_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(&a.tile_state[0], ep | (2ull << BT_FLAG_SHIFT) | total, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&a.tile_state[tile], ep | (1ull << BT_FLAG_SHIFT) | total, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            int64_t j = (int64_t)tile - 1;
            int spins = 0;
            while (j >= 0) {
                const uint64_t st = __hip_atomic_load(&a.tile_state[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t sep = (uint32_t)(st >> BT_EPOCH_SHIFT);
                const uint32_t flag = (uint32_t)(st >> BT_FLAG_SHIFT) & 3u;
                if (sep != a.epoch || flag == 0) {
                    if (++spins > a.spin_limit) {
                        __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (int64_t)(st & BT_VALUE_MASK);
                if (flag == 2) break;
                j--;
            }
            __hip_atomic_store(&a.tile_state[tile], ep | (2ull << BT_FLAG_SHIFT) | ((uint64_t)excl + total),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_excl = excl;
        a.tile_incl[tile] = excl + (int64_t)total;
    }
    __syncthreads();
    const int64_t pos = s_excl + woff + wpre;
    if (keep && pos < a.m) {
        a.ku[pos] = eu;
        a.kv[pos] = ev;
        a.kw[pos] = kwv;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned bt_grid(int64_t items, int per) {
    int64_t g = (items + per - 1) / per;
    return (unsigned)(g < 1 ? 1 : g);
}

void bt_sketch(hipStream_t s, Tiles rt, const int32_t *iptr, const int32_t *ieid, const double *ib, int k,
               SketchCols cols, int uniform, double *Y, Tiles et, const int32_t *edge_off, uint64_t *S) {
    if (S && !uniform && et.count > 0) {
        hipLaunchKernelGGL(kb_sketch_signs, dim3(et.count), dim3(256), 0, s, et, edge_off, k, cols, 0, S);
        hipLaunchKernelGGL(kb_sketch<true>, dim3(rt.count * kSketchSplit), dim3(64 * rt.waves), 0, s, rt, iptr, ieid,
                           ib, k, cols, uniform, 0, edge_off, S, Y);
    } else {
        hipLaunchKernelGGL(kb_sketch<false>, dim3(rt.count * kSketchSplit), dim3(64 * rt.waves), 0, s, rt, iptr, ieid,
                           ib, k, cols, uniform, 0, edge_off, S, Y);
    }
}

template <typename TA, typename TB>
void bt_tile_partial(hipStream_t s, Tiles rt, int k, const TA *A, const TB *B, const double *diag, int op,
                     double *partial) {
    // all 16 waves whatever the tuned tile width: the reduction reads one
    // block per tile with nothing else to hide its load latency
    hipLaunchKernelGGL((kb_tile_partial<TA, TB>), dim3(rt.count), dim3(64 * kMaxTileWaves), 0, s, rt, k, A, B, diag,
                       op, partial);
}
template void bt_tile_partial<double, double>(hipStream_t, Tiles, int, const double *, const double *,
                                              const double *, int, double *);
template void bt_tile_partial<float, float>(hipStream_t, Tiles, int, const float *, const float *, const double *,
                                            int, double *);
template void bt_tile_partial<_Float16, _Float16>(hipStream_t, Tiles, int, const _Float16 *, const _Float16 *,
                                                  const double *, int, double *);

void bt_pcg_init(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p_bb, const double *p_rz,
                 PcgScalars sc, const float *psc) {
    hipLaunchKernelGGL(kb_pcg_init, dim3(gk), dim3(64), 0, s, k, gtile, p_bb, p_rz, sc, psc);
}
void bt_prec_scale(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p_rr, const double *rr,
                   const int32_t *row_off, float *sc, float *inv) {
    hipLaunchKernelGGL(kb_prec_scale, dim3(gk), dim3(64), 0, s, k, gtile, p_rr, rr, row_off, sc, inv);
}
void bt_pcg_alpha(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, PcgScalars sc, int *nactive,
                  int *flag) {
    hipLaunchKernelGGL(kb_pcg_alpha, dim3(gk), dim3(64), 0, s, k, gtile, p, sc, nactive, flag);
}
void bt_pcg_check(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, int rule, double tol,
                  PcgScalars sc, int *nactive) {
    hipLaunchKernelGGL(kb_pcg_check, dim3(gk), dim3(64), 0, s, k, gtile, p, rule, tol, sc, nactive);
}
void bt_pcg_beta(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, PcgScalars sc,
                 const float *psc) {
    hipLaunchKernelGGL(kb_pcg_beta, dim3(gk), dim3(64), 0, s, k, gtile, p, sc, psc);
}
void bt_graph_sums(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, double *out) {
    hipLaunchKernelGGL(kb_graph_sums, dim3(gk), dim3(64), 0, s, k, gtile, p, out);
}
void bt_small_comp_sums(hipStream_t s, int nsmall, const int32_t *small_id, const int32_t *cptr, const int32_t *cvert,
                        int k, const double *Y, double *csum) {
    if (nsmall <= 0) return;
    hipLaunchKernelGGL(kb_small_comp_sums, dim3(bt_grid(nsmall, 4)), dim3(256), 0, s, nsmall, small_id, cptr, cvert, k,
                       Y, csum);
}
void bt_comp_means(hipStream_t s, int ncomp, int k, const int32_t *comp_graph, const int32_t *comp_off,
                   const int32_t *cptr, const double *gsum, double *csum, double *cmean) {
    hipLaunchKernelGGL(kb_comp_means, dim3(bt_grid((int64_t)ncomp * k, 256)), dim3(256), 0, s, ncomp, k, comp_graph,
                       comp_off, cptr, gsum, csum, cmean);
}
template <typename TS, typename T>
void bt_project(hipStream_t s, Tiles rt, int k, const TS *Y, const int32_t *comp_of_row, const double *cmean, T *R,
                T *B, T *X) {
    hipLaunchKernelGGL((kb_project<TS, T>), dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, k, Y, comp_of_row, cmean, R, B, X);
}
template void bt_project<double, double>(hipStream_t, Tiles, int, const double *, const int32_t *, const double *,
                                         double *, double *, double *);
template void bt_project<float, float>(hipStream_t, Tiles, int, const float *, const int32_t *, const double *,
                                       float *, float *, float *);
// Row-block launches (update_px: no reduction, so the mapping leaves every
// value's bits alone): thread i of a tile block works on column i % k of
// every (blockDim / k)-th row, so at least k threads are needed.
static int rowblock_threads(const Tiles &rt, int k) {
    const int w = 64 * rt.waves, need = (k + 63) / 64 * 64;
    return w > need ? w : need;
}
template <typename T, typename TZ>
void bt_update_px(hipStream_t s, Tiles rt, int k, const double *alpha, const double *beta, const TZ *Z,
                  const float *zinv, T *P, T *X) {
    hipLaunchKernelGGL((kb_update_px<T, TZ>), dim3(rt.count), dim3(rowblock_threads(rt, k)), 0, s, rt, k, alpha, beta,
                       Z, zinv, P, X);
}
template void bt_update_px<double, double>(hipStream_t, Tiles, int, const double *, const double *, const double *,
                                           const float *, double *, double *);
template void bt_update_px<float, float>(hipStream_t, Tiles, int, const double *, const double *, const float *,
                                         const float *, float *, float *);
template void bt_update_px<float, _Float16>(hipStream_t, Tiles, int, const double *, const double *,
                                            const _Float16 *, const float *, float *, float *);
void bt_update_px_h(hipStream_t s, Tiles rt, int k, const double *alpha, const double *beta, const _Float16 *Z,
                    const float *zinv, float *P, float *X) {
    if (k % 4 == 0 && k >= 256) {
        const int ke = k / 4, th = std::max(64 * rt.waves, (ke + 63) / 64 * 64);
        hipLaunchKernelGGL((kb_update_px_h<f16x4, f32x4>), dim3(rt.count), dim3(th), 0, s, rt, k, alpha, beta,
                           reinterpret_cast<const f16x4 *>(Z), zinv, reinterpret_cast<f32x4 *>(P),
                           reinterpret_cast<f32x4 *>(X));
    } else {
        const int ke = k / 2, th = std::max(64 * rt.waves, (ke + 63) / 64 * 64);
        hipLaunchKernelGGL((kb_update_px_h<f16x2, f32x2>), dim3(rt.count), dim3(th), 0, s, rt, k, alpha, beta,
                           reinterpret_cast<const f16x2 *>(Z), zinv, reinterpret_cast<f32x2 *>(P),
                           reinterpret_cast<f32x2 *>(X));
    }
}
void bt_tile_rz_h(hipStream_t s, Tiles rt, int k, const _Float16 *W, const double *diag, double *partial) {
    if (k % 4 == 0 && k >= 256)
        hipLaunchKernelGGL((kb_tile_rz_h<f16x4, f32x4>), dim3(rt.count), dim3(64 * kMaxTileWaves), 0, s, rt, k,
                           reinterpret_cast<const f16x4 *>(W), diag, partial);
    else
        hipLaunchKernelGGL((kb_tile_rz_h<f16x2, f32x2>), dim3(rt.count), dim3(64 * kMaxTileWaves), 0, s, rt, k,
                           reinterpret_cast<const f16x2 *>(W), diag, partial);
}
template <typename T>
void bt_update_x(hipStream_t s, Tiles rt, int k, const double *alpha, const T *P, T *X) {
    hipLaunchKernelGGL(kb_update_x<T>, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, k, alpha, P, X);
}
template void bt_update_x<double>(hipStream_t, Tiles, int, const double *, const double *, double *);
template void bt_update_x<float>(hipStream_t, Tiles, int, const double *, const float *, float *);
void bt_spmm_z(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const float *val, int k,
               const _Float16 *Z, const float *zinv, const double *beta, const float *P, float *Q, double *partial) {
    if (k % 4 == 0 && k >= 256)
        hipLaunchKernelGGL((kb_spmm_z<f16x4, f32x4>), dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, rowptr, col, val,
                           k, reinterpret_cast<const f16x4 *>(Z), zinv, beta, reinterpret_cast<const f32x4 *>(P),
                           reinterpret_cast<f32x4 *>(Q), partial);
    else
        hipLaunchKernelGGL((kb_spmm_z<f16x2, f32x2>), dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, rowptr, col, val,
                           k, reinterpret_cast<const f16x2 *>(Z), zinv, beta, reinterpret_cast<const f32x2 *>(P),
                           reinterpret_cast<f32x2 *>(Q), partial);
}
template <typename T>
void bt_spmm_dot(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const T *val, int k, const T *P,
                 T *Q, double *partial) {
    if constexpr (sizeof(T) == 4) {
        if (k % 2 == 0 && k >= 128) {
            // FP32 with k even and at least 128: two columns per lane
            hipLaunchKernelGGL((kb_spmm_dot<float, f32x2>), dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, rowptr,
                               col, val, k, reinterpret_cast<const f32x2 *>(P), reinterpret_cast<f32x2 *>(Q), partial);
            return;
        }
    }
    hipLaunchKernelGGL((kb_spmm_dot<T, T>), dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, rowptr, col, val, k, P, Q,
                       partial);
}
template void bt_spmm_dot<double>(hipStream_t, Tiles, const int32_t *, const int32_t *, const double *, int,
                                  const double *, double *, double *);
template void bt_spmm_dot<float>(hipStream_t, Tiles, const int32_t *, const int32_t *, const float *, int,
                                 const float *, float *, double *);
template <typename T>
void bt_update_r_rr(hipStream_t s, Tiles rt, int k, const double *alpha, const T *Q, T *R, double *partial) {
    hipLaunchKernelGGL(kb_update_r_rr<T>, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, k, alpha, Q, R, partial);
}
template void bt_update_r_rr<double>(hipStream_t, Tiles, int, const double *, const double *, double *, double *);
template void bt_update_r_rr<float>(hipStream_t, Tiles, int, const double *, const float *, float *, double *);
void bt_residual64(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const double *val, int k,
                   const double *X, const double *B, double *out, double *partial, const int *mask) {
    hipLaunchKernelGGL(kb_residual64, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, rowptr, col, val, k, X, B, mask,
                       out, partial);
}
void bt_convert_masked(hipStream_t s, Tiles rt, int k, const double *src, const int *keep, float *dst) {
    hipLaunchKernelGGL(kb_convert_masked, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, k, src, keep, dst);
}
void bt_accumulate(hipStream_t s, Tiles rt, int k, const float *d, double *x) {
    hipLaunchKernelGGL(kb_accumulate, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, k, d, x);
}
void bt_reliable_init(hipStream_t s, int gk, PcgScalars sc, double *rr_ref, int *flag, int *status) {
    hipLaunchKernelGGL(kb_reliable_init, dim3(bt_grid(gk, 256)), dim3(256), 0, s, gk, sc, rr_ref, flag, status);
}
void bt_reliable_check(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, int rule, double tol,
                       double delta2, bool defer, PcgScalars sc, const double *rr_ref, int *flag, int *status) {
    hipLaunchKernelGGL(kb_reliable_check, dim3(gk), dim3(64), 0, s, k, gtile, p, rule, tol, delta2, defer ? 1 : 0, sc,
                       rr_ref, flag, status);
}
void bt_reliable_group(hipStream_t s, int gk, int k, const int32_t *gs, const int32_t *ge, PcgScalars sc, int *flag,
                       int *status) {
    hipLaunchKernelGGL(kb_reliable_group, dim3(bt_grid(gk, 256)), dim3(256), 0, s, gk, k, gs, ge, sc, flag, status);
}
void bt_reliable_fold(hipStream_t s, Tiles rt, int k, const int *flag, const double *alpha, const float *P,
                      float *X32, double *X64, float *XF) {
    hipLaunchKernelGGL(kb_reliable_fold, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, k, flag, alpha, P, X32, X64,
                       XF);
}
void bt_reliable_apply(hipStream_t s, int gk, int k, const int32_t *gtile, const double *p, int rule, double tol,
                       PcgScalars sc, double *rr_ref, int *flag, int *status) {
    hipLaunchKernelGGL(kb_reliable_apply, dim3(gk), dim3(64), 0, s, k, gtile, p, rule, tol, sc, rr_ref, flag, status);
}
void bt_reliable_residual(hipStream_t s, Tiles rt, const int32_t *rowptr, const int32_t *col, const double *val,
                          int k, const int *flag, const float *X, double *R, float *R32, double *partial) {
    if (k % 4 == 0 && k >= 256)
        hipLaunchKernelGGL(kb_reliable_residual4, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, rowptr, col, val, k,
                           flag, reinterpret_cast<const f32x4 *>(X), reinterpret_cast<f64x4 *>(R), R32, partial);
    else
        hipLaunchKernelGGL(kb_reliable_residual, dim3(rt.count), dim3(64 * rt.waves), 0, s, rt, rowptr, col, val, k,
                           flag, X, R, R32, partial);
}
// New diagonal of a sampled Laplacian, as Triplet::insert / process_diagonal
// / to_csc and the sparse add build it (src/sparsifier.rs:48-87, :645-691):
// each edge e contributes delta_e = -change_e if kept (change = w - w /
// sqrt(p)), -w if dropped, to both endpoints; vertex x sums its edges'
// deltas in edge order from zero and adds the sum to its old diagonal.
__global__ __launch_bounds__(256) void kb_apply_diag(int32_t n, const int32_t *__restrict__ rowptr,
                                                     const int32_t *__restrict__ col, const double *__restrict__ val,
                                                     const int32_t *__restrict__ iptr,
                                                     const int32_t *__restrict__ ieid, const double *__restrict__ ew,
                                                     const double *__restrict__ p, const uint8_t *__restrict__ mask,
                                                     double *__restrict__ diag_new) {
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    double s = 0.0;
    for (int q = iptr[x]; q < iptr[x + 1]; q++) {
        const int e = ieid[q];
        const double true_value = ew[e];
        double delta;
        if (mask[e]) {
            const double target = true_value / sqrt(p[e]);
            const double change = true_value - target;
            delta = change * -1.0;
        } else {
            delta = true_value * -1.0;
        }
        s = s + delta;
    }
    double old = 0.0;
    for (int q = rowptr[x]; q < rowptr[x + 1]; q++)
        if (col[q] == x) old = val[q];
    diag_new[x] = old + s;
}

void bt_apply_diag(hipStream_t s, int32_t n, const int32_t *rowptr, const int32_t *col, const double *val,
                   const int32_t *iptr, const int32_t *ieid, const double *ew, const double *p, const uint8_t *mask,
                   double *diag_new) {
    if (n > 0)
        hipLaunchKernelGGL(kb_apply_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, rowptr, col, val,
                           iptr, ieid, ew, p, mask, diag_new);
}

void bt_er_sample(hipStream_t s, const BatchSampleArgs &a, int ntiles) {
    hipLaunchKernelGGL(kb_er_sample, dim3(ntiles), dim3(256), 0, s, a);
}

}  // namespace ss
