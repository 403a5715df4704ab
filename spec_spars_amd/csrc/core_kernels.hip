// Dense cores of the preconditioner: per graph, the last c rows of its
// elimination order (a suffix, so the rows outside never depend on them in
// the forward solve and the core rows never depend on the rows outside in
// the backward solve) form the unit-lower block L_cc of the factor.  The
// block is inverted once on the device when the factor is uploaded,
// M = L_cc^{-1}, and each preconditioner application multiplies by it:
//   forward   y_c = M   (r_c - L_co y_o)      (the head launch subtracts L_co y_o)
//   backward  z_c = M^T (D_c^+ y_c)
// so a core costs one dense product per direction instead of one step per
// 64 rows.  L = I - N with N >= 0 strictly lower, so M = I + N + N^2 + ...
// has entries in [0, 1] and the inversion only ever adds terms of one sign.
//
// Layout per graph g: the core is padded to cp = a multiple of kCT rows
// (padding rows: identity in M, zero right-hand side); M is row-major
// cp x cp at dense_off[g]; the gathered right-hand sides Tc are rows
// [tc_off[g], tc_off[g] + cp) of a (sum cp) x k block.  Every sum runs in an
// order fixed by the graph alone, so a graph's bits do not depend on the
// batch it is in.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ss {

constexpr int kCT = kCoreTile;   // tile rows and columns

// L_cc, strictly lower part, FP64: dense[pos[e]] = val[e].
__global__ __launch_bounds__(256) void k_core_scatter(int64_t count, const int64_t *__restrict__ pos,
                                                      const double *__restrict__ val, double *__restrict__ dense) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * blockDim.x)
        dense[pos[e]] = val[e];
}

// Inverse of every diagonal tile b of one graph's L_cc (unit lower):
// thread j substitutes column j of the inverse, x_i = -sum_{j<=m<i} L[i][m] x_m.
__global__ __launch_bounds__(kCT) void k_core_inv_diag(int cp, const double *__restrict__ L, double *__restrict__ M) {
    __shared__ double ls[kCT][kCT + 1];
    __shared__ double xs[kCT][kCT + 1];
    const int b = blockIdx.x, j = threadIdx.x;
    const size_t o = (size_t)b * kCT;
    for (int i = 0; i < kCT; i++) ls[i][j] = j < i ? L[(o + i) * cp + o + j] : 0.0;
    __syncthreads();
    for (int i = 0; i < kCT; i++) {
        double x;
        if (i < j) {
            x = 0.0;
        } else if (i == j) {
            x = 1.0;
        } else {
            double s = 0.0;
            for (int m = j; m < i; m++) s = __builtin_fma(ls[i][m], xs[m][j], s);
            x = -s;
        }
        xs[i][j] = x;
    }
    for (int i = 0; i < kCT; i++) M[(o + i) * cp + o + j] = xs[i][j];
}

__global__ __launch_bounds__(256) void k_to_f32(int64_t count, const double *__restrict__ a, float *__restrict__ b) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * blockDim.x)
        b[e] = (float)a[e];
}

__global__ __launch_bounds__(256) void k_to_f16(int64_t count, const double *__restrict__ a,
                                                _Float16 *__restrict__ b) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * blockDim.x)
        b[e] = (_Float16)a[e];
}

// Tc[r] = forward: out[vert[r]] (the head's r_c - L_co y_o);
//         backward: in[vert[r]] / D (D == 0 keeps the value, as the level
//         solves do); zero on padding rows (vert < 0).
template <typename T, typename V, typename TC, bool BWD>
__global__ __launch_bounds__(256) void k_core_gather(int64_t rows, const int32_t *__restrict__ vert,
                                                     const double *__restrict__ diag, int k,
                                                     const V *__restrict__ src, TC *__restrict__ tc) {
    const int64_t total = rows * k;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / k;
        const int j = (int)(e - r * k);
        const int v = vert[r];
        T x = (T)0;
        if (v >= 0) {
            x = (T)src[(size_t)v * k + j];
            if (BWD) {
                const double d = diag[v];
                if (d != 0.0) x = (T)((double)x / d);
            }
        }
        tc[e] = (TC)x;
    }
}

// One 16 x 16 x 4 matrix-core step: acc += A (16 x 4) B (4 x 16), lane l
// holding A[l % 16][l / 16] and B[l / 16][l % 16]; acc[i] is the block's
// column l % 16, row acc_row<T>(l / 16, i): 4 (l / 16) + i in FP32,
// (l / 16) + 4 i in FP64 (the two instructions deal their outputs out
// differently; checked on the device).
typedef double f64x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
template <typename T> __device__ __forceinline__ int acc_row(int lk, int i);
template <> __device__ __forceinline__ int acc_row<float>(int lk, int i) { return 4 * lk + i; }
template <> __device__ __forceinline__ int acc_row<double>(int lk, int i) { return lk + 4 * i; }
template <typename T> struct Acc4;
template <> struct Acc4<float> { typedef f32x4 type; };
template <> struct Acc4<double> { typedef f64x4 type; };

// One piece of a core product: rows [r0, r0 + kCT) of graph g, inner range
// [m0, m1) -> partial tile part[piece] (kCT x k).  Forward A = M, backward
// A = M^T.  The workgroup holds the tile's 64 rows by all k columns (k
// rounded up to KC = 64 RB): its four waves each take RB of the four 16-row
// blocks by the four 16-column blocks of one 64-column chunk, on the matrix
// cores (16 x 16 x 4), so every A tile is read from memory once whatever k
// is.  The inner dimension is staged in LDS kAK rows at a time, double
// buffered: the next stage's loads are in flight while the current one is
// multiplied, with one barrier per stage.  Each output adds its terms in
// ascending inner order, four at a time, for every k and RB.
constexpr int kAK = 16;   // inner rows per stage of k_core_apply
template <typename T, bool BWD, int RB>
__global__ __launch_bounds__(256) void k_core_apply(const CorePiece *__restrict__ pieces,
                                                    const int32_t *__restrict__ cpad,
                                                    const int64_t *__restrict__ dense_off,
                                                    const int64_t *__restrict__ tc_off, const T *__restrict__ M,
                                                    int k, const T *__restrict__ tc, T *__restrict__ part) {
    static_assert(RB == 1 || RB == 2 || RB == 4, "row blocks per wave");
    typedef typename Acc4<T>::type AccT;
    constexpr int KC = 64 * RB;              // columns held
    constexpr int NA = kCT * kAK / 256;      // A elements per thread per stage
    constexpr int NB = kAK * KC / 256;       // B elements per thread per stage
    __shared__ T as[2][kAK][kCT + 4];        // as[.][m][r]: A(r, m)
    __shared__ T bs[2][kAK][KC + 4];         // bs[.][m][c]: Tc(m, c)
    const CorePiece pc = pieces[blockIdx.x];
    const int cp = cpad[pc.graph];
    const T *Mg = M + dense_off[pc.graph];
    const T *tg = tc + (size_t)tc_off[pc.graph] * k;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = (wv % (4 / RB)) * 16 * RB;   // this wave's first row
    const int wc = (wv / (4 / RB)) * 64;        // and first column
    const int li = lane & 15, lk = lane >> 4;
    T ra[NA], rb[NB];
    // stage loads: A element e = t + 256 u; forward A(r, m) = M[r0 + r][m]
    // (rows of kAK contiguous entries), backward A(r, m) = M[m][r0 + r]
    // (rows of kCT contiguous entries); B element e: Tc row m0 + e / KC,
    // column e % KC (zero past k)
    auto load = [&](int m0) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int e = t + 256 * u;
            ra[u] = BWD ? Mg[(size_t)(m0 + e / kCT) * cp + pc.r0 + e % kCT]
                        : Mg[(size_t)(pc.r0 + e / kAK) * cp + m0 + e % kAK];
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int e = t + 256 * u, c = e % KC;
            rb[u] = c < k ? tg[(size_t)(m0 + e / KC) * k + c] : (T)0;
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int e = t + 256 * u;
            if (BWD)
                as[b][e / kCT][e % kCT] = ra[u];
            else
                as[b][e % kAK][e / kAK] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int e = t + 256 * u;
            bs[b][e / KC][e % KC] = rb[u];
        }
    };
    AccT acc[RB][4];
#pragma unroll
    for (int x = 0; x < RB; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = AccT{0, 0, 0, 0};
    load(pc.m0);
    int b = 0;
    for (int m0 = pc.m0; m0 < pc.m1; m0 += kAK, b ^= 1) {
        // buffer b was last read two stages ago: every wave has passed the
        // previous stage's barrier since
        store(b);
        __syncthreads();
        if (m0 + kAK < pc.m1) load(m0 + kAK);
#pragma unroll
        for (int m4 = 0; m4 < kAK; m4 += 4) {
            T a[RB], bb[4];
#pragma unroll
            for (int x = 0; x < RB; x++) a[x] = as[b][m4 + lk][wr + 16 * x + li];
#pragma unroll
            for (int y = 0; y < 4; y++) bb[y] = bs[b][m4 + lk][wc + 16 * y + li];
#pragma unroll
            for (int x = 0; x < RB; x++)
#pragma unroll
                for (int y = 0; y < 4; y++) acc[x][y] = mfma16(a[x], bb[y], acc[x][y]);
        }
    }
    T *po = part + (size_t)blockIdx.x * kCT * k;
#pragma unroll
    for (int x = 0; x < RB; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int c = wc + 16 * y + li;
            if (c < k)
#pragma unroll
                for (int i = 0; i < 4; i++) po[(size_t)(wr + 16 * x + acc_row<T>(lk, i)) * k + c] = acc[x][y][i];
        }
}

// The dense-core products of the mixed path with FP16 preconditioner
// rows: M and the gathered rows in FP16, FP32 accumulation on the matrix
// cores (16 x 16 x 16 FP16), partials in FP32.  The pieces, tiles and
// double-buffered stages of k_core_apply, kAKH inner rows per stage, held
// inner-index-contiguous in LDS so that a lane's four inner values for one
// matrix-core step are one 8-byte read: lane l holds A(row l % 16, inner
// 4 (l / 16) .. + 3) and Tc(the same inner, column l % 16); the outputs land
// as in the FP32 kernel (row 4 (l / 16) + i, column l % 16).  Each output
// adds its terms in ascending inner order, sixteen at a time, for every k
// and RB.
constexpr int kAKH = 32;   // inner rows per stage of k_core_apply_h
constexpr int kAKHPad = 8; // LDS row padding (halves): rows stay 8-byte aligned
template <bool BWD, int RB>
__global__ __launch_bounds__(256) void k_core_apply_h(const CorePiece *__restrict__ pieces,
                                                      const int32_t *__restrict__ cpad,
                                                      const int64_t *__restrict__ dense_off,
                                                      const int64_t *__restrict__ tc_off,
                                                      const _Float16 *__restrict__ M, int k,
                                                      const _Float16 *__restrict__ tc, float *__restrict__ part) {
    static_assert(RB == 1 || RB == 2 || RB == 4, "row blocks per wave");
    static_assert((kAKH + kAKHPad) % 4 == 0 && kAKH % 16 == 0, "8-byte LDS reads, whole matrix-core steps");
    constexpr int KC = 64 * RB;                 // columns held
    constexpr int NA = kCT * kAKH / 256;        // A elements per thread per stage
    constexpr int NB = kAKH * KC / 256;         // B elements per thread per stage
    __shared__ __align__(16) _Float16 as[2][kCT][kAKH + kAKHPad];   // as[.][r][m]: A(r, m)
    __shared__ __align__(16) _Float16 bs[2][KC][kAKH + kAKHPad];    // bs[.][c][m]: Tc(m, c)
    const CorePiece pc = pieces[blockIdx.x];
    const int cp = cpad[pc.graph];
    const _Float16 *Mg = M + dense_off[pc.graph];
    const _Float16 *tg = tc + (size_t)tc_off[pc.graph] * k;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int wr = (wv % (4 / RB)) * 16 * RB;   // this wave's first row
    const int wc = (wv / (4 / RB)) * 64;        // and first column
    const int li = lane & 15, lk = lane >> 4;
    _Float16 ra[NA], rb[NB];
    // stage loads: forward A(r, m) = M[r0 + r][m0 + m] (element e: r = e /
    // kAKH, m = e % kAKH), backward A(r, m) = M[m0 + m][r0 + r] (m = e /
    // kCT, r = e % kCT); B element e: Tc row m0 + e / KC, column e % KC
    // (zero past k)
    auto load = [&](int m0) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int e = t + 256 * u;
            ra[u] = BWD ? Mg[(size_t)(m0 + e / kCT) * cp + pc.r0 + e % kCT]
                        : Mg[(size_t)(pc.r0 + e / kAKH) * cp + m0 + e % kAKH];
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int e = t + 256 * u, c = e % KC;
            rb[u] = c < k ? tg[(size_t)(m0 + e / KC) * k + c] : (_Float16)0;
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int e = t + 256 * u;
            if (BWD)
                as[b][e % kCT][e / kCT] = ra[u];
            else
                as[b][e / kAKH][e % kAKH] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int e = t + 256 * u;
            bs[b][e % KC][e / KC] = rb[u];
        }
    };
    f32x4 acc[RB][4];
#pragma unroll
    for (int x = 0; x < RB; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) acc[x][y] = f32x4{0, 0, 0, 0};
    load(pc.m0);
    int b = 0;
    for (int m0 = pc.m0; m0 < pc.m1; m0 += kAKH, b ^= 1) {
        // buffer b was last read two stages ago: every wave has passed the
        // previous stage's barrier since
        store(b);
        __syncthreads();
        if (m0 + kAKH < pc.m1) load(m0 + kAKH);
#pragma unroll
        for (int mm = 0; mm < kAKH; mm += 16) {
            f16x4 a[RB], bb[4];
#pragma unroll
            for (int x = 0; x < RB; x++) a[x] = *reinterpret_cast<const f16x4 *>(&as[b][wr + 16 * x + li][mm + 4 * lk]);
#pragma unroll
            for (int y = 0; y < 4; y++) bb[y] = *reinterpret_cast<const f16x4 *>(&bs[b][wc + 16 * y + li][mm + 4 * lk]);
#pragma unroll
            for (int x = 0; x < RB; x++)
#pragma unroll
                for (int y = 0; y < 4; y++)
                    acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x16f16(a[x], bb[y], acc[x][y], 0, 0, 0);
        }
    }
    float *po = part + (size_t)blockIdx.x * kCT * k;
#pragma unroll
    for (int x = 0; x < RB; x++)
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const int c = wc + 16 * y + li;
            if (c < k)
#pragma unroll
                for (int i = 0; i < 4; i++) po[(size_t)(wr + 16 * x + 4 * lk + i) * k + c] = acc[x][y][i];
        }
}

// One block row of the inverse: for every tile column jt < it,
//   M_{it,jt} = -Minv_{it,it} * sum_{jt <= mt < it} L_{it,mt} M_{mt,jt}
// (block forward substitution of L M = I, column tiles independent).  A
// workgroup per jt: its four waves each hold 16 rows x 64 columns of the
// 64 x 64 tile on the matrix cores (16 x 16 x 4 FP64).  The inner range is
// staged in LDS kAK rows at a time, double buffered; the sum tile is then
// multiplied by the diagonal tile's inverse (k_core_inv_diag) the same way.
__global__ __launch_bounds__(256) void k_core_inv_row(int it, int cp, const double *__restrict__ L,
                                                      double *__restrict__ M) {
    constexpr int NA = kCT * kAK / 256, NB = kAK * kCT / 256;
    __shared__ double as[2][kAK][kCT + 4];   // as[.][m][r] = L(r0 + r, m0 + m)
    __shared__ double bs[2][kAK][kCT + 4];   // bs[.][m][c] = M(m0 + m, c0 + c)
    __shared__ double ts[kCT][kCT + 4];      // the sum tile
    __shared__ double ms[kCT][kCT + 4];      // ms[m][r] = Minv_{it,it}(r, m)
    const int jt = blockIdx.x, t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = lane & 15, lk = lane >> 4, wr = 16 * wv;
    const size_t r0 = (size_t)it * kCT, c0 = (size_t)jt * kCT;
    f64x4 acc[4];
#pragma unroll
    for (int y = 0; y < 4; y++) acc[y] = f64x4{0, 0, 0, 0};
    double ra[NA], rb[NB];
    auto load = [&](size_t m0) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int e = t + 256 * u;
            ra[u] = L[(r0 + e / kAK) * cp + m0 + e % kAK];
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int e = t + 256 * u;
            rb[u] = M[(m0 + e / kCT) * cp + c0 + e % kCT];
        }
    };
    auto store = [&](int b) {
#pragma unroll
        for (int u = 0; u < NA; u++) {
            const int e = t + 256 * u;
            as[b][e % kAK][e / kAK] = ra[u];
        }
#pragma unroll
        for (int u = 0; u < NB; u++) {
            const int e = t + 256 * u;
            bs[b][e / kCT][e % kCT] = rb[u];
        }
    };
    // the diagonal tile's inverse, transposed into LDS
    for (int e = t; e < kCT * kCT; e += 256) ms[e % kCT][e / kCT] = M[(r0 + e / kCT) * cp + r0 + e % kCT];
    load(c0);
    int b = 0;
    for (size_t m0 = c0; m0 < r0; m0 += kAK, b ^= 1) {
        store(b);
        __syncthreads();
        if (m0 + kAK < r0) load(m0 + kAK);
#pragma unroll
        for (int m4 = 0; m4 < kAK; m4 += 4) {
            const double a = as[b][m4 + lk][wr + li];
#pragma unroll
            for (int y = 0; y < 4; y++) acc[y] = mfma16(a, bs[b][m4 + lk][16 * y + li], acc[y]);
        }
    }
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
        for (int i = 0; i < 4; i++) ts[wr + acc_row<double>(lk, i)][16 * y + li] = acc[y][i];
    __syncthreads();
    f64x4 out[4];
#pragma unroll
    for (int y = 0; y < 4; y++) out[y] = f64x4{0, 0, 0, 0};
#pragma unroll 4
    for (int m4 = 0; m4 < kCT; m4 += 4) {
        const double a = ms[m4 + lk][wr + li];
#pragma unroll
        for (int y = 0; y < 4; y++) out[y] = mfma16(a, ts[m4 + lk][16 * y + li], out[y]);
    }
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
        for (int i = 0; i < 4; i++) M[(r0 + wr + acc_row<double>(lk, i)) * cp + c0 + 16 * y + li] = -out[y][i];
}

// out[vert[r]] = sum of the row's pieces, in piece order.
template <typename T, typename V>
__global__ __launch_bounds__(256) void k_core_reduce(int32_t tiles, const int32_t *__restrict__ tile_piece,
                                                     const int32_t *__restrict__ tile_row,
                                                     const int32_t *__restrict__ vert, int k,
                                                     const T *__restrict__ part, V *__restrict__ out) {
    const int64_t total = (int64_t)tiles * kCT * k;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t tr = e / ((int64_t)kCT * k);
        const int rem = (int)(e - tr * kCT * k);   // row-in-tile * k + column
        const int v = vert[tile_row[tr] + rem / k];
        if (v < 0) continue;
        const int p0 = tile_piece[tr], p1 = tile_piece[tr + 1];
        T s = part[(size_t)p0 * kCT * k + rem];
        for (int p = p0 + 1; p < p1; p++) s = s + part[(size_t)p * kCT * k + rem];
        out[(size_t)v * k + rem % k] = (V)s;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned stream_grid(int64_t items) {
    int64_t g = (items + 255) / 256;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return (unsigned)g;
}

void launch_core_scatter(hipStream_t s, int64_t count, const int64_t *pos, const double *val, double *dense) {
    if (count > 0)
        hipLaunchKernelGGL(k_core_scatter, dim3(stream_grid(count)), dim3(256), 0, s, count, pos, val, dense);
}

void launch_core_invert(hipStream_t s, int cp, const double *L, double *M) {
    const int nb = cp / kCT;
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_core_inv_diag, dim3(nb), dim3(kCT), 0, s, cp, L, M);
    for (int it = 1; it < nb; it++)
        hipLaunchKernelGGL(k_core_inv_row, dim3(it), dim3(256), 0, s, it, cp, L, M);
}

void launch_to_f32(hipStream_t s, int64_t count, const double *a, float *b) {
    if (count > 0) hipLaunchKernelGGL(k_to_f32, dim3(stream_grid(count)), dim3(256), 0, s, count, a, b);
}

void launch_to_f16(hipStream_t s, int64_t count, const double *a, _Float16 *b) {
    if (count > 0) hipLaunchKernelGGL(k_to_f16, dim3(stream_grid(count)), dim3(256), 0, s, count, a, b);
}

template <typename T, typename V, bool BWD>
void launch_core_apply(hipStream_t s, const CoreApply &c, const T *M, int k, const double *diag, const V *src,
                       T *tc, T *part, V *out) {
    if (c.pieces <= 0) return;
    hipLaunchKernelGGL((k_core_gather<T, V, T, BWD>), dim3(stream_grid(c.rows * k)), dim3(256), 0, s, c.rows, c.vert,
                       diag, k, src, tc);
    const CorePiece *pc = BWD ? c.piece_bwd : c.piece_fwd;
    if (k <= 64)
        hipLaunchKernelGGL((k_core_apply<T, BWD, 1>), dim3((unsigned)c.pieces), dim3(256), 0, s, pc, c.cpad,
                           c.dense_off, c.tc_off, M, k, tc, part);
    else if (k <= 128)
        hipLaunchKernelGGL((k_core_apply<T, BWD, 2>), dim3((unsigned)c.pieces), dim3(256), 0, s, pc, c.cpad,
                           c.dense_off, c.tc_off, M, k, tc, part);
    else
        hipLaunchKernelGGL((k_core_apply<T, BWD, 4>), dim3((unsigned)c.pieces), dim3(256), 0, s, pc, c.cpad,
                           c.dense_off, c.tc_off, M, k, tc, part);
    hipLaunchKernelGGL((k_core_reduce<T, V>), dim3(stream_grid((int64_t)c.tiles * kCT * k)), dim3(256), 0, s,
                       c.tiles, BWD ? c.tile_piece_bwd : c.tile_piece_fwd, c.tile_row, c.vert, k, part, out);
}
#define SS_INST_CORE_APPLY(T, V, B)                                                                          \
    template void launch_core_apply<T, V, B>(hipStream_t, const CoreApply &, const T *, int, const double *,  \
                                             const V *, T *, T *, V *);
SS_INST_CORE_APPLY(double, double, false)
SS_INST_CORE_APPLY(double, double, true)
SS_INST_CORE_APPLY(float, float, false)
SS_INST_CORE_APPLY(float, float, true)
SS_INST_CORE_APPLY(float, _Float16, false)
SS_INST_CORE_APPLY(float, _Float16, true)

template <bool BWD>
void launch_core_apply_h(hipStream_t s, const CoreApply &c, const _Float16 *M, int k, const double *diag,
                         const _Float16 *src, _Float16 *tc, float *part, _Float16 *out) {
    if (c.pieces <= 0) return;
    hipLaunchKernelGGL((k_core_gather<float, _Float16, _Float16, BWD>), dim3(stream_grid(c.rows * k)), dim3(256), 0,
                       s, c.rows, c.vert, diag, k, src, tc);
    const CorePiece *pc = BWD ? c.piece_bwd : c.piece_fwd;
    if (k <= 64)
        hipLaunchKernelGGL((k_core_apply_h<BWD, 1>), dim3((unsigned)c.pieces), dim3(256), 0, s, pc, c.cpad,
                           c.dense_off, c.tc_off, M, k, tc, part);
    else if (k <= 128)
        hipLaunchKernelGGL((k_core_apply_h<BWD, 2>), dim3((unsigned)c.pieces), dim3(256), 0, s, pc, c.cpad,
                           c.dense_off, c.tc_off, M, k, tc, part);
    else
        hipLaunchKernelGGL((k_core_apply_h<BWD, 4>), dim3((unsigned)c.pieces), dim3(256), 0, s, pc, c.cpad,
                           c.dense_off, c.tc_off, M, k, tc, part);
    hipLaunchKernelGGL((k_core_reduce<float, _Float16>), dim3(stream_grid((int64_t)c.tiles * kCT * k)), dim3(256), 0,
                       s, c.tiles, BWD ? c.tile_piece_bwd : c.tile_piece_fwd, c.tile_row, c.vert, k, part, out);
}
template void launch_core_apply_h<false>(hipStream_t, const CoreApply &, const _Float16 *, int, const double *,
                                         const _Float16 *, _Float16 *, float *, _Float16 *);
template void launch_core_apply_h<true>(hipStream_t, const CoreApply &, const _Float16 *, int, const double *,
                                        const _Float16 *, _Float16 *, float *, _Float16 *);

}  // namespace ss
