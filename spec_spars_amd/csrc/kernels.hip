// gfx950 kernels of the sparsification hot path: the single-graph sketch
// and sampling kernels behind ss_sketch / ss_er_sample, and the
// preconditioner's triangular solves (level launches, small-level runs,
// dense cores) shared by every resident graph and batch.
//
// Dense n x k blocks are row-major (node i's k values contiguous); a row's
// k values sit on the 64 lanes of a wavefront (64-column chunks for larger
// k), so a node's k-vector moves as one coalesced 64*s-byte access.
#include <hip/hip_runtime.h>

#include "hash_dev.h"
#include "kernels.h"

namespace ss {

#define SS_LANE (threadIdx.x & 63)
#define SS_WAVE_IN_BLOCK (threadIdx.x >> 6)

__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// acc - v * x with one rounding.  Every accumulate of the solves is an
// explicit fused multiply-add: the compiler only contracts a multiply and a
// subtract that land in one basic block, so leaving it to contraction would
// round some terms twice depending on how a loop was laid out.
__device__ __forceinline__ double fnms(double v, double x, double acc) { return __builtin_fma(-v, x, acc); }
__device__ __forceinline__ float fnms(float v, float x, float acc) { return __builtin_fmaf(-v, x, acc); }

// ---------------------------------------------------------------------------
// (a) sketch, gather form: Y[x, j] = sum over incident edges e of x, in
// increasing e, of  fl(scale * sign(e, j)) * b(x, e),  b = +-sqrt(w).
// Matches the sequential edge loop of jl_sketch_colwise
// (src/sparsifier.rs:490-504) term for term; fp-contract is off so each
// product rounds before the add.
// ---------------------------------------------------------------------------
template <typename OutT>
__global__ __launch_bounds__(256) void k_sketch(int32_t n, const int32_t *__restrict__ iptr,
                                                const int32_t *__restrict__ ieid,
                                                const double *__restrict__ ib, int k,
                                                uint64_t seed, double scale, int uniform,
                                                int variant, OutT *__restrict__ Y) {
#pragma clang fp contract(off)
    const int row = wave_uniform(blockIdx.x * 4 + SS_WAVE_IN_BLOCK);
    if (row >= n) return;
    const int q0 = iptr[row], q1 = iptr[row + 1];
    for (int jb = 0; jb < k; jb += 64) {
        const int j = jb + SS_LANE;
        double acc = 0.0;
        for (int q = q0; q < q1; q++) {
            const int e = ieid[q];
            const double b = ib[q];
            const double sv = sign_value((uint64_t)e, (uint64_t)j, seed, uniform, variant) * scale;
            const double term = sv * b;
            acc = acc + term;
        }
        if (j < k) Y[(size_t)row * k + j] = (OutT)acc;
    }
}

__global__ __launch_bounds__(256) void k_copy_f64(int64_t count, const double *__restrict__ a,
                                                  double *__restrict__ b) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (int64_t)gridDim.x * blockDim.x)
        b[t] = a[t];
}

__global__ void k_fill_u32(int64_t count, uint32_t *p, uint32_t v) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (int64_t)gridDim.x * blockDim.x)
        p[t] = v;
}

// ---------------------------------------------------------------------------
// (c) level-scheduled triangular solves: one launch per level.
// Schedule positions [t0, t1) hold the rows of one level (sorted by level,
// so every row they read was finished by an earlier launch on the stream);
// lanes over the k columns, a row's arithmetic fixed by solve_shape.h.
//   forward  (BWD=false): y_i = r_i - sum_{j<i} L[i,j] y_j
//   backward (BWD=true):  z_i = y_i / D_i - sum_{j>i} L[j,i] z_j
//                         (division skipped where D_i == 0, as
//                          include/custom_cg.hpp:227-233 does)
//
// A row of the solve is stored as ke = k / n elements of type V, each
// holding n = Lane<V>::n adjacent columns: V = T (n = 1), or for FP32
// arithmetic two columns per element, as FP32 (f32x2) or FP16 (f16x2), or
// four as FP16 (f16x4), so one lane moves two or four columns with one 4-
// or 8-byte access.  A lane's value
// is Lane<V>::type; the arithmetic of every column is the same whatever n
// is.  The solve's input (forward: the PCG residual; backward: the forward
// result) is read the same way, as elements of type VI.
// ---------------------------------------------------------------------------
template <typename V> struct Lane;
template <> struct Lane<double> { typedef double type; static constexpr int n = 1; };
template <> struct Lane<float> { typedef float type; static constexpr int n = 1; };
template <> struct Lane<f32x2> { typedef f32x2 type; static constexpr int n = 2; };
template <> struct Lane<f16x2> { typedef f32x2 type; static constexpr int n = 2; };
template <> struct Lane<f32x4> { typedef f32x4 type; static constexpr int n = 4; };
template <> struct Lane<f16x4> { typedef f32x4 type; static constexpr int n = 4; };

// An element as stored -> the lane value.  The gather loops hold the
// stored elements and widen them where they are used, so all of a group's
// loads are issued before the first wait.  Stored elements that need
// widening (FP16) are never loaded under a branch: a slot past the row's
// entries loads an entry of the group that exists and its term is
// selected away, so the compiler cannot tie a widening, and a wait, to
// each load.
__device__ __forceinline__ double cvt(double x) { return x; }
__device__ __forceinline__ float cvt(float x) { return x; }
__device__ __forceinline__ f32x2 cvt(f32x2 x) { return x; }
__device__ __forceinline__ f32x2 cvt(f16x2 x) { return __builtin_convertvector(x, f32x2); }
__device__ __forceinline__ f32x4 cvt(f32x4 x) { return x; }
__device__ __forceinline__ f32x4 cvt(f16x4 x) { return __builtin_convertvector(x, f32x4); }
template <typename V> __device__ __forceinline__ typename Lane<V>::type ldv(const V *p) { return cvt(*p); }
template <typename V> constexpr bool widened() { return sizeof(typename Lane<V>::type) > sizeof(V); }
__device__ __forceinline__ void stv(double *p, double v) { *p = v; }
__device__ __forceinline__ void stv(float *p, float v) { *p = v; }
__device__ __forceinline__ void stv(f32x2 *p, f32x2 v) { *p = v; }
__device__ __forceinline__ void stv(f16x2 *p, f32x2 v) { *p = __builtin_convertvector(v, f16x2); }
__device__ __forceinline__ void stv(f32x4 *p, f32x4 v) { *p = v; }
__device__ __forceinline__ void stv(f16x4 *p, f32x4 v) { *p = __builtin_convertvector(v, f16x4); }

// Two-column lanes: each column its own fused multiply-add / rounding.
__device__ __forceinline__ f32x2 fnms(float v, f32x2 x, f32x2 acc) {
    f32x2 r;
    r.x = __builtin_fmaf(-v, x.x, acc.x);
    r.y = __builtin_fmaf(-v, x.y, acc.y);
    return r;
}
__device__ __forceinline__ f32x4 fnms(float v, f32x4 x, f32x4 acc) {
    f32x4 r;
    r.x = __builtin_fmaf(-v, x.x, acc.x);
    r.y = __builtin_fmaf(-v, x.y, acc.y);
    r.z = __builtin_fmaf(-v, x.z, acc.z);
    r.w = __builtin_fmaf(-v, x.w, acc.w);
    return r;
}
__device__ __forceinline__ double div_d(double x, double d) { return x / d; }
__device__ __forceinline__ float div_d(float x, double d) { return (float)((double)x / d); }
__device__ __forceinline__ f32x2 div_d(f32x2 x, double d) {
    f32x2 r;
    r.x = (float)((double)x.x / d);
    r.y = (float)((double)x.y / d);
    return r;
}
__device__ __forceinline__ f32x4 div_d(f32x4 x, double d) {
    f32x4 r;
    r.x = (float)((double)x.x / d);
    r.y = (float)((double)x.y / d);
    r.z = (float)((double)x.z / d);
    r.w = (float)((double)x.w / d);
    return r;
}

// Column scale of element e of a row of graph g (FP16 rows): n floats at
// s[g * k + n e].
template <typename LV> __device__ __forceinline__ LV scale_at(const float *s, size_t base, int e);
template <> __device__ __forceinline__ float scale_at<float>(const float *s, size_t base, int e) {
    return s[base + e];
}
template <> __device__ __forceinline__ f32x2 scale_at<f32x2>(const float *s, size_t base, int e) {
    return reinterpret_cast<const f32x2 *>(s + base)[e];
}
template <> __device__ __forceinline__ f32x4 scale_at<f32x4>(const float *s, size_t base, int e) {
    return reinterpret_cast<const f32x4 *>(s + base)[e];
}
template <> __device__ __forceinline__ double scale_at<double>(const float *s, size_t base, int e) {
    return (double)s[base + e];
}

// acc -= sum_q val[q] * X[col[q]*ke + e] over q in [q0, q1) with stride
// `step`, 8 gathers in flight, terms subtracted in q order.
template <typename T, typename V>
__device__ __forceinline__ typename Lane<V>::type row_dot_sub(typename Lane<V>::type acc, int q0, int q1, int step,
                                                              const int32_t *__restrict__ col,
                                                              const T *__restrict__ val, const V *X, int ke, int e) {
    int q = q0;
    for (; q + 7 * step < q1; q += 8 * step) {
        V x[8];
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            x[u] = X[(size_t)col[q + u * step] * ke + e];
            v[u] = val[q + u * step];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc = fnms(v[u], cvt(x[u]), acc);
    }
    for (; q < q1; q += step) acc = fnms(val[q], ldv(X + (size_t)col[q] * ke + e), acc);
    return acc;
}

// Start value of element e of row i: forward r_i; backward y_i / D_i (D_i
// == 0 keeps y_i).  With rs.s (FP16 rows), the forward start is r_i times
// its columns' scales rs.s[g * k + column] (g = the row's graph), so every
// row of the solve holds the scaled vector; the backward reads rows already
// scaled.
template <typename LV, bool BWD, typename VI>
__device__ __forceinline__ LV row_start(int i, int e, int k, int ke, double d, const VI *in, RowScale rs) {
    LV acc = ldv(in + (size_t)i * ke + e);
    if (BWD && d != 0.0) acc = div_d(acc, d);
    if (!BWD && rs.s) acc = acc * scale_at<LV>(rs.s, (size_t)(rs.rg ? rs.rg[i] : 0) * k, e);
    return acc;
}

// Pairwise tree fold of v[0..N) in place; the result lands in v[0].
template <typename T, int N>
__device__ __forceinline__ void tree_fold(T (&v)[N]) {
#pragma unroll
    for (int step = 1; step < N; step *= 2)
#pragma unroll
        for (int u = 0; u + step < N; u += 2 * step) v[u] = v[u] + v[u + step];
}

// Tree fold of sm[first .. first + N) (element lane).
template <typename T, int N>
__device__ __forceinline__ T lds_tree_n(T (*sm)[64], int first, int lane) {
    T v[N];
#pragma unroll
    for (int u = 0; u < N; u++) v[u] = sm[first + u][lane];
    tree_fold<T, N>(v);
    return v[0];
}
template <typename T>
__device__ __forceinline__ T lds_tree(T (*sm)[64], int first, int n, int lane) {
    switch (n) {
        case 1: return sm[first][lane];
        case 2: return lds_tree_n<T, 2>(sm, first, lane);
        case 4: return lds_tree_n<T, 4>(sm, first, lane);
        case 8: return lds_tree_n<T, 8>(sm, first, lane);
        default: return lds_tree_n<T, 16>(sm, first, lane);
    }
}

// Lane l's value of v (l uniform; the lane need not be active).
__device__ __forceinline__ int lane_get(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float lane_get(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_get(double v, int l) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Parts of a row whose entries sit one per lane (entry q in lane q & 63 of
// c0/v0 below 64, of c1/v1 from 64): gathered 16 at a time, entry q into
// part q % P.  Lanes past the row's end hold zero values, and their terms
// (zero times zero) leave a part exact.
template <typename T, typename V, int P>
__device__ __forceinline__ typename Lane<V>::type wave_parts(int len, int c0, int c1, T v0, T v1, const V *X, int ke,
                                                             int ee) {
    typedef typename Lane<V>::type LV;
    constexpr int GW = 16;
    LV part[P];
#pragma unroll
    for (int u = 0; u < P; u++) part[u] = LV();
    for (int e0 = 0; e0 < len; e0 += GW) {
        const bool hi = e0 >= 64;
        const int cs = hi ? c1 : c0;
        const T vs = hi ? v1 : v0;
        V x[GW];
        const int cf = lane_get(cs, e0 & 63);   // entry e0 exists
#pragma unroll
        for (int u = 0; u < GW; u++) {
            const int c = lane_get(cs, (e0 + u) & 63);
            if constexpr (widened<V>())
                x[u] = X[(size_t)(e0 + u < len ? c : cf) * ke + ee];
            else
                x[u] = e0 + u < len ? X[(size_t)c * ke + ee] : V();
        }
#pragma unroll
        for (int u = 0; u < GW; u++)
            part[u % P] = fnms(lane_get(vs, (e0 + u) & 63), e0 + u < len ? cvt(x[u]) : LV(), part[u % P]);
    }
    tree_fold<LV, P>(part);
    return part[0];
}

// Tree fold of parts [s0, s0 + NA) of segment [a, b) split in P parts, on
// one wave: slot l of a 64-slot round holds entry a + s0 + l % NA + P (l / NA)
// (so part s0 + u takes its entries in order), the slots' entries load one
// per lane, then gather GW at a time.  Slots past the segment's end hold
// zero values, whose terms leave a part exact.
template <typename T, typename V, int NA, int GW = 16>
__device__ __forceinline__ typename Lane<V>::type wave_subtree(int a, int b, int P, int s0,
                                                               const int32_t *__restrict__ col,
                                                               const T *__restrict__ val, const V *X, int ke, int ee,
                                                               int lane) {
    typedef typename Lane<V>::type LV;
    static_assert(NA == 1 || NA == 2 || NA == 4 || NA == 8 || NA == 16, "parts per wave");
    static_assert((GW == 8 || GW == 16) && NA <= GW, "gathers in flight: a whole number of slots per part");
    LV part[NA];
#pragma unroll
    for (int u = 0; u < NA; u++) part[u] = LV();
    const int a0 = a + s0;
    for (int base = 0; a0 + P * (base / NA) < b; base += 64) {
        const int l = base + lane;
        const int q = a0 + l % NA + P * (l / NA);
        const int c = q < b ? col[q] : 0;
        const T v = q < b ? val[q] : (T)0;
        for (int e0 = 0; e0 < 64 && a0 + P * ((base + e0) / NA) < b; e0 += GW) {
            V x[GW];
            const int cf = lane_get(c, e0);   // slot e0's entry exists (the loop's test)
#pragma unroll
            for (int u = 0; u < GW; u++) {
                const int cu = lane_get(c, e0 + u);
                const int lu = base + e0 + u;
                const bool on = a0 + lu % NA + P * (lu / NA) < b;
                if constexpr (widened<V>())
                    x[u] = X[(size_t)(on ? cu : cf) * ke + ee];
                else
                    x[u] = on ? X[(size_t)cu * ke + ee] : V();
            }
#pragma unroll
            for (int u = 0; u < GW; u++) {
                const int lu = base + e0 + u;
                const bool on = a0 + lu % NA + P * (lu / NA) < b;
                part[u % NA] = fnms(lane_get(v, e0 + u), on ? cvt(x[u]) : LV(), part[u % NA]);
            }
        }
    }
    tree_fold<LV, NA>(part);
    return part[0];
}

// na in {1, 2, 4, 8, 16}, na <= MAXNA (the bound keeps wider variants out
// of a kernel's registers).
template <typename T, typename V, int MAXNA>
__device__ __forceinline__ typename Lane<V>::type wave_subtree_n(int na, int a, int b, int P, int s0,
                                                                 const int32_t *__restrict__ col,
                                                                 const T *__restrict__ val, const V *X, int ke,
                                                                 int ee, int lane) {
    constexpr int M = MAXNA;
    if (na == 1 || M == 1) return wave_subtree<T, V, 1>(a, b, P, s0, col, val, X, ke, ee, lane);
    if (na == 2 || M == 2) return wave_subtree<T, V, (M < 2 ? M : 2)>(a, b, P, s0, col, val, X, ke, ee, lane);
    if (na == 4 || M == 4) return wave_subtree<T, V, (M < 4 ? M : 4)>(a, b, P, s0, col, val, X, ke, ee, lane);
    if (na == 8 || M == 8) return wave_subtree<T, V, (M < 8 ? M : 8)>(a, b, P, s0, col, val, X, ke, ee, lane);
    return wave_subtree<T, V, M>(a, b, P, s0, col, val, X, ke, ee, lane);
}

// One row of at most kWaveRowMax entries on one wave: the row's entries are
// loaded one per lane, so a single dependent load separates the row record
// from all of its gathers.
template <typename T, typename VI, typename V, bool BWD>
__device__ __forceinline__ void wave_row(int t, const int32_t *__restrict__ rows, const int32_t *__restrict__ ptr,
                                         const int32_t *__restrict__ col, const T *__restrict__ val,
                                         const double *__restrict__ diag, int k, int ke, const VI *__restrict__ in,
                                         V *out, RowScale rs, int lane) {
    typedef typename Lane<V>::type LV;
    const int i = wave_uniform(rows[t]);
    const int q0 = wave_uniform(ptr[t]), len = wave_uniform(ptr[t + 1]) - q0;
    const int c0 = lane < len ? col[q0 + lane] : 0;
    const T v0 = lane < len ? val[q0 + lane] : (T)0;
    const int c1 = 64 + lane < len ? col[q0 + 64 + lane] : 0;
    const T v1 = 64 + lane < len ? val[q0 + 64 + lane] : (T)0;
    const double d = BWD ? diag[i] : 1.0;
    const int P = seg_parts(len);
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + lane, ee = e < ke ? e : ke - 1;
        LV acc = row_start<LV, BWD>(i, ee, k, ke, d, in, rs);
        if (P == 1) {
            for (int q = 0; q < len; q++)
                acc = fnms(lane_get(v0, q), ldv(out + (size_t)lane_get(c0, q) * ke + ee), acc);
        } else if (P == 4) {
            acc = acc + wave_parts<T, V, 4>(len, c0, c1, v0, v1, out, ke, ee);
        } else {
            acc = acc + wave_parts<T, V, 8>(len, c0, c1, v0, v1, out, ke, ee);
        }
        if (e < ke) stv(out + (size_t)i * ke + e, acc);
    }
}

// Up to R <= kTinyGroup consecutive rows of at most kTinyRow entries each on
// one wave (positions [tb, tb + cnt)): their row records load as scalars (the
// schedule arrays carry kTinyGroup entries of padding), their entries
// (contiguous, at most 32) one per lane, every gather of the group is in
// flight before the first row finishes, and the stores come last.
template <typename T, typename VI, typename V, bool BWD, int R>
__device__ __forceinline__ void wave_tiny_rows(int tb, int cnt, const int32_t *__restrict__ rows,
                                               const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                                               const T *__restrict__ val, const double *__restrict__ diag, int k,
                                               int ke, const VI *__restrict__ in, V *out, RowScale rs, int lane) {
    typedef typename Lane<V>::type LV;
    static_assert(R >= 1 && R <= kTinyGroup, "at most kTinyGroup tiny rows at a time");
    constexpr int E = kTinyRow;
    int ir[R], pr[R + 1];
#pragma unroll
    for (int r = 0; r < R; r++) ir[r] = rows[tb + r];
#pragma unroll
    for (int r = 0; r <= R; r++) pr[r] = ptr[tb + r];
    const int base = pr[0];
    const int ne = ptr[tb + cnt] - base;   // <= R * E
    int ar[R], nr[R];
    double dr[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        ar[r] = pr[r] - base;
        nr[r] = r < cnt ? pr[r + 1] - pr[r] : 0;
        dr[r] = BWD ? diag[ir[r]] : 1.0;
    }
    if (ne == 0) {
        // no entries in the group: every row is its start value
        for (int eb = 0; eb < ke; eb += 64) {
            const int e = eb + lane;
            if (e >= ke) break;
            LV st[R];
#pragma unroll
            for (int r = 0; r < R; r++) st[r] = r < cnt ? row_start<LV, BWD>(ir[r], e, k, ke, dr[r], in, rs) : LV();
#pragma unroll
            for (int r = 0; r < R; r++)
                if (r < cnt) stv(out + (size_t)ir[r] * ke + e, st[r]);
        }
        return;
    }
    const int lc = lane < ne ? col[base + lane] : 0;
    const T lv = lane < ne ? val[base + lane] : (T)0;
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + lane, ee = e < ke ? e : ke - 1;
        LV st[R];
        V x[R][E];
        const int cf = lane_get(lc, 0);   // the group's first entry (ne > 0)
#pragma unroll
        for (int r = 0; r < R; r++) {
            st[r] = r < cnt ? row_start<LV, BWD>(ir[r], ee, k, ke, dr[r], in, rs) : LV();
#pragma unroll
            for (int q = 0; q < E; q++) {
                const int c = lane_get(lc, (ar[r] + q) & 63);
                if constexpr (widened<V>())
                    x[r][q] = out[(size_t)(q < nr[r] ? c : cf) * ke + ee];
                else
                    x[r][q] = q < nr[r] ? out[(size_t)c * ke + ee] : V();
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            LV acc = st[r];
#pragma unroll
            for (int q = 0; q < E; q++) {
                const T v = q < nr[r] ? lane_get(lv, (ar[r] + q) & 63) : (T)0;
                acc = fnms(v, q < nr[r] ? cvt(x[r][q]) : LV(), acc);
            }
            if (r < cnt && e < ke) stv(out + (size_t)ir[r] * ke + e, acc);
        }
    }
}

// A level whose rows have no entries (the first forward level: the rows
// that depend on nothing): every row is its start value.  Thread i of a
// block copies element i % ke of rows i / ke, i / ke + 256 / ke, ..., four
// rows in flight, so a block pass moves whole rows as one contiguous run.
template <typename T, typename VI, typename V, bool BWD>
__global__ __launch_bounds__(256) void k_level_copy(int32_t t0, int32_t nrows, const int32_t *__restrict__ rows,
                                                    const double *__restrict__ diag, int k,
                                                    const VI *__restrict__ in, V *__restrict__ out, RowScale rs) {
    typedef typename Lane<V>::type LV;
    const int ke = k / Lane<V>::n;
    // thread i copies element i % ke of every (256 / ke)-th row, 4 rows in flight
    const int RP = (int)blockDim.x / ke, i = (int)threadIdx.x;
    if (i >= RP * ke) return;
    const int e = i % ke, ro = i / ke;
    const int64_t step = (int64_t)gridDim.x * RP * 4;
    for (int64_t tb = (int64_t)blockIdx.x * RP * 4 + ro; tb < nrows; tb += step) {
        int r[4];
        LV v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) r[u] = tb + u * RP < nrows ? rows[t0 + tb + u * RP] : 0;
#pragma unroll
        for (int u = 0; u < 4; u++)
            v[u] = tb + u * RP < nrows ? row_start<LV, BWD>(r[u], e, k, ke, BWD ? diag[r[u]] : 1.0, in, rs) : LV();
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (tb + u * RP < nrows) stv(out + (size_t)r[u] * ke + e, v[u]);
    }
}

// Rows of more than kWgRowMax entries (positions [t0, t0 + count)), split
// in P = 64 or 128 parts folded as one adjacent-pair tree (tree_fold's), so
// any split of the parts into aligned power-of-two groups gives the same
// bits.  hm == 1: one workgroup per row, wave w folding parts [w P/16,
// (w+1) P/16).  hm > 1: H = P / 16 workgroups per row (hm launched per
// row), workgroup h folding parts [16h, 16h + 16) with one wave per part and
// storing that subtree of every element in its scratch slot (lane values,
// full precision); the last of the row's workgroups to arrive (one
// device-scope counter per row, reset by that workgroup) folds the H
// subtrees, adds them to the start value and stores the row.
template <typename T, typename VI, typename V, bool BWD>
__global__ __launch_bounds__(64 * kLevelWaves) void k_trisolve_huge(int32_t t0, int32_t hm,
                                                                    const int32_t *__restrict__ rows,
                                                                    const int32_t *__restrict__ ptr,
                                                                    const int32_t *__restrict__ col,
                                                                    const T *__restrict__ val,
                                                                    const double *__restrict__ diag, int k,
                                                                    const VI *__restrict__ in, V *out,
                                                                    typename Lane<V>::type *scratch, int *counter,
                                                                    RowScale rs) {
    typedef typename Lane<V>::type LV;
    const int ke = k / Lane<V>::n;
    __shared__ LV sm[kWgParts][64];
    const int lane = SS_LANE, wv = SS_WAVE_IN_BLOCK;
    const int slot = (int)blockIdx.x / hm, h = (int)blockIdx.x - slot * hm;
    const int t = t0 + slot;
    const int i = wave_uniform(rows[t]);
    const int q0 = wave_uniform(ptr[t]), q1 = wave_uniform(ptr[t + 1]);
    const int P = seg_parts(q1 - q0);
    const int H = row_wgs(q1 - q0);
    const double d = BWD ? diag[i] : 1.0;
    if (hm == 1) {
        const int na = P / kWgParts;
        for (int eb = 0; eb < ke; eb += 64) {
            const int e = eb + lane, ee = e < ke ? e : ke - 1;
            if constexpr (sizeof(LV) == 8) {
                // 8-byte lane values: 8 parts as two subtrees of 4 (the
                // same tree), 8 gathers in flight, which keeps the kernel
                // within the 128 registers a 1,024-thread workgroup leaves
                // each lane
                if (na == 4)
                    sm[wv][lane] = wave_subtree<T, V, 4, 8>(q0, q1, P, wv * 4, col, val, out, ke, ee, lane);
                else
                    sm[wv][lane] = wave_subtree<T, V, 4, 8>(q0, q1, P, wv * 8, col, val, out, ke, ee, lane) +
                                   wave_subtree<T, V, 4, 8>(q0, q1, P, wv * 8 + 4, col, val, out, ke, ee, lane);
            } else {
                sm[wv][lane] = wave_subtree_n<T, V, kMaxWg>(na, q0, q1, P, wv * na, col, val, out, ke, ee, lane);
            }
            __syncthreads();
            if (wv == 0 && e < ke) {
                const LV acc = row_start<LV, BWD>(i, ee, k, ke, d, in, rs);
                stv(out + (size_t)i * ke + e, acc + lds_tree_n<LV, kWgParts>(sm, 0, lane));
            }
            __syncthreads();
        }
        return;
    }
    if (h >= H) return;   // uniform over the workgroup
    LV *sc = scratch + ((size_t)slot * kMaxWg + h) * ke;
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + lane, ee = e < ke ? e : ke - 1;
        sm[wv][lane] = wave_subtree<T, V, 1, sizeof(LV) == 8 ? 8 : 16>(q0, q1, P, h * kWgParts + wv, col, val, out,
                                                                     ke, ee, lane);
        __syncthreads();
        if (wv == 0 && e < ke) sc[e] = lds_tree_n<LV, kWgParts>(sm, 0, lane);
        __syncthreads();
    }
    if (wv != 0) return;
    // release this workgroup's subtrees before its arrival; only the last
    // arrival acquires the others'
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    int old = 0;
    if (lane == 0) old = __hip_atomic_fetch_add(counter + slot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    old = __shfl(old, 0, 64);
    if (old != H - 1) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const LV *rsub = scratch + (size_t)slot * kMaxWg * ke;
    for (int eb = 0; eb < ke; eb += 64) {
        const int e = eb + lane, ee = e < ke ? e : ke - 1;
        LV acc = row_start<LV, BWD>(i, ee, k, ke, d, in, rs);
        LV fold;
        if (H == 4) {
            LV v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) v[u] = rsub[(size_t)u * ke + ee];
            tree_fold<LV, 4>(v);
            fold = v[0];
        } else {
            LV v[kMaxWg];
#pragma unroll
            for (int u = 0; u < kMaxWg; u++) v[u] = rsub[(size_t)u * ke + ee];
            tree_fold<LV, kMaxWg>(v);
            fold = v[0];
        }
        acc = acc + fold;
        if (e < ke) stv(out + (size_t)i * ke + e, acc);
    }
    if (lane == 0) counter[slot] = 0;
}

// One level of a triangular solve: schedule positions [t0, t1); the first
// nlong of them take a block each (its W = blockDim / 64 waves share the
// row's parts), the next nmed one wave each, the rest (tiny rows) rpw
// consecutive rows per wave, TG at a time (the group size sets how many
// gathers a wave holds in flight, and so its registers).
template <typename T, typename VI, typename V, bool BWD, int TG>
__global__ __launch_bounds__(64 * kLevelWaves) void k_trisolve_level(int32_t t0, int32_t nlong, int32_t nmed,
                                                                     int32_t t1, int rpw, int xcd_slices,
                                                                     const int32_t *__restrict__ rows,
                                                                     const int32_t *__restrict__ ptr,
                                                                     const int32_t *__restrict__ col,
                                                                     const T *__restrict__ val,
                                                                     const double *__restrict__ diag, int k,
                                                                     const VI *__restrict__ in, V *__restrict__ out,
                                                                     RowScale rs) {
    typedef typename Lane<V>::type LV;
    const int ke = k / Lane<V>::n;
    const int lane = SS_LANE, wv = SS_WAVE_IN_BLOCK, W = (int)(blockDim.x >> 6);
    if ((int)blockIdx.x < nlong) {
        // a long row: gg of the W waves hold P / gg consecutive parts each
        // (subtrees of the row's part tree), wave 0 folds them
        __shared__ LV sm[kLevelWaves][64];
        const int t = t0 + blockIdx.x;
        const int i = wave_uniform(rows[t]);
        const int q0 = wave_uniform(ptr[t]), q1 = wave_uniform(ptr[t + 1]);
        const double d = BWD ? diag[i] : 1.0;
        const int P = seg_parts(q1 - q0);
        const int gg = P < W ? P : W, na = P / gg;
        for (int eb = 0; eb < ke; eb += 64) {
            const int e = eb + lane, ee = e < ke ? e : ke - 1;
            LV acc = wv == 0 ? row_start<LV, BWD>(i, ee, k, ke, d, in, rs) : LV();
            if (P == 1) {
                if (wv == 0) acc = row_dot_sub<T, V>(acc, q0, q1, 1, col, val, out, ke, ee);
            } else if (wv < gg) {
                const int s0 = wv * na;
                sm[wv][lane] = wave_subtree_n<T, V, kWgParts / 4>(na, q0, q1, P, s0, col, val, out, ke, ee, lane);
            }
            __syncthreads();
            if (wv == 0) {
                if (P > 1) acc = acc + lds_tree<LV>(sm, 0, gg, lane);
                if (e < ke) stv(out + (size_t)i * ke + e, acc);
            }
            __syncthreads();
        }
        return;
    }
    // wave work: nmed single rows, then groups of rpw tiny rows.  Workgroups
    // go to the 8 XCDs round robin; with xcd_slices block b works on chunk
    // (b % 8) * per + b / 8, so each XCD owns one contiguous slice of the
    // level's rows (with rows in elimination order: one region of the
    // graph, the same region at every level)
    const int nb = (int)gridDim.x - nlong, b = (int)blockIdx.x - nlong;
    const int per = nb / kXcds;
    const int chunk = xcd_slices ? (b % kXcds) * per + b / kXcds : b;
    const int w = wave_uniform(chunk * W + wv);
    if (w < nmed) {
        wave_row<T, VI, V, BWD>(t0 + nlong + w, rows, ptr, col, val, diag, k, ke, in, out, rs, lane);
        return;
    }
    const int tb = t0 + nlong + nmed + (w - nmed) * rpw;
    for (int g0 = 0; g0 < rpw; g0 += TG) {
        const int t = tb + g0;
        if (t >= t1) break;
        int cnt = rpw - g0 < TG ? rpw - g0 : TG;
        if (t + cnt > t1) cnt = t1 - t;
        wave_tiny_rows<T, VI, V, BWD, TG>(t, cnt, rows, ptr, col, val, diag, k, ke, in, out, rs, lane);
    }
}

// A run of small levels [l0, l1) walked by one workgroup of 16 waves with a
// workgroup barrier between levels (all waves share one CU, so the barrier's
// workgroup-scope fence orders the row stores before the next level's
// loads).  A level of r rows gives each row g = 16 / r waves (rounded down
// to a power of two), which split the row's parts between them.
//
// The schedule of the run (level bounds, row records, entries) does not
// depend on the solve, so it is staged in LDS a chunk of levels at a time:
// three dependent loads per chunk instead of per level, leaving each level
// one dependent load (its gathers) before its barrier.  A chunk is the most
// levels whose rows and entries fit the stage; one small level always fits.
constexpr int kStageLevels = 256, kStageRows = 512, kStageNnz = 2 * kSmallNnz;
static_assert(kStageRows >= kSmallRows && kStageNnz >= kSmallNnz, "one small level fits the stage");
template <typename T, typename VI, typename V, bool BWD>
__global__ __launch_bounds__(1024) void k_tri_small(int32_t l0, int32_t l1, const int32_t *__restrict__ level_ptr,
                                                    const int32_t *__restrict__ rows,
                                                    const int32_t *__restrict__ ptr,
                                                    const int32_t *__restrict__ col, const T *__restrict__ val,
                                                    const double *__restrict__ diag, int k,
                                                    const VI *__restrict__ in, V *out, RowScale rs) {
    typedef typename Lane<V>::type LV;
    const int ke = k / Lane<V>::n;
    __shared__ LV sm[16][64];
    __shared__ int32_t s_lp[kStageLevels + 1];
    __shared__ int32_t s_rows[kStageRows];
    __shared__ int32_t s_ptr[kStageRows + 1];
    __shared__ int32_t s_col[kStageNnz];
    __shared__ T s_val[kStageNnz];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c0 = l0; c0 < l1;) {
        // level bounds, then row records, then entries of the chunk
        const int nl = l1 - c0 < kStageLevels ? l1 - c0 : kStageLevels;
        for (int u = threadIdx.x; u <= nl; u += blockDim.x) s_lp[u] = level_ptr[c0 + u];
        __syncthreads();
        const int tb = s_lp[0];
        const int np = s_lp[nl] - tb < kStageRows ? s_lp[nl] - tb : kStageRows;
        for (int u = threadIdx.x; u <= np; u += blockDim.x) {
            s_ptr[u] = ptr[tb + u];
            if (u < np) s_rows[u] = rows[tb + u];
        }
        __syncthreads();
        const int eb0 = s_ptr[0];
        // levels 1 .. nc end inside the stage (a prefix: bounds only grow)
        const int u = (int)threadIdx.x + 1;
        const bool fits = u <= nl && s_lp[u] - tb <= np && s_ptr[s_lp[u] - tb] - eb0 <= kStageNnz;
        const int nc = __syncthreads_count(fits);
        if (nc == 0) return;   // not reached: every level of a run is small, and one small level fits
        const int ne = s_ptr[s_lp[nc] - tb] - eb0;
        for (int q = threadIdx.x; q < ne; q += blockDim.x) {
            s_col[q] = col[eb0 + q];
            s_val[q] = val[eb0 + q];
        }
        __syncthreads();
        for (int l = 0; l < nc; l++) {
            const int pb = s_lp[l] - tb, pe = s_lp[l + 1] - tb;
            const int r = pe - pb;                  // r <= kSmallRows == 16
            if (r <= 0) continue;
            int g = 16 / r;
            g = 1 << (31 - __builtin_clz(g));
            const int slot = wv / g, sub = wv - slot * g;
            const int pt = pb + slot;
            const bool act = pt < pe;
            const int i = act ? s_rows[pt] : 0;
            const int q0 = act ? s_ptr[pt] - eb0 : 0, q1 = act ? s_ptr[pt + 1] - eb0 : 0;
            const double d = (BWD && act) ? diag[i] : 1.0;
            const int P = act ? seg_parts(q1 - q0) : 1;
            const int gg = P < g ? P : g;             // waves that hold parts, P / gg each
            for (int eb = 0; eb < ke; eb += 64) {
                const int e = eb + lane, ee = e < ke ? e : ke - 1;
                LV acc = LV();
                if (act && sub == 0) acc = row_start<LV, BWD>(i, ee, k, ke, d, in, rs);
                if (act && P == 1 && sub == 0) acc = row_dot_sub<T, V>(acc, q0, q1, 1, s_col, s_val, out, ke, ee);
                if (act && P > 1 && sub < gg)
                    sm[slot * g + sub][lane] = wave_subtree_n<T, V, kWgParts>(P / gg, q0, q1, P, sub * (P / gg),
                                                                              s_col, s_val, out, ke, ee, lane);
                __syncthreads();
                if (act && P > 1 && sub == 0) acc = acc + lds_tree<LV>(sm, slot * g, gg, lane);
                if (act && sub == 0 && e < ke) stv(out + (size_t)i * ke + e, acc);
                __syncthreads();   // also orders this level's stores before the next level's loads
            }
        }
        c0 += nc;
    }
}


// ---------------------------------------------------------------------------
// (d) fused effective resistance + probability + coin + sample + compaction.
// One thread per edge (canonical order), 256 edges per tile.  Tiles take a
// ticket so that every tile a tile waits on has started; the running kept
// count moves through a decoupled look-back over per-tile words holding
// (epoch, status, count) in one 64-bit value, so no fence is needed.
//   d_e^2 = sum_j (Z_u - Z_v)^2, j in order     (src/sparsifier.rs:575-578)
//   p_e = min(1, beta * -1 * L_uv * d_e.powi(2) / k)  (src/sparsifier.rs:553)
//   keep = coin < p; kept weight -(L_uv + (w - w/sqrt(p)))
//                                                (src/sparsifier.rs:639-661)
// ---------------------------------------------------------------------------
#define SS_TILE_EPOCH_SHIFT 40
#define SS_TILE_FLAG_SHIFT 38
#define SS_TILE_VALUE_MASK ((1ull << 38) - 1)

template <typename TZ>
__global__ __launch_bounds__(256) void k_er_sample(ErSampleArgs<TZ> a) {
#pragma clang fp contract(off)
    __shared__ uint32_t s_tile;
    __shared__ int s_wcnt[4];
    __shared__ int64_t s_excl;
    if (threadIdx.x == 0) s_tile = atomicAdd(a.ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const int64_t e = (int64_t)tile * 256 + threadIdx.x;
    bool keep = false;
    int32_t eu = 0, ev = 0;
    double kwv = 0.0;
    if (e < a.m) {
        eu = a.eu[e];
        ev = a.ev[e];
        const double w = a.ew_is_lval ? -1.0 * a.ew[e] : a.ew[e];
        const double lval = -1.0 * w;
        double p = 0.0;
        // an edge that is not u < v < n with L_uv < 0 (or a probability
        // outside [0, 1]) is flagged and dropped before Z is touched
        bool bad = eu < 0 || ev >= a.n || eu >= ev || !(w > 0.0);
        if (!bad && a.p_in) {
            p = a.p_in[e];
            bad = !(p >= 0.0 && p <= 1.0);
        }
        if (bad) {
            __hip_atomic_store(a.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            p = 0.0;
        } else if (!a.p_in) {
            const TZ *zu = a.Z + (size_t)eu * a.k;
            const TZ *zv = a.Z + (size_t)ev * a.k;
            double s = 0.0;
            for (int j = 0; j < a.k; j++) {
                const double t = (double)zu[j] - (double)zv[j];
                s = s + t * t;
            }
            const double d = sqrt(s);
            if (a.d_out) a.d_out[e] = d;
            const double d2 = d * d;
            const double x = a.beta * -1.0 * lval * d2 / (double)a.k;
            p = 1.0 * (x < 1.0 ? x : 1.0);
        }
        if (a.p_out) a.p_out[e] = p;
        const double c = coin_value(a.ck, (uint64_t)e);
        keep = c < p;
        if (a.mask) a.mask[e] = keep ? 1 : 0;
        if (keep) {
            const double target = w / sqrt(p);
            const double change = w - target;
            const double lnew = lval + change;
            kwv = -1.0 * lnew;
        }
    }
    const int lane = SS_LANE, wv = SS_WAVE_IN_BLOCK;
    const uint64_t bal = __ballot(keep);
    const int wpre = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_wcnt[wv] = __popcll(bal);
    __syncthreads();
    int woff = 0;
    for (int w2 = 0; w2 < wv; w2++) woff += s_wcnt[w2];
    if (threadIdx.x == 0) {
        const uint64_t total = (uint64_t)(s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3]);
        const uint64_t ep = (uint64_t)a.epoch << SS_TILE_EPOCH_SHIFT;
        int64_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(&a.tile_state[0], ep | (2ull << SS_TILE_FLAG_SHIFT) | total,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&a.tile_state[tile], ep | (1ull << SS_TILE_FLAG_SHIFT) | total,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int64_t j = (int64_t)tile - 1;
            int spins = 0;
            while (j >= 0) {
                const uint64_t st =
                    __hip_atomic_load(&a.tile_state[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const uint32_t sep = (uint32_t)(st >> SS_TILE_EPOCH_SHIFT);
                const uint32_t flag = (uint32_t)(st >> SS_TILE_FLAG_SHIFT) & 3u;
                if (sep != a.epoch || flag == 0) {
                    if (++spins > a.spin_limit) {
                        __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                excl += (int64_t)(st & SS_TILE_VALUE_MASK);
                if (flag == 2) break;
                j--;
            }
            __hip_atomic_store(&a.tile_state[tile],
                               ep | (2ull << SS_TILE_FLAG_SHIFT) | ((uint64_t)excl + total),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_excl = excl;
        if ((int64_t)tile == a.ntiles - 1) *a.kept_total = excl + (int64_t)total;
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
// Host launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(int64_t items, int per_block, unsigned cap = 0) {
    int64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (cap && g > cap) g = cap;
    return (unsigned)g;
}
static const unsigned kStreamGrid = 2048;   // grid-stride kernels

template <typename OutT>
void launch_sketch(hipStream_t s, int32_t n, const int32_t *iptr, const int32_t *ieid, const double *ib,
                   int k, uint64_t seed, double scale, int uniform, int variant, OutT *Y) {
    hipLaunchKernelGGL(k_sketch<OutT>, dim3(grid_for(n, 4)), dim3(256), 0, s, n, iptr, ieid, ib, k, seed,
                       scale, uniform, variant, Y);
}
template void launch_sketch<double>(hipStream_t, int32_t, const int32_t *, const int32_t *, const double *,
                                    int, uint64_t, double, int, int, double *);
template void launch_sketch<float>(hipStream_t, int32_t, const int32_t *, const int32_t *, const double *,
                                   int, uint64_t, double, int, int, float *);

void launch_copy_f64(hipStream_t s, int64_t count, const double *a, double *b) {
    hipLaunchKernelGGL(k_copy_f64, dim3(grid_for(count, 256, 1 << 20)), dim3(256), 0, s, count, a, b);
}

void launch_fill_u32(hipStream_t s, int64_t count, uint32_t *p, uint32_t v) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(count, 256, kStreamGrid)), dim3(256), 0, s, count, p, v);
}

template <typename T, typename TI, typename V>
void launch_core_head(hipStream_t s, int32_t t0, int32_t nhuge, int32_t t1, const int32_t *rows, const int32_t *ptr,
                      const int32_t *col, const T *val, const double *diag, int k, const TI *in, V *out,
                      RowScale rs) {
    // every core row as a long row of a level, a workgroup each (the core
    // rows are many): the rows of more than kWgRowMax entries first
    launch_trisolve_level<T, TI, V, false>(s, t0, nhuge, 1, t1 - t0 - nhuge, 0, t1, kLevelWaves, 1, 0, rows, ptr,
                                           col, val, diag, k, in, out, nullptr, nullptr, rs);
}
#define SS_INST_HEAD(T, TI, V)                                                                                  \
    template void launch_core_head<T, TI, V>(hipStream_t, int32_t, int32_t, int32_t, const int32_t *,            \
                                             const int32_t *, const int32_t *, const T *, const double *, int,   \
                                             const TI *, V *, RowScale);
SS_INST_HEAD(double, double, double)
SS_INST_HEAD(float, float, float)
SS_INST_HEAD(float, f32x2, f32x2)
SS_INST_HEAD(float, f32x2, f16x2)
SS_INST_HEAD(float, f32x4, f16x4)

template <typename T, typename TI, typename V, bool BWD>
void launch_tri_small(hipStream_t s, int32_t l0, int32_t l1, const int32_t *level_ptr, const int32_t *rows,
                      const int32_t *ptr, const int32_t *col, const T *val, const double *diag, int k, const TI *in,
                      V *out, RowScale rs) {
    if (l1 <= l0) return;
    hipLaunchKernelGGL((k_tri_small<T, TI, V, BWD>), dim3(1), dim3(1024), 0, s, l0, l1, level_ptr, rows, ptr, col,
                       val, diag, k, in, out, rs);
}
#define SS_INST_SMALL(T, TI, V, B)                                                                              \
    template void launch_tri_small<T, TI, V, B>(hipStream_t, int32_t, int32_t, const int32_t *, const int32_t *,  \
                                                const int32_t *, const int32_t *, const T *, const double *,    \
                                                int, const TI *, V *, RowScale);
SS_INST_SMALL(double, double, double, false)
SS_INST_SMALL(double, double, double, true)
SS_INST_SMALL(float, float, float, false)
SS_INST_SMALL(float, float, float, true)
SS_INST_SMALL(float, f32x2, f32x2, false)
SS_INST_SMALL(float, f32x2, f32x2, true)
SS_INST_SMALL(float, f32x2, f16x2, false)
SS_INST_SMALL(float, f16x2, f16x2, true)
SS_INST_SMALL(float, f32x4, f16x4, false)
SS_INST_SMALL(float, f16x4, f16x4, true)

template <typename T, typename TI, typename V, bool BWD>
void launch_trisolve_level(hipStream_t s, int32_t t0, int32_t nhuge, int32_t hm, int32_t nlong, int32_t nmed,
                           int32_t t1, int waves, int rpw, int xcd, const int32_t *rows, const int32_t *ptr,
                           const int32_t *col, const T *val, const double *diag, int k, const TI *in, V *out,
                           void *scratch, int *counter, RowScale rs) {
    const int ke = k / Lane<V>::n;
    if (nhuge > 0) {
        hipLaunchKernelGGL((k_trisolve_huge<T, TI, V, BWD>), dim3((unsigned)nhuge * (unsigned)hm),
                           dim3(64 * kWgParts), 0, s, t0, hm, rows, ptr, col, val, diag, k, in, out,
                           static_cast<typename Lane<V>::type *>(scratch), counter, rs);
        t0 += nhuge;
    }
    if (t1 <= t0) return;
    if (waves == 0) {   // a level of rows without entries
        const int32_t nrows = t1 - t0;
        hipLaunchKernelGGL((k_level_copy<T, TI, V, BWD>), dim3(grid_for(nrows, 4 * (256 / ke), 1 << 16)), dim3(256),
                           0, s, t0, nrows, rows, diag, k, in, out, rs);
        return;
    }
    if (rpw < 1) rpw = 1;
    if (waves < 4 || waves > kLevelWaves || (waves & (waves - 1))) waves = 4;
    const int64_t tiny_rows = (int64_t)t1 - t0 - nlong - nmed;
    const int64_t wave_work = (int64_t)nmed + (tiny_rows + rpw - 1) / rpw;
    int64_t blocks = (wave_work + waves - 1) / waves;
    // XCD slices: a whole number of blocks per XCD (blocks past the rows
    // return at once), only when every XCD gets several
    const int xcd_slices = xcd && blocks >= 4 * kXcds;
    if (xcd_slices) blocks = (blocks + kXcds - 1) / kXcds * kXcds;
    const unsigned grid = (unsigned)nlong + (unsigned)blocks;
    if (rpw == 1)
        hipLaunchKernelGGL((k_trisolve_level<T, TI, V, BWD, 1>), dim3(grid), dim3(64 * waves), 0, s, t0, nlong, nmed,
                           t1, rpw, xcd_slices, rows, ptr, col, val, diag, k, in, out, rs);
    else if (rpw == 2)
        hipLaunchKernelGGL((k_trisolve_level<T, TI, V, BWD, 2>), dim3(grid), dim3(64 * waves), 0, s, t0, nlong, nmed,
                           t1, rpw, xcd_slices, rows, ptr, col, val, diag, k, in, out, rs);
    else
        hipLaunchKernelGGL((k_trisolve_level<T, TI, V, BWD, kTinyGroup>), dim3(grid), dim3(64 * waves), 0, s, t0,
                           nlong, nmed, t1, rpw, xcd_slices, rows, ptr, col, val, diag, k, in, out, rs);
}
#define SS_INST_TRI(T, TI, V, B)                                                                                \
    template void launch_trisolve_level<T, TI, V, B>(hipStream_t, int32_t, int32_t, int32_t, int32_t, int32_t,   \
                                                     int32_t, int, int, int, const int32_t *, const int32_t *,   \
                                                     const int32_t *, const T *, const double *, int, const TI *, \
                                                     V *, void *, int *, RowScale);
SS_INST_TRI(double, double, double, false)
SS_INST_TRI(double, double, double, true)
SS_INST_TRI(float, float, float, false)
SS_INST_TRI(float, float, float, true)
SS_INST_TRI(float, f32x2, f32x2, false)
SS_INST_TRI(float, f32x2, f32x2, true)
SS_INST_TRI(float, f32x2, f16x2, false)
SS_INST_TRI(float, f16x2, f16x2, true)
SS_INST_TRI(float, f32x4, f16x4, false)
SS_INST_TRI(float, f16x4, f16x4, true)


template <typename TZ>
void launch_er_sample(hipStream_t s, const ErSampleArgs<TZ> &a) {
    hipLaunchKernelGGL(k_er_sample<TZ>, dim3((unsigned)a.ntiles), dim3(256), 0, s, a);
}
template void launch_er_sample<double>(hipStream_t, const ErSampleArgs<double> &);
template void launch_er_sample<float>(hipStream_t, const ErSampleArgs<float> &);

}  // namespace ss
