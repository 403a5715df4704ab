/*
 * oracle.c -- CPU restatement of the spec_spars hot path.  TEST
 * INFRASTRUCTURE ONLY (checker and CPU baseline); never linked into the
 * shipped library.  Built with -ffp-contract=off so that every product and
 * sum rounds separately, as the reference's Rust/C++ does on x86-64.
 *
 * Pinning status of each function is listed in oracle.h.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * CityHash64WithSeed for the 16-byte key written by
 *   input1.hash(&mut h); input2.hash(&mut h); h.finish() as u32
 * (src/sparsifier.rs:397-407, :410-420).  `u64::hash` writes the 8 native
 * (little-endian) bytes, fasthash's Hasher64 buffers them and hashes the
 * 16-byte buffer with CityHash64WithSeed(buf, 16, seed) on finish().
 * The algorithm below is Google CityHash's published one (v1.1 HashLen0to16
 * for 8 <= len <= 16; v1.0.3 selectable with variant = 1).
 * ==================================================================== */
static const uint64_t CH_K2 = 0x9ae16a3b2f90404fULL;

static uint64_t rotr64(uint64_t v, int s) {
    return s == 0 ? v : ((v >> s) | (v << (64 - s)));
}

uint64_t or_hash128to64(uint64_t lo, uint64_t hi) {
    const uint64_t kmul = 0x9ddfea08eb382d69ULL;
    uint64_t a = (lo ^ hi) * kmul;
    a ^= a >> 47;
    uint64_t b = (hi ^ a) * kmul;
    b ^= b >> 47;
    return b * kmul;
}

static uint64_t hash_len16_mul(uint64_t u, uint64_t v, uint64_t mul) {
    uint64_t a = (u ^ v) * mul;
    a ^= a >> 47;
    uint64_t b = (v ^ a) * mul;
    b ^= b >> 47;
    return b * mul;
}

uint64_t or_city64_seed16(uint64_t e, uint64_t j, uint64_t seed, int variant) {
    uint64_t h;
    if (variant == 0) {              /* v1.1: HashLen0to16, len = 16 */
        uint64_t mul = CH_K2 + 32;
        uint64_t a = e + CH_K2;
        uint64_t b = j;
        uint64_t c = rotr64(b, 37) * mul + a;
        uint64_t d = (rotr64(a, 25) + b) * mul;
        h = hash_len16_mul(c, d, mul);
    } else {                         /* v1.0.3: HashLen0to16, len = 16 */
        uint64_t a = e, b = j;
        h = or_hash128to64(a, rotr64(b + 16, 16)) ^ b;
    }
    /* CityHash64WithSeed(s,len,seed) = HashLen16(CityHash64(s,len)-k2, seed) */
    return or_hash128to64(h - CH_K2, seed);
}

/* Rademacher: ((x as i64 >> 31) * 2 - 1) with x = h as u32
 * (src/sparsifier.rs:379-382); uniform: 2*((x+0.5)/2^32)-1
 * (src/sparsifier.rs:385-394); times jl_scaling_factor
 * (src/sparsifier.rs:444-458). */
double or_sketch_value(uint64_t e, uint64_t j, uint64_t seed, double scale,
                       int uniform, int variant) {
    uint32_t x = (uint32_t)or_city64_seed16(e, j, seed, variant);
    double h;
    if (uniform) {
        double u = ((double)x + 0.5) / 4294967296.0;
        h = 2.0 * u - 1.0;
    } else {
        int64_t s = (((int64_t)x) >> 31) * 2 - 1;
        h = (double)s;
    }
    return h * scale;
}

/* Y = B Q (src/sparsifier.rs:490-504).  Column e of the incidence matrix B
 * holds -sqrt(w) at v1 and +sqrt(w) at v2 (src/sparsifier.rs:95-110); the
 * column is visited in ascending row order and each term is the rounded
 * product q_j * B[row,e] added to Y[row,j], edges in order.  The reference
 * sums per-block partial Y's in thread-arrival order; for m <= n there is
 * one block and this loop is the same arithmetic. */
void or_sketch(int64_t m, const int32_t *v1, const int32_t *v2, const double *w,
               int64_t n, int32_t k, uint64_t seed, double scale, int uniform,
               int variant, double *Y) {
    memset(Y, 0, sizeof(double) * (size_t)n * (size_t)k);
    double *q = (double *)malloc(sizeof(double) * (size_t)k);
    for (int64_t e = 0; e < m; e++) {
        for (int32_t j = 0; j < k; j++)
            q[j] = or_sketch_value((uint64_t)e, (uint64_t)j, seed, scale,
                                   uniform, variant);
        double nv = sqrt(-1.0 * (-w[e])) * -1.0;   /* value at v1 */
        int32_t rows[2];
        double vals[2];
        if (v1[e] < v2[e]) {
            rows[0] = v1[e]; vals[0] = nv;
            rows[1] = v2[e]; vals[1] = -1.0 * nv;
        } else {
            rows[0] = v2[e]; vals[0] = -1.0 * nv;
            rows[1] = v1[e]; vals[1] = nv;
        }
        for (int t = 0; t < 2; t++) {
            double *yr = Y + (size_t)rows[t] * (size_t)k;
            for (int32_t j = 0; j < k; j++) yr[j] += q[j] * vals[t];
        }
    }
    free(q);
}

/* The same sum over a subset of the edges: edge eid[t] = (v1[t], v2[t],
 * w[t]) for t in [0, count), ids ascending, added into Y (not cleared).  A
 * row that every incident edge of is in the subset gets exactly its
 * or_sketch value; the reference splits the edges into blocks of n the same
 * way, one thread each (src/sparsifier.rs:509-529).                      */
void or_sketch_edges(int64_t count, const int64_t *eid, const int32_t *v1, const int32_t *v2,
                     const double *w, int64_t n, int32_t k, uint64_t seed, double scale,
                     int uniform, int variant, double *Y) {
    (void)n;
    double *q = (double *)malloc(sizeof(double) * (size_t)k);
    for (int64_t t = 0; t < count; t++) {
        const uint64_t e = (uint64_t)eid[t];
        for (int32_t j = 0; j < k; j++)
            q[j] = or_sketch_value(e, (uint64_t)j, seed, scale, uniform, variant);
        double nv = sqrt(-1.0 * (-w[t])) * -1.0;   /* value at v1 */
        int32_t rows[2];
        double vals[2];
        if (v1[t] < v2[t]) {
            rows[0] = v1[t]; vals[0] = nv;
            rows[1] = v2[t]; vals[1] = -1.0 * nv;
        } else {
            rows[0] = v2[t]; vals[0] = -1.0 * nv;
            rows[1] = v1[t]; vals[1] = nv;
        }
        for (int r = 0; r < 2; r++) {
            double *yr = Y + (size_t)rows[r] * (size_t)k;
            for (int32_t j = 0; j < k; j++) yr[j] += q[j] * vals[r];
        }
    }
    free(q);
}

/* P A P^T of a CSC matrix for the permutation rank (new index of old index
 * i is rank[i]), rows sorted within each column: the Laplacian in a
 * factor's position space.  Output arrays sized n + 1 and nnz.            */
void or_permute_csc(int64_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals,
                    const int32_t *rank, int32_t *pcolptr, int32_t *prowidx, double *pvals) {
    pcolptr[0] = 0;
    for (int64_t c = 0; c < n; c++) pcolptr[rank[c] + 1] = colptr[c + 1] - colptr[c];
    for (int64_t i = 0; i < n; i++) pcolptr[i + 1] += pcolptr[i];
    for (int64_t c = 0; c < n; c++) {
        const int32_t o = pcolptr[rank[c]];
        const int32_t len = colptr[c + 1] - colptr[c];
        for (int32_t t = 0; t < len; t++) {
            prowidx[o + t] = rank[rowidx[colptr[c] + t]];
            pvals[o + t] = vals[colptr[c] + t];
        }
        /* insertion sort: a Laplacian column is short except at hubs */
        for (int32_t a = 1; a < len; a++) {
            const int32_t r = prowidx[o + a];
            const double x = pvals[o + a];
            int32_t b = a - 1;
            while (b >= 0 && prowidx[o + b] > r) {
                prowidx[o + b + 1] = prowidx[o + b];
                pvals[o + b + 1] = pvals[o + b];
                b--;
            }
            prowidx[o + b + 1] = r;
            pvals[o + b + 1] = x;
        }
    }
}

/* ======================================================================
 * Laplacian (src/sparsifier.rs:48-87, :693-709): each edge contributes -w
 * at (v1,v2) and (v2,v1); diag[v] accumulates w in insertion order; the
 * triplets are converted to CSC (duplicates summed, exact zeros dropped,
 * as sprs' add does on the next step).
 * ==================================================================== */
typedef struct { int32_t row; int64_t order; double val; } or_ent;

static int cmp_ent(const void *a, const void *b) {
    const or_ent *x = (const or_ent *)a, *y = (const or_ent *)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->order != y->order) return x->order < y->order ? -1 : 1;
    return 0;
}

int64_t or_laplacian(int64_t n, int64_t m, const int32_t *v1, const int32_t *v2,
                     const double *w, int32_t *colptr, int32_t *rowidx,
                     double *vals) {
    double *diag = (double *)calloc((size_t)n, sizeof(double));
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t e = 0; e < m; e++) {
        if (v1[e] < 0 || v1[e] >= n || v2[e] < 0 || v2[e] >= n) {
            free(diag); free(cnt); return -1;
        }
        if (v1[e] == v2[e]) continue;
        cnt[v1[e] + 1]++; cnt[v2[e] + 1]++;
        diag[v1[e]] += w[e];
        diag[v2[e]] += w[e];
    }
    for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    /* entries per column: off-diagonals + 1 diagonal */
    int64_t total = cnt[n] + n;
    or_ent *ent = (or_ent *)malloc(sizeof(or_ent) * (size_t)(total ? total : 1));
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t c = 0; c <= n; c++) pos[c] = cnt[c] + c;   /* + diag slots */
    int64_t *start = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    memcpy(start, pos, sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t e = 0; e < m; e++) {
        if (v1[e] == v2[e]) continue;
        or_ent a = { v1[e], 2 * e, -1.0 * w[e] };
        ent[pos[v2[e]]++] = a;
        or_ent b = { v2[e], 2 * e + 1, -1.0 * w[e] };
        ent[pos[v1[e]]++] = b;
    }
    for (int64_t c = 0; c < n; c++) {
        or_ent d = { (int32_t)c, 2 * m + c, diag[c] };
        ent[pos[c]++] = d;
    }
    int64_t nnz = 0;
    colptr[0] = 0;
    for (int64_t c = 0; c < n; c++) {
        int64_t s = start[c], t = start[c + 1];
        qsort(ent + s, (size_t)(t - s), sizeof(or_ent), cmp_ent);
        int64_t i = s;
        while (i < t) {
            int32_t r = ent[i].row;
            double acc = ent[i].val;
            int64_t j = i + 1;
            while (j < t && ent[j].row == r) { acc += ent[j].val; j++; }
            if (acc != 0.0) { rowidx[nnz] = r; vals[nnz] = acc; nnz++; }
            i = j;
        }
        colptr[c + 1] = (int32_t)nnz;
    }
    free(diag); free(cnt); free(ent); free(pos); free(start);
    return nnz;
}

/* Upper-triangle iteration `for (value,(row,col)) in L.iter() if row < col`
 * (src/sparsifier.rs:547-548, :571-572, :648-649): columns ascending, rows
 * ascending within a column. */
int64_t or_upper_edges(int64_t n, const int32_t *colptr, const int32_t *rowidx,
                       const double *vals, int32_t *u, int32_t *v, double *lval) {
    int64_t e = 0;
    for (int64_t c = 0; c < n; c++)
        for (int32_t t = colptr[c]; t < colptr[c + 1]; t++)
            if (rowidx[t] < c) {
                u[e] = rowidx[t]; v[e] = (int32_t)c; lval[e] = vals[t]; e++;
            }
    return e;
}

/* ======================================================================
 * Level sets.  Forward: the DP of src/example.cc:531-542 over the factor
 * CSR (row i of U lists column i of L, diagonal first), 1-based.
 * Backward: the same DP reversed for the L^T solve.
 * ==================================================================== */
void or_levels_fwd(int64_t n, const int32_t *urowptr, const int32_t *ucol,
                   int32_t *level) {
    for (int64_t i = 0; i < n; i++) level[i] = 1;
    for (int64_t i = 0; i < n; i++)
        for (int32_t t = urowptr[i]; t < urowptr[i + 1]; t++) {
            int32_t c = ucol[t];
            if (c == i) continue;
            if (level[i] + 1 > level[c]) level[c] = level[i] + 1;
        }
}

void or_levels_bwd(int64_t n, const int32_t *urowptr, const int32_t *ucol,
                   int32_t *level) {
    for (int64_t i = n - 1; i >= 0; i--) {
        int32_t l = 1;
        for (int32_t t = urowptr[i]; t < urowptr[i + 1]; t++) {
            int32_t c = ucol[t];
            if (c == i) continue;
            if (level[c] + 1 > l) l = level[c] + 1;
        }
        level[i] = l;
    }
}

/* ======================================================================
 * Preconditioned CG (include/custom_cg.hpp:95-274).
 * M^{-1} r = U^{-1} D^+ U^{-T} r with U the factor CSR (unit diagonal):
 * trsv(TRANSPOSE, U), divide by D skipping zeros, trsv(NON_TRANSPOSE, U)
 * (include/custom_cg.hpp:224-235).
 * ==================================================================== */
void or_precond_apply(int64_t n, const int32_t *urowptr, const int32_t *ucol,
                      const double *uval, const double *diag, const double *r,
                      double *z) {
    double *t = (double *)malloc(sizeof(double) * (size_t)(n ? n : 1));
    memcpy(t, r, sizeof(double) * (size_t)n);
    /* U^T t = r : column-oriented forward substitution */
    for (int64_t i = 0; i < n; i++) {
        double dii = 1.0;
        for (int32_t q = urowptr[i]; q < urowptr[i + 1]; q++)
            if (ucol[q] == i) dii = uval[q];
        t[i] = t[i] / dii;
        for (int32_t q = urowptr[i]; q < urowptr[i + 1]; q++)
            if (ucol[q] != i) t[ucol[q]] -= uval[q] * t[i];
    }
    for (int64_t i = 0; i < n; i++)
        if (diag[i] != 0) t[i] = t[i] / diag[i];
    /* U z = t : backward substitution */
    for (int64_t i = n - 1; i >= 0; i--) {
        double acc = t[i], dii = 1.0;
        for (int32_t q = urowptr[i]; q < urowptr[i + 1]; q++) {
            if (ucol[q] == i) dii = uval[q];
            else acc -= uval[q] * z[ucol[q]];
        }
        z[i] = acc / dii;
    }
    free(t);
}

static void spmv(int64_t n, const int32_t *rp, const int32_t *ci,
                 const double *v, const double *x, double *y) {
    for (int64_t i = 0; i < n; i++) {
        double acc = 0.0;
        for (int32_t t = rp[i]; t < rp[i + 1]; t++) acc += v[t] * x[ci[t]];
        y[i] = acc;
    }
}

static double dot(int64_t n, const double *a, const double *b) {
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int or_pcg(int64_t n, const int32_t *arowptr, const int32_t *acol,
           const double *aval, const int32_t *urowptr, const int32_t *ucol,
           const double *uval, const double *diag, const double *b_in,
           double *x, int stop_rule, double tol, int max_iter, double *relres) {
    size_t sz = sizeof(double) * (size_t)(n ? n : 1);
    double *b = (double *)malloc(sz), *r = (double *)malloc(sz);
    double *z = (double *)malloc(sz), *p = (double *)malloc(sz);
    double *q = (double *)malloc(sz);
    /* enforce_zero_sum_vector (include/custom_cg.hpp:80-93) */
    double sum = 0.0;
    for (int64_t i = 0; i < n; i++) sum += b_in[i];
    double mean = sum / (double)n;
    for (int64_t i = 0; i < n; i++) b[i] = b_in[i] - mean;
    for (int64_t i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; }
    double bnorm = sqrt(dot(n, b, b));
    int it = 0;
    if (bnorm > 0.0) {
        or_precond_apply(n, urowptr, ucol, uval, diag, r, z);
        memcpy(p, z, sz);
        double rz = dot(n, r, z);
        while (it < max_iter) {
            spmv(n, arowptr, acol, aval, p, q);
            double pq = dot(n, p, q);
            if (pq == 0.0) break;
            double alpha = rz / pq;
            for (int64_t i = 0; i < n; i++) {
                x[i] += alpha * p[i];
                r[i] -= alpha * q[i];
            }
            it++;
            double rr = dot(n, r, r);
            /* user stop test (include/custom_cg.hpp:213-222) or relative */
            if (stop_rule == 0 ? (sqrt(rr) <= sqrt(tol))
                               : (sqrt(rr) <= tol * bnorm))
                break;
            or_precond_apply(n, urowptr, ucol, uval, diag, r, z);
            double rzn = dot(n, r, z);
            double beta = rzn / rz;
            rz = rzn;
            for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        }
    }
    /* true residual (include/custom_cg.hpp:245-258) */
    spmv(n, arowptr, acol, aval, x, q);
    double rn = 0.0;
    for (int64_t i = 0; i < n; i++) { double d = b[i] - q[i]; rn += d * d; }
    *relres = bnorm > 0.0 ? sqrt(rn) / bnorm : 0.0;
    free(b); free(r); free(z); free(p); free(q);
    return it;
}

/* k sequential solves (src/example.cc:552-569) with the row-major n x k
 * layout of FlattenedVec (src/lib.rs:116-142, src/example.cc:589-618). */
int or_solve_lap(int64_t n, int32_t k, const double *Y, const int32_t *arowptr,
                 const int32_t *acol, const double *aval, const int32_t *urowptr,
                 const int32_t *ucol, const double *uval, const double *diag,
                 int stop_rule, double tol, int max_iter, double *Z,
                 int32_t *iters, double *relres) {
    size_t sz = sizeof(double) * (size_t)(n ? n : 1);
    double *b = (double *)malloc(sz), *x = (double *)malloc(sz);
    int worst = 0;
    for (int32_t j = 0; j < k; j++) {
        for (int64_t i = 0; i < n; i++) b[i] = Y[(size_t)i * k + j];
        double rr;
        int it = or_pcg(n, arowptr, acol, aval, urowptr, ucol, uval, diag, b, x,
                        stop_rule, tol, max_iter, &rr);
        for (int64_t i = 0; i < n; i++) Z[(size_t)i * k + j] = x[i];
        if (iters) iters[j] = it;
        if (relres) relres[j] = rr;
        if (it > worst) worst = it;
    }
    free(b); free(x);
    return worst;
}

/* ======================================================================
 * Effective resistance and probabilities.
 * d_e = sqrt(sum_j (Z[u,j]-Z[v,j])^2), j sequential (src/sparsifier.rs:565-583)
 * p_e = min(1, beta * -1.0 * L_uv * d_e.powi(2) / k)  (src/sparsifier.rs:542-562)
 * ==================================================================== */
void or_diff_norms(int64_t m, const int32_t *u, const int32_t *v, int32_t k,
                   const double *Z, double *d) {
    for (int64_t e = 0; e < m; e++) {
        const double *zu = Z + (size_t)u[e] * k, *zv = Z + (size_t)v[e] * k;
        double s = 0.0;
        for (int32_t j = 0; j < k; j++) {
            double t = zu[j] - zv[j];
            s += t * t;
        }
        d[e] = sqrt(s);
    }
}

void or_probs(int64_t m, const double *lval, const double *d, double beta,
              int32_t k, double *p) {
    for (int64_t e = 0; e < m; e++) {
        double d2 = d[e] * d[e];
        double x = beta * -1.0 * lval[e] * d2 / (double)k;
        p[e] = 1.0 * (x < 1.0 ? x : 1.0);
    }
}

/* ======================================================================
 * Coins: StdRng::seed_from_u64(seed).gen_range(0.0..1.0), one per edge
 * (src/sparsifier.rs:625-633).  StdRng = ChaCha12 (rand_chacha 0.3) with a
 * 64-bit block counter and 64-bit stream id 0; seed_from_u64 expands the
 * u64 with PCG32 into the 32-byte key (rand_core 0.6); each f64 draw takes
 * one next_u64 = keystream words 2e (low) and 2e+1 (high) and maps it to
 * (u >> 12) * 2^-52, never rejecting (rand 0.8.5 UniformFloat::sample_single).
 * ==================================================================== */
#define ROTL32(v, c) (((v) << (c)) | ((v) >> (32 - (c))))
#define QR(a, b, c, d)                                   \
    do {                                                 \
        a += b; d ^= a; d = ROTL32(d, 16);               \
        c += d; b ^= c; b = ROTL32(b, 12);               \
        a += b; d ^= a; d = ROTL32(d, 8);                \
        c += d; b ^= c; b = ROTL32(b, 7);                \
    } while (0)

void or_chacha_block(const uint32_t key[8], uint64_t counter, uint64_t stream,
                     int rounds, uint32_t out[16]) {
    uint32_t s[16];
    s[0] = 0x61707865u; s[1] = 0x3320646eu; s[2] = 0x79622d32u; s[3] = 0x6b206574u;
    for (int i = 0; i < 8; i++) s[4 + i] = key[i];
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
    s[14] = (uint32_t)stream;  s[15] = (uint32_t)(stream >> 32);
    uint32_t x[16];
    memcpy(x, s, sizeof x);
    for (int r = 0; r < rounds; r += 2) {
        QR(x[0], x[4], x[8], x[12]);  QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]); QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]); QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);  QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

void or_stdrng_seed(uint64_t seed, uint32_t key[8]) {
    const uint64_t MUL = 6364136223846793005ULL;
    const uint64_t INC = 11634580027462260723ULL;
    uint64_t st = seed;
    for (int i = 0; i < 8; i++) {
        st = st * MUL + INC;
        uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27);
        uint32_t rot = (uint32_t)(st >> 59);
        key[i] = (xs >> rot) | (xs << ((32 - rot) & 31));
    }
}

void or_coins(uint64_t seed, int64_t first, int64_t count, double *c) {
    uint32_t key[8], blk[16];
    or_stdrng_seed(seed, key);
    int64_t cur = -1;
    for (int64_t i = 0; i < count; i++) {
        int64_t e = first + i;
        int64_t b = e / 8;
        if (b != cur) { or_chacha_block(key, (uint64_t)b, 0, 12, blk); cur = b; }
        int w = (int)(e % 8) * 2;
        uint64_t u = (uint64_t)blk[w] | ((uint64_t)blk[w + 1] << 32);
        c[i] = (double)(u >> 12) * (1.0 / 4503599627370496.0);
    }
}

/* ======================================================================
 * sample_and_reweight + apply_reweightings (src/sparsifier.rs:635-691):
 * kept iff c < p; kept: target = w/sqrt(p), change = w - target, new
 * L_uv = L_uv + change, diagonals get -change; dropped: L_uv + w = 0
 * (removed), diagonals get -w.  Diagonal changes sum in edge order.
 * ==================================================================== */
int64_t or_sample_reweight(int64_t m, int64_t n, const int32_t *u,
                           const int32_t *v, const double *lval,
                           const double *p, const double *c, uint8_t *mask,
                           int32_t *ku, int32_t *kv, double *kw, double *ddiag) {
    int64_t kept = 0;
    if (ddiag) memset(ddiag, 0, sizeof(double) * (size_t)n);
    for (int64_t e = 0; e < m; e++) {
        double true_value = lval[e] * -1.0;
        int keep = c[e] < p[e];
        mask[e] = (uint8_t)keep;
        if (keep) {
            double target = true_value / sqrt(p[e]);
            double change = true_value - target;
            double x = change * -1.0;
            if (ddiag) { ddiag[u[e]] += x; ddiag[v[e]] += x; }
            double lnew = lval[e] + x * -1.0;
            ku[kept] = u[e]; kv[kept] = v[e]; kw[kept] = -1.0 * lnew;
            kept++;
        } else {
            double change = true_value * -1.0;
            if (ddiag) { ddiag[u[e]] += change; ddiag[v[e]] += change; }
        }
    }
    return kept;
}
