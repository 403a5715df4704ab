// Host approximate-Cholesky factor of a graph Laplacian and the level
// analysis its triangular solves are scheduled by.
//
// Algorithm (the reference's randomized elimination, natural order):
//   src/example.cc:90-324   perform_factorization_device
//   src/example.cc:331-471  factorization_driver (assembly of L and D)
//   include/auxilliary.hpp:90-173 search_for_updates (fill gather)
//   include/pre_process.hpp:734-812 make_lower_triangular (|values|)
// Column c gathers its original lower entries plus the fill edges earlier
// columns sent to it, merges equal rows, sorts by value, and for every entry
// i but the last samples a partner j > i with probability proportional to
// its value; the fill edge (max row, min row) of weight
// v_i (tot - cum_i) / tot goes to column min row.  L[r,c] = -v_r/tot,
// D[c] = tot; a column with no lower entries gets D = 0.
//
// Differences, all deliberate: columns are eliminated in increasing order on
// one thread (every fill lands in a later column, so this order is always
// legal); each column draws from its own splitmix64 stream keyed by
// (seed, column); duplicate rows are merged in ascending value order and
// the value sort breaks ties by row.  The factor is therefore a pure
// function of (Laplacian, seed), independent of scheduling.
#include "ss_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace ss {

namespace {

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

struct ColumnRng {
    uint64_t s;
    ColumnRng(uint64_t seed, int64_t col)
        : s(mix64(seed ^ 0x6a09e667f3bcc909ULL) ^ mix64((uint64_t)col + 1)) {}
    double uniform() {           // [0,1) with 53 random bits
        s += 0x9e3779b97f4a7c15ULL;
        return (double)(mix64(s) >> 11) * (1.0 / 9007199254740992.0);
    }
};

struct Ent {
    int32_t row;
    double val;
};

struct Fill {
    int32_t row;
    int32_t next;   // index into the fill pool, -1 ends the list
    double val;
};

}  // namespace

int factor_build(int64_t n, const int32_t *colptr, const int32_t *rowidx,
                 const double *vals, uint64_t seed, Factor &F) {
    auto t0 = std::chrono::steady_clock::now();
    if (n <= 0 || n > INT32_MAX - 1) return fail(SS_ERR_INVALID, "factor: bad n=%lld", (long long)n);
    F.n = n;
    std::vector<int32_t> head((size_t)n, -1);
    std::vector<Fill> pool;
    pool.reserve((size_t)colptr[n]);
    // Per-column output: rows and values of L below the diagonal.
    std::vector<int64_t> cstart((size_t)n + 1, 0);
    std::vector<int32_t> crow;
    std::vector<double> cval;
    crow.reserve((size_t)colptr[n]);
    cval.reserve((size_t)colptr[n]);
    F.diag.assign((size_t)n, 0.0);

    std::vector<Ent> work;
    std::vector<double> cum;
    work.reserve(64);
    for (int64_t c = 0; c < n; c++) {
        work.clear();
        for (int32_t t = colptr[c]; t < colptr[c + 1]; t++) {
            int32_t r = rowidx[t];
            if (r < 0 || r >= n) return fail(SS_ERR_INVALID, "factor: row index %d out of range", r);
            if (r > c) work.push_back({r, std::fabs(vals[t])});
        }
        for (int32_t f = head[c]; f >= 0; f = pool[f].next) work.push_back({pool[f].row, pool[f].val});
        cstart[c] = (int64_t)crow.size();
        if (work.empty()) {            // no lower entries: D = 0 (grounds the component)
            F.diag[c] = 0.0;
            cstart[c + 1] = cstart[c];
            continue;
        }
        // merge equal rows, summing in ascending value order
        std::sort(work.begin(), work.end(), [](const Ent &a, const Ent &b) {
            return a.row != b.row ? a.row < b.row : a.val < b.val;
        });
        size_t cnt = 0;
        for (size_t i = 0; i < work.size(); i++) {
            if (cnt > 0 && work[cnt - 1].row == work[i].row) work[cnt - 1].val += work[i].val;
            else work[cnt++] = work[i];
        }
        work.resize(cnt);
        std::sort(work.begin(), work.end(), [](const Ent &a, const Ent &b) {
            return a.val != b.val ? a.val < b.val : a.row < b.row;
        });
        cum.resize(cnt);
        double run = 0.0;
        for (size_t i = 0; i < cnt; i++) { run += work[i].val; cum[i] = run; }
        const double tot = run;
        ColumnRng rng(seed, c);
        for (size_t i = 0; i + 1 < cnt; i++) {
            double rest = tot - cum[i];
            double target = rng.uniform() * rest;
            // first j > i with cum[j] - cum[i] >= target
            size_t lo = i + 1, hi = cnt;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (cum[mid] - cum[i] < target) lo = mid + 1; else hi = mid;
            }
            size_t j = lo < cnt ? lo : cnt - 1;
            int32_t a = work[i].row, b = work[j].row;
            int32_t frow = std::max(a, b), fcol = std::min(a, b);
            double fval = work[i].val * rest / tot;
            if (pool.size() >= (size_t)INT32_MAX) return fail(SS_ERR_CAPACITY, "factor: fill pool exceeds 2^31 entries");
            pool.push_back({frow, head[fcol], fval});
            head[fcol] = (int32_t)(pool.size() - 1);
        }
        F.diag[c] = tot;
        std::sort(work.begin(), work.end(), [](const Ent &a, const Ent &b) { return a.row < b.row; });
        for (size_t i = 0; i < cnt; i++) {
            crow.push_back(work[i].row);
            cval.push_back(-(work[i].val / tot));
        }
        cstart[c + 1] = (int64_t)crow.size();
        // the fill lists of column c are no longer needed
        head[c] = -1;
    }
    int64_t nnz_lower = (int64_t)crow.size();
    if (nnz_lower + n > (int64_t)INT32_MAX) return fail(SS_ERR_CAPACITY, "factor: nnz %lld exceeds int32", (long long)(nnz_lower + n));
    // U CSR = columns of L, diagonal first (src/example.cc:422-470)
    F.urowptr.assign((size_t)n + 1, 0);
    F.ucol.resize((size_t)(nnz_lower + n));
    F.uval.resize((size_t)(nnz_lower + n));
    int64_t p = 0;
    for (int64_t c = 0; c < n; c++) {
        F.urowptr[c] = (int32_t)p;
        F.ucol[p] = (int32_t)c; F.uval[p] = 1.0; p++;
        for (int64_t t = cstart[c]; t < cstart[c + 1]; t++) {
            F.ucol[p] = crow[t]; F.uval[p] = cval[t]; p++;
        }
    }
    F.urowptr[n] = (int32_t)p;
    F.nnz = p;
    factor_levels(F);
    F.ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return SS_OK;
}

// Forward levels: the DP of src/example.cc:531-542 (row i pushes level+1 to
// every row of column i).  Backward levels: the same DP on L^T, i.e. row i
// of the L^T solve waits for every row of column i.
void factor_levels(Factor &F) {
    const int64_t n = F.n;
    F.lvl_fwd.assign((size_t)n, 1);
    F.lvl_bwd.assign((size_t)n, 1);
    for (int64_t i = 0; i < n; i++)
        for (int32_t t = F.urowptr[i]; t < F.urowptr[i + 1]; t++) {
            int32_t c = F.ucol[t];
            if (c != i && F.lvl_fwd[i] + 1 > F.lvl_fwd[c]) F.lvl_fwd[c] = F.lvl_fwd[i] + 1;
        }
    for (int64_t i = n - 1; i >= 0; i--) {
        int32_t l = 1;
        for (int32_t t = F.urowptr[i]; t < F.urowptr[i + 1]; t++) {
            int32_t c = F.ucol[t];
            if (c != i && F.lvl_bwd[c] + 1 > l) l = F.lvl_bwd[c] + 1;
        }
        F.lvl_bwd[i] = l;
    }
    F.depth_fwd = 0;
    F.depth_bwd = 0;
    for (int64_t i = 0; i < n; i++) {
        F.depth_fwd = std::max(F.depth_fwd, F.lvl_fwd[i]);
        F.depth_bwd = std::max(F.depth_bwd, F.lvl_bwd[i]);
    }
}

// Level-ordered CSR for the two solves.  Position t of the forward schedule
// holds row perm[t] (rows sorted by (level, row)); its entries are the
// strictly-lower row of L (pull form).  The backward schedule holds the
// strictly-upper rows of L^T, i.e. the off-diagonal part of U's rows.
void factor_schedules(const Factor &F, Schedule &fwd, Schedule &bwd) {
    const int64_t n = F.n;
    auto order = [&](const std::vector<int32_t> &lvl, Schedule &S) {
        S.n = n;
        S.depth = 0;
        for (int64_t i = 0; i < n; i++) S.depth = std::max<int32_t>(S.depth, lvl[i]);
        // counting sort by level, stable in row id
        std::vector<int64_t> cnt((size_t)S.depth + 2, 0);
        for (int64_t i = 0; i < n; i++) cnt[lvl[i]]++;
        S.level_ptr.assign((size_t)S.depth + 1, 0);
        int64_t acc = 0;
        for (int32_t l = 1; l <= S.depth; l++) { S.level_ptr[l - 1] = (int32_t)acc; acc += cnt[l]; }
        S.level_ptr[S.depth] = (int32_t)acc;
        std::vector<int64_t> pos((size_t)S.depth + 1);
        for (int32_t l = 1; l <= S.depth; l++) pos[l] = S.level_ptr[l - 1];
        S.rows.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) S.rows[pos[lvl[i]]++] = (int32_t)i;
    };
    order(F.lvl_fwd, fwd);
    order(F.lvl_bwd, bwd);

    // strictly-lower rows of L: transpose of U's off-diagonal part
    std::vector<int32_t> lcnt((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; i++)
        for (int32_t t = F.urowptr[i]; t < F.urowptr[i + 1]; t++)
            if (F.ucol[t] != i) lcnt[F.ucol[t] + 1]++;
    for (int64_t i = 0; i < n; i++) lcnt[i + 1] += lcnt[i];
    std::vector<int32_t> lcol((size_t)lcnt[n]);
    std::vector<double> lval((size_t)lcnt[n]);
    {
        std::vector<int32_t> fillp(lcnt.begin(), lcnt.end() - 1);
        for (int64_t i = 0; i < n; i++)      // i ascending => each row's columns ascending
            for (int32_t t = F.urowptr[i]; t < F.urowptr[i + 1]; t++) {
                int32_t r = F.ucol[t];
                if (r == i) continue;
                lcol[fillp[r]] = (int32_t)i;
                lval[fillp[r]] = F.uval[t];
                fillp[r]++;
            }
    }
    auto pack = [&](Schedule &S, const int32_t *ptr, const int32_t *col, const double *val, bool skip_diag) {
        S.ptr.assign((size_t)n + 1, 0);
        int64_t total = 0;
        for (int64_t t = 0; t < n; t++) {
            int32_t i = S.rows[t];
            S.ptr[t] = (int32_t)total;
            for (int32_t q = ptr[i]; q < ptr[i + 1]; q++)
                if (!(skip_diag && col[q] == i)) total++;
        }
        S.ptr[n] = (int32_t)total;
        S.col.resize((size_t)total);
        S.val.resize((size_t)total);
        int64_t p = 0;
        for (int64_t t = 0; t < n; t++) {
            int32_t i = S.rows[t];
            for (int32_t q = ptr[i]; q < ptr[i + 1]; q++) {
                if (skip_diag && col[q] == i) continue;
                S.col[p] = col[q];
                S.val[p] = val[q];
                p++;
            }
        }
    };
    pack(fwd, lcnt.data(), lcol.data(), lval.data(), false);
    pack(bwd, F.urowptr.data(), F.ucol.data(), F.uval.data(), true);
}

}  // namespace ss
