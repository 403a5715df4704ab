// Internal declarations shared by the host C++ and the HIP sources.
#pragma once
#include <utility>
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/spec_spars_hip.h"

namespace ss {

// ---- errors: thread-local last message, returned status ----------------
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
inline int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
inline int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    set_error("%s", buf);
    return code;
}

#define SS_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t _e = (call);                                                         \
        if (_e != hipSuccess)                                                           \
            return ::ss::fail(SS_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(_e), \
                              __FILE__, __LINE__);                                      \
    } while (0)

#define SS_TRY(expr)              \
    do {                          \
        int _s = (expr);          \
        if (_s != SS_OK) return _s; \
    } while (0)

// ---- host factor --------------------------------------------------------
// Elimination orders (the randomized elimination itself is the same).
enum Ordering { ORDER_NATURAL = 0, ORDER_RANDOM = 1, ORDER_ROUNDS = 2 };

// The factor is stored in elimination-position space: position p is vertex
// order[p]; row p of U lists column p of L (positions > p, diagonal first),
// the reference precond_M layout.  D and the levels are per position.
struct Factor {
    int64_t n = 0;
    int64_t nnz = 0;                  // incl. unit diagonal
    int32_t ordering = ORDER_NATURAL;
    int32_t rounds = 0;               // elimination rounds (ORDER_ROUNDS; a dense-tail vertex is a round)
    int64_t tail = 0;                 // ORDER_ROUNDS: positions [n - tail, n) are the dense tail
    std::vector<int32_t> order;       // position -> vertex
    std::vector<int32_t> urowptr;     // reference precond_M layout
    std::vector<int32_t> ucol;
    std::vector<double> uval;
    std::vector<double> diag;         // D (0 on each component's last column)
    std::vector<int32_t> lvl_fwd, lvl_bwd;   // 1-based
    int32_t depth_fwd = 0, depth_bwd = 0;
    double ms_build = 0.0;
};

// Level-ordered pull-form CSR of one triangular solve over the rows outside
// the dense cores.  Forward positions [core_t0, n) then hold the core rows
// in elimination order with their entries outside the core.
struct Schedule {
    int64_t n = 0;                    // scheduled rows (positions)
    int32_t depth = 0;
    int32_t core_t0 = 0;
    int32_t core_huge = 0;            // forward core rows of more than kWgRowMax entries (placed first)
    std::vector<int32_t> rows;        // vertex finished at position t
    std::vector<int32_t> level_ptr;   // positions of level l: [level_ptr[l], level_ptr[l+1])
    std::vector<int32_t> level_long;  // rows of level l with more than long_row entries (placed first)
    std::vector<int32_t> level_huge;  // the first of them: rows of more than kWgRowMax entries
    std::vector<int32_t> level_med;   // then rows taking one wave each (more than kTinyRow entries); tiny rows last
    // Runs of consecutive small levels (at least two levels of at most
    // small_rows rows each), each walked by one workgroup: [first, second).
    std::vector<std::pair<int32_t, int32_t>> runs;
    std::vector<int32_t> ptr;         // entries of position t: [ptr[t], ptr[t+1])
    std::vector<int32_t> col;
    std::vector<double> val;
};

// Rounds order: the dense tail starts once at most kTailMaxRows vertices
// are live, a round took fewer than 1/kTailFraction of them and their mean
// live degree is at least 1/kTailDensity of their count.
constexpr int64_t kTailMaxRows = 32768;
constexpr int64_t kTailFraction = 64;
constexpr uint64_t kTailDensity = 16;

int factor_build(int64_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals,
                 uint64_t seed, int ordering, Factor &F);
// Host threads of the factorization (0 = default); results do not depend on it.
void factor_set_threads(int threads);
int factor_threads();
void factor_levels(Factor &F);
// First 0-based forward level of a graph's dense core: the trailing run of
// levels with at most small_rows rows each, cut to at most max_rows rows.
int32_t forward_tail_start(const Factor &F, int32_t small_rows, int64_t max_rows);
// Level-ordered schedules in vertex ids of the positions outside the dense
// cores (core[p] != 0); checks that every dependency sits at a strictly
// lower level.  long_row and small_rows must be the kernels' kLongRow /
// kSmallRows.
int factor_schedules(const Factor &F, const std::vector<char> &core, Schedule &fwd, Schedule &bwd, int32_t long_row,
                     int32_t small_rows);

// ---- host graph structure -----------------------------------------------
// Canonical edges (u < v, sorted by (v, u)) -> full Laplacian CSR with the
// diagonal and, for every off-diagonal entry, the id of its edge.
struct HostGraph {
    int64_t n = 0, m = 0;
    std::vector<int32_t> rowptr, col, eid;   // eid = -1 on the diagonal
    std::vector<double> val;
};
int graph_check_edges(int64_t n, int64_t m, const int32_t *u, const int32_t *v, const double *w);
// diag (n values, may be NULL) replaces the diagonal summed from the weights.
int graph_build(int64_t n, int64_t m, const int32_t *u, const int32_t *v, const double *w,
                HostGraph &G, const double *diag = nullptr);

}  // namespace ss
