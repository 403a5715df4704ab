// Internal declarations shared by the host C++ and the HIP sources.
#pragma once
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
    int32_t rounds = 0;               // elimination rounds (ORDER_ROUNDS)
    std::vector<int32_t> order;       // position -> vertex
    std::vector<int32_t> urowptr;     // reference precond_M layout
    std::vector<int32_t> ucol;
    std::vector<double> uval;
    std::vector<double> diag;         // D (0 on each component's last column)
    std::vector<int32_t> lvl_fwd, lvl_bwd;   // 1-based
    int32_t depth_fwd = 0, depth_bwd = 0;
    double ms_build = 0.0;
};

// Level-ordered pull-form CSR of one triangular solve.
struct Schedule {
    int64_t n = 0;
    int32_t depth = 0;
    std::vector<int32_t> rows;        // row processed at position t
    std::vector<int32_t> level_ptr;   // positions of level l: [ptr[l-1], ptr[l])
    std::vector<int32_t> level_long;  // rows of level l with more than long_row entries (placed first)
    // Small-level run handled by one workgroup: levels [small_lo, small_hi)
    // (0-based).  Forward: the trailing run; each of its rows lists the
    // entries of earlier (head) levels first, split[t] of them.  Backward:
    // the leading run.
    int32_t small_lo = 0, small_hi = 0;
    // Forward rows: entries [ptr[t], ptr[t] + split[t]) form the first
    // segment (the columns in the head levels of the row's graph), the rest
    // the second; split[t] = row length outside the graph's own tail run.
    std::vector<int32_t> split;
    std::vector<int32_t> ptr;         // entries of position t: [ptr[t], ptr[t+1])
    std::vector<int32_t> col;
    std::vector<double> val;
};

int factor_build(int64_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals,
                 uint64_t seed, int ordering, Factor &F);
void factor_levels(Factor &F);
// Small-level run of one graph's forward schedule: the first 0-based level
// of its trailing run of levels with at most small_rows rows.
int32_t forward_tail_start(const Factor &F, int32_t small_rows);
// Level-ordered schedules in vertex ids; checks that every dependency sits
// at a strictly lower level.  Forward rows list their entries sorted by the
// level of the column (stable), so the entries of earlier levels lead.
// tail_lo (per position, may be NULL = forward_tail_start(F) everywhere)
// is the tail start of the position's own graph, which fixes the segments.
// long_row and small_rows must be the kLongRow / kSmallRows of the kernels.
int factor_schedules(const Factor &F, Schedule &fwd, Schedule &bwd, const std::vector<int32_t> *tail_lo,
                     int32_t long_row, int32_t small_rows);

// ---- host graph structure -----------------------------------------------
// Canonical edges (u < v, sorted by (v, u)) -> full Laplacian CSR with the
// diagonal and, for every off-diagonal entry, the id of its edge.
struct HostGraph {
    int64_t n = 0, m = 0;
    std::vector<int32_t> rowptr, col, eid;   // eid = -1 on the diagonal
    std::vector<double> val;
};
int graph_check_edges(int64_t n, int64_t m, const int32_t *u, const int32_t *v, const double *w);
int graph_build(int64_t n, int64_t m, const int32_t *u, const int32_t *v, const double *w,
                HostGraph &G);

}  // namespace ss
