// Host construction of the device graph's structure from a canonical edge
// list.  The Laplacian follows Triplet::insert / process_diagonal / to_csc
// (src/sparsifier.rs:48-87): off-diagonals -w, diag[v] accumulated in edge
// order.  Canonical order = the CSC upper triangle (column = larger id
// ascending, then row ascending; src/sparsifier.rs:547-548), which is also
// the insertion order InputStream::run_stream produces (src/stream.rs:44-48).
#include <cmath>

#include "ss_internal.h"

namespace ss {

int graph_check_edges(int64_t n, int64_t m, const int32_t *u, const int32_t *v, const double *w) {
    if (n <= 0 || n >= INT32_MAX) return fail(SS_ERR_INVALID, "graph: bad node count %lld", (long long)n);
    if (m < 0) return fail(SS_ERR_INVALID, "graph: negative edge count");
    if (n + 2 * m >= (int64_t)INT32_MAX) return fail(SS_ERR_CAPACITY, "graph: n + 2m exceeds int32 indexing");
    for (int64_t e = 0; e < m; e++) {
        if (u[e] < 0 || v[e] >= n || u[e] >= v[e])
            return fail(SS_ERR_INVALID, "graph: edge %lld = (%d,%d) is not u < v < n", (long long)e, u[e], v[e]);
        if (!(w[e] > 0.0) || !std::isfinite(w[e]))
            return fail(SS_ERR_INVALID, "graph: edge %lld has weight %g (must be finite and > 0)", (long long)e, w[e]);
        if (e > 0 && (v[e] < v[e - 1] || (v[e] == v[e - 1] && u[e] <= u[e - 1])))
            return fail(SS_ERR_INVALID, "graph: edge %lld breaks canonical (column, row) order or repeats an edge", (long long)e);
    }
    return SS_OK;
}

int graph_build(int64_t n, int64_t m, const int32_t *u, const int32_t *v, const double *w, HostGraph &G,
                const double *diag_in) {
    SS_TRY(graph_check_edges(n, m, u, v, w));
    G.n = n;
    G.m = m;
    std::vector<int32_t> lower((size_t)n, 0), deg((size_t)n, 0);
    std::vector<double> diag((size_t)n, 0.0);
    for (int64_t e = 0; e < m; e++) {
        lower[v[e]]++;
        deg[u[e]]++;
        deg[v[e]]++;
        diag[u[e]] += w[e];
        diag[v[e]] += w[e];
    }
    // a Laplacian handed over as a matrix keeps its own diagonal (after a
    // sparsification pass it is the old diagonal plus the changes, summed
    // in another order than the weights)
    if (diag_in)
        for (int64_t i = 0; i < n; i++) diag[i] = diag_in[i];
    G.rowptr.assign((size_t)n + 1, 0);
    for (int64_t i = 0; i < n; i++) G.rowptr[i + 1] = G.rowptr[i] + deg[i] + 1;
    const size_t nnz = (size_t)G.rowptr[n];
    G.col.resize(nnz);
    G.val.resize(nnz);
    G.eid.resize(nnz);
    std::vector<int32_t> lo_pos((size_t)n), hi_pos((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        lo_pos[i] = G.rowptr[i];
        int32_t d = G.rowptr[i] + lower[i];
        G.col[d] = (int32_t)i;
        G.val[d] = diag[i];
        G.eid[d] = -1;
        hi_pos[i] = d + 1;
    }
    for (int64_t e = 0; e < m; e++) {
        int32_t a = u[e], b = v[e];
        int32_t p = lo_pos[b]++;          // row b, column a < b
        G.col[p] = a; G.val[p] = -1.0 * w[e]; G.eid[p] = (int32_t)e;
        int32_t q = hi_pos[a]++;          // row a, column b > a
        G.col[q] = b; G.val[q] = -1.0 * w[e]; G.eid[q] = (int32_t)e;
    }
    return SS_OK;
}

}  // namespace ss
