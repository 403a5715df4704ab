// Device-resident batch of graphs (a single graph is a batch of one),
// multi-RHS PCG driver and the sparsify pipeline.
//
// The PCG is the method MKL's dcg runs for one column at a time in
// include/custom_cg.hpp:95-274 (x0 = 0, b projected to zero mean, stop test
// after every x/r update, M^{-1} = L^{-T} D^+ L^{-1}), run here on all k
// columns of all G graphs at once: every (graph, column) pair keeps its own
// alpha/beta/rz in FP64 and a convergence mask that freezes it (alpha = 0)
// once its stop rule holds.  The projection removes the mean of every
// connected component (the reference removes one global mean, the same
// thing on a connected graph).  Every reduction is segmented per
// (graph, column) over row tiles that never straddle two graphs, and summed
// in a fixed order, so a graph's results do not depend on what else is in
// the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <type_traits>
#include <type_traits>

#include "device.h"

namespace ss {

static const int kSpinLimit = 1 << 22;   // polls per look-back wait (~seconds)

// SPEC_SPARS_VERBOSE set: the host-side phases of an upload on stderr
struct UploadClock {
    const char *what;
    const bool on;
    std::chrono::steady_clock::time_point t;
    std::string line;
    explicit UploadClock(const char *w)
        : what(w), on(std::getenv("SPEC_SPARS_VERBOSE") != nullptr), t(std::chrono::steady_clock::now()) {}
    void mark(const char *phase) {
        if (!on) return;
        const auto now = std::chrono::steady_clock::now();
        char buf[96];
        std::snprintf(buf, sizeof buf, " %s %.2fs", phase, std::chrono::duration<double>(now - t).count());
        line += buf;
        t = now;
    }
    ~UploadClock() {
        if (on) std::fprintf(stderr, "[upload] %s:%s\n", what, line.c_str());
    }
};

// ---------------------------------------------------------------------------
// per-kernel timing
// ---------------------------------------------------------------------------
static hipEvent_t prof_event(ss_graph *g) {
    if (g->ev_used == g->ev_pool.size()) {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        g->ev_pool.push_back(e);
    }
    return g->ev_pool[g->ev_used++];
}

ProfScope::ProfScope(ss_graph *g_, int kind_, int launches_) : g(g_), kind(kind_), launches(launches_) {
    if (g->prof && (a = prof_event(g))) (void)hipEventRecord(a, g->ctx->stream);
}

ProfScope::~ProfScope() {
    if (!g->prof || !a) return;
    hipEvent_t b = prof_event(g);
    if (!b) return;
    (void)hipEventRecord(b, g->ctx->stream);
    if (launches > 0) g->prof_recs.push_back({kind, launches, a, b});
}

int prof_resolve(ss_graph *g) {
    if (!g->prof) return SS_OK;
    if (!g->prof_recs.empty()) SS_HIP(hipEventSynchronize(g->prof_recs.back().b));
    for (auto &r : g->prof_recs) {
        float ms = 0.f;
        SS_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        g->prof_ms[r.kind] += ms;
        g->prof_cnt[r.kind] += r.launches;
    }
    g->prof_recs.clear();
    g->ev_used = 0;
    return SS_OK;
}

// ---------------------------------------------------------------------------
// host structure helpers
// ---------------------------------------------------------------------------
// Incidence lists of B (m x n, one row per edge): per vertex, the edges that
// touch it with B's entry.  B[v1,e] = sqrt(w) * -1, B[v2,e] = -1 * that
// (src/sparsifier.rs:97-110, the edge's smaller endpoint is v1).
void build_incidence(int64_t n, int64_t m, const int32_t *v1, const int32_t *v2, const double *w,
                     std::vector<int32_t> &iptr, std::vector<int32_t> &ieid, std::vector<double> &ib) {
    iptr.assign((size_t)n + 1, 0);
    for (int64_t e = 0; e < m; e++) {
        iptr[v1[e] + 1]++;
        iptr[v2[e] + 1]++;
    }
    for (int64_t i = 0; i < n; i++) iptr[i + 1] += iptr[i];
    ieid.resize((size_t)2 * m);
    ib.resize((size_t)2 * m);
    std::vector<int32_t> pos(iptr.begin(), iptr.end() - 1);
    for (int64_t e = 0; e < m; e++) {
        const double nv = std::sqrt(-1.0 * (-w[e])) * -1.0;
        const int32_t p1 = pos[v1[e]]++;
        ieid[p1] = (int32_t)e;
        ib[p1] = nv;
        const int32_t p2 = pos[v2[e]]++;
        ieid[p2] = (int32_t)e;
        ib[p2] = -1.0 * nv;
    }
}

// Connected components of one graph (union-find over its CSR): comp[v] is
// the component index, the largest component first (ties: the one holding
// the smaller vertex), the others ordered by their smallest vertex.
static void graph_components(const HostGraph &H, std::vector<int32_t> &comp, std::vector<int32_t> &sizes) {
    const int64_t n = H.n;
    std::vector<int32_t> parent((size_t)n);
    std::iota(parent.begin(), parent.end(), 0);
    auto find = [&](int32_t x) {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    };
    for (int64_t r = 0; r < n; r++)
        for (int32_t q = H.rowptr[r]; q < H.rowptr[r + 1]; q++) {
            const int32_t a = find((int32_t)r), b = find(H.col[q]);
            if (a != b) parent[std::max(a, b)] = std::min(a, b);
        }
    std::vector<int32_t> root_id((size_t)n, -1);
    std::vector<int64_t> cnt;
    for (int64_t v = 0; v < n; v++) {
        const int32_t r = find((int32_t)v);
        if (root_id[r] < 0) {
            root_id[r] = (int32_t)cnt.size();
            cnt.push_back(0);
        }
        cnt[root_id[r]]++;
    }
    const int32_t nc = (int32_t)cnt.size();
    int32_t mainc = 0;
    for (int32_t c = 1; c < nc; c++)
        if (cnt[c] > cnt[mainc]) mainc = c;
    std::vector<int32_t> remap((size_t)nc);
    int32_t next = 1;
    for (int32_t c = 0; c < nc; c++) remap[c] = c == mainc ? 0 : next++;
    comp.resize((size_t)n);
    sizes.assign((size_t)nc, 0);
    for (int64_t v = 0; v < n; v++) {
        comp[v] = remap[root_id[find((int32_t)v)]];
        sizes[comp[v]]++;
    }
}

template <typename T>
static int h2d(hipStream_t s, T *dst, const T *src, size_t count) {
    if (count == 0) return SS_OK;
    SS_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    return SS_OK;
}

// ---------------------------------------------------------------------------
// batch upload
// ---------------------------------------------------------------------------
int batch_upload(ss_context *ctx, const std::vector<const HostGraph *> &graphs, ss_graph **out) {
    SS_HIP(hipSetDevice(ctx->device));
    const int G = (int)graphs.size();
    if (G < 1) return fail(SS_ERR_INVALID, "a batch needs at least one graph");
    std::unique_ptr<ss_graph> g(new ss_graph());
    g->ctx = ctx;
    g->G = G;
    g->tile_waves = ctx->tune.tile_waves;
    g->row_off.assign((size_t)G + 1, 0);
    g->edge_off.assign((size_t)G + 1, 0);
    g->rtile_off.assign((size_t)G + 1, 0);
    g->etile_off.assign((size_t)G + 1, 0);
    for (int i = 0; i < G; i++) {
        g->row_off[i + 1] = g->row_off[i] + graphs[i]->n;
        g->edge_off[i + 1] = g->edge_off[i] + graphs[i]->m;
        const int64_t tr = row_tile_rows(graphs[i]->n);
        const int64_t rt = g->rtile_off[i] + (graphs[i]->n + tr - 1) / tr;
        const int64_t et = g->etile_off[i] + (graphs[i]->m + kEdgeTile - 1) / kEdgeTile;
        if (rt >= INT32_MAX || et >= INT32_MAX) return fail(SS_ERR_CAPACITY, "batch: too many tiles");
        g->rtile_off[i + 1] = (int32_t)rt;
        g->etile_off[i + 1] = (int32_t)et;
    }
    g->N = g->row_off[G];
    g->M = g->edge_off[G];
    if (g->N + 2 * g->M >= (int64_t)INT32_MAX)
        return fail(SS_ERR_CAPACITY, "batch: n + 2m = %lld exceeds int32 indexing", (long long)(g->N + 2 * g->M));
    g->nrt = g->rtile_off[G];
    g->net = g->etile_off[G];
    const int64_t N = g->N, M = g->M, nnz = N + 2 * M;
    UploadClock clk("graph upload");

    // tiles
    std::vector<int32_t> rt_start((size_t)g->nrt + 1), rt_graph((size_t)std::max(g->nrt, 1)),
        et_start((size_t)g->net + 1), et_graph((size_t)std::max(g->net, 1));
    for (int i = 0; i < G; i++) {
        for (int32_t t = g->rtile_off[i]; t < g->rtile_off[i + 1]; t++) {
            rt_start[t] = (int32_t)(g->row_off[i] + (int64_t)(t - g->rtile_off[i]) * row_tile_rows(graphs[i]->n));
            rt_graph[t] = i;
        }
        for (int32_t t = g->etile_off[i]; t < g->etile_off[i + 1]; t++) {
            et_start[t] = (int32_t)(g->edge_off[i] + (int64_t)(t - g->etile_off[i]) * kEdgeTile);
            et_graph[t] = i;
        }
    }
    rt_start[g->nrt] = (int32_t)N;
    et_start[g->net] = (int32_t)M;

    // block-diagonal Laplacian, canonical edges, incidence lists, components
    std::vector<int32_t> rowptr((size_t)N + 1), col((size_t)nnz), iptr((size_t)N + 1), ieid((size_t)2 * M);
    std::vector<double> val((size_t)nnz), ib((size_t)2 * M), ew((size_t)M);
    std::vector<int32_t> eu((size_t)M), ev((size_t)M), comp_of_row((size_t)N), comp_graph, cptr(1, 0), cvert,
        small_id, comp_off((size_t)G + 1, 0), edge_off32((size_t)G + 1);
    g->h_comp_count.assign((size_t)G, 0);
    for (int i = 0; i < G; i++) {
        const HostGraph &H = *graphs[i];
        const int64_t ro = g->row_off[i], eo = g->edge_off[i];
        const int64_t po = ro + 2 * eo;   // first nonzero of graph i
        for (int64_t r = 0; r <= H.n; r++) rowptr[ro + r] = (int32_t)(po + H.rowptr[r]);
        for (size_t q = 0; q < H.col.size(); q++) {
            col[po + q] = (int32_t)(H.col[q] + ro);
            val[po + q] = H.val[q];
        }
        // canonical edges back from the CSR's upper part (local ids)
        std::vector<int32_t> lu((size_t)H.m), lv((size_t)H.m);
        std::vector<double> lw((size_t)H.m);
        for (int64_t r = 0; r < H.n; r++)
            for (int32_t q = H.rowptr[r]; q < H.rowptr[r + 1]; q++)
                if (H.col[q] > r) {
                    const int32_t e = H.eid[q];
                    lu[e] = (int32_t)r;
                    lv[e] = H.col[q];
                    lw[e] = -1.0 * H.val[q];
                }
        for (int64_t e = 0; e < H.m; e++) {
            eu[eo + e] = (int32_t)(lu[e] + ro);
            ev[eo + e] = (int32_t)(lv[e] + ro);
            ew[eo + e] = lw[e];
        }
        std::vector<int32_t> li, le;
        std::vector<double> lb;
        build_incidence(H.n, H.m, lu.data(), lv.data(), lw.data(), li, le, lb);
        for (int64_t r = 0; r <= H.n; r++) iptr[ro + r] = (int32_t)(2 * eo + li[r]);
        for (int64_t q = 0; q < 2 * H.m; q++) {
            ieid[2 * eo + q] = le[q];   // the edge's index within its graph: the sign hash input
            ib[2 * eo + q] = lb[q];
        }
        std::vector<int32_t> comp, sizes;
        graph_components(H, comp, sizes);
        const int32_t cbase = (int32_t)comp_graph.size();
        comp_off[i] = cbase;
        g->h_comp_count[i] = (int32_t)sizes.size();
        std::vector<std::vector<int32_t>> members(sizes.size());
        for (int64_t v = 0; v < H.n; v++) {
            comp_of_row[ro + v] = cbase + comp[v];
            members[comp[v]].push_back((int32_t)(ro + v));
        }
        for (size_t c = 0; c < sizes.size(); c++) {
            comp_graph.push_back(i);
            cvert.insert(cvert.end(), members[c].begin(), members[c].end());
            cptr.push_back((int32_t)cvert.size());
            if (c > 0) small_id.push_back(cbase + (int32_t)c);
        }
        edge_off32[i] = (int32_t)eo;
    }
    comp_off[G] = (int32_t)comp_graph.size();
    edge_off32[G] = (int32_t)M;
    // SpMM tiles: row tile t of nnz_t entries becomes ceil(nnz_t /
    // kSpmmTileNnz) tiles (at most kSpmmMaxSplit) of equal row counts
    std::vector<int32_t> st_start, st_graph;
    g->stile_off.assign((size_t)G + 1, 0);
    for (int i = 0; i < G; i++) {
        g->stile_off[i] = (int32_t)st_start.size();
        for (int32_t t = g->rtile_off[i]; t < g->rtile_off[i + 1]; t++) {
            const int32_t a = rt_start[t], b = rt_start[t + 1];
            const int64_t nnz_t = (int64_t)rowptr[b] - rowptr[a];
            const int64_t split = std::min<int64_t>(kSpmmMaxSplit, std::max<int64_t>(1, (nnz_t + kSpmmTileNnz - 1) / kSpmmTileNnz));
            const int32_t per = (int32_t)((b - a + split - 1) / split);
            for (int32_t r = a; r < b; r += per) {
                st_start.push_back(r);
                st_graph.push_back(i);
            }
        }
    }
    g->stile_off[G] = (int32_t)st_start.size();
    g->nst = (int)st_start.size();
    st_start.push_back((int32_t)N);
    if (st_graph.empty()) st_graph.push_back(0);
    g->ncomp = (int)comp_graph.size();
    g->nsmall = (int)small_id.size();
    std::vector<float> valf(val.begin(), val.end());
    clk.mark("host arrays");

    auto &A = g->arena_struct;
    SS_TRY(A.alloc(&g->rt_start, rt_start.size()));
    SS_TRY(A.alloc(&g->rt_graph, rt_graph.size()));
    SS_TRY(A.alloc(&g->et_start, et_start.size()));
    SS_TRY(A.alloc(&g->et_graph, et_graph.size()));
    SS_TRY(A.alloc(&g->gtile, g->rtile_off.size()));
    SS_TRY(A.alloc(&g->st_start, st_start.size()));
    SS_TRY(A.alloc(&g->st_graph, st_graph.size()));
    SS_TRY(A.alloc(&g->sgtile, g->stile_off.size()));
    SS_TRY(A.alloc(&g->edge_off_dev, edge_off32.size()));
    std::vector<int32_t> row_off32(g->row_off.begin(), g->row_off.end());
    SS_TRY(A.alloc(&g->row_off_dev, row_off32.size()));
    std::vector<int32_t> row_graph;
    if (G > 1) {
        row_graph.resize((size_t)N);
        for (int gi = 0; gi < G; gi++)
            std::fill(row_graph.begin() + g->row_off[gi], row_graph.begin() + g->row_off[gi + 1], gi);
        SS_TRY(A.alloc(&g->row_graph, (size_t)N));
    }
    SS_TRY(A.alloc(&g->eu, M));
    SS_TRY(A.alloc(&g->ev, M));
    SS_TRY(A.alloc(&g->ew, M));
    SS_TRY(A.alloc(&g->rowptr, N + 1));
    SS_TRY(A.alloc(&g->col, nnz));
    SS_TRY(A.alloc(&g->val, nnz));
    SS_TRY(A.alloc(&g->valf, nnz));
    SS_TRY(A.alloc(&g->iptr, N + 1));
    SS_TRY(A.alloc(&g->ieid, 2 * M));
    SS_TRY(A.alloc(&g->ib, 2 * M));
    SS_TRY(A.alloc(&g->comp_of_row, N));
    SS_TRY(A.alloc(&g->comp_graph, comp_graph.size()));
    SS_TRY(A.alloc(&g->comp_off, comp_off.size()));
    SS_TRY(A.alloc(&g->cptr, cptr.size()));
    SS_TRY(A.alloc(&g->cvert, cvert.size()));
    SS_TRY(A.alloc(&g->small_id, small_id.size()));
    SS_TRY(A.alloc(&g->status, 3));
    SS_TRY(A.alloc(&g->p, M));
    SS_TRY(A.alloc(&g->dnorm, M));
    SS_TRY(A.alloc(&g->mask, M));
    SS_TRY(A.alloc(&g->ku, M));
    SS_TRY(A.alloc(&g->kv, M));
    SS_TRY(A.alloc(&g->kw, M));
    SS_TRY(A.alloc(&g->tile_state, (size_t)g->net + 1));
    SS_TRY(A.alloc(&g->tile_incl, (size_t)g->net + 1));
    SS_TRY(A.alloc(&g->ticket, 1));
    auto &P = g->arena_params;
    SS_TRY(P.alloc(&g->p_colseed, (size_t)G * kMaxK));
    SS_TRY(P.alloc(&g->p_colscale, (size_t)G * kMaxK));
    SS_TRY(P.alloc(&g->p_colj, (size_t)G * kMaxK));
    SS_TRY(P.alloc(&g->p_betas, G));
    SS_TRY(P.alloc(&g->p_keys, (size_t)8 * G));
    hipStream_t s = ctx->stream;
    SS_TRY(h2d(s, g->rt_start, rt_start.data(), rt_start.size()));
    SS_TRY(h2d(s, g->rt_graph, rt_graph.data(), rt_graph.size()));
    SS_TRY(h2d(s, g->et_start, et_start.data(), et_start.size()));
    SS_TRY(h2d(s, g->et_graph, et_graph.data(), et_graph.size()));
    SS_TRY(h2d(s, g->gtile, g->rtile_off.data(), g->rtile_off.size()));
    SS_TRY(h2d(s, g->st_start, st_start.data(), st_start.size()));
    SS_TRY(h2d(s, g->st_graph, st_graph.data(), st_graph.size()));
    SS_TRY(h2d(s, g->sgtile, g->stile_off.data(), g->stile_off.size()));
    SS_TRY(h2d(s, g->edge_off_dev, edge_off32.data(), edge_off32.size()));
    SS_TRY(h2d(s, g->row_off_dev, row_off32.data(), row_off32.size()));
    if (G > 1) SS_TRY(h2d(s, g->row_graph, row_graph.data(), (size_t)N));
    SS_TRY(h2d(s, g->eu, eu.data(), M));
    SS_TRY(h2d(s, g->ev, ev.data(), M));
    SS_TRY(h2d(s, g->ew, ew.data(), M));
    SS_TRY(h2d(s, g->rowptr, rowptr.data(), N + 1));
    SS_TRY(h2d(s, g->col, col.data(), nnz));
    SS_TRY(h2d(s, g->val, val.data(), nnz));
    SS_TRY(h2d(s, g->valf, valf.data(), nnz));
    SS_TRY(h2d(s, g->iptr, iptr.data(), N + 1));
    SS_TRY(h2d(s, g->ieid, ieid.data(), 2 * M));
    SS_TRY(h2d(s, g->ib, ib.data(), 2 * M));
    SS_TRY(h2d(s, g->comp_of_row, comp_of_row.data(), N));
    SS_TRY(h2d(s, g->comp_graph, comp_graph.data(), comp_graph.size()));
    SS_TRY(h2d(s, g->comp_off, comp_off.data(), comp_off.size()));
    SS_TRY(h2d(s, g->cptr, cptr.data(), cptr.size()));
    SS_TRY(h2d(s, g->cvert, cvert.data(), cvert.size()));
    SS_TRY(h2d(s, g->small_id, small_id.data(), small_id.size()));
    SS_HIP(hipMemsetAsync(g->status, 0, sizeof(int) * 3, s));
    SS_HIP(hipMemsetAsync(g->tile_state, 0, sizeof(uint64_t) * ((size_t)g->net + 1), s));
    SS_HIP(hipStreamSynchronize(s));
    clk.mark("copies");
    g->kept_start.assign((size_t)G, 0);
    g->kept_count.assign((size_t)G, 0);
    ctx->live_graphs++;
    *out = g.release();
    return SS_OK;
}

// One factor of the block-diagonal system from the graphs' own factors:
// positions of graph i shift by its row offset; levels stay, since no entry
// couples two graphs.
static void factor_concat(const std::vector<const Factor *> &fs, const std::vector<int64_t> &row_off, Factor &F) {
    F = Factor();
    F.n = row_off.back();
    F.order.resize((size_t)F.n);
    F.diag.resize((size_t)F.n);
    F.lvl_fwd.resize((size_t)F.n);
    F.lvl_bwd.resize((size_t)F.n);
    F.urowptr.assign((size_t)F.n + 1, 0);
    int64_t nnz = 0;
    for (const Factor *f : fs) nnz += f->nnz;
    F.ucol.resize((size_t)nnz);
    F.uval.resize((size_t)nnz);
    int64_t q = 0;
    for (size_t i = 0; i < fs.size(); i++) {
        const Factor &f = *fs[i];
        const int64_t o = row_off[i];
        for (int64_t p = 0; p < f.n; p++) {
            F.order[o + p] = (int32_t)(f.order[p] + o);
            F.diag[o + p] = f.diag[p];
            F.lvl_fwd[o + p] = f.lvl_fwd[p];
            F.lvl_bwd[o + p] = f.lvl_bwd[p];
            F.urowptr[o + p] = (int32_t)q;
            for (int32_t t = f.urowptr[p]; t < f.urowptr[p + 1]; t++) {
                F.ucol[q] = (int32_t)(f.ucol[t] + o);
                F.uval[q] = f.uval[t];
                q++;
            }
        }
        F.depth_fwd = std::max(F.depth_fwd, f.depth_fwd);
        F.depth_bwd = std::max(F.depth_bwd, f.depth_bwd);
    }
    F.urowptr[F.n] = (int32_t)q;
    F.nnz = q;
}

// Rows of a factor's dense core: its dense tail (rounds order), else the
// last positions whose forward levels belong to its trailing run of small
// levels (at most kCoreMaxRows).  Always a suffix of the elimination order.
static int64_t core_rows_of(const Factor &f) {
    if (f.tail > 0) return std::min<int64_t>(f.tail, kTailMaxRows);
    const int32_t lo = forward_tail_start(f, kSmallRows, kCoreMaxRows);
    int64_t c = 0;
    while (c < f.n && c < kCoreMaxRows && f.lvl_fwd[(size_t)(f.n - 1 - c)] > lo) c++;
    return c;
}

// Invert every graph's core block of L on the device: its strictly lower
// entries (the entries of the core columns, whose rows are all core rows)
// are scattered into a zeroed FP64 cp x cp block, inverted tile row by tile
// row, and copied to FP32.
static int core_invert(ss_graph *g, const Factor &F, const std::vector<int32_t> &core_c,
                       const std::vector<int32_t> &cpad, const std::vector<int64_t> &dense_off) {
    const int64_t total = dense_off[g->G];
    if (total == 0) return SS_OK;
    hipStream_t s = g->ctx->stream;
    std::vector<int64_t> pos;
    std::vector<double> val;
    for (int i = 0; i < g->G; i++) {
        const int64_t c = core_c[i];
        if (c == 0) continue;
        const int64_t start = g->row_off[i + 1] - c;   // first core position
        for (int64_t p = start; p < g->row_off[i + 1]; p++)
            for (int32_t t = F.urowptr[p]; t < F.urowptr[p + 1]; t++) {
                const int64_t q = F.ucol[t];
                if (q == p) continue;
                if (q < start || q >= g->row_off[i + 1])
                    return fail(SS_ERR_INVALID, "core: column %lld has row %lld outside its core", (long long)p,
                                (long long)q);
                pos.push_back(dense_off[i] + (q - start) * cpad[i] + (p - start));
                val.push_back(F.uval[t]);
            }
    }
    DevArena tmp;
    double *L = nullptr, *dval = nullptr;
    int64_t *dpos = nullptr;
    SS_TRY(tmp.alloc(&L, (size_t)total));
    SS_TRY(tmp.alloc(&dpos, pos.size()));
    SS_TRY(tmp.alloc(&dval, val.size()));
    SS_HIP(hipMemsetAsync(L, 0, sizeof(double) * (size_t)total, s));
    SS_HIP(hipMemsetAsync(g->Md, 0, sizeof(double) * (size_t)total, s));
    SS_TRY(h2d(s, dpos, pos.data(), pos.size()));
    SS_TRY(h2d(s, dval, val.data(), val.size()));
    launch_core_scatter(s, (int64_t)pos.size(), dpos, dval, L);
    for (int i = 0; i < g->G; i++)
        if (cpad[i] > 0) launch_core_invert(s, cpad[i], L + dense_off[i], g->Md + dense_off[i]);
    launch_to_f32(s, total, g->Md, g->Mf);
    launch_to_f16(s, total, g->Md, g->Mh);
    SS_HIP(hipGetLastError());
    SS_HIP(hipStreamSynchronize(s));   // the host vectors and tmp are freed on return
    return SS_OK;
}

// Gathered right-hand sides and product pieces of the dense cores for k
// columns at a precision.
static int core_ensure_work(ss_graph *g, int k, int prec) {
    if (g->core.pieces == 0) return SS_OK;
    if (g->core_wprec == prec && k <= g->core_wcap) return SS_OK;
    g->arena_core_work.release();
    g->core_wcap = 0;
    g->core_wprec = -1;
    const size_t es = prec == SS_PREC_FP64 ? 8 : 4;
    unsigned char *tc = nullptr, *part = nullptr;
    SS_TRY(g->arena_core_work.alloc(&tc, (size_t)g->core.rows * k * es));
    SS_TRY(g->arena_core_work.alloc(&part, (size_t)g->core.pieces * kCoreTile * k * es));
    g->core_tc = tc;
    g->core_part = part;
    g->core_wcap = k;
    g->core_wprec = prec;
    return SS_OK;
}

int batch_upload_factors(ss_graph *g, const std::vector<const Factor *> &factors) {
    if ((int)factors.size() != g->G) return fail(SS_ERR_INVALID, "need %d factors, got %zu", g->G, factors.size());
    for (int i = 0; i < g->G; i++)
        if (!factors[i] || factors[i]->n != g->row_off[i + 1] - g->row_off[i])
            return fail(SS_ERR_INVALID, "factor %d does not match graph %d (n=%lld)", i, i,
                        (long long)(g->row_off[i + 1] - g->row_off[i]));
    SS_HIP(hipSetDevice(g->ctx->device));
    UploadClock clk("factor upload");
    Factor Fc;
    const Factor *Fp = factors[0];
    if (g->G > 1) {
        factor_concat(factors, g->row_off, Fc);
        Fp = &Fc;
    }
    const Factor &F = *Fp;
    // each graph's dense core: the last core_rows_of(factor) positions of
    // its elimination order
    std::vector<char> core((size_t)F.n, 0);
    std::vector<int32_t> core_c((size_t)g->G, 0), cpad((size_t)g->G, 0), core_vert;
    std::vector<int64_t> dense_off((size_t)g->G + 1, 0), tc_off((size_t)g->G + 1, 0);
    for (int i = 0; i < g->G; i++) {
        const int64_t ni = g->row_off[i + 1] - g->row_off[i];
        const int64_t c = core_rows_of(*factors[i]);
        core_c[i] = (int32_t)c;
        cpad[i] = (int32_t)((c + kCoreTile - 1) / kCoreTile * kCoreTile);
        for (int64_t p = g->row_off[i] + ni - c; p < g->row_off[i + 1]; p++) {
            core[p] = 1;
            core_vert.push_back(F.order[p]);
        }
        for (int64_t r = c; r < cpad[i]; r++) core_vert.push_back(-1);
        dense_off[i + 1] = dense_off[i] + (int64_t)cpad[i] * cpad[i];
        tc_off[i + 1] = tc_off[i] + cpad[i];
    }
    Schedule fw, bw;
    clk.mark("cores");
    SS_TRY(factor_schedules(F, core, fw, bw, kLongRow, kSmallRows));
    clk.mark("schedules");
    // product pieces: per row tile of each core, inner ranges of at most
    // kCoreSeg rows -- forward [0, r0 + kCoreTile), backward [r0, cp)
    std::vector<CorePiece> pf, pb;
    std::vector<int32_t> tile_row, tpf(1, 0), tpb(1, 0);
    for (int i = 0; i < g->G; i++) {
        for (int32_t r0 = 0; r0 < cpad[i]; r0 += kCoreTile) {
            for (int32_t m0 = 0; m0 < r0 + kCoreTile; m0 += kCoreSeg)
                pf.push_back({i, r0, m0, std::min(m0 + kCoreSeg, r0 + kCoreTile)});
            for (int32_t m0 = r0; m0 < cpad[i]; m0 += kCoreSeg) pb.push_back({i, r0, m0, std::min(m0 + kCoreSeg, cpad[i])});
            tile_row.push_back((int32_t)(tc_off[i] + r0));
            tpf.push_back((int32_t)pf.size());
            tpb.push_back((int32_t)pb.size());
        }
    }
    if (pf.size() != pb.size()) return fail(SS_ERR_INVALID, "core pieces: %zu forward, %zu backward", pf.size(), pb.size());
    std::vector<double> vdiag((size_t)F.n);
    for (int64_t p = 0; p < F.n; p++) vdiag[F.order[p]] = F.diag[p];
    g->arena_factor.release();
    g->has_factor = false;
    auto &A = g->arena_factor;
    const int64_t n = g->N;
    const int64_t fn = fw.n, bn = bw.n;
    const int64_t fnz = (int64_t)fw.col.size(), bnz = (int64_t)bw.col.size();
    // the row and pointer arrays carry kTinyGroup entries of padding (a
    // group of tiny rows reads its records without bounds checks)
    SS_TRY(A.alloc(&g->f_rows, fn + kTinyGroup));
    SS_TRY(A.alloc(&g->f_ptr, fn + 1 + kTinyGroup));
    SS_TRY(A.alloc(&g->f_col, fnz));
    SS_TRY(A.alloc(&g->f_val, fnz));
    SS_TRY(A.alloc(&g->f_valf, fnz));
    SS_TRY(A.alloc(&g->b_rows, bn + kTinyGroup));
    SS_TRY(A.alloc(&g->b_ptr, bn + 1 + kTinyGroup));
    SS_TRY(A.alloc(&g->b_col, bnz));
    SS_TRY(A.alloc(&g->b_val, bnz));
    SS_TRY(A.alloc(&g->b_valf, bnz));
    SS_TRY(A.alloc(&g->diag, n));
    SS_TRY(A.alloc(&g->f_lp_dev, fw.level_ptr.size()));
    SS_TRY(A.alloc(&g->b_lp_dev, bw.level_ptr.size()));
    const int64_t dense_total = dense_off[g->G];
    int32_t *d_vert = nullptr, *d_cpad = nullptr, *d_tile_row = nullptr, *d_tpf = nullptr, *d_tpb = nullptr;
    int64_t *d_dense_off = nullptr, *d_tc_off = nullptr;
    CorePiece *d_pf = nullptr, *d_pb = nullptr;
    SS_TRY(A.alloc(&d_vert, core_vert.size()));
    SS_TRY(A.alloc(&d_cpad, cpad.size()));
    SS_TRY(A.alloc(&d_dense_off, dense_off.size()));
    SS_TRY(A.alloc(&d_tc_off, tc_off.size()));
    SS_TRY(A.alloc(&d_pf, pf.size()));
    SS_TRY(A.alloc(&d_pb, pb.size()));
    SS_TRY(A.alloc(&d_tile_row, tile_row.size()));
    SS_TRY(A.alloc(&d_tpf, tpf.size()));
    SS_TRY(A.alloc(&d_tpb, tpb.size()));
    SS_TRY(A.alloc(&g->Md, dense_total));
    SS_TRY(A.alloc(&g->Mf, dense_total));
    SS_TRY(A.alloc(&g->Mh, dense_total));
    hipStream_t s = g->ctx->stream;
    std::vector<float> fvf(fw.val.begin(), fw.val.end()), bvf(bw.val.begin(), bw.val.end());
    std::vector<int32_t> frows(fw.rows), fptr(fw.ptr), brows(bw.rows), bptr(bw.ptr);
    frows.resize((size_t)fn + kTinyGroup, 0);
    fptr.resize((size_t)fn + 1 + kTinyGroup, fw.ptr[fn]);
    brows.resize((size_t)bn + kTinyGroup, 0);
    bptr.resize((size_t)bn + 1 + kTinyGroup, bw.ptr[bn]);
    SS_TRY(h2d(s, g->f_rows, frows.data(), fn + kTinyGroup));
    SS_TRY(h2d(s, g->f_ptr, fptr.data(), fn + 1 + kTinyGroup));
    SS_TRY(h2d(s, g->f_col, fw.col.data(), fnz));
    SS_TRY(h2d(s, g->f_val, fw.val.data(), fnz));
    SS_TRY(h2d(s, g->f_valf, fvf.data(), fnz));
    SS_TRY(h2d(s, g->b_rows, brows.data(), bn + kTinyGroup));
    SS_TRY(h2d(s, g->b_ptr, bptr.data(), bn + 1 + kTinyGroup));
    SS_TRY(h2d(s, g->b_col, bw.col.data(), bnz));
    SS_TRY(h2d(s, g->b_val, bw.val.data(), bnz));
    SS_TRY(h2d(s, g->b_valf, bvf.data(), bnz));
    SS_TRY(h2d(s, g->diag, vdiag.data(), n));
    SS_TRY(h2d(s, g->f_lp_dev, fw.level_ptr.data(), fw.level_ptr.size()));
    SS_TRY(h2d(s, g->b_lp_dev, bw.level_ptr.data(), bw.level_ptr.size()));
    SS_TRY(h2d(s, d_vert, core_vert.data(), core_vert.size()));
    SS_TRY(h2d(s, d_cpad, cpad.data(), cpad.size()));
    SS_TRY(h2d(s, d_dense_off, dense_off.data(), dense_off.size()));
    SS_TRY(h2d(s, d_tc_off, tc_off.data(), tc_off.size()));
    SS_TRY(h2d(s, d_pf, pf.data(), pf.size()));
    SS_TRY(h2d(s, d_pb, pb.data(), pb.size()));
    SS_TRY(h2d(s, d_tile_row, tile_row.data(), tile_row.size()));
    SS_TRY(h2d(s, d_tpf, tpf.data(), tpf.size()));
    SS_TRY(h2d(s, d_tpb, tpb.data(), tpb.size()));
    SS_HIP(hipStreamSynchronize(s));
    clk.mark("copies");
    SS_TRY(core_invert(g, F, core_c, cpad, dense_off));
    SS_HIP(hipStreamSynchronize(s));
    clk.mark("core inverse");
    g->core.rows = tc_off[g->G];
    g->core.pieces = (int32_t)pf.size();
    g->core.tiles = (int32_t)tile_row.size();
    g->core.vert = d_vert;
    g->core.cpad = d_cpad;
    g->core.dense_off = d_dense_off;
    g->core.tc_off = d_tc_off;
    g->core.piece_fwd = d_pf;
    g->core.piece_bwd = d_pb;
    g->core.tile_row = d_tile_row;
    g->core.tile_piece_fwd = d_tpf;
    g->core.tile_piece_bwd = d_tpb;
    g->core_wcap = 0;
    g->core_wprec = -1;
    g->arena_core_work.release();
    g->f_level_ptr = fw.level_ptr;
    g->b_level_ptr = bw.level_ptr;
    g->f_level_long = fw.level_long;
    g->b_level_long = bw.level_long;
    g->f_level_med = fw.level_med;
    g->b_level_med = bw.level_med;
    g->f_runs = fw.runs;
    g->b_runs = bw.runs;
    g->f_core_t0 = fw.core_t0;
    g->f_nrows = (int32_t)fn;
    g->core_max = 0;
    g->core_total = 0;
    int core_graphs = 0;
    for (int i = 0; i < g->G; i++) {
        g->core_max = std::max(g->core_max, core_c[i]);
        g->core_total += core_c[i];
        core_graphs += core_c[i] > 0;
    }
    g->core_entries = dense_total;
    // a level's long rows of up to kWgRowMax entries get a block of as many
    // waves as their longest row has parts (4..16; 0 marks a level without
    // entries); its rows of more than kWgRowMax entries get as many
    // workgroups each as the longest of them needs, folding through scratch
    // slots and counters sized for the most such rows one launch holds
    int64_t nhuge = 1;
    auto max_len = [](const Schedule &S, int64_t t0, int64_t t1) {
        int32_t m = 0;
        for (int64_t t = t0; t < t1; t++) m = std::max(m, S.ptr[t + 1] - S.ptr[t]);
        return m;
    };
    for (int dir = 0; dir < 2; dir++) {
        const Schedule &S = dir == 0 ? fw : bw;
        std::vector<int32_t> &out = dir == 0 ? g->f_level_waves : g->b_level_waves;
        std::vector<int32_t> &hm = dir == 0 ? g->f_level_hm : g->b_level_hm;
        out.assign((size_t)S.depth, 4);
        hm.assign((size_t)S.depth, 1);
        for (int32_t l = 0; l < S.depth; l++) {
            const int64_t t0 = S.level_ptr[l], th = t0 + S.level_huge[l];
            out[l] = std::min(kLevelWaves, std::max(4, seg_parts(max_len(S, th, t0 + S.level_long[l]))));
            hm[l] = row_wgs(max_len(S, t0, th));
            nhuge = std::max<int64_t>(nhuge, S.level_huge[l]);
            if (S.ptr[S.level_ptr[l + 1]] == S.ptr[S.level_ptr[l]]) out[l] = 0;   // no entries: a copy launch
        }
    }
    g->f_level_huge = fw.level_huge;
    g->b_level_huge = bw.level_huge;
    g->core_huge = fw.core_huge;
    SS_TRY(A.alloc(&g->huge_counter, (size_t)nhuge));
    SS_TRY(A.alloc(&g->huge_scratch, (size_t)nhuge * kMaxWg * kMaxK));
    SS_HIP(hipMemsetAsync(g->huge_counter, 0, sizeof(int) * (size_t)nhuge, s));
    SS_HIP(hipStreamSynchronize(s));
    for (int dir = 0; dir < 2; dir++) {
        const Schedule &S = dir == 0 ? fw : bw;
        ss_schedule_info &I = g->sched_info[dir];
        I = ss_schedule_info{};
        I.levels = S.depth;
        std::vector<char> in_run((size_t)S.depth, 0);
        for (const auto &r : S.runs)
            for (int32_t l = r.first; l < r.second; l++) in_run[l] = 1;
        for (int32_t l = 0; l < S.depth; l++) I.small_levels += in_run[l];
        I.level_launches = S.depth - I.small_levels;
        I.small_runs = (int32_t)S.runs.size();
        for (int32_t l = 0; l < S.depth; l++) {
            const bool sm = in_run[l] != 0;
            const int64_t rows = S.level_ptr[l + 1] - S.level_ptr[l];
            const int64_t nz = S.ptr[S.level_ptr[l + 1]] - S.ptr[S.level_ptr[l]];
            (sm ? I.small_rows : I.level_rows) += rows;
            (sm ? I.small_nnz : I.level_nnz) += nz;
        }
        I.core_rows = g->core_total;
        I.core_entries = dense_total;
        I.core_outside_nnz = dir == 0 ? (int64_t)(fw.ptr[fn] - fw.ptr[fw.core_t0]) : 0;
        I.core_steps = g->core.pieces;
        I.core_graphs = core_graphs;
    }
    g->factor_nnz = F.nnz;
    g->depth_fwd = F.depth_fwd;
    g->depth_bwd = F.depth_bwd;
    g->factor_epoch++;
    g->gc64.reset();
    g->gc32.reset();
    g->has_factor = true;
    return SS_OK;
}

// Point the per-(graph, column) scalars at the scalar buffer for k columns.
static void bind_scalars(ss_graph *g, int k) {
    const size_t gk = (size_t)g->G * k;
    g->sc.rz = g->sc_buf;
    g->sc.alpha = g->sc_buf + gk;
    g->sc.beta = g->sc_buf + 2 * gk;
    g->sc.rr = g->sc_buf + 3 * gk;
    g->sc.bnorm = g->sc_buf + 4 * gk;
    g->sc.active = g->sc_int;
    g->sc.iters = g->sc_int + gk;
    g->rel_ref = g->sc_buf + 5 * gk;
    g->rel_flag = g->sc_int + 2 * gk;
    if (g->prec_sc) g->prec_inv = g->prec_sc + gk;
}

size_t graph_work_bytes(const ss_graph *g, int k, int prec) {
    if (g->wprec == prec && k <= g->wcap) return 0;
    const int cap = g->wprec == prec ? std::max(k, g->wcap) : k;
    const size_t nk = (size_t)g->N * (size_t)cap, gk = (size_t)g->G * cap;
    size_t b = nk * 8 * 2;                                  // Y64, X64
    b += prec == SS_PREC_FP64 ? nk * 8 * 6 : nk * (8 * 2 + 4 * 2 + 4 * 6);
    b += ((size_t)std::max(std::max(g->nrt, g->nst), 1) + (size_t)std::max(g->nrt, 1)) * cap * 8 + gk * 8 +
         (size_t)g->ncomp * cap * 8 * 2;
    b += gk * (6 * 8 + 3 * 4 + 4 + (prec == SS_PREC_FP64 ? 0 : 2 * 4)) + (size_t)cap * 8;
    // the dense cores' gathered block and product pieces
    b += ((size_t)g->core.rows + (size_t)g->core.pieces * kCoreTile) * cap * (prec == SS_PREC_FP64 ? 8 : 4);
    return b;
}

// Work blocks for k columns at a precision.  Blocks allocated for more
// columns at the same precision are reused (the kernels take k as the row
// stride), so a resident graph serving requests of different k keeps one
// allocation.
int graph_ensure_work(ss_graph *g, int k, int prec) {
    if (k < 1 || k > kMaxK) return fail(SS_ERR_INVALID, "k=%d outside 1..%d", k, kMaxK);
    if (prec != SS_PREC_FP64 && prec != SS_PREC_MIXED) return fail(SS_ERR_INVALID, "unknown precision %d", prec);
    if (g->wprec == prec && k <= g->wcap) {
        g->wk = k;
        bind_scalars(g, k);
        return SS_OK;
    }
    const int cap = g->wprec == prec ? std::max(k, g->wcap) : k;
    SS_HIP(hipSetDevice(g->ctx->device));
    g->gc64.reset();
    g->gc32.reset();
    g->arena_work.release();
    g->wk = g->wcap = 0;
    g->wprec = -1;
    g->Y64 = g->X64 = g->R64 = g->B64 = nullptr;
    g->Y32 = g->X32 = nullptr;
    g->prec_sc = g->prec_inv = nullptr;
    g->w64 = PcgWork<double>();
    g->w32 = PcgWork<float>();
    auto &A = g->arena_work;
    const size_t nk = (size_t)g->N * (size_t)cap, gk = (size_t)g->G * cap;
    SS_TRY(A.alloc(&g->Y64, nk));
    SS_TRY(A.alloc(&g->X64, nk));
    auto alloc_work = [&](auto &w) -> int {
        SS_TRY(A.alloc(&w.B, nk));
        SS_TRY(A.alloc(&w.R, nk));
        SS_TRY(A.alloc(&w.P, nk));
        SS_TRY(A.alloc(&w.Q, nk));
        SS_TRY(A.alloc(&w.Zp, nk));
        SS_TRY(A.alloc(&w.W, nk));
        return SS_OK;
    };
    if (prec == SS_PREC_FP64) {
        SS_TRY(alloc_work(g->w64));
    } else {
        SS_TRY(A.alloc(&g->B64, nk));
        SS_TRY(A.alloc(&g->R64, nk));
        SS_TRY(A.alloc(&g->Y32, nk));
        SS_TRY(A.alloc(&g->X32, nk));
        SS_TRY(alloc_work(g->w32));
        SS_TRY(A.alloc(&g->prec_sc, 2 * gk));
    }
    SS_TRY(A.alloc(&g->partial, (size_t)std::max(std::max(g->nrt, g->nst), 1) * cap));
    SS_TRY(A.alloc(&g->partial2, (size_t)std::max(g->nrt, 1) * cap));
    SS_TRY(A.alloc(&g->gsum, gk));
    SS_TRY(A.alloc(&g->csum, (size_t)g->ncomp * cap));
    SS_TRY(A.alloc(&g->cmean, (size_t)g->ncomp * cap));
    SS_TRY(A.alloc(&g->sc_buf, 6 * gk));
    SS_TRY(A.alloc(&g->sc_int, 3 * gk));
    SS_TRY(A.alloc(&g->colkeep, gk));
    SS_TRY(A.alloc(&g->rel_gs, (size_t)cap));
    SS_TRY(A.alloc(&g->rel_ge, (size_t)cap));
    g->wk = k;
    g->wcap = cap;
    g->wprec = prec;
    bind_scalars(g, k);
    return SS_OK;
}

// ---------------------------------------------------------------------------
// preconditioner and PCG
// ---------------------------------------------------------------------------
template <typename T> static const T *fvals(const ss_graph *g);
template <> const double *fvals<double>(const ss_graph *g) { return g->f_val; }
template <> const float *fvals<float>(const ss_graph *g) { return g->f_valf; }
template <typename T> static const T *bvals(const ss_graph *g);
template <> const double *bvals<double>(const ss_graph *g) { return g->b_val; }
template <> const float *bvals<float>(const ss_graph *g) { return g->b_valf; }
template <typename T> static const T *avals(const ss_graph *g);
template <> const double *avals<double>(const ss_graph *g) { return g->val; }
template <> const float *avals<float>(const ss_graph *g) { return g->valf; }

static SketchCols sketch_cols(const ss_graph *g) { return SketchCols{g->p_colseed, g->p_colscale, g->p_colj}; }
static Tiles rtiles(const ss_graph *g) { return Tiles{g->rt_start, g->rt_graph, g->nrt, g->tile_waves}; }
static Tiles stiles(const ss_graph *g) { return Tiles{g->st_start, g->st_graph, g->nst, g->tile_waves}; }

template <typename T> static const T *core_M(const ss_graph *g);
template <> const double *core_M<double>(const ss_graph *g) { return g->Md; }
template <> const float *core_M<float>(const ss_graph *g) { return g->Mf; }

// One direction of the sparse part of the preconditioner: per-level launches
// between the runs of small levels, one single-workgroup launch per run.
// Arithmetic in T; start values read as TI (the solve's input), rows
// written and gathered as V (T, or pairs of columns: FP32, or FP16 scaled
// by rs).
template <typename T, typename TI, typename V, bool BWD>
static void solve_levels(ss_graph *g, int k, const TI *in, V *out, RowScale rs) {
    hipStream_t s = g->ctx->stream;
    const int rpw = g->ctx->tune.level_rows_per_wave;
    const std::vector<int32_t> &lp = BWD ? g->b_level_ptr : g->f_level_ptr;
    const std::vector<int32_t> &nlong = BWD ? g->b_level_long : g->f_level_long;
    const std::vector<int32_t> &nmed = BWD ? g->b_level_med : g->f_level_med;
    const std::vector<int32_t> &nhuge = BWD ? g->b_level_huge : g->f_level_huge;
    const std::vector<int32_t> &hms = BWD ? g->b_level_hm : g->f_level_hm;
    const std::vector<int32_t> &waves = BWD ? g->b_level_waves : g->f_level_waves;
    const auto &runs = BWD ? g->b_runs : g->f_runs;
    const int32_t *rows = BWD ? g->b_rows : g->f_rows, *ptr = BWD ? g->b_ptr : g->f_ptr;
    const int32_t *col = BWD ? g->b_col : g->f_col, *lp_dev = BWD ? g->b_lp_dev : g->f_lp_dev;
    const T *val = BWD ? bvals<T>(g) : fvals<T>(g);
    const int32_t depth = (int32_t)lp.size() - 1;
    size_t r = 0;
    for (int32_t l = 0; l < depth;) {
        if (r < runs.size() && runs[r].first == l) {
            const int32_t e = runs[r++].second;
            ProfScope ps(g, BWD ? SS_KERNEL_TRI_BWD_SMALL : SS_KERNEL_TRI_FWD_SMALL);
            launch_tri_small<T, TI, V, BWD>(s, l, e, lp_dev, rows, ptr, col, val, g->diag, k, in, out, rs);
            l = e;
            continue;
        }
        // the level launches up to the next run are timed as one group
        const int32_t e = r < runs.size() ? runs[r].first : depth;
        ProfScope ps(g, BWD ? SS_KERNEL_TRI_BWD : SS_KERNEL_TRI_FWD, e - l);
        for (; l < e; l++)
            launch_trisolve_level<T, TI, V, BWD>(s, lp[l], nhuge[l], hms[l], nlong[l] - nhuge[l], nmed[l], lp[l + 1],
                                                 waves[l], rpw, g->ctx->tune.xcd_slices, rows, ptr, col, val,
                                                 g->diag, k, in, out, g->huge_scratch, g->huge_counter, rs);
    }
}

// The dense cores of one direction: forward = their rows' entries outside
// the cores (the head), then y_c = M (that); backward = z_c = M^T (D^+ y_c).
// A row-major n x k block of elements V seen as k scalars per row (the
// dense-core and reduction kernels take one column per thread).
template <typename V> struct ScalarOf { typedef V type; };
template <> struct ScalarOf<f32x2> { typedef float type; };
template <> struct ScalarOf<f16x2> { typedef _Float16 type; };
template <> struct ScalarOf<f16x4> { typedef _Float16 type; };

template <typename T, typename TI, typename V, bool BWD>
static void solve_cores(ss_graph *g, int k, const TI *in, V *out, RowScale rs) {
    typedef typename ScalarOf<V>::type S;
    if (g->core_total == 0) return;
    hipStream_t s = g->ctx->stream;
    ProfScope ps(g, BWD ? SS_KERNEL_CORE_BWD : SS_KERNEL_CORE_FWD, BWD ? 3 : 4);
    T *tc = static_cast<T *>(g->core_tc), *part = static_cast<T *>(g->core_part);
    S *outs = reinterpret_cast<S *>(out);
    if constexpr (!BWD)
        launch_core_head<T, TI, V>(s, g->f_core_t0, g->core_huge, g->f_nrows, g->f_rows, g->f_ptr, g->f_col,
                                   fvals<T>(g), g->diag, k, in, out, rs);
    const S *src = BWD ? reinterpret_cast<const S *>(in) : outs;
    if constexpr (std::is_same<S, _Float16>::value)
        // FP16 rows: FP16 products on the matrix cores
        launch_core_apply_h<BWD>(s, g->core, g->Mh, k, g->diag, src, static_cast<_Float16 *>(g->core_tc),
                                 static_cast<float *>(g->core_part), outs);
    else
        launch_core_apply<T, S, BWD>(s, g->core, core_M<T>(g), k, g->diag, src, tc, part, outs);
}

// Zp = L^{-T} D^+ L^{-1} R (W holds the forward result y).  rz_partial
// receives the per-tile partials of r . Zp = y^T D^+ y.  Forward: the
// sparse levels, then the dense cores (which only read finished rows);
// backward: the cores first (their rows only need core rows), then the
// sparse levels.  V = FP16 pairs (mixed precision): W and Zp hold the
// rows scaled by rs (the partials then carry the scale squared).
template <typename T, typename VI, typename V>
static int precond_v(ss_graph *g, int k, const VI *R, V *W, V *Zp, double *rz_partial, RowScale rs) {
    typedef typename ScalarOf<V>::type S;
    solve_levels<T, VI, V, false>(g, k, R, W, rs);
    solve_cores<T, VI, V, false>(g, k, R, W, rs);
    {
        ProfScope ps(g, SS_KERNEL_VECTOR);
        const S *ws = reinterpret_cast<const S *>(W);
        if constexpr (std::is_same<S, _Float16>::value)
            bt_tile_rz_h(g->ctx->stream, rtiles(g), k, ws, g->diag, rz_partial);
        else
            bt_tile_partial<S, S>(g->ctx->stream, rtiles(g), k, ws, ws, g->diag, 2, rz_partial);
    }
    solve_cores<T, V, V, true>(g, k, W, Zp, rs);
    solve_levels<T, V, V, true>(g, k, W, Zp, rs);
    SS_HIP(hipGetLastError());
    return SS_OK;
}
// FP64, or FP32 rows: two columns per lane where k is even and at least
// 128 (the same arithmetic per column, half the gather instructions; a
// narrower row would leave most of a wave's lanes idle at twice the
// registers).
template <typename T>
static int precond(ss_graph *g, int k, const T *R, T *W, T *Zp, double *rz_partial) {
    const RowScale none{nullptr, nullptr};
    if constexpr (std::is_same<T, float>::value) {
        if (k % 2 == 0 && k >= 128)
            return precond_v<float, f32x2, f32x2>(g, k, reinterpret_cast<const f32x2 *>(R),
                                                  reinterpret_cast<f32x2 *>(W), reinterpret_cast<f32x2 *>(Zp),
                                                  rz_partial, none);
    }
    return precond_v<T, T, T>(g, k, R, W, Zp, rz_partial, none);
}

// The mixed path's FP16 preconditioner rows, two columns per lane (k
// even), in the FP32 W and Zp blocks (which hold twice the room), scaled per
// (graph, column) by prec_sc.
static RowScale prec_scale(const ss_graph *g) { return RowScale{g->prec_sc, g->row_graph}; }
static int precond_half(ss_graph *g, int k, const PcgWork<float> &w) {
    if (k % 4 == 0 && k >= 256)   // four columns per lane: one 8-byte access
        return precond_v<float, f32x4, f16x4>(g, k, reinterpret_cast<const f32x4 *>(w.R),
                                              reinterpret_cast<f16x4 *>(w.W), reinterpret_cast<f16x4 *>(w.Zp),
                                              g->partial2, prec_scale(g));
    return precond_v<float, f32x2, f16x2>(g, k, reinterpret_cast<const f32x2 *>(w.R), reinterpret_cast<f16x2 *>(w.W),
                                          reinterpret_cast<f16x2 *>(w.Zp), g->partial2, prec_scale(g));
}

static int read_status(ss_graph *g, int *nactive, int *nflagged = nullptr) {
    ss_context *c = g->ctx;
    SS_HIP(hipMemcpyAsync(c->h_status, g->status, 3 * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SS_HIP(hipStreamSynchronize(c->stream));
    if (c->h_status[1] != 0) {
        (void)hipMemsetAsync(g->status, 0, 3 * sizeof(int), c->stream);
        (void)hipStreamSynchronize(c->stream);
        return fail(SS_ERR_TIMEOUT, "device wait exceeded its spin limit (compaction look-back)");
    }
    if (nactive) *nactive = c->h_status[0];
    if (nflagged) *nflagged = c->h_status[2];
    return SS_OK;
}

// Per-component means of the FP64 block Y into g->cmean.
static void component_means(ss_graph *g, int k, const double *Y) {
    hipStream_t s = g->ctx->stream;
    bt_tile_partial<double, double>(s, rtiles(g), k, Y, Y, nullptr, 1, g->partial);
    bt_graph_sums(s, g->G * k, k, g->gtile, g->partial, g->gsum);
    bt_small_comp_sums(s, g->nsmall, g->small_id, g->cptr, g->cvert, k, Y, g->csum);
    bt_comp_means(s, g->ncomp, k, g->comp_graph, g->comp_off, g->cptr, g->gsum, g->csum, g->cmean);
}

// One PCG iteration on every (graph, column): P = Z + beta P together with
// the previous iteration's X += alpha P; Q = A P; alpha; R -= alpha Q; stop
// test.  With tail, also the next preconditioner application and beta
// (what the next iteration needs).  The solve ends with the last X update.
// With delta2 >= 0 (reliable updates), the check flags columns for a
// residual replacement instead of stopping them (defer: a column at its
// stop target waits for its replacement group).
template <typename T>
static int pcg_iteration(ss_graph *g, int k, T *X, const PcgWork<T> &w, int rule, double tol, bool tail,
                         double delta2 = -1.0, bool half = false, bool defer = false) {
    hipStream_t s = g->ctx->stream;
    const Tiles rt = rtiles(g);
    const int gk = g->G * k;
    {
        ProfScope ps(g, SS_KERNEL_VECTOR);
        // P = Z + beta P; X += alpha P of the previous iteration
        if constexpr (std::is_same<T, float>::value) {
            if (half)
                bt_update_px_h(s, rt, k, g->sc.alpha, g->sc.beta, reinterpret_cast<const _Float16 *>(w.Zp),
                               g->prec_inv, w.P, X);
            else
                bt_update_px<T, T>(s, rt, k, g->sc.alpha, g->sc.beta, w.Zp, nullptr, w.P, X);
        } else {
            bt_update_px<T, T>(s, rt, k, g->sc.alpha, g->sc.beta, w.Zp, nullptr, w.P, X);
        }
    }
    {
        ProfScope ps(g, SS_KERNEL_SPMM);
        // Q = A P, P.Q; with FP16 preconditioner rows as A Z + beta Q
        if constexpr (std::is_same<T, float>::value) {
            if (half)
                bt_spmm_z(s, stiles(g), g->rowptr, g->col, g->valf, k, reinterpret_cast<const _Float16 *>(w.Zp),
                          g->prec_inv, g->sc.beta, w.P, w.Q, g->partial);
            else
                bt_spmm_dot<T>(s, stiles(g), g->rowptr, g->col, avals<T>(g), k, w.P, w.Q, g->partial);
        } else {
            bt_spmm_dot<T>(s, stiles(g), g->rowptr, g->col, avals<T>(g), k, w.P, w.Q, g->partial);
        }
    }
    bt_pcg_alpha(s, gk, k, g->sgtile, g->partial, g->sc, g->status, delta2 >= 0.0 ? g->rel_flag : nullptr);
    {
        ProfScope ps(g, SS_KERNEL_VECTOR);
        bt_update_r_rr<T>(s, rt, k, g->sc.alpha, w.Q, w.R, g->partial);
    }
    if (delta2 >= 0.0)
        bt_reliable_check(s, gk, k, g->gtile, g->partial, rule, tol, delta2, defer, g->sc, g->rel_ref,
                          g->rel_flag, g->status);
    else
        bt_pcg_check(s, gk, k, g->gtile, g->partial, rule, tol, g->sc, g->status);
    if (tail) {
        SS_TRY(precond<T>(g, k, w.R, w.W, w.Zp, g->partial2));
        bt_pcg_beta(s, gk, k, g->gtile, g->partial2, g->sc);
    }
    SS_HIP(hipGetLastError());
    return SS_OK;
}

// Kernel launches of one PCG iteration with its preconditioner application
// (an upper bound: per-level launches of both directions, small-level runs,
// dense-core steps, vector updates and reductions).
static int64_t launches_per_iteration(const ss_graph *g) {
    const int64_t levels = (int64_t)g->f_level_ptr.size() + (int64_t)g->b_level_ptr.size();
    return levels + 8 + 16;
}
constexpr int64_t kGraphMaxNodes = 1024;

template <typename T> static GraphCache &graph_cache(ss_graph *g);
template <> GraphCache &graph_cache<double>(ss_graph *g) { return g->gc64; }
template <> GraphCache &graph_cache<float>(ss_graph *g) { return g->gc32; }

// PCG on all k columns of all graphs: X ~ A^+ (rhs - component means).
// Vectors in precision T, dots and scalars in FP64.  With mean_src the
// component means of that FP64 block are removed (rhs itself in the FP64
// solve); without it rhs is used as given.
template <typename T>
static int pcg_run(ss_graph *g, int k, const T *rhs, const double *mean_src, T *X, const PcgWork<T> &w, int rule,
                   double tol, int max_iter, std::vector<int32_t> &iters) {
    hipStream_t s = g->ctx->stream;
    const Tiles rt = rtiles(g);
    const int gk = g->G * k;
    {
        ProfScope ps(g, SS_KERNEL_VECTOR);
        if (mean_src)
            component_means(g, k, mean_src);
        else
            launch_fill_u32(s, (int64_t)2 * g->ncomp * k, reinterpret_cast<uint32_t *>(g->cmean), 0u);
        bt_project<T, T>(s, rt, k, rhs, g->comp_of_row, g->cmean, w.R, w.B, X);
    }
    SS_TRY(precond<T>(g, k, w.R, w.W, w.Zp, g->partial2));
    bt_tile_partial<T, T>(s, rt, k, w.B, w.B, nullptr, 0, g->partial);
    bt_pcg_init(s, gk, k, g->gtile, g->partial, g->partial2, g->sc);   // beta = 0: P starts as Z
    // captured iterations only for schedules of at most kGraphMaxNodes
    // launches per capture: a deep factor (thousands of level launches per
    // iteration) is launched directly
    const int m = g->prof || g->ctx->tune.graph_iters * launches_per_iteration(g) > kGraphMaxNodes
                      ? 0
                      : g->ctx->tune.graph_iters;
    int it = 0;
    if (m > 0) {
        // whole iterations replayed from a captured graph, the host reading
        // the active count after each replay; an iteration after every
        // column converged changes nothing (alpha = 0 freezes X and R)
        GraphCache &gc = graph_cache<T>(g);
        const GraphKey key{k, rule, tol, m, (const void *)X, (const void *)w.R, g->factor_epoch,
                           g->ctx->tune.level_rows_per_wave, g->tile_waves};
        if (!gc.exec || !(gc.key == key)) {
            gc.reset();
            int rc = SS_OK;
            SS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
            for (int r = 0; r < m && rc == SS_OK; r++) rc = pcg_iteration<T>(g, k, X, w, rule, tol, true);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &graph);
            if (rc != SS_OK || e != hipSuccess || !graph) {
                if (graph) (void)hipGraphDestroy(graph);
                (void)hipGetLastError();
                return rc != SS_OK ? rc : fail(SS_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
            }
            const hipError_t ei = hipGraphInstantiate(&gc.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                gc.exec = nullptr;
                return fail(SS_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
            }
            gc.key = key;
        }
        for (; it + m <= max_iter; it += m) {
            SS_HIP(hipGraphLaunch(gc.exec, s));
            int nactive = 0;
            SS_TRY(read_status(g, &nactive));
            if (nactive == 0) {
                it = max_iter;
                break;
            }
        }
    }
    for (; it < max_iter; it++) {
        SS_TRY(pcg_iteration<T>(g, k, X, w, rule, tol, false));
        int nactive = 0;
        SS_TRY(read_status(g, &nactive));
        if (nactive == 0) break;
        SS_TRY(precond<T>(g, k, w.R, w.W, w.Zp, g->partial2));
        bt_pcg_beta(s, gk, k, g->gtile, g->partial2, g->sc);
    }
    {
        // the last iteration's X += alpha P (alpha is zero for the columns
        // that stopped earlier: their last update rode with update_px)
        ProfScope ps(g, SS_KERNEL_VECTOR);
        bt_update_x<T>(s, rt, k, g->sc.alpha, w.P, X);
    }
    SS_TRY(read_status(g, nullptr));
    iters.resize(gk);
    SS_HIP(hipMemcpyAsync(iters.data(), g->sc.iters, sizeof(int32_t) * gk, hipMemcpyDeviceToHost, s));
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

// Mixed precision with reliable updates (FP64 B64 projected, X64 = 0 and
// R64 = B64 on entry; bn2 = |B64|^2 per (graph, column)).  One FP32 PCG
// (FP32 SpMM and triangular solves, FP64 dots and scalars) runs on
// (float) B64; a column whose recursive residual has fallen to delta^2
// times its reference (|r|^2 at its last replacement) or to the stop
// target folds its FP32 iterate x into X64 and replaces its residual by the
// FP64 true residual B64 - A X64 (kept current in R64 as R64 - A x, so the
// replacement gathers FP32 rows), keeping its search direction unless the
// true residual jumped above the recursive one; it stops when that true
// residual meets the stop rule.  Unlike restarting the FP32 PCG after every
// refinement, the Krylov space built so far is kept.  With replacement
// groups (grouped), a group's columns replace together, and a column at its
// stop target waits (iterating on) until every active column of its group
// is there or the group replaces for drift: one verifying round per group
// instead of one per iteration in which some of its columns arrive.
// rounds = replacement rounds (launches of the FP64 residual).
// SPEC_SPARS_VERBOSE >= 2: one line per replacement round on stderr
static const bool g_trace_rounds = [] {
    const char *v = std::getenv("SPEC_SPARS_VERBOSE");
    return v && std::atoi(v) >= 2;
}();

static int pcg_reliable(ss_graph *g, int k, int rule, double tol, double delta, int max_iter,
                        const std::vector<double> &bn2, std::vector<int32_t> &iters, int &rounds, bool &all_done,
                        bool grouped, bool half) {
    hipStream_t s = g->ctx->stream;
    const Tiles rt = rtiles(g);
    const int gk = g->G * k;
    const PcgWork<float> &w = g->w32;
    const double delta2 = delta * delta;
    {
        ProfScope ps(g, SS_KERNEL_VECTOR);
        bt_convert_masked(s, rt, k, g->R64, nullptr, g->Y32);
        launch_fill_u32(s, (int64_t)2 * g->ncomp * k, reinterpret_cast<uint32_t *>(g->cmean), 0u);
        bt_project<float, float>(s, rt, k, g->Y32, g->comp_of_row, g->cmean, w.R, w.B, g->X32);
    }
    bt_tile_partial<float, float>(s, rt, k, w.B, w.B, nullptr, 0, g->partial);
    if (half) {
        // FP16 preconditioner rows: R (= B) scaled to unit mean square
        bt_prec_scale(s, gk, k, g->gtile, g->partial, nullptr, g->row_off_dev, g->prec_sc, g->prec_inv);
        SS_TRY(precond_half(g, k, w));
    } else {
        SS_TRY(precond<float>(g, k, w.R, w.W, w.Zp, g->partial2));
    }
    bt_pcg_init(s, gk, k, g->gtile, g->partial, g->partial2, g->sc, half ? g->prec_sc : nullptr);
    SS_HIP(hipMemcpyAsync(g->rel_ref, bn2.data(), sizeof(double) * gk, hipMemcpyHostToDevice, s));
    bt_reliable_init(s, gk, g->sc, g->rel_ref, g->rel_flag, g->status);
    rounds = 0;
    all_done = false;
    for (int it = 0; it < max_iter; it++) {
        SS_TRY(pcg_iteration<float>(g, k, g->X32, w, rule, tol, false, delta2, half, grouped));
        if (grouped) bt_reliable_group(s, gk, k, g->rel_gs, g->rel_ge, g->sc, g->rel_flag, g->status);
        int nactive = 0, nflag = 0;
        SS_TRY(read_status(g, &nactive, &nflag));
        if (nflag > 0) {
            rounds++;
            if (g_trace_rounds) std::fprintf(stderr, "[reliable] iteration %d: %d of %d active columns replaced\n", it + 1, nflag, nactive);
            {
                // the folded iterate goes to Y32, free once the right-hand
                // sides were projected
                ProfScope ps(g, SS_KERNEL_VECTOR);
                bt_reliable_fold(s, rt, k, g->rel_flag, g->sc.alpha, w.P, g->X32, g->X64, g->Y32);
            }
            {
                ProfScope ps(g, SS_KERNEL_RESIDUAL);
                bt_reliable_residual(s, stiles(g), g->rowptr, g->col, g->val, k, g->rel_flag, g->Y32, g->R64, w.R,
                                     g->partial);
            }
            bt_reliable_apply(s, gk, k, g->sgtile, g->partial, rule, tol, g->sc, g->rel_ref, g->rel_flag, g->status);
            SS_TRY(read_status(g, &nactive));
        }
        if (nactive == 0) {
            all_done = true;
            break;
        }
        if (half) {
            // |R|^2 of this iteration's check (the replaced columns': their
            // FP64 residual's)
            bt_prec_scale(s, gk, k, g->gtile, nullptr, g->sc.rr, g->row_off_dev, g->prec_sc, g->prec_inv);
            SS_TRY(precond_half(g, k, w));
        } else {
            SS_TRY(precond<float>(g, k, w.R, w.W, w.Zp, g->partial2));
        }
        bt_pcg_beta(s, gk, k, g->gtile, g->partial2, g->sc, half ? g->prec_sc : nullptr);
    }
    {
        // what the FP32 iterate holds past the last replacement (nothing
        // for the columns that stopped: alpha is zero and X32 was cleared)
        ProfScope ps(g, SS_KERNEL_VECTOR);
        bt_update_x<float>(s, rt, k, g->sc.alpha, w.P, g->X32);
        bt_accumulate(s, rt, k, g->X32, g->X64);
    }
    SS_TRY(read_status(g, nullptr));
    iters.resize(gk);
    SS_HIP(hipMemcpyAsync(iters.data(), g->sc.iters, sizeof(int32_t) * gk, hipMemcpyDeviceToHost, s));
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

// |v|^2 per (graph, column) of an FP64 block, on the host.
static int block_norms2(ss_graph *g, int k, const double *V, std::vector<double> &out) {
    hipStream_t s = g->ctx->stream;
    const int gk = g->G * k;
    bt_tile_partial<double, double>(s, rtiles(g), k, V, V, nullptr, 0, g->partial);
    bt_graph_sums(s, gk, k, g->gtile, g->partial, g->gsum);
    out.resize(gk);
    SS_HIP(hipMemcpyAsync(out.data(), g->gsum, sizeof(double) * gk, hipMemcpyDeviceToHost, s));
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

// FP64 true residual R = B - A X and its |R|^2 per (graph, column).
static int residual_norms(ss_graph *g, int k, const double *X, const double *B, double *R, std::vector<double> &rn2) {
    hipStream_t s = g->ctx->stream;
    const int gk = g->G * k;
    {
        ProfScope ps(g, SS_KERNEL_RESIDUAL);
        bt_residual64(s, stiles(g), g->rowptr, g->col, g->val, k, X, B, R, g->partial);
    }
    bt_graph_sums(s, gk, k, g->sgtile, g->partial, g->gsum);
    rn2.resize(gk);
    SS_HIP(hipMemcpyAsync(rn2.data(), g->gsum, sizeof(double) * gk, hipMemcpyDeviceToHost, s));
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

int graph_solve(ss_graph *g, int k, const ss_solve_opts &o, SolveStats &st, const std::vector<int32_t> *groups) {
    if (!g->has_factor) return fail(SS_ERR_INVALID, "no factor uploaded");
    if (g->wk != k) return fail(SS_ERR_INVALID, "work not allocated for k=%d", k);
    if (o.max_iter < 0 || !(o.tol >= 0.0)) return fail(SS_ERR_INVALID, "bad solve options");
    SS_HIP(hipSetDevice(g->ctx->device));
    SS_TRY(core_ensure_work(g, k, g->wprec));
    hipStream_t s = g->ctx->stream;
    const int gk = g->G * k;
    std::vector<double> bn2, rn2;
    st.iters.assign(gk, 0);
    st.relres.assign(gk, 0.0);
    st.refinements = 0;
    if (g->wprec == SS_PREC_FP64) {
        SS_TRY(pcg_run<double>(g, k, g->Y64, g->Y64, g->X64, g->w64, o.stop_rule, o.tol, o.max_iter, st.iters));
        SS_TRY(block_norms2(g, k, g->w64.B, bn2));
        // true residual b - A x (include/custom_cg.hpp:245-258)
        SS_TRY(residual_norms(g, k, g->X64, g->w64.B, g->w64.Q, rn2));
    } else {
        // FP64 outer refinement around an FP32 PCG.
        const Tiles rt = rtiles(g);
        component_means(g, k, g->Y64);
        bt_project<double, double>(s, rt, k, g->Y64, g->comp_of_row, g->cmean, g->R64, g->B64, g->X64);
        SS_TRY(block_norms2(g, k, g->B64, bn2));
        if (!(o.inner_tol >= 0.0 && o.inner_tol < 1.0)) return fail(SS_ERR_INVALID, "inner_tol %g outside [0, 1)", o.inner_tol);
        if (o.mixed_mode != SS_MIXED_RELIABLE && o.mixed_mode != SS_MIXED_RESTART &&
            o.mixed_mode != SS_MIXED_RELIABLE_F16)
            return fail(SS_ERR_INVALID, "unknown mixed_mode %d", o.mixed_mode);
        if (o.mixed_mode == SS_MIXED_RELIABLE || o.mixed_mode == SS_MIXED_RELIABLE_F16) {
            bool grouped = false;
            if (groups) {
                if ((int)groups->size() != k) return fail(SS_ERR_INVALID, "replacement groups: %zu for k=%d", groups->size(), k);
                std::vector<int32_t> gs((size_t)k), ge((size_t)k);
                for (int j = 0; j < k; j++) {
                    if (j > 0 && (*groups)[j] < (*groups)[j - 1])
                        return fail(SS_ERR_INVALID, "replacement groups must not decrease");
                    gs[j] = j > 0 && (*groups)[j] == (*groups)[j - 1] ? gs[j - 1] : j;
                }
                for (int j = k - 1; j >= 0; j--)
                    ge[j] = j + 1 < k && (*groups)[j] == (*groups)[j + 1] ? ge[j + 1] : j + 1;
                for (int j = 0; j < k; j++) grouped |= ge[j] - gs[j] > 1;
                SS_TRY(h2d(s, g->rel_gs, gs.data(), (size_t)k));
                SS_TRY(h2d(s, g->rel_ge, ge.data(), (size_t)k));
                SS_HIP(hipStreamSynchronize(s));   // gs / ge leave scope
            }
            bool all_done = false;
            SS_TRY(pcg_reliable(g, k, o.stop_rule, o.tol, o.inner_tol > 0.0 ? o.inner_tol : 3e-3, o.max_iter, bn2,
                                st.iters, st.refinements, all_done, grouped,
                                o.mixed_mode == SS_MIXED_RELIABLE_F16 && k % 2 == 0));
            if (all_done) {
                // every column stopped on its FP64 true residual, which the
                // last replacement left in rel_ref (a zero right-hand side
                // keeps its zero); X64 has not changed since
                rn2.resize(gk);
                SS_HIP(hipMemcpyAsync(rn2.data(), g->rel_ref, sizeof(double) * gk, hipMemcpyDeviceToHost, s));
                SS_HIP(hipStreamSynchronize(s));
            } else {
                SS_TRY(residual_norms(g, k, g->X64, g->B64, g->R64, rn2));
            }
            for (int i = 0; i < gk; i++) st.relres[i] = bn2[i] > 0.0 ? std::sqrt(rn2[i] / bn2[i]) : 0.0;
            return SS_OK;
        }
        // Restarts: each FP32 sweep cuts the FP64 residual by ~1e-3 (well
        // above FP32's stagnation level); sweeps repeat until the FP64 stop
        // rule holds.
        const double inner_tol = std::max(o.inner_tol > 0.0 ? o.inner_tol : 1e-3, o.stop_rule == SS_STOP_REL ? o.tol : 0.0);
        const int inner_max = std::min(o.max_iter, 300);
        std::vector<int32_t> it_sweep, keep(gk);
        rn2 = bn2;
        for (int sweep = 0; sweep < 16; sweep++) {
            bool all_done = true;
            for (int i = 0; i < gk; i++) {
                const double r = std::sqrt(rn2[i]);
                const bool done = rn2[i] == 0.0 ||
                                  (o.stop_rule == SS_STOP_REL ? r <= o.tol * std::sqrt(bn2[i]) : rn2[i] <= o.tol);
                keep[i] = done ? 0 : 1;
                if (!done) all_done = false;
            }
            if (all_done) break;
            SS_HIP(hipMemcpyAsync(g->colkeep, keep.data(), sizeof(int) * gk, hipMemcpyHostToDevice, s));
            bt_convert_masked(s, rt, k, g->R64, g->colkeep, g->Y32);
            SS_TRY(pcg_run<float>(g, k, g->Y32, nullptr, g->X32, g->w32, SS_STOP_REL, inner_tol, inner_max,
                                  it_sweep));
            for (int i = 0; i < gk; i++) st.iters[i] += it_sweep[i];
            bt_accumulate(s, rt, k, g->X32, g->X64);
            SS_TRY(residual_norms(g, k, g->X64, g->B64, g->R64, rn2));
            st.refinements++;
        }
    }
    for (int i = 0; i < gk; i++) st.relres[i] = bn2[i] > 0.0 ? std::sqrt(rn2[i] / bn2[i]) : 0.0;
    return SS_OK;
}

// Median device time (ms) of reps runs of f on the graph's stream.
template <typename F>
static int time_median(ss_graph *g, int reps, F &&f, double &ms) {
    ss_context *c = g->ctx;
    std::vector<float> t;
    SS_TRY(f());   // warm
    for (int r = 0; r < reps; r++) {
        SS_HIP(hipEventRecord(c->ev[6], c->stream));
        SS_TRY(f());
        SS_HIP(hipEventRecord(c->ev[7], c->stream));
        SS_HIP(hipEventSynchronize(c->ev[7]));
        float e = 0.f;
        SS_HIP(hipEventElapsedTime(&e, c->ev[6], c->ev[7]));
        t.push_back(e);
    }
    std::sort(t.begin(), t.end());
    ms = t[t.size() / 2];
    return SS_OK;
}

template <typename T>
static int autotune_t(ss_graph *g, int k, const PcgWork<T> &w, int reps, ss_tuning_state &best, double &ms_level,
                      double &ms_tile) {
    hipStream_t s = g->ctx->stream;
    const bool prof = g->prof;
    g->prof = false;
    // a realistic right-hand side: the sketch, projected (FP64) and converted
    std::vector<GraphParams> prm((size_t)g->G);
    for (auto &p : prm) p.scale = 1.0 / std::sqrt((double)k);
    SS_TRY(graph_set_params(g, prm, k));
    bt_sketch(s, rtiles(g), g->iptr, g->ieid, g->ib, k, sketch_cols(g), 0, g->Y64);
    component_means(g, k, g->Y64);
    if (std::is_same<T, double>::value) {
        bt_project<double, double>(s, rtiles(g), k, g->Y64, g->comp_of_row, g->cmean, (double *)w.R, (double *)w.B,
                                   g->X64);
    } else {
        bt_project<double, double>(s, rtiles(g), k, g->Y64, g->comp_of_row, g->cmean, g->R64, g->B64, g->X64);
        bt_convert_masked(s, rtiles(g), k, g->R64, nullptr, (float *)w.R);
        // the mixed path's preconditioner (FP16 rows) is the one timed
        bt_tile_partial<float, float>(s, rtiles(g), k, (const float *)w.R, (const float *)w.R, nullptr, 0,
                                      g->partial);
        bt_prec_scale(s, g->G * k, k, g->gtile, g->partial, nullptr, g->row_off_dev, g->prec_sc, g->prec_inv);
    }
    const ss_tuning_state orig = g->ctx->tune;
    const int orig_waves = g->tile_waves;
    ms_level = 1e30;
    for (int xs : {0, 1})
        for (int rpw : {1, 2, 4}) {
            g->ctx->tune.level_rows_per_wave = rpw;
            g->ctx->tune.xcd_slices = xs;
            double ms = 0;
            SS_TRY(time_median(g, reps, [&]() {
                if constexpr (std::is_same<T, float>::value) {
                    if (k % 2 == 0) return precond_half(g, k, w);
                }
                return precond<T>(g, k, w.R, w.W, w.Zp, g->partial2);
            }, ms));
            if (ms < ms_level) {
                ms_level = ms;
                best.level_rows_per_wave = rpw;
                best.xcd_slices = xs;
            }
        }
    ms_tile = 1e30;
    launch_fill_u32(s, (int64_t)2 * g->G * k, reinterpret_cast<uint32_t *>(g->sc.beta), 0u);
    launch_fill_u32(s, (int64_t)2 * g->G * k, reinterpret_cast<uint32_t *>(g->sc.alpha), 0u);
    for (int tw : {2, 4, 8, 16}) {
        g->tile_waves = tw;
        const Tiles rt = rtiles(g);
        double ms = 0;
        SS_TRY(time_median(g, reps, [&]() {
            bt_update_px<T, T>(s, rt, k, g->sc.alpha, g->sc.beta, w.Zp, nullptr, w.P, w.B);
            bt_spmm_dot<T>(s, stiles(g), g->rowptr, g->col, avals<T>(g), k, w.P, w.Q, g->partial);
            bt_update_r_rr<T>(s, rt, k, g->sc.alpha, w.Q, w.W, g->partial);
            SS_HIP(hipGetLastError());
            return SS_OK;
        }, ms));
        if (ms < ms_tile) {
            ms_tile = ms;
            best.tile_waves = tw;
        }
    }
    // hipGraph replay is not a candidate: it never beat direct launches
    // (the solves are kernel-bound), and a capture of a deep factor holds
    // thousands of nodes (see pcg_run's node cap)
    best.graph_iters = 0;
    g->tile_waves = orig_waves;
    g->ctx->tune = orig;
    g->prof = prof;
    return SS_OK;
}

int graph_autotune(ss_graph *g, int k, int prec, int reps, ss_tuning_state &best, double &ms_level,
                   double &ms_tile) {
    if (!g->has_factor) return fail(SS_ERR_INVALID, "no factor uploaded");
    if (reps < 1 || reps > 100) return fail(SS_ERR_INVALID, "reps=%d outside 1..100", reps);
    SS_TRY(graph_ensure_work(g, k, prec));
    SS_HIP(hipSetDevice(g->ctx->device));
    SS_TRY(core_ensure_work(g, k, prec));
    best = g->ctx->tune;
    if (prec == SS_PREC_FP64) return autotune_t<double>(g, k, g->w64, reps, best, ms_level, ms_tile);
    return autotune_t<float>(g, k, g->w32, reps, best, ms_level, ms_tile);
}

int graph_set_params(ss_graph *g, const std::vector<GraphParams> &prm, int k) {
    if ((int)prm.size() != g->G) return fail(SS_ERR_INVALID, "need %d parameter sets, got %zu", g->G, prm.size());
    if (k < 1 || k > kMaxK) return fail(SS_ERR_INVALID, "k=%d outside 1..%d", k, kMaxK);
    const size_t gk = (size_t)g->G * k;
    std::vector<uint64_t> seeds(gk);
    std::vector<double> scales(gk), betas(g->G);
    std::vector<int32_t> js(gk);
    std::vector<uint32_t> keys((size_t)8 * g->G);
    for (int i = 0; i < g->G; i++) {
        for (int j = 0; j < k; j++) {
            seeds[(size_t)i * k + j] = prm[i].sketch_seed;
            scales[(size_t)i * k + j] = prm[i].scale;
            js[(size_t)i * k + j] = j;
        }
        betas[i] = prm[i].beta;
        const CoinKey ck = coin_key_from_u64(prm[i].sampling_seed);
        for (int q = 0; q < 8; q++) keys[(size_t)8 * i + q] = ck.k[q];
    }
    SS_HIP(hipSetDevice(g->ctx->device));
    hipStream_t s = g->ctx->stream;
    SS_TRY(h2d(s, g->p_colseed, seeds.data(), gk));
    SS_TRY(h2d(s, g->p_colscale, scales.data(), gk));
    SS_TRY(h2d(s, g->p_colj, js.data(), gk));
    SS_TRY(h2d(s, g->p_betas, betas.data(), betas.size()));
    SS_TRY(h2d(s, g->p_keys, keys.data(), keys.size()));
    // the host vectors go out of scope: wait for the copies
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

int graph_set_columns(ss_graph *g, int k, const std::vector<uint64_t> &seeds, const std::vector<double> &scales,
                      const std::vector<int32_t> &js) {
    if (g->G != 1) return fail(SS_ERR_INVALID, "per-column parameters need a batch of one");
    if (k < 1 || k > kMaxK || (int)seeds.size() != k || (int)scales.size() != k || (int)js.size() != k)
        return fail(SS_ERR_INVALID, "per-column parameters: k=%d", k);
    SS_HIP(hipSetDevice(g->ctx->device));
    hipStream_t s = g->ctx->stream;
    SS_TRY(h2d(s, g->p_colseed, seeds.data(), (size_t)k));
    SS_TRY(h2d(s, g->p_colscale, scales.data(), (size_t)k));
    SS_TRY(h2d(s, g->p_colj, js.data(), (size_t)k));
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

int graph_set_sampling(ss_graph *g, int i, double beta, uint64_t sampling_seed) {
    if (i < 0 || i >= g->G) return fail(SS_ERR_INVALID, "graph %d outside the batch of %d", i, g->G);
    const CoinKey ck = coin_key_from_u64(sampling_seed);
    SS_HIP(hipSetDevice(g->ctx->device));
    hipStream_t s = g->ctx->stream;
    SS_TRY(h2d(s, g->p_betas + i, &beta, 1));
    SS_TRY(h2d(s, g->p_keys + (size_t)8 * i, ck.k, 8));
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

int graph_sketch(ss_graph *g, int k, int uniform, int prec) {
    SS_TRY(graph_ensure_work(g, k, prec));
    SS_HIP(hipSetDevice(g->ctx->device));
    ProfScope ps(g, SS_KERNEL_SKETCH);
    // the signs' bits in X64 (the solve's output, free until the solve) when
    // they fit: ceil(k / 64) words per edge in N * cap doubles
    const size_t nw = ((size_t)k + 63) / 64;
    uint64_t *S = !uniform && g->M > 0 && (size_t)g->M * nw <= (size_t)g->N * (size_t)g->wcap
                      ? reinterpret_cast<uint64_t *>(g->X64)
                      : nullptr;
    bt_sketch(g->ctx->stream, rtiles(g), g->iptr, g->ieid, g->ib, k, sketch_cols(g), uniform, g->Y64,
              Tiles{g->et_start, g->et_graph, g->net, 4}, g->edge_off_dev, S);
    SS_HIP(hipGetLastError());
    return SS_OK;
}

int graph_sample(ss_graph *g, int k, const double *p_in_dev, bool write_d) {
    return graph_sample_cols(g, k, 0, k, p_in_dev, write_d, g->ku, g->kv, g->kw, g->kept_start, g->kept_count);
}

int graph_sample_cols(ss_graph *g, int k, int zoff, int ldz, const double *p_in_dev, bool write_d, int32_t *ku,
                      int32_t *kv, double *kw, std::vector<int64_t> &kept_start, std::vector<int64_t> &kept_count) {
    SS_HIP(hipSetDevice(g->ctx->device));
    hipStream_t s = g->ctx->stream;
    kept_start.assign((size_t)g->G, 0);
    kept_count.assign((size_t)g->G, 0);
    if (g->M == 0) return SS_OK;
    if (!p_in_dev && (g->wk != ldz || zoff < 0 || k < 1 || zoff + k > ldz))
        return fail(SS_ERR_INVALID, "no solution columns [%d, %d) in a block of %d", zoff, zoff + k, g->wk);
    if (g->samp_epoch >= (1u << 24) - 2) {
        SS_HIP(hipMemsetAsync(g->tile_state, 0, sizeof(uint64_t) * ((size_t)g->net + 1), s));
        g->samp_epoch = 0;
    }
    g->samp_epoch++;
    launch_fill_u32(s, 1, g->ticket, 0u);
    BatchSampleArgs a;
    a.m = g->M;
    a.et = Tiles{g->et_start, g->et_graph, g->net, 4};
    a.edge_off = g->edge_off_dev;
    a.eu = g->eu;
    a.ev = g->ev;
    a.ew = g->ew;
    a.Z = g->X64 + zoff;
    a.k = k;
    a.ldz = ldz;
    a.beta = g->p_betas;
    a.keys = g->p_keys;
    a.p_in = p_in_dev;
    a.p_out = g->p;
    a.d_out = write_d ? g->dnorm : nullptr;
    a.mask = g->mask;
    a.ku = ku;
    a.kv = kv;
    a.kw = kw;
    a.ticket = g->ticket;
    a.tile_state = g->tile_state;
    a.tile_incl = g->tile_incl;
    a.epoch = g->samp_epoch;
    a.err = g->status + 1;
    a.spin_limit = kSpinLimit;
    {
        ProfScope ps(g, SS_KERNEL_ER_SAMPLE);
        bt_er_sample(s, a, g->net);
    }
    SS_HIP(hipGetLastError());
    std::vector<int64_t> incl((size_t)g->net);
    SS_HIP(hipMemcpyAsync(incl.data(), g->tile_incl, sizeof(int64_t) * g->net, hipMemcpyDeviceToHost, s));
    SS_TRY(read_status(g, nullptr));
    for (int i = 0; i < g->G; i++) {
        const int32_t t0 = g->etile_off[i], t1 = g->etile_off[i + 1];
        const int64_t start = t0 > 0 ? incl[t0 - 1] : 0;
        const int64_t end = t1 > t0 ? incl[t1 - 1] : start;
        kept_start[i] = start;
        kept_count[i] = end - start;
    }
    return SS_OK;
}

int graph_apply_reweight(ss_graph *g, double *diag_out) {
    if (g->G != 1) return fail(SS_ERR_INVALID, "apply_reweight takes a single graph");
    if (!diag_out) return fail(SS_ERR_INVALID, "null output");
    SS_HIP(hipSetDevice(g->ctx->device));
    hipStream_t s = g->ctx->stream;
    double *dn = nullptr;
    DevArena tmp;
    SS_TRY(tmp.alloc(&dn, (size_t)g->N));
    bt_apply_diag(s, (int32_t)g->N, g->rowptr, g->col, g->val, g->iptr, g->ieid, g->ew, g->p, g->mask, dn);
    SS_HIP(hipGetLastError());
    SS_HIP(hipMemcpyAsync(diag_out, dn, sizeof(double) * (size_t)g->N, hipMemcpyDeviceToHost, s));
    SS_HIP(hipStreamSynchronize(s));
    return SS_OK;
}

}  // namespace ss
