// Device-resident state: context (device + stream) and the batch of graphs
// (a single graph is a batch of one).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "batch_kernels.h"
#include "kernels.h"
#include "ss_internal.h"

// Launch shapes picked by ss_context_autotune (see ss_tuning in the C ABI).
struct ss_tuning_state {
    int tile_waves = 4;          // wavefronts per row-tile block (1, 2, 4, 8, 16)
    int level_rows_per_wave = 1; // rows per wavefront of the short rows in the per-level solve launches
    int graph_iters = 0;         // PCG iterations per captured hipGraph replay (0 = direct launches)
    int xcd_slices = 0;          // per-level launches: each XCD takes one contiguous slice of a level's rows
};

struct ss_context {
    int device = 0;
    ss_tuning_state tune;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    int *h_status = nullptr;      // pinned: {nactive, err, flagged}
    int64_t *h_noalloc = nullptr; // pinned: {kept, err} of ss_er_sample_noalloc
    int live_graphs = 0;          // graphs created on this context and not yet destroyed
};

namespace ss {

// Device address of the allocation-free sampling scratch array, once per
// device with that device current (noalloc.hip).
int noalloc_setup(int device);

// A set of device allocations released together.
struct DevArena {
    std::vector<void *> ptrs;
    size_t bytes = 0;
    template <typename T>
    int alloc(T **p, size_t count) {
        void *q = nullptr;
        size_t b = count * sizeof(T);
        if (b == 0) b = sizeof(T);
        hipError_t e = hipMalloc(&q, b);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(SS_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", b, hipGetErrorString(e));
        }
        ptrs.push_back(q);
        bytes += b;
        *p = (T *)q;
        return SS_OK;
    }
    void release() {
        for (void *q : ptrs) (void)hipFree(q);
        ptrs.clear();
        bytes = 0;
    }
    ~DevArena() { release(); }
};

// What a captured PCG iteration graph bakes in; a different key recaptures.
struct GraphKey {
    int k = 0, rule = 0;
    double tol = 0.0;
    int iters = 0;
    const void *X = nullptr, *R = nullptr;
    uint64_t factor_epoch = 0;
    int rpw = 0, tile_waves = 0;
    bool operator==(const GraphKey &o) const {
        return k == o.k && rule == o.rule && tol == o.tol && iters == o.iters && X == o.X && R == o.R &&
               factor_epoch == o.factor_epoch && rpw == o.rpw && tile_waves == o.tile_waves;
    }
};
struct GraphCache {
    GraphKey key;
    hipGraphExec_t exec = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    ~GraphCache() { reset(); }
};

// Work blocks of one PCG run in precision T (N x k each).
template <typename T>
struct PcgWork {
    T *B = nullptr, *R = nullptr, *P = nullptr, *Q = nullptr, *Zp = nullptr, *W = nullptr;
};

}  // namespace ss

struct ss_graph {
    ss_context *ctx = nullptr;
    int G = 0;                       // graphs in the batch
    int tile_waves = 4;              // fixed at creation: the reductions' shape depends on it
    int64_t N = 0, M = 0;            // total rows / edges
    std::vector<int64_t> row_off, edge_off;       // host, G+1
    std::vector<int32_t> rtile_off, etile_off;    // host, G+1: tile ranges per graph
    int nrt = 0, net = 0;
    ss::DevArena arena_struct, arena_factor, arena_work, arena_params;
    // tiles
    int32_t *rt_start = nullptr, *rt_graph = nullptr, *et_start = nullptr, *et_graph = nullptr;
    int32_t *gtile = nullptr;        // device copy of rtile_off
    // SpMM tiles: each row tile split into up to kSpmmMaxSplit tiles so that
    // none holds much more than kSpmmTileNnz matrix entries (a power-law
    // graph's hub rows sit in a few row tiles); per graph ranges on the host
    // (stile_off) and the device (sgtile)
    std::vector<int32_t> stile_off;
    int nst = 0;
    int32_t *st_start = nullptr, *st_graph = nullptr, *sgtile = nullptr;
    int32_t *edge_off_dev = nullptr; // device copy of edge_off (int32)
    int32_t *row_off_dev = nullptr;  // device copy of row_off (int32)
    int32_t *row_graph = nullptr;    // per row: its graph (batches of more than one graph; else null)
    // canonical edges (global vertex ids)
    int32_t *eu = nullptr, *ev = nullptr;
    double *ew = nullptr;
    // block-diagonal Laplacian CSR (with diagonal), FP64 and FP32 values
    int32_t *rowptr = nullptr, *col = nullptr;
    double *val = nullptr;
    float *valf = nullptr;
    // incidence lists for the sketch: per node, (local edge id, +-sqrt(w))
    int32_t *iptr = nullptr, *ieid = nullptr;
    double *ib = nullptr;
    // connected components: comp_of_row, per comp graph / vertex list;
    // per graph the main (largest) component comes first
    int ncomp = 0, nsmall = 0;
    int32_t *comp_of_row = nullptr, *comp_graph = nullptr, *comp_off = nullptr, *cptr = nullptr, *cvert = nullptr,
            *small_id = nullptr;
    std::vector<int32_t> h_comp_count;   // components per graph
    // factor schedules (of the concatenated factor)
    bool has_factor = false;
    uint64_t factor_epoch = 0;             // bumped on every factor upload
    ss::GraphCache gc64, gc32;             // captured PCG iterations (FP64 / FP32 work)
    int64_t factor_nnz = 0;
    int32_t depth_fwd = 0, depth_bwd = 0;
    int32_t *f_rows = nullptr, *f_ptr = nullptr, *f_col = nullptr;
    int32_t *b_rows = nullptr, *b_ptr = nullptr, *b_col = nullptr;
    double *f_val = nullptr, *b_val = nullptr, *diag = nullptr;
    float *f_valf = nullptr, *b_valf = nullptr;
    std::vector<int32_t> f_level_ptr, b_level_ptr;   // host copies: launch bounds per level
    std::vector<int32_t> f_level_long, b_level_long; // long rows leading each level
    std::vector<int32_t> f_level_med, b_level_med;   // then rows taking one wave each
    std::vector<int32_t> f_level_huge, b_level_huge; // the first long rows: more than kWgRowMax entries
    std::vector<int32_t> f_level_hm, b_level_hm;     // workgroups per such row (the level's most)
    int32_t core_huge = 0;                           // the same for the forward core rows
    double *huge_scratch = nullptr;   // per long row of more than kWgRowMax entries: kMaxWg subtrees x kMaxK values
    int *huge_counter = nullptr;      // per such row: workgroups arrived (zero between launches)
    std::vector<int32_t> f_level_waves, b_level_waves;   // wavefronts per block of each level launch
    std::vector<std::pair<int32_t, int32_t>> f_runs, b_runs; // runs of small levels (single-workgroup launches)
    int32_t *f_lp_dev = nullptr, *b_lp_dev = nullptr;   // level_ptr on the device
    // dense cores: per graph, the last rows of its elimination order (its
    // factor's dense tail, else its trailing small forward levels), applied
    // as the inverse of their block of L (core_kernels.hip)
    int32_t f_core_t0 = 0, f_nrows = 0;   // forward positions [f_core_t0, f_nrows): core rows' outside entries
    int core_max = 0;                     // largest core (rows)
    int64_t core_total = 0;               // core rows of the batch
    int64_t core_entries = 0;             // entries of the padded inverses (sum of cp^2)
    ss::CoreApply core;                   // device arrays of the core products
    double *Md = nullptr;                 // inverses, FP64, FP32 and FP16 (the mixed path's FP16 rows)
    float *Mf = nullptr;
    _Float16 *Mh = nullptr;
    ss::DevArena arena_core_work;         // gathered right-hand sides and product pieces
    int core_wcap = 0, core_wprec = -1;   // columns / precision they are sized for
    void *core_tc = nullptr, *core_part = nullptr;
    ss_schedule_info sched_info[2] = {};
    // status
    int *status = nullptr;           // {nactive, err, flagged (reliable updates)}
    // work (allocated for wk columns at wprec)
    int wk = 0, wcap = 0, wprec = -1;   // columns in use, columns allocated, precision
    double *Y64 = nullptr, *X64 = nullptr, *R64 = nullptr, *B64 = nullptr;
    float *Y32 = nullptr, *X32 = nullptr;
    ss::PcgWork<double> w64;
    ss::PcgWork<float> w32;
    double *partial = nullptr, *partial2 = nullptr;
    double *gsum = nullptr, *csum = nullptr, *cmean = nullptr;
    double *sc_buf = nullptr;
    int *sc_int = nullptr, *colkeep = nullptr;
    // mixed precision with FP16 preconditioner rows: per (graph, column)
    // the rows' scale s (G x k), then 1 / s (G x k)
    float *prec_sc = nullptr, *prec_inv = nullptr;
    ss::PcgScalars sc{};
    double *rel_ref = nullptr;       // per (graph, column): |r|^2 at the last reliable update
    int *rel_flag = nullptr;         // per (graph, column): replacement due
    int32_t *rel_gs = nullptr, *rel_ge = nullptr;   // per column: its replacement group [gs, ge)
    // per-call parameters (per graph)
    uint64_t *p_colseed = nullptr;   // per (graph, column), G x kMaxK: sketch seed,
    double *p_colscale = nullptr;    // scale,
    int32_t *p_colj = nullptr;       // and the column index the signs hash
    double *p_betas = nullptr;
    uint32_t *p_keys = nullptr;
    // sampling
    double *p = nullptr, *dnorm = nullptr;
    uint8_t *mask = nullptr;
    int32_t *ku = nullptr, *kv = nullptr;
    double *kw = nullptr;
    uint64_t *tile_state = nullptr;
    int64_t *tile_incl = nullptr;
    uint32_t *ticket = nullptr;
    uint32_t samp_epoch = 0;
    std::vector<int64_t> kept_start, kept_count;   // per graph, last sample
    // several requests sampled from one run (ss_graph_sparsify_many):
    // request r's kept edges at many_k*[r * M ..], many_kept[r] of them
    ss::DevArena arena_many;
    int many_cap = 0;
    int32_t *many_ku = nullptr, *many_kv = nullptr;
    double *many_kw = nullptr;
    std::vector<int64_t> many_kept;
    // per-kernel HIP-event timing (ss_graph_profile)
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    struct ProfRec {
        int kind, launches;
        hipEvent_t a, b;
    };
    std::vector<ProfRec> prof_recs;
    double prof_ms[SS_KERNEL_COUNT] = {};
    int64_t prof_cnt[SS_KERNEL_COUNT] = {};
    int last_k = 0;
};

namespace ss {
// A scope covering `launches` kernel launches of one kind.
struct ProfScope {
    ss_graph *g;
    int kind;
    int launches;
    hipEvent_t a = nullptr;
    ProfScope(ss_graph *g_, int kind_, int launches_ = 1);
    ~ProfScope();
};
int prof_resolve(ss_graph *g);

// Incidence lists of B for the sketch: per vertex, (edge id, B entry).
void build_incidence(int64_t n, int64_t m, const int32_t *v1, const int32_t *v2, const double *w,
                     std::vector<int32_t> &iptr, std::vector<int32_t> &ieid, std::vector<double> &ib);
// Upload a batch of graphs (host structures built by graph_build).
int batch_upload(ss_context *ctx, const std::vector<const HostGraph *> &graphs, ss_graph **out);
// Concatenate the graphs' factors into one schedule and upload it.
int batch_upload_factors(ss_graph *g, const std::vector<const Factor *> &factors);
int graph_ensure_work(ss_graph *g, int k, int prec);
// Device bytes graph_ensure_work(g, k, prec) would allocate (0 if none).
size_t graph_work_bytes(const ss_graph *g, int k, int prec);

struct SolveStats {
    std::vector<int32_t> iters;      // [G*k]
    std::vector<double> relres;      // [G*k]
    int32_t refinements = 0;
};
// Solve L Z = Y - (component means) for every graph; Y64 must hold the
// right-hand sides; the FP64 solution lands in X64.
// groups (mixed precision, reliable updates): k non-decreasing group ids
// of the columns (the requests of a sparsify pass), whose columns replace
// their residuals together; NULL: every column on its own.
int graph_solve(ss_graph *g, int k, const ss_solve_opts &o, SolveStats &st,
                const std::vector<int32_t> *groups = nullptr);
// Per-graph parameters of the next sketch / sample.
struct GraphParams {
    uint64_t sketch_seed = 0, sampling_seed = 0;
    double scale = 1.0, beta = 1.0;
};
// Per-graph parameters for k columns (column j of graph i hashes index j
// with graph i's sketch seed and scale).
int graph_set_params(ss_graph *g, const std::vector<GraphParams> &prm, int k);
// A batch of one with several requests side by side: column c hashes
// index js[c] with seeds[c] and scale scales[c].
int graph_set_columns(ss_graph *g, int k, const std::vector<uint64_t> &seeds, const std::vector<double> &scales,
                      const std::vector<int32_t> &js);
// Sampling parameters of graph i: beta and the coin stream's seed.
int graph_set_sampling(ss_graph *g, int i, double beta, uint64_t sampling_seed);
int graph_sketch(ss_graph *g, int k, int uniform, int prec);
// Time launch-shape candidates on g (see ss_context_autotune).
int graph_autotune(ss_graph *g, int k, int prec, int reps, ss_tuning_state &best, double &ms_level,
                   double &ms_tile);
int graph_sample(ss_graph *g, int k, const double *p_in_dev, bool write_d);
// The last sample applied to the Laplacian as apply_reweightings does
// (src/sparsifier.rs:682-691): the new diagonal, old diagonal plus each
// vertex's changes summed in edge order, into diag_out (host, n values).
int graph_apply_reweight(ss_graph *g, double *diag_out);
// Sampling from columns [zoff, zoff + k) of the n x ldz solution block into
// the given kept arrays (capacity M); per graph start and count.
int graph_sample_cols(ss_graph *g, int k, int zoff, int ldz, const double *p_in_dev, bool write_d, int32_t *ku,
                      int32_t *kv, double *kw, std::vector<int64_t> &kept_start, std::vector<int64_t> &kept_count);
}  // namespace ss
