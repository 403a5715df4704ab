// Device-resident state: context (device + stream) and graph.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.h"
#include "ss_internal.h"

struct ss_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    int *h_status = nullptr;      // pinned: {nactive, err}
};

namespace ss {

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

// Work blocks of one PCG run in precision T (n x k each).
template <typename T>
struct PcgWork {
    T *B = nullptr, *R = nullptr, *P = nullptr, *P2 = nullptr, *Q = nullptr, *Zp = nullptr, *W = nullptr;
};

}  // namespace ss

struct ss_graph {
    ss_context *ctx = nullptr;
    int64_t n = 0, m = 0;
    ss::DevArena arena_struct, arena_factor, arena_work;
    // canonical edges
    int32_t *eu = nullptr, *ev = nullptr;
    double *ew = nullptr;
    // Laplacian CSR (with diagonal), FP64 and FP32 values
    int32_t *rowptr = nullptr, *col = nullptr;
    double *val = nullptr;
    float *valf = nullptr;
    // incidence lists for the sketch: per node, (edge id, +-sqrt(w)) by edge
    int32_t *iptr = nullptr, *ieid = nullptr;
    double *ib = nullptr;
    // factor schedules
    bool has_factor = false;
    int64_t factor_nnz = 0;
    int32_t depth_fwd = 0, depth_bwd = 0;
    int32_t *f_rows = nullptr, *f_ptr = nullptr, *f_col = nullptr;
    int32_t *b_rows = nullptr, *b_ptr = nullptr, *b_col = nullptr;
    double *f_val = nullptr, *b_val = nullptr, *diag = nullptr;
    float *f_valf = nullptr, *b_valf = nullptr;
    int64_t f_nnz = 0, b_nnz = 0;
    std::vector<int32_t> f_level_ptr, b_level_ptr;   // host copies: launch bounds per level
    std::vector<int32_t> f_level_long, b_level_long; // long rows leading each level
    int32_t f_small_lo = 0, f_small_hi = 0, b_small_lo = 0, b_small_hi = 0;
    int32_t *f_lp_dev = nullptr, *b_lp_dev = nullptr, *f_split = nullptr;
    ss_schedule_info sched_info[2] = {};
    // sync state
    uint32_t *flags = nullptr;
    uint32_t epoch = 0;
    int *status = nullptr;          // {nactive, err}
    // work (allocated for wk columns at wprec)
    int wk = 0, wprec = -1;
    double *Y64 = nullptr, *X64 = nullptr, *R64 = nullptr, *B64 = nullptr;
    float *Y32 = nullptr, *X32 = nullptr;
    ss::PcgWork<double> w64;
    ss::PcgWork<float> w32;
    double *partial = nullptr, *partial2 = nullptr, *colbuf = nullptr;
    double *sc_buf = nullptr;
    int *sc_int = nullptr;
    ss::PcgScalars sc{};
    // sampling
    double *p = nullptr, *dnorm = nullptr;
    uint8_t *mask = nullptr;
    int32_t *ku = nullptr, *kv = nullptr;
    double *kw = nullptr;
    uint64_t *tile_state = nullptr;
    uint32_t *ticket = nullptr;
    int64_t *kept = nullptr;
    uint32_t samp_epoch = 0;
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
    int64_t last_kept = 0;
    int last_k = 0;
    int last_prec = 0;
};

namespace ss {
// Records HIP events around one launch on the graph's stream when profiling.
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
int graph_upload(ss_context *ctx, const HostGraph &G, ss_graph **out);
int graph_ensure_work(ss_graph *g, int k, int prec);
int graph_upload_factor(ss_graph *g, const Factor &F);
struct SolveStats {
    int32_t max_iters = 0;
    int32_t refinements = 0;
    double max_relres = 0.0;
    std::vector<int32_t> iters;
    std::vector<double> relres;
};
// Solve L Z = Y - mean for the graph's resident sketch (Y64 or Y32); the
// FP64 solution lands in X64.
int graph_solve(ss_graph *g, int k, const ss_solve_opts &o, SolveStats &st);
int graph_sketch(ss_graph *g, int k, uint64_t seed, double scale, int uniform, int prec);
int graph_sample(ss_graph *g, int k, double beta, uint64_t sampling_seed, const double *p_in_dev,
                 bool write_d);
}  // namespace ss
