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
#include "solve_shape.h"
#include "host_pool.h"

#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstring>
#include <sys/mman.h>

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
    // The value the (i+1)-th call of uniform() returns, without the calls
    // before it (the stream is counter-based).
    double at(uint64_t i) const {
        return (double)(mix64(s + (i + 1) * 0x9e3779b97f4a7c15ULL) >> 11) * (1.0 / 9007199254740992.0);
    }
};

struct Ent {
    int32_t row;
    double val;
};

// A large zero-initialised array from anonymous memory (pages fault in as
// they are first touched, by whichever thread touches them), backed by huge
// pages where the system allows: the dense tail's random updates then miss
// the TLB far less often.
template <typename T>
class ZeroedBuffer {
public:
    explicit ZeroedBuffer(size_t n) : bytes_(n * sizeof(T)) {
        void *p = mmap(nullptr, bytes_ ? bytes_ : 1, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
        if (p == MAP_FAILED) return;
        (void)madvise(p, bytes_, MADV_HUGEPAGE);
        p_ = static_cast<T *>(p);
    }
    ~ZeroedBuffer() {
        if (p_) munmap(p_, bytes_ ? bytes_ : 1);
    }
    ZeroedBuffer(const ZeroedBuffer &) = delete;
    ZeroedBuffer &operator=(const ZeroedBuffer &) = delete;
    T *data() const { return p_; }

private:
    size_t bytes_;
    T *p_ = nullptr;
};

struct Fill {
    int32_t row;
    int32_t next;   // index into the fill pool, -1 ends the list
    double val;
};

}  // namespace

// Eliminate one vertex whose live lower neighbours (row, |L_rc| + fill) are
// in `work`, distinct rows in ascending order: sort by value, sample one
// fill edge per entry but the last, scale.  Fill edges go to emit(a, b,
// value); the column of L (rows ascending, values -v/tot) to `col`.
// Returns D = tot (0 if empty).
template <typename Emit>
static double eliminate_merged(std::vector<Ent> &work, std::vector<double> &cum, ColumnRng &rng,
                               std::vector<Ent> &col, Emit emit) {
    const size_t cnt = work.size();
    col.assign(work.begin(), work.end());   // rows ascending
    if (cnt == 0) return 0.0;
    std::sort(work.begin(), work.end(), [](const Ent &a, const Ent &b) {
        return a.val != b.val ? a.val < b.val : a.row < b.row;
    });
    cum.resize(cnt);
    double run = 0.0;
    for (size_t i = 0; i < cnt; i++) { run += work[i].val; cum[i] = run; }
    const double tot = run;
    for (size_t i = 0; i + 1 < cnt; i++) {
        const double rest = tot - cum[i];
        const double target = rng.uniform() * rest;
        // first j > i with cum[j] - cum[i] >= target (std::lower_bound in the
        // reference, src/example.cc:266-269)
        size_t lo = i + 1, hi = cnt;
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if (cum[mid] - cum[i] < target) lo = mid + 1; else hi = mid;
        }
        const size_t j = lo < cnt ? lo : cnt - 1;
        emit(work[i].row, work[j].row, work[i].val * rest / tot);
    }
    for (Ent &e : col) e.val = -(e.val / tot);
    return tot;
}

// The same for a work list that may repeat rows: equal rows are merged
// first, their values summed in ascending order.
template <typename Emit>
static double eliminate(std::vector<Ent> &work, std::vector<double> &cum, ColumnRng &rng, std::vector<Ent> &col,
                        Emit emit) {
    std::sort(work.begin(), work.end(), [](const Ent &a, const Ent &b) {
        return a.row != b.row ? a.row < b.row : a.val < b.val;
    });
    size_t cnt = 0;
    for (size_t i = 0; i < work.size(); i++) {
        if (cnt > 0 && work[cnt - 1].row == work[i].row) work[cnt - 1].val += work[i].val;
        else work[cnt++] = work[i];
    }
    work.resize(cnt);
    return eliminate_merged(work, cum, rng, col, emit);
}

// Natural order on the given CSC (positions == indices of the CSC).
static int factor_natural(int64_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals,
                          uint64_t seed, std::vector<int64_t> &cstart, std::vector<int32_t> &crow,
                          std::vector<double> &cval, std::vector<double> &diag) {
    std::vector<int32_t> head((size_t)n, -1);
    std::vector<Fill> pool;
    pool.reserve((size_t)colptr[n]);
    cstart.assign((size_t)n + 1, 0);
    crow.clear();
    cval.clear();
    crow.reserve((size_t)colptr[n]);
    cval.reserve((size_t)colptr[n]);
    diag.assign((size_t)n, 0.0);
    std::vector<Ent> work, col;
    std::vector<double> cum;
    bool overflow = false;
    for (int64_t c = 0; c < n; c++) {
        work.clear();
        for (int32_t t = colptr[c]; t < colptr[c + 1]; t++) {
            int32_t r = rowidx[t];
            if (r < 0 || r >= n) return fail(SS_ERR_INVALID, "factor: row index %d out of range", r);
            if (r > c) work.push_back({r, std::fabs(vals[t])});
        }
        for (int32_t f = head[c]; f >= 0; f = pool[f].next) work.push_back({pool[f].row, pool[f].val});
        head[c] = -1;
        ColumnRng rng(seed, c);
        // fill edge (max, min) goes to column min (src/example.cc:270-286)
        diag[c] = eliminate(work, cum, rng, col, [&](int32_t a, int32_t b, double fv) {
            if (pool.size() >= (size_t)INT32_MAX) { overflow = true; return; }
            const int32_t fr = std::max(a, b), fc = std::min(a, b);
            pool.push_back({fr, head[fc], fv});
            head[fc] = (int32_t)(pool.size() - 1);
        });
        if (overflow) return fail(SS_ERR_CAPACITY, "factor: fill pool exceeds 2^31 entries");
        for (const Ent &e : col) { crow.push_back(e.row); cval.push_back(e.val); }
        cstart[c + 1] = (int64_t)crow.size();
    }
    return SS_OK;
}

// Rounds: every round eliminates an independent set of the current
// fill-augmented graph, picked greedily in (live degree, random) order among
// the vertices of degree at most twice the round's minimum.  A round's
// vertices never depend on each other, so every round is one level of both
// triangular solves; low degree first keeps fill (and PCG iterations) at or
// below the natural order's.
//
// Dense tail: once at most kTailMaxRows vertices are live and a round takes
// fewer than 1/kTailFraction of them (the remaining graph is nearly dense:
// the hubs of a power-law graph, a grid's last separators), the rest are
// eliminated one at a time in (live degree, random) order.  They form the
// trailing block of the factor, which the device applies as a dense
// inverse instead of level by level.
//
// Parallel and deterministic: every step of a round either works on
// disjoint data per vertex (adjacency compaction and keys, eliminations of
// the independent set) or is a fixed sequential rule (candidate sort, the
// greedy pick).  Fill edges are bucketed by (producing thread, owner of the
// target vertex) and appended by the owner; the order inside an adjacency
// list never matters (eliminate() sorts by (row, value) and merges equal
// rows), so the factor is the same for every thread count.
static int factor_rounds(int64_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals,
                         uint64_t seed, int threads, std::vector<int32_t> &order, std::vector<int64_t> &vstart,
                         std::vector<int32_t> &crow, std::vector<double> &cval, std::vector<double> &vdiag,
                         int32_t &rounds, int64_t &tail_rows) {
    for (int64_t t = 0; t < colptr[n]; t++)
        if (rowidx[t] < 0 || rowidx[t] >= n) return fail(SS_ERR_INVALID, "factor: row index %d out of range", rowidx[t]);
    HostPool pool(threads);
    const int T = pool.size();
    std::vector<std::vector<Ent>> adj((size_t)n);
    pool.run([&](int t) {
        int64_t lo, hi;
        HostPool::range(n, T, t, lo, hi);
        for (int64_t c = lo; c < hi; c++) {
            adj[c].reserve((size_t)(colptr[c + 1] - colptr[c]));
            for (int32_t q = colptr[c]; q < colptr[c + 1]; q++)
                if (rowidx[q] != c) adj[c].push_back({rowidx[q], std::fabs(vals[q])});
        }
    });
    std::vector<char> elim((size_t)n, 0);
    std::vector<uint64_t> key((size_t)n);
    std::vector<int32_t> live((size_t)n), sel, cand;
    std::vector<int32_t> blocked((size_t)n, 0);   // == round + 1: a neighbour was picked
    for (int64_t i = 0; i < n; i++) live[i] = (int32_t)i;
    order.clear();
    order.reserve((size_t)n);
    vstart.assign((size_t)n, 0);
    std::vector<int64_t> vcount((size_t)n, 0);
    crow.clear();
    cval.clear();
    // the factor holds about 2-3x the Laplacian's entries: room reserved up
    // front (untouched pages cost nothing) spares the copies of regrowth
    crow.reserve((size_t)colptr[n] * 3);
    cval.reserve((size_t)colptr[n] * 3);
    vdiag.assign((size_t)n, 0.0);
    rounds = 0;
    tail_rows = 0;
    // per thread: scratch of eliminate(), the columns it produced (in sel
    // order), fill edges bucketed by owner thread, candidate lists
    struct Fill3 {
        int32_t target, other;
        double val;
    };
    struct PerThread {
        std::vector<Ent> work, col, cols;
        std::vector<double> cum;
        std::vector<int64_t> ccount;
        std::vector<std::vector<Fill3>> fills;
        std::vector<int32_t> cand, next;
        uint64_t mindeg = ~0ULL, sumdeg = 0;
    };
    std::vector<PerThread> pt((size_t)T);
    for (auto &x : pt) x.fills.resize((size_t)T);
    auto owner = [T](int32_t v) { return (int)(((uint32_t)v >> 4) % (uint32_t)T); };
    // SPEC_SPARS_VERBOSE set: a progress line on stderr about every 20 s
    const char *venv = std::getenv("SPEC_SPARS_VERBOSE");
    const bool verbose = venv != nullptr, trace = venv && std::atoi(venv) >= 2;
    auto last_report = std::chrono::steady_clock::now();
    const auto t_start = last_report;
    int64_t last_sel = (int64_t)n;
    std::vector<int64_t> lsplit;
    // SPEC_SPARS_VERBOSE >= 2: seconds per step kind, summed over rounds
    double st[7] = {0, 0, 0, 0, 0, 0, 0};
    auto lap = [&](int i, std::chrono::steady_clock::time_point &a) {
        if (!trace) return;
        const auto b = std::chrono::steady_clock::now();
        st[i] += std::chrono::duration<double>(b - a).count();
        a = b;
    };
    while (!live.empty()) {
        auto tl = trace ? std::chrono::steady_clock::now() : std::chrono::steady_clock::time_point();
        if (verbose && std::chrono::steady_clock::now() - last_report > std::chrono::seconds(20)) {
            last_report = std::chrono::steady_clock::now();
            std::fprintf(stderr, "[factor] round %d: %zu of %lld vertices left, %zu factor entries\n", rounds,
                         live.size(), (long long)n, crow.size());
            std::fflush(stderr);
        }
        const int64_t L = (int64_t)live.size();
        const int active = L >= 4096 ? T : 1;
        const uint64_t rkey = mix64(seed ^ (0x94d049bb133111ebULL * (uint64_t)(rounds + 1)));
        // 1. drop eliminated neighbours, key = (live degree, random); the
        // live set is split by list lengths, not by count (the hubs, which
        // hold most entries, have low ids)
        lsplit.assign((size_t)active + 1, L);
        lsplit[0] = 0;
        if (active > 1) {
            int64_t total = 0;
            for (int64_t x = 0; x < L; x++) total += 1 + (int64_t)adj[live[x]].size();
            int64_t run = 0;
            int t = 1;
            for (int64_t x = 0; x < L && t < active; x++) {
                run += 1 + (int64_t)adj[live[x]].size();
                while (t < active && run >= total * t / active) lsplit[(size_t)t++] = x + 1;
            }
        }
        pool.run([&](int t) {
            const int64_t lo = lsplit[(size_t)t], hi = lsplit[(size_t)t + 1];
            uint64_t md = ~0ULL, sd = 0;
            for (int64_t x = lo; x < hi; x++) {
                const int32_t v = live[x];
                auto &a = adj[v];
                size_t w = 0;
                for (const Ent &e : a)
                    if (!elim[e.row]) a[w++] = e;
                a.resize(w);
                const uint64_t deg = std::min<uint64_t>(w, 0xffffff);
                key[v] = (deg << 40) | (mix64(rkey ^ (uint64_t)v) & ((1ULL << 40) - 1));
                md = std::min(md, deg);
                sd += deg;
            }
            pt[t].mindeg = md;
            pt[t].sumdeg = sd;
        }, active);
        lap(0, tl);
        uint64_t mindeg = ~0ULL, sumdeg = 0;
        for (int t = 0; t < active; t++) {
            mindeg = std::min(mindeg, pt[t].mindeg);
            sumdeg += pt[t].sumdeg;
        }
        if (trace && L <= 2 * kTailMaxRows)
            std::fprintf(stderr, "[factor] round %d live %lld last_sel %lld mindeg %llu avgdeg %.1f\n", rounds,
                         (long long)L, (long long)last_sel, (unsigned long long)mindeg, (double)sumdeg / (double)L);
        // dense enough: the live graph's mean degree is at least 1/kTailDensity of its size
        if (L <= kTailMaxRows && last_sel * kTailFraction < L && sumdeg * kTailDensity >= (uint64_t)(L * L)) {
            // dense tail: the rest one at a time in key order
            const auto t_tail = std::chrono::steady_clock::now();
            if (verbose) {
                std::fprintf(stderr, "[factor] rounds done: %.1f s\n",
                             std::chrono::duration<double>(t_tail - t_start).count());
                std::fflush(stderr);
            }
            // tail ids in vertex order; every live list sorted by (row,
            // value) and merged once, so the sequential eliminations below
            // see lists in an order independent of the thread count
            std::vector<int32_t> tid((size_t)n, -1);
            for (int64_t x = 0; x < L; x++) tid[live[x]] = (int32_t)x;   // live is ascending
            pool.run([&](int t) {
                int64_t lo, hi;
                HostPool::range(L, active, t, lo, hi);
                for (int64_t x = lo; x < hi; x++) {
                    auto &a = adj[live[x]];
                    std::sort(a.begin(), a.end(), [](const Ent &p, const Ent &q) {
                        return p.row != q.row ? p.row < q.row : p.val < q.val;
                    });
                    size_t c = 0;
                    for (size_t i = 0; i < a.size(); i++) {
                        if (c > 0 && a[c - 1].row == a[i].row) a[c - 1].val += a[i].val;
                        else a[c++] = a[i];
                    }
                    a.resize(c);
                }
            }, active);
            cand.assign(live.begin(), live.end());
            std::sort(cand.begin(), cand.end(),
                      [&](int32_t x, int32_t y) { return key[x] != key[y] ? key[x] < key[y] : x < y; });
            // the tail graph as a dense symmetric matrix over tail ids (in
            // vertex order): its live entries accumulate fills in
            // elimination order, and a row scan lists a vertex's live
            // neighbours with rows ascending
            ZeroedBuffer<double> W((size_t)L * (size_t)L);
            if (!W.data()) return fail(SS_ERR_NOMEM, "factor: dense tail of %lld vertices", (long long)L);
            pool.run([&](int t) {
                int64_t lo, hi;
                HostPool::range(L, active, t, lo, hi);
                for (int64_t x = lo; x < hi; x++) {
                    double *row = W.data() + (size_t)x * L;
                    for (const Ent &e : adj[live[x]]) row[tid[e.row]] = e.val;
                    std::vector<Ent>().swap(adj[live[x]]);
                }
            }, active);
            std::vector<char> tdone((size_t)L, 0);
            // One elimination per step, every step on all threads, phases
            // separated by spin barriers (a step is tens of microseconds):
            //   A  each thread scans its slice of the vertex's W row (live
            //      entries, rows ascending) and sorts a copy by (value, row);
            //   M  the sorted slices merge pairwise, a tree of log2 T levels;
            //   C  thread 0: prefix sums, D, the factor column's place;
            //   S  each thread samples the partners of its slice of entries
            //      (draw i is the i-th of the vertex's stream, computed
            //      directly) and writes its slice of the factor column;
            //   F  each thread adds the fills that land in its rows of W.
            // The factor is the one the sequential loop gives: the same
            // sort (a total order), sums in entry order, draws and fills.
            const int TT = pool.size();
            SpinBarrier bar(TT);
            std::vector<std::vector<Ent>> scan((size_t)TT), run((size_t)TT);
            std::vector<int64_t> soff((size_t)TT + 1, 0);
            std::vector<Ent> mA, mB;
            std::vector<double> cum;
            struct TailFill {
                int32_t a, b;   // tail ids
                double val;
            };
            std::vector<TailFill> tf;
            double tot = 0.0;
            int64_t cnt = 0, cbase = 0;
            const Ent *sorted = nullptr;
            // SPEC_SPARS_VERBOSE >= 2: time of the phases (thread 0's view)
            double t_scan = 0, t_sample = 0, t_fill = 0;
            using clk = std::chrono::steady_clock;
            auto secs = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); };
            pool.run([&](int t) {
                int64_t clo, chi, rlo, rhi;
                HostPool::range(L, TT, t, clo, chi);   // W columns scanned
                HostPool::range(L, TT, t, rlo, rhi);   // W rows updated
                for (const int32_t v : cand) {
                    const int32_t tv = tid[v];
                    const double *row = W.data() + (size_t)tv * L;
                    const clk::time_point c0 = trace && t == 0 ? clk::now() : clk::time_point();
                    // A
                    auto &sl = scan[t];
                    sl.clear();
                    for (int64_t x = clo; x < chi; x++)
                        if (row[x] != 0.0 && !tdone[x]) sl.push_back({live[x], row[x]});
                    run[t].assign(sl.begin(), sl.end());
                    std::sort(run[t].begin(), run[t].end(), [](const Ent &a, const Ent &b) {
                        return a.val != b.val ? a.val < b.val : a.row < b.row;
                    });
                    bar.wait();
                    // M: the slices' offsets, then pairwise merges
                    if (t == 0) {
                        soff[0] = 0;
                        for (int u = 0; u < TT; u++) soff[u + 1] = soff[u] + (int64_t)scan[u].size();
                        cnt = soff[TT];
                        if ((int64_t)mA.size() < cnt) {
                            mA.resize((size_t)cnt);
                            mB.resize((size_t)cnt);
                        }
                    }
                    bar.wait();
                    std::copy(run[t].begin(), run[t].end(), mA.begin() + soff[t]);
                    bar.wait();
                    Ent *src = mA.data(), *dst = mB.data();
                    for (int width = 1; width < TT; width *= 2) {
                        const int first = 2 * width * t;   // this thread's pair of run groups
                        if (first < TT) {
                            const int mid = std::min(first + width, TT), last = std::min(first + 2 * width, TT);
                            std::merge(src + soff[first], src + soff[mid], src + soff[mid], src + soff[last],
                                       dst + soff[first], [](const Ent &a, const Ent &b) {
                                           return a.val != b.val ? a.val < b.val : a.row < b.row;
                                       });
                        }
                        std::swap(src, dst);
                        bar.wait();
                    }
                    const clk::time_point c1 = trace && t == 0 ? clk::now() : clk::time_point();
                    // C
                    if (t == 0) {
                        sorted = src;
                        order.push_back(v);
                        cum.resize((size_t)std::max<int64_t>(cnt, 1));
                        double r = 0.0;
                        for (int64_t i = 0; i < cnt; i++) { r += sorted[i].val; cum[(size_t)i] = r; }
                        tot = r;
                        vdiag[v] = cnt > 0 ? tot : 0.0;
                        cbase = (int64_t)crow.size();
                        crow.resize((size_t)(cbase + cnt));
                        cval.resize((size_t)(cbase + cnt));
                        vstart[v] = cbase;
                        vcount[v] = cnt;
                        tf.resize((size_t)std::max<int64_t>(cnt - 1, 0));
                    }
                    bar.wait();
                    // S: partners of entries [lo, hi) of the cnt - 1 that
                    // sample one, and this thread's slice of the column
                    if (cnt > 0) {
                        const ColumnRng rng(seed, v);
                        int64_t lo, hi;
                        HostPool::range(cnt - 1, TT, t, lo, hi);
                        for (int64_t i = lo; i < hi; i++) {
                            const double rest = tot - cum[(size_t)i];
                            const double target = rng.at((uint64_t)i) * rest;
                            int64_t a = i + 1, b = cnt;
                            while (a < b) {
                                const int64_t mid = (a + b) / 2;
                                if (cum[(size_t)mid] - cum[(size_t)i] < target) a = mid + 1; else b = mid;
                            }
                            const int64_t j = a < cnt ? a : cnt - 1;
                            tf[(size_t)i] = {tid[sorted[i].row], tid[sorted[j].row], sorted[i].val * rest / tot};
                        }
                        const int64_t o = cbase + soff[t];
                        for (size_t i = 0; i < sl.size(); i++) {
                            crow[(size_t)o + i] = sl[i].row;
                            cval[(size_t)o + i] = -(sl[i].val / tot);
                        }
                    }
                    bar.wait();
                    const clk::time_point c2 = trace && t == 0 ? clk::now() : clk::time_point();
                    // F: one elimination adds to each entry of W at most
                    // once (its fill pairs are distinct), so the additions
                    // may go in any order: threads take their rows
                    double *w = W.data();
                    for (const TailFill &f : tf) {
                        if (f.a >= rlo && f.a < rhi) w[(size_t)f.a * L + f.b] += f.val;
                        if (f.b >= rlo && f.b < rhi) w[(size_t)f.b * L + f.a] += f.val;
                    }
                    if (t == 0) {
                        elim[v] = 1;
                        tdone[tv] = 1;
                        rounds++;
                    }
                    bar.wait();
                    if (trace && t == 0) {
                        const clk::time_point c3 = clk::now();
                        t_scan += secs(c0, c1);
                        t_sample += secs(c1, c2);
                        t_fill += secs(c2, c3);
                    }
                }
            });
            if (trace)
                std::fprintf(stderr, "[factor] dense tail: scans + sorts + merges %.1f s, sampling %.1f s, fill additions %.1f s\n",
                             t_scan, t_sample, t_fill);
            tail_rows = (int64_t)cand.size();
            live.clear();
            if (verbose) {
                std::fprintf(stderr, "[factor] dense tail of %lld vertices from round %d: %.1f s\n",
                             (long long)tail_rows, rounds - (int32_t)tail_rows,
                             std::chrono::duration<double>(std::chrono::steady_clock::now() - t_tail).count());
                std::fflush(stderr);
            }
            break;
        }
        // 2. candidates of degree <= 2x the minimum, in key order
        const uint64_t capdeg = std::max<uint64_t>(2 * mindeg, 2);
        pool.run([&](int t) {
            int64_t lo, hi;
            HostPool::range(L, active, t, lo, hi);
            pt[t].cand.clear();
            for (int64_t x = lo; x < hi; x++)
                if ((key[live[x]] >> 40) <= capdeg) pt[t].cand.push_back(live[x]);
        }, active);
        cand.clear();
        for (int t = 0; t < active; t++) cand.insert(cand.end(), pt[t].cand.begin(), pt[t].cand.end());
        std::sort(cand.begin(), cand.end(),
                  [&](int32_t x, int32_t y) { return key[x] != key[y] ? key[x] < key[y] : x < y; });
        lap(1, tl);
        // 3. greedy maximal independent set (sequential, in key order)
        sel.clear();
        for (int32_t v : cand) {
            if (blocked[v] == rounds + 1) continue;
            sel.push_back(v);
            for (const Ent &e : adj[v]) blocked[e.row] = rounds + 1;
        }
        std::sort(sel.begin(), sel.end());
        last_sel = (int64_t)sel.size();
        lap(2, tl);
        // 4. eliminate the set: thread t takes a contiguous range of sel
        const int64_t S = (int64_t)sel.size();
        const int sact = S >= 256 ? T : 1;
        pool.run([&](int t) {
            PerThread &P = pt[t];
            int64_t lo, hi;
            HostPool::range(S, sact, t, lo, hi);
            P.cols.clear();
            P.ccount.clear();
            for (auto &f : P.fills) f.clear();
            for (int64_t x = lo; x < hi; x++) {
                const int32_t v = sel[x];
                P.work.assign(adj[v].begin(), adj[v].end());
                ColumnRng rng(seed, v);
                vdiag[v] = eliminate(P.work, P.cum, rng, P.col, [&](int32_t a, int32_t b, double fv) {
                    P.fills[owner(a)].push_back({a, b, fv});
                    P.fills[owner(b)].push_back({b, a, fv});
                });
                P.ccount.push_back((int64_t)P.col.size());
                P.cols.insert(P.cols.end(), P.col.begin(), P.col.end());
                elim[v] = 1;
                std::vector<Ent>().swap(adj[v]);
            }
        }, sact);
        lap(3, tl);
        // columns in sel order, each thread copying its own
        {
            const int64_t obase = (int64_t)order.size();
            order.resize((size_t)(obase + S));
            std::vector<int64_t> cb((size_t)sact + 1);
            cb[0] = (int64_t)crow.size();
            for (int t = 0; t < sact; t++) cb[(size_t)t + 1] = cb[(size_t)t] + (int64_t)pt[t].cols.size();
            crow.resize((size_t)cb[(size_t)sact]);
            cval.resize((size_t)cb[(size_t)sact]);
            pool.run([&](int t) {
                int64_t lo, hi;
                HostPool::range(S, sact, t, lo, hi);
                int64_t off = cb[(size_t)t];
                for (int64_t x = lo; x < hi; x++) {
                    const int32_t v = sel[x];
                    order[(size_t)(obase + x)] = v;
                    vstart[v] = off;
                    vcount[v] = pt[t].ccount[x - lo];
                    off += vcount[v];
                }
                const auto &cols = pt[t].cols;
                for (size_t q = 0; q < cols.size(); q++) {
                    crow[(size_t)cb[(size_t)t] + q] = cols[q].row;
                    cval[(size_t)cb[(size_t)t] + q] = cols[q].val;
                }
            }, sact);
        }
        lap(4, tl);
        // 5. fill edges, appended by the owner of their target
        pool.run([&](int o) {
            for (int t = 0; t < sact; t++)
                for (const Fill3 &f : pt[t].fills[o]) adj[f.target].push_back({f.other, f.val});
        });
        lap(5, tl);
        // 6. the live set shrinks (order kept)
        pool.run([&](int t) {
            int64_t lo, hi;
            HostPool::range(L, active, t, lo, hi);
            pt[t].next.clear();
            for (int64_t x = lo; x < hi; x++)
                if (!elim[live[x]]) pt[t].next.push_back(live[x]);
        }, active);
        int64_t w = 0;
        for (int t = 0; t < active; t++) {
            std::copy(pt[t].next.begin(), pt[t].next.end(), live.begin() + w);
            w += (int64_t)pt[t].next.size();
        }
        live.resize((size_t)w);
        lap(6, tl);
        rounds++;
    }
    if (trace)
        std::fprintf(stderr, "[factor] rounds: degrees+keys %.1f s, candidates %.1f s, greedy set %.1f s, "
                     "eliminations %.1f s, column copy %.1f s, fill exchange %.1f s, live set %.1f s\n",
                     st[0], st[1], st[2], st[3], st[4], st[5], st[6]);
    // vstart/vcount are per vertex; pack them as [start, start+count)
    std::vector<int64_t> packed((size_t)n * 2);
    for (int64_t v = 0; v < n; v++) { packed[2 * v] = vstart[v]; packed[2 * v + 1] = vcount[v]; }
    vstart.swap(packed);
    return SS_OK;
}

static int g_factor_threads = 0;   // 0: HostPool::default_threads()

void factor_set_threads(int threads) { g_factor_threads = threads < 0 ? 0 : threads; }
int factor_threads() { return g_factor_threads > 0 ? g_factor_threads : HostPool::default_threads(); }

int factor_build(int64_t n, const int32_t *colptr, const int32_t *rowidx, const double *vals, uint64_t seed,
                 int ordering, Factor &F) {
    auto t0 = std::chrono::steady_clock::now();
    if (n <= 0 || n > INT32_MAX - 1) return fail(SS_ERR_INVALID, "factor: bad n=%lld", (long long)n);
    if (ordering < ORDER_NATURAL || ordering > ORDER_ROUNDS) return fail(SS_ERR_INVALID, "factor: unknown ordering %d", ordering);
    F.n = n;
    F.ordering = ordering;
    F.rounds = 0;
    F.tail = 0;
    std::vector<int32_t> rank((size_t)n);
    // per-position column of L in position ids: start, rows, values
    std::vector<int64_t> pstart((size_t)n + 1, 0);
    std::vector<int32_t> prow;
    std::vector<double> pval;
    if (ordering == ORDER_ROUNDS) {
        std::vector<int64_t> vsc;
        std::vector<int32_t> crow;
        std::vector<double> cval, vdiag;
        SS_TRY(factor_rounds(n, colptr, rowidx, vals, seed, factor_threads(), F.order, vsc, crow, cval, vdiag,
                             F.rounds, F.tail));
        const auto t_elim = std::chrono::steady_clock::now();
        if (std::getenv("SPEC_SPARS_VERBOSE")) {
            std::fprintf(stderr, "[factor] eliminated at %.1f s\n", std::chrono::duration<double>(t_elim - t0).count());
            std::fflush(stderr);
        }
        for (int64_t p = 0; p < n; p++) rank[F.order[p]] = (int32_t)p;
        // U rows (columns of L in position ids, diagonal first, rows
        // ascending) assembled in parallel over positions
        F.diag.resize((size_t)n);
        F.urowptr.assign((size_t)n + 1, 0);
        int64_t tot = 0;
        for (int64_t p = 0; p < n; p++) {
            F.urowptr[p] = (int32_t)std::min<int64_t>(tot, INT32_MAX);
            tot += 1 + vsc[2 * (size_t)F.order[p] + 1];
        }
        if (tot > (int64_t)INT32_MAX) return fail(SS_ERR_CAPACITY, "factor: nnz %lld exceeds int32", (long long)tot);
        F.urowptr[n] = (int32_t)tot;
        F.ucol.resize((size_t)tot);
        F.uval.resize((size_t)tot);
        HostPool pool(factor_threads());
        // positions split by entries, not by count: the dense tail's
        // columns at the end hold most of them
        auto entry_split = [&](int t) {
            const int64_t want = tot * t / pool.size();
            return (int64_t)(std::lower_bound(F.urowptr.begin(), F.urowptr.begin() + n, (int32_t)want) -
                             F.urowptr.begin());
        };
        pool.run([&](int t) {
            const int64_t lo = entry_split(t), hi = t + 1 == pool.size() ? n : entry_split(t + 1);
            std::vector<Ent> tmp;
            for (int64_t p = lo; p < hi; p++) {
                const int32_t v = F.order[p];
                F.diag[p] = vdiag[v];
                tmp.clear();
                for (int64_t q = vsc[2 * v]; q < vsc[2 * v] + vsc[2 * v + 1]; q++) tmp.push_back({rank[crow[q]], cval[q]});
                std::sort(tmp.begin(), tmp.end(), [](const Ent &a, const Ent &b) { return a.row < b.row; });
                int64_t o = F.urowptr[p];
                F.ucol[o] = (int32_t)p;
                F.uval[o] = 1.0;
                for (const Ent &e : tmp) { o++; F.ucol[o] = e.row; F.uval[o] = e.val; }
            }
        });
        F.nnz = tot;
        const auto t_asm = std::chrono::steady_clock::now();
        factor_levels(F);
        F.ms_build = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (std::getenv("SPEC_SPARS_VERBOSE")) {
            std::fprintf(stderr, "[factor] assembled at %.1f s, levels %.1f s, total %.1f s\n",
                         std::chrono::duration<double>(t_asm - t0).count(),
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - t_asm).count(),
                         F.ms_build / 1000.0);
            std::fflush(stderr);
        }
        return SS_OK;
    } else {
        F.order.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) F.order[i] = (int32_t)i;
        if (ordering == ORDER_RANDOM) {
            // Fisher-Yates with the factor seed
            ColumnRng rng(seed ^ 0x2545f4914f6cdd1dULL, -1);
            for (int64_t i = n - 1; i > 0; i--) {
                int64_t j = (int64_t)(rng.uniform() * (double)(i + 1));
                if (j > i) j = i;
                std::swap(F.order[i], F.order[j]);
            }
        }
        for (int64_t p = 0; p < n; p++) rank[F.order[p]] = (int32_t)p;
        if (ordering == ORDER_NATURAL) {
            SS_TRY(factor_natural(n, colptr, rowidx, vals, seed, pstart, prow, pval, F.diag));
        } else {
            // the permuted Laplacian P L P^T as CSC
            std::vector<int32_t> pc((size_t)n + 1, 0);
            for (int64_t c = 0; c < n; c++) pc[rank[c] + 1] = colptr[c + 1] - colptr[c];
            for (int64_t i = 0; i < n; i++) pc[i + 1] += pc[i];
            std::vector<int32_t> pr((size_t)colptr[n]);
            std::vector<double> pv((size_t)colptr[n]);
            std::vector<Ent> tmp;
            for (int64_t c = 0; c < n; c++) {
                tmp.clear();
                for (int32_t t = colptr[c]; t < colptr[c + 1]; t++) {
                    if (rowidx[t] < 0 || rowidx[t] >= n) return fail(SS_ERR_INVALID, "factor: row index %d out of range", rowidx[t]);
                    tmp.push_back({rank[rowidx[t]], vals[t]});
                }
                std::sort(tmp.begin(), tmp.end(), [](const Ent &a, const Ent &b) { return a.row < b.row; });
                int32_t o = pc[rank[c]];
                for (size_t q = 0; q < tmp.size(); q++) { pr[o + q] = tmp[q].row; pv[o + q] = tmp[q].val; }
            }
            SS_TRY(factor_natural(n, pc.data(), pr.data(), pv.data(), seed, pstart, prow, pval, F.diag));
        }
    }
    const int64_t nnz_lower = (int64_t)prow.size();
    if (nnz_lower + n > (int64_t)INT32_MAX) return fail(SS_ERR_CAPACITY, "factor: nnz %lld exceeds int32", (long long)(nnz_lower + n));
    // U CSR = columns of L, diagonal first (src/example.cc:422-470)
    F.urowptr.assign((size_t)n + 1, 0);
    F.ucol.resize((size_t)(nnz_lower + n));
    F.uval.resize((size_t)(nnz_lower + n));
    int64_t p = 0;
    for (int64_t c = 0; c < n; c++) {
        F.urowptr[c] = (int32_t)p;
        F.ucol[p] = (int32_t)c; F.uval[p] = 1.0; p++;
        for (int64_t t = pstart[c]; t < pstart[c + 1]; t++) { F.ucol[p] = prow[t]; F.uval[p] = pval[t]; p++; }
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

// Level-ordered CSR for the two solves, in vertex ids.  Position t of the
// forward schedule holds the vertex at elimination position perm[t]
// (positions sorted by (level, position)); its entries are that row of L in
// pull form (the vertices it waits for).  The backward schedule holds the
// rows of L^T, i.e. the off-diagonal part of U's rows.  Every dependency
// must have a strictly lower level -- the sync-free solve relies on it.
int32_t forward_tail_start(const Factor &F, int32_t small_rows, int64_t max_rows) {
    std::vector<int64_t> cnt((size_t)F.depth_fwd + 1, 0);
    for (int64_t p = 0; p < F.n; p++) cnt[F.lvl_fwd[p]]++;
    int32_t lo = F.depth_fwd;   // 0-based levels [lo, depth) form the run
    int64_t rows = 0;
    while (lo > 0 && cnt[lo] <= small_rows && rows + cnt[lo] <= max_rows) rows += cnt[lo--];
    return lo;
}

int factor_schedules(const Factor &F, const std::vector<char> &core, Schedule &fwd, Schedule &bwd, int32_t long_row,
                     int32_t small_rows) {
    const int64_t n = F.n;
    if ((int64_t)core.size() != n) return fail(SS_ERR_INVALID, "schedule: core flags sized %zu, n=%lld", core.size(), (long long)n);
    // strictly-lower rows of L (position space): transpose of U's off-diagonal;
    // a core row keeps only its entries in columns outside the core
    std::vector<int32_t> lcnt((size_t)n + 1, 0), blen((size_t)n, 0);
    for (int64_t i = 0; i < n; i++)
        for (int32_t t = F.urowptr[i]; t < F.urowptr[i + 1]; t++) {
            const int32_t r = F.ucol[t];
            if (r == i) continue;
            blen[i]++;
            if (!(core[r] && core[i])) lcnt[r + 1]++;
        }
    for (int64_t i = 0; i < n; i++) lcnt[i + 1] += lcnt[i];
    std::vector<int32_t> lcol((size_t)lcnt[n]);
    std::vector<double> lval((size_t)lcnt[n]);
    {
        std::vector<int32_t> fillp(lcnt.begin(), lcnt.end() - 1);
        for (int64_t i = 0; i < n; i++)      // i ascending => each row's columns ascending
            for (int32_t t = F.urowptr[i]; t < F.urowptr[i + 1]; t++) {
                const int32_t r = F.ucol[t];
                if (r == i || (core[r] && core[i])) continue;
                lcol[fillp[r]] = (int32_t)i;
                lval[fillp[r]] = F.uval[t];
                fillp[r]++;
            }
    }
    std::vector<int32_t> flen((size_t)n);
    for (int64_t i = 0; i < n; i++) flen[i] = lcnt[i + 1] - lcnt[i];

    // non-core positions by (level, position), levels renumbered densely;
    // within a level, long rows first (they get a block each)
    auto order = [&](const std::vector<int32_t> &lvl, const std::vector<int32_t> &len, Schedule &S,
                     std::vector<int32_t> &slot) {
        int32_t maxl = 0;
        for (int64_t i = 0; i < n; i++)
            if (!core[i]) maxl = std::max(maxl, lvl[i]);
        std::vector<int64_t> cnt((size_t)maxl + 2, 0);
        for (int64_t i = 0; i < n; i++)
            if (!core[i]) cnt[lvl[i]]++;
        std::vector<int32_t> dense((size_t)maxl + 2, -1);
        S.depth = 0;
        for (int32_t l = 1; l <= maxl; l++)
            if (cnt[l]) dense[l] = S.depth++;
        S.level_ptr.assign((size_t)S.depth + 1, 0);
        for (int32_t l = 1; l <= maxl; l++)
            if (cnt[l]) S.level_ptr[dense[l] + 1] = (int32_t)cnt[l];
        for (int32_t l = 0; l < S.depth; l++) S.level_ptr[l + 1] += S.level_ptr[l];
        std::vector<int64_t> pos(S.level_ptr.begin(), S.level_ptr.end());
        slot.assign((size_t)S.level_ptr[S.depth], 0);
        for (int64_t i = 0; i < n; i++)
            if (!core[i]) slot[pos[dense[lvl[i]]]++] = (int32_t)i;
        // then the rows taking one wave each, then the tiny rows; the long
        // rows of more than kWgRowMax entries (several workgroups each) lead
        S.level_long.assign((size_t)S.depth, 0);
        S.level_huge.assign((size_t)S.depth, 0);
        S.level_med.assign((size_t)S.depth, 0);
        for (int32_t l = 0; l < S.depth; l++) {
            auto b = slot.begin() + S.level_ptr[l], e = slot.begin() + S.level_ptr[l + 1];
            const int32_t wave_max = e - b >= kWaveLevelRows ? std::max(long_row, (int32_t)kWaveRowMax) : long_row;
            auto mid = std::stable_partition(b, e, [&](int32_t p) { return len[p] > wave_max; });
            auto tiny = std::stable_partition(mid, e, [&](int32_t p) { return len[p] > kTinyRow; });
            auto huge = std::stable_partition(b, mid, [&](int32_t p) { return len[p] > kWgRowMax; });
            S.level_long[l] = (int32_t)(mid - b);
            S.level_huge[l] = (int32_t)(huge - b);
            S.level_med[l] = (int32_t)(tiny - mid);
        }
    };
    // backward levels of the rows outside the cores, whose rows (solved
    // first, densely) count as level 0
    std::vector<int32_t> blvl((size_t)n, 0);
    for (int64_t i = n - 1; i >= 0; i--) {
        if (core[i]) continue;
        int32_t l = 1;
        for (int32_t t = F.urowptr[i]; t < F.urowptr[i + 1]; t++) {
            const int32_t c = F.ucol[t];
            if (c != i && !core[c] && blvl[c] + 1 > l) l = blvl[c] + 1;
        }
        blvl[i] = l;
    }
    std::vector<int32_t> fslot, bslot;
    order(F.lvl_fwd, flen, fwd, fslot);
    order(blvl, blen, bwd, bslot);
    // forward: the core rows follow the levels in elimination order (their
    // entries outside the core; the rest is the dense core solve)
    fwd.core_t0 = (int32_t)fslot.size();
    for (int64_t i = 0; i < n; i++)
        if (core[i]) fslot.push_back((int32_t)i);
    fwd.core_huge = (int32_t)(std::stable_partition(fslot.begin() + fwd.core_t0, fslot.end(),
                                                    [&](int32_t p) { return flen[p] > kWgRowMax; }) -
                              (fslot.begin() + fwd.core_t0));
    bwd.core_t0 = (int32_t)bslot.size();

    auto pack = [&](Schedule &S, const std::vector<int32_t> &slot, const std::vector<int32_t> &lvl,
                    const int32_t *ptr, const int32_t *col, const double *val, bool skip_diag) -> int {
        const int64_t rows = (int64_t)slot.size();
        S.n = rows;
        S.ptr.assign((size_t)rows + 1, 0);
        S.rows.resize((size_t)rows);
        int64_t total = 0;
        for (int64_t t = 0; t < rows; t++) {
            const int32_t i = slot[t];
            S.ptr[t] = (int32_t)total;
            for (int32_t q = ptr[i]; q < ptr[i + 1]; q++)
                if (!(skip_diag && col[q] == i)) total++;
        }
        S.ptr[rows] = (int32_t)total;
        S.col.resize((size_t)total);
        S.val.resize((size_t)total);
        int64_t p = 0;
        for (int64_t t = 0; t < rows; t++) {
            const int32_t i = slot[t];
            S.rows[t] = F.order[i];
            for (int32_t q = ptr[i]; q < ptr[i + 1]; q++) {
                if (skip_diag && col[q] == i) continue;
                if (lvl[col[q]] >= lvl[i])
                    return fail(SS_ERR_INVALID, "schedule: position %d depends on position %d of level >= its own", i, col[q]);
                S.col[p] = F.order[col[q]];
                S.val[p] = val[q];
                p++;
            }
        }
        return SS_OK;
    };
    SS_TRY(pack(fwd, fslot, F.lvl_fwd, lcnt.data(), lcol.data(), lval.data(), false));
    SS_TRY(pack(bwd, bslot, blvl, F.urowptr.data(), F.ucol.data(), F.uval.data(), true));

    // every maximal run of at least two small levels goes to one workgroup
    for (Schedule *S : {&fwd, &bwd}) {
        // (and no row of more than kWgRowMax entries: the workgroup folds at
        // most kWgParts parts of a row)
        auto small = [&](int32_t l) {
            if (S->level_ptr[l + 1] - S->level_ptr[l] > small_rows ||
                S->ptr[S->level_ptr[l + 1]] - S->ptr[S->level_ptr[l]] > kSmallNnz)
                return false;
            for (int32_t t = S->level_ptr[l]; t < S->level_ptr[l + 1]; t++)
                if (S->ptr[t + 1] - S->ptr[t] > kWgRowMax) return false;
            return true;
        };
        S->runs.clear();
        for (int32_t l = 0; l < S->depth;) {
            if (!small(l)) { l++; continue; }
            int32_t e = l;
            while (e < S->depth && small(e)) e++;
            if (e - l >= 2) S->runs.push_back({l, e});
            l = e;
        }
    }
    return SS_OK;
}

}  // namespace ss
