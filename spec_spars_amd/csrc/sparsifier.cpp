// C++ host layer mirroring the reference's Sparsifier (src/sparsifier.rs)
// over the C ABI, plus its C wrapper (include/spec_spars_host.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/spec_spars.hpp"
#include "../../include/spec_spars_host.h"
#include "ss_internal.h"

namespace spec_spars {

using ss::fail;

SparsifierParameters SparsifierParameters::new_default(bool bench) {
    SparsifierParameters p;   // member defaults are new_default's values
    p.bench = bench;
    return p;
}

const double *CscMatrix::get(int32_t row, int32_t col) const {
    if (col < 0 || col >= n) return nullptr;
    auto b = rowidx.begin() + colptr[col], e = rowidx.begin() + colptr[col + 1];
    auto it = std::lower_bound(b, e, row);
    if (it == e || *it != row) return nullptr;
    return &vals[it - rowidx.begin()];
}

int Triplet::insert(int32_t v1, int32_t v2, double value) {
    // src/sparsifier.rs:48-70
    if (v1 < 0 || v2 < 0 || v1 >= num_nodes || v2 >= num_nodes)
        return fail(SS_ERR_INVALID, "Triplet::insert: vertex (%d,%d) outside 0..%d", v1, v2, num_nodes);
    if (v1 != v2) {
        row_indices.push_back(v1);
        col_indices.push_back(v2);
        values.push_back(value * -1.0);
        row_indices.push_back(v2);
        col_indices.push_back(v1);
        values.push_back(value * -1.0);
        diagonal[v1] += value;
        diagonal[v2] += value;
    }
    return SS_OK;
}

void Triplet::process_diagonal() {
    // src/sparsifier.rs:72-80
    for (int32_t i = 0; i < num_nodes; i++) {
        row_indices.push_back(i);
        col_indices.push_back(i);
        values.push_back(diagonal[i]);
    }
}

CscMatrix Triplet::to_csc() const {
    // src/sparsifier.rs:83-87 (sprs TriMat::to_csc: duplicates summed)
    CscMatrix A;
    A.n = num_nodes;
    const size_t t = values.size();
    std::vector<int32_t> cnt((size_t)num_nodes + 1, 0);
    for (size_t i = 0; i < t; i++) cnt[col_indices[i] + 1]++;
    for (int32_t c = 0; c < num_nodes; c++) cnt[c + 1] += cnt[c];
    std::vector<size_t> idx(t);
    {
        std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
        for (size_t i = 0; i < t; i++) idx[pos[col_indices[i]]++] = i;
    }
    A.colptr.assign((size_t)num_nodes + 1, 0);
    for (int32_t c = 0; c < num_nodes; c++) {
        auto b = idx.begin() + cnt[c], e = idx.begin() + cnt[c + 1];
        std::stable_sort(b, e, [&](size_t x, size_t y) { return row_indices[x] < row_indices[y]; });
        for (auto it = b; it != e;) {
            const int32_t r = row_indices[*it];
            double acc = values[*it];
            ++it;
            while (it != e && row_indices[*it] == r) acc += values[*it++];
            A.rowidx.push_back(r);
            A.vals.push_back(acc);
        }
        A.colptr[c + 1] = (int32_t)A.rowidx.size();
    }
    return A;
}

void Triplet::delete_state() {
    col_indices.clear();
    row_indices.clear();
    values.clear();
    diagonal.assign((size_t)num_nodes, 0.0);
}

CscMatrix csc_add(const CscMatrix &a, const CscMatrix &b) {
    CscMatrix r;
    r.n = a.n;
    r.colptr.assign((size_t)a.n + 1, 0);
    r.rowidx.reserve(a.nnz() + b.nnz());
    r.vals.reserve(a.nnz() + b.nnz());
    for (int32_t c = 0; c < a.n; c++) {
        int32_t i = a.colptr[c], ie = a.colptr[c + 1], j = b.colptr[c], je = b.colptr[c + 1];
        while (i < ie || j < je) {
            int32_t row;
            double v;
            if (j >= je || (i < ie && a.rowidx[i] < b.rowidx[j])) { row = a.rowidx[i]; v = a.vals[i++]; }
            else if (i >= ie || b.rowidx[j] < a.rowidx[i]) { row = b.rowidx[j]; v = b.vals[j++]; }
            else { row = a.rowidx[i]; v = a.vals[i++] + b.vals[j++]; }
            if (v != 0.0) { r.rowidx.push_back(row); r.vals.push_back(v); }
        }
        r.colptr[c + 1] = (int32_t)r.rowidx.size();
    }
    return r;
}

Sparsifier::Sparsifier(ss_context *ctx_, int32_t n, const SparsifierParameters &p)
    : ctx(ctx_), num_nodes(n), new_entries(n), params(p) {
    // src/sparsifier.rs:262-297
    epsilon = p.epsilon;
    beta_constant = p.beta_constant;
    row_constant = p.row_constant;
    beta = std::pow(p.epsilon, -2.0) * (double)p.beta_constant * std::log((double)n);
    threshold = (size_t)n * (size_t)std::llround(beta) * (size_t)p.row_constant;
    verbose = p.verbose;
    jl_factor = p.jl_factor;
    jl_dim = (int32_t)std::ceil(std::log2((double)n) * p.jl_factor);
    jl_scaling_factor = p.jl_scaling_factor;
    sketch_seed = p.sketch_seed;
    sampling_seed = p.sampling_seed;
    sketch_uniform = p.sketch_uniform;
    streaming = p.streaming;
    current_laplacian.n = n;
    current_laplacian.colptr.assign((size_t)n + 1, 0);
}

int Sparsifier::insert(int32_t v1, int32_t v2, double value) {
    // src/sparsifier.rs:357-376
    if (v1 < 0 || v2 < 0 || v1 >= num_nodes || v2 >= num_nodes)
        return fail(SS_ERR_INVALID, "Sparsifier::insert: vertex (%d,%d) outside 0..%d", v1, v2, num_nodes);
    if (value != 0.0) {
        int rc = new_entries.insert(v1, v2, value);
        if (rc != SS_OK) return rc;
    }
    if (streaming && num_edges() > threshold) {
        triggered++;
        return sparsify(false, nullptr);
    }
    return SS_OK;
}

size_t Sparsifier::num_edges() const {
    // src/sparsifier.rs:346-353: (nnz - stored diagonal) / 2
    size_t diag = 0;
    for (int32_t c = 0; c < num_nodes; c++)
        if (current_laplacian.get(c, c)) diag++;
    return (current_laplacian.nnz() - diag) / 2;
}

int Sparsifier::check_diagonal() const {
    // src/sparsifier.rs:806-817: every column sums to 0 (1e-10), symmetric
    const CscMatrix &A = current_laplacian;
    for (int32_t c = 0; c < num_nodes; c++) {
        double s = 0.0;
        for (int32_t q = A.colptr[c]; q < A.colptr[c + 1]; q++) s += A.vals[q];
        if (std::fabs(s) > 1e-10) return fail(SS_ERR_INVALID, "check_diagonal: column %d sums to %g", c, s);
        for (int32_t q = A.colptr[c]; q < A.colptr[c + 1]; q++) {
            const double *t = A.get(c, A.rowidx[q]);
            if (!t || *t != A.vals[q]) return fail(SS_ERR_INVALID, "check_diagonal: not symmetric at (%d,%d)", A.rowidx[q], c);
        }
    }
    return SS_OK;
}

int Sparsifier::form_laplacian(bool check) {
    // src/sparsifier.rs:693-709
    new_entries.process_diagonal();
    CscMatrix fresh = new_entries.to_csc();
    new_entries.delete_state();
    current_laplacian = csc_add(current_laplacian, fresh);
    return check ? check_diagonal() : SS_OK;
}

int Sparsifier::jl_sketch(std::vector<double> &Y) const {
    // the incidence matrix of new_entries (src/sparsifier.rs:90-114): entry
    // pairs (v1,v2,-w),(v2,v1,-w) -> edge (v1, v2, w) in insertion order
    const size_t m = new_entries.values.size() / 2;
    std::vector<int32_t> v1(m), v2(m);
    std::vector<double> w(m);
    for (size_t i = 0; i < m; i++) {
        v1[i] = new_entries.row_indices[2 * i];
        v2[i] = new_entries.col_indices[2 * i];
        w[i] = -1.0 * new_entries.values[2 * i];
    }
    Y.assign((size_t)num_nodes * jl_dim, 0.0);
    if (m == 0) return SS_OK;
    return ss_sketch(ctx, v1.data(), v2.data(), w.data(), (int64_t)m, num_nodes, jl_dim, sketch_seed,
                     jl_scaling_factor, sketch_uniform ? 1 : 0, Y.data());
}

void Sparsifier::upper_edges(std::vector<int32_t> &u, std::vector<int32_t> &v, std::vector<double> &lval) const {
    const CscMatrix &A = current_laplacian;
    for (int32_t c = 0; c < num_nodes; c++)
        for (int32_t q = A.colptr[c]; q < A.colptr[c + 1]; q++)
            if (A.rowidx[q] < c) {
                u.push_back(A.rowidx[q]);
                v.push_back(c);
                lval.push_back(A.vals[q]);
            }
}

static ss_solve_opts solve_opts_of(const SparsifierParameters &p) {
    ss_solve_opts o;
    ss_solve_opts_default(&o);
    o.factor_seed = p.factor_seed;
    o.stop_rule = p.stop_rule;
    o.tol = p.tol;
    o.max_iter = p.max_iter;
    o.precision = p.precision;
    o.ordering = p.ordering;
    return o;
}

int Sparsifier::get_probs(size_t length, const std::vector<double> &Y, std::vector<double> &probs) {
    // src/sparsifier.rs:586-622: run_solve_lap, then diff norms and probabilities
    const CscMatrix &A = current_laplacian;
    std::vector<double> Z((size_t)num_nodes * jl_dim);
    ss_solve_opts o = solve_opts_of(params);
    int rc = ss_solve_lap(ctx, Y.data(), num_nodes, jl_dim, A.colptr.data(), A.rowidx.data(), A.vals.data(),
                          (int64_t)A.nnz(), &o, Z.data(), nullptr);
    if (rc != SS_OK) return rc;
    std::vector<int32_t> u, v;
    std::vector<double> lval;
    upper_edges(u, v, lval);
    if (u.size() != length) return fail(SS_ERR_INVALID, "get_probs: %zu edges, expected %zu", u.size(), length);
    probs.assign(length, 1.0);
    std::vector<uint8_t> mask(length);
    int64_t kept = 0;
    if (length == 0) return SS_OK;
    return ss_er_sample(ctx, Z.data(), num_nodes, jl_dim, u.data(), v.data(), lval.data(), (int64_t)length, beta,
                        sampling_seed, nullptr, probs.data(), nullptr, mask.data(), &kept, nullptr, nullptr,
                        nullptr);
}

int Sparsifier::reweight_from_mask(const std::vector<int32_t> &u, const std::vector<int32_t> &v,
                                   const std::vector<double> &lval, const std::vector<double> &p,
                                   const std::vector<uint8_t> &mask, Triplet &rw) const {
    // src/sparsifier.rs:645-677
    rw = Triplet(num_nodes);
    for (size_t e = 0; e < u.size(); e++) {
        const double true_value = lval[e] * -1.0;
        int rc;
        if (mask[e]) {
            const double target = true_value / std::sqrt(p[e]);
            const double change = true_value - target;
            rc = rw.insert(u[e], v[e], change * -1.0);
        } else {
            rc = rw.insert(u[e], v[e], true_value * -1.0);
        }
        if (rc != SS_OK) return rc;
    }
    return SS_OK;
}

int Sparsifier::sample_and_reweight(const std::vector<double> &probs, Triplet &reweightings) {
    std::vector<int32_t> u, v;
    std::vector<double> lval;
    upper_edges(u, v, lval);
    if (probs.size() != u.size()) return fail(SS_ERR_INVALID, "sample_and_reweight: %zu probabilities for %zu edges",
                                              probs.size(), u.size());
    std::vector<uint8_t> mask(u.size());
    int64_t kept = 0;
    if (!u.empty()) {
        int rc = ss_er_sample(ctx, nullptr, num_nodes, std::max(jl_dim, 1), u.data(), v.data(), lval.data(),
                              (int64_t)u.size(), beta, sampling_seed, probs.data(), nullptr, nullptr, mask.data(),
                              &kept, nullptr, nullptr, nullptr);
        if (rc != SS_OK) return rc;
    }
    return reweight_from_mask(u, v, lval, probs, mask, reweightings);
}

int Sparsifier::apply_reweightings(Triplet &rw, bool report, SparsificationStats *stats) {
    // src/sparsifier.rs:682-691
    rw.process_diagonal();
    CscMatrix delta = rw.to_csc();
    const size_t before = num_edges();
    current_laplacian = csc_add(current_laplacian, delta);
    const size_t after = num_edges();
    if (report && verbose)
        printf("edges before sparsification: %zu, after: %zu (%.2f%%)\n", before, after,
               before ? 100.0 * (double)after / (double)before : 0.0);
    if (stats) {
        stats->original_num_edges = before;
        stats->final_num_edges = after;
    }
    return SS_OK;
}

int Sparsifier::sparsify(bool check, SparsificationStats *stats) {
    // src/sparsifier.rs:722-770.  The probabilities and the sample come out
    // of one fused GPU pass (ER + coin + compare), so the coins are drawn
    // with the probabilities just computed, as flip_coins would.
    std::vector<double> Y;
    int rc = jl_sketch(Y);
    if (rc != SS_OK) return rc;
    if ((rc = form_laplacian(check)) != SS_OK) return rc;
    const size_t num_nnz = num_edges();
    std::vector<int32_t> u, v;
    std::vector<double> lval;
    upper_edges(u, v, lval);
    if (params.resident && num_nnz) return sparsify_resident(Y, stats);
    std::vector<double> probs(num_nnz, 1.0);
    std::vector<uint8_t> mask(num_nnz, 1);
    if (num_nnz) {
        const CscMatrix &A = current_laplacian;
        std::vector<double> Z((size_t)num_nodes * jl_dim);
        ss_solve_opts o = solve_opts_of(params);
        if ((rc = ss_solve_lap(ctx, Y.data(), num_nodes, jl_dim, A.colptr.data(), A.rowidx.data(), A.vals.data(),
                               (int64_t)A.nnz(), &o, Z.data(), nullptr)) != SS_OK)
            return rc;
        int64_t kept = 0;
        if ((rc = ss_er_sample(ctx, Z.data(), num_nodes, jl_dim, u.data(), v.data(), lval.data(), (int64_t)num_nnz,
                               beta, sampling_seed, nullptr, probs.data(), nullptr, mask.data(), &kept, nullptr,
                               nullptr, nullptr)) != SS_OK)
            return rc;
    }
    Triplet rw(num_nodes);
    if ((rc = reweight_from_mask(u, v, lval, probs, mask, rw)) != SS_OK) return rc;
    if ((rc = apply_reweightings(rw, true, stats)) != SS_OK) return rc;
    return check ? check_diagonal() : SS_OK;
}

int Sparsifier::sparsify_resident(const std::vector<double> &Y, SparsificationStats *stats) {
    // The same pass as sparsify() from the solve on: the Laplacian and Y go
    // to the device, the factor is the host's (same seed and order); the
    // solution stays on the device for the fused ER/coin/sample pass, and
    // apply_reweightings runs there too.  Only the kept edges and the new
    // diagonal come back.
    if (!ctx) return fail(SS_ERR_INVALID, "sparsify: no device context");
    const CscMatrix &A = current_laplacian;
    const size_t before = num_edges();
    struct Handles {
        ss_graph *g = nullptr;
        ss_factor *f = nullptr;
        ~Handles() {
            if (g) ss_graph_destroy(g);
            if (f) ss_factor_destroy(f);
        }
    } h;
    int rc;
    if ((rc = ss_graph_create_csc(ctx, num_nodes, A.colptr.data(), A.rowidx.data(), A.vals.data(), &h.g)) != SS_OK)
        return rc;
    ss_solve_opts o = solve_opts_of(params);
    if ((rc = ss_factor_build_ordered(num_nodes, A.colptr.data(), A.rowidx.data(), A.vals.data(), o.factor_seed,
                                      o.ordering, &h.f)) != SS_OK)
        return rc;
    if ((rc = ss_graph_set_factor(h.g, h.f)) != SS_OK) return rc;
    if ((rc = ss_batch_write_rhs(h.g, 0, jl_dim, o.precision, Y.data())) != SS_OK) return rc;
    if ((rc = ss_batch_solve(h.g, jl_dim, &o, nullptr, nullptr, nullptr)) != SS_OK) return rc;
    int64_t kept = 0;
    if ((rc = ss_graph_sample(h.g, jl_dim, beta, sampling_seed, &kept)) != SS_OK) return rc;
    std::vector<double> diag((size_t)num_nodes);
    if ((rc = ss_graph_apply_reweight(h.g, diag.data())) != SS_OK) return rc;
    std::vector<int32_t> ku((size_t)std::max<int64_t>(kept, 1)), kv(ku.size());
    std::vector<double> kw(ku.size());
    if ((rc = ss_graph_read_kept(h.g, kept, ku.data(), kv.data(), kw.data(), &kept)) != SS_OK) return rc;
    // the new Laplacian: kept off-diagonals -kw (canonical order, so each
    // column's rows come out sorted), the diagonal where it is non-zero
    std::vector<int32_t> cnt((size_t)num_nodes + 1, 0);
    for (int64_t e = 0; e < kept; e++) {
        cnt[(size_t)ku[e] + 1]++;
        cnt[(size_t)kv[e] + 1]++;
    }
    for (int32_t c = 0; c < num_nodes; c++) cnt[(size_t)c + 1] += cnt[(size_t)c] + (diag[(size_t)c] != 0.0);
    CscMatrix H;
    H.n = num_nodes;
    H.colptr = cnt;
    H.rowidx.resize((size_t)cnt[(size_t)num_nodes]);
    H.vals.resize(H.rowidx.size());
    // column c: its rows below c (edges (u, c), u ascending), then the
    // diagonal, then its rows above c (edges (c, v), v ascending)
    std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t e = 0; e < kept; e++) {
        const int32_t c = kv[e];
        H.rowidx[(size_t)pos[c]] = ku[e];
        H.vals[(size_t)pos[c]++] = -1.0 * kw[e];
    }
    for (int32_t c = 0; c < num_nodes; c++)
        if (diag[(size_t)c] != 0.0) {
            H.rowidx[(size_t)pos[c]] = c;
            H.vals[(size_t)pos[c]++] = diag[(size_t)c];
        }
    for (int64_t e = 0; e < kept; e++) {
        const int32_t c = ku[e];
        H.rowidx[(size_t)pos[c]] = kv[e];
        H.vals[(size_t)pos[c]++] = -1.0 * kw[e];
    }
    current_laplacian.colptr.swap(H.colptr);
    current_laplacian.rowidx.swap(H.rowidx);
    current_laplacian.vals.swap(H.vals);
    const size_t after = num_edges();
    if (verbose)
        printf("edges before sparsification: %zu, after: %zu (%.2f%%)\n", before, after,
               before ? 100.0 * (double)after / (double)before : 0.0);
    if (stats) {
        stats->original_num_edges = before;
        stats->final_num_edges = after;
    }
    return SS_OK;
}

}  // namespace spec_spars

// ---------------------------------------------------------------------------
// C wrapper
// ---------------------------------------------------------------------------
using spec_spars::Sparsifier;
using spec_spars::SparsifierParameters;
using ss::fail;

#include "host_handles.h"

static SparsifierParameters to_params(const ssh_params *p) {
    SparsifierParameters q = SparsifierParameters::new_default(false);
    if (!p) return q;
    q.epsilon = p->epsilon;
    q.beta_constant = p->beta_constant;
    q.row_constant = p->row_constant;
    q.verbose = p->verbose != 0;
    q.jl_factor = p->jl_factor;
    q.jl_scaling_factor = p->jl_scaling_factor;
    q.sketch_seed = p->sketch_seed;
    q.sampling_seed = p->sampling_seed;
    q.sketch_uniform = p->sketch_uniform != 0;
    q.streaming = p->streaming != 0;
    q.factor_seed = p->factor_seed;
    q.ordering = p->ordering;
    q.precision = p->precision;
    q.stop_rule = p->stop_rule;
    q.max_iter = p->max_iter;
    q.tol = p->tol;
    q.resident = p->resident != 0;
    return q;
}

extern "C" {

void ssh_params_default(ssh_params *p, int bench) {
    if (!p) return;
    SparsifierParameters q = SparsifierParameters::new_default(bench != 0);
    p->epsilon = q.epsilon;
    p->beta_constant = q.beta_constant;
    p->row_constant = q.row_constant;
    p->verbose = q.verbose;
    p->jl_factor = q.jl_factor;
    p->jl_scaling_factor = q.jl_scaling_factor;
    p->sketch_seed = q.sketch_seed;
    p->sampling_seed = q.sampling_seed;
    p->sketch_uniform = q.sketch_uniform;
    p->streaming = q.streaming;
    p->factor_seed = q.factor_seed;
    p->ordering = q.ordering;
    p->precision = q.precision;
    p->stop_rule = q.stop_rule;
    p->max_iter = q.max_iter;
    p->tol = q.tol;
    p->resident = q.resident;
}

int ssh_sparsifier_create(ss_context *ctx, int32_t num_nodes, const ssh_params *p, ssh_sparsifier **out) {
    // ctx may be NULL: the host-only methods (insert, form_laplacian,
    // check_diagonal, sample bookkeeping) work without a GPU; the GPU steps
    // then report SS_ERR_INVALID.
    if (!out) return fail(SS_ERR_INVALID, "null argument");
    if (num_nodes < 2) return fail(SS_ERR_INVALID, "num_nodes must be >= 2");
    try {
        SparsifierParameters q = to_params(p);
        ssh_sparsifier *s = new ssh_sparsifier{Sparsifier(ctx, num_nodes, q)};
        if (s->sp.jl_dim < 1 || s->sp.jl_dim > 256) {
            int k = s->sp.jl_dim;
            delete s;
            return fail(SS_ERR_INVALID, "jl_dim %d outside 1..256", k);
        }
        *out = s;
        return SS_OK;
    } catch (const std::bad_alloc &) {
        return fail(SS_ERR_NOMEM, "host allocation failed");
    }
}

int ssh_sparsifier_destroy(ssh_sparsifier *s) {
    delete s;
    return SS_OK;
}

int ssh_sparsifier_insert_many(ssh_sparsifier *s, int64_t count, const int32_t *v1, const int32_t *v2,
                               const double *value) {
    if (!s || (count > 0 && (!v1 || !v2 || !value))) return fail(SS_ERR_INVALID, "null argument");
    try {
        for (int64_t i = 0; i < count; i++) {
            int rc = s->sp.insert(v1[i], v2[i], value[i]);
            if (rc != SS_OK) return rc;
        }
        return SS_OK;
    } catch (const std::bad_alloc &) {
        return fail(SS_ERR_NOMEM, "host allocation failed");
    }
}

int ssh_sparsifier_form_laplacian(ssh_sparsifier *s, int check) {
    if (!s) return fail(SS_ERR_INVALID, "null sparsifier");
    return s->sp.form_laplacian(check != 0);
}

int ssh_sparsifier_check_diagonal(const ssh_sparsifier *s) {
    if (!s) return fail(SS_ERR_INVALID, "null sparsifier");
    return s->sp.check_diagonal();
}

int ssh_sparsifier_num_edges(const ssh_sparsifier *s, int64_t *out) {
    if (!s || !out) return fail(SS_ERR_INVALID, "null argument");
    *out = (int64_t)s->sp.num_edges();
    return SS_OK;
}

int ssh_sparsifier_sparsify(ssh_sparsifier *s, int check, int64_t *before, int64_t *after) {
    if (!s) return fail(SS_ERR_INVALID, "null sparsifier");
    try {
        spec_spars::SparsificationStats st;
        int rc = s->sp.sparsify(check != 0, &st);
        if (before) *before = (int64_t)st.original_num_edges;
        if (after) *after = (int64_t)st.final_num_edges;
        return rc;
    } catch (const std::bad_alloc &) {
        return fail(SS_ERR_NOMEM, "host allocation failed");
    }
}

int ssh_sparsifier_sample_apply(ssh_sparsifier *s, const double *probs, int64_t len, int64_t *before,
                                int64_t *after) {
    if (!s || (len > 0 && !probs)) return fail(SS_ERR_INVALID, "null argument");
    try {
        std::vector<double> p(probs, probs + len);
        spec_spars::Triplet rw;
        int rc = s->sp.sample_and_reweight(p, rw);
        if (rc != SS_OK) return rc;
        spec_spars::SparsificationStats st;
        rc = s->sp.apply_reweightings(rw, false, &st);
        if (before) *before = (int64_t)st.original_num_edges;
        if (after) *after = (int64_t)st.final_num_edges;
        return rc;
    } catch (const std::bad_alloc &) {
        return fail(SS_ERR_NOMEM, "host allocation failed");
    }
}

int ssh_sparsifier_laplacian(const ssh_sparsifier *s, int64_t *nnz, int32_t *colptr, int32_t *rowidx,
                             double *vals) {
    if (!s) return fail(SS_ERR_INVALID, "null sparsifier");
    const auto &A = s->sp.current_laplacian;
    if (nnz) *nnz = (int64_t)A.nnz();
    if (colptr) memcpy(colptr, A.colptr.data(), sizeof(int32_t) * A.colptr.size());
    if (rowidx) memcpy(rowidx, A.rowidx.data(), sizeof(int32_t) * A.nnz());
    if (vals) memcpy(vals, A.vals.data(), sizeof(double) * A.nnz());
    return SS_OK;
}

int ssh_sparsifier_info(const ssh_sparsifier *s, double *beta, int32_t *jl_dim, int64_t *threshold) {
    if (!s) return fail(SS_ERR_INVALID, "null sparsifier");
    if (beta) *beta = s->sp.beta;
    if (jl_dim) *jl_dim = s->sp.jl_dim;
    if (threshold) *threshold = (int64_t)s->sp.threshold;
    return SS_OK;
}

int ssh_sparsifier_sketch(const ssh_sparsifier *s, double *Y) {
    if (!s || !Y) return fail(SS_ERR_INVALID, "null argument");
    std::vector<double> y;
    int rc = s->sp.jl_sketch(y);
    if (rc == SS_OK) memcpy(Y, y.data(), sizeof(double) * y.size());
    return rc;
}

}  // extern "C"

extern "C" int ssh_stream_run(const ssh_stream *st, ss_context *ctx, const ssh_params *p, int test, int writeout,
                              const char *output_file, ssh_sparsifier **out, int64_t *before, int64_t *after) {
    if (!st || !out) return fail(SS_ERR_INVALID, "null argument");
    try {
        SparsifierParameters q = to_params(p);
        if (output_file) q.output_file = output_file;
        ssh_sparsifier *s = new ssh_sparsifier{Sparsifier(ctx, st->is.num_nodes, q)};
        spec_spars::SparsificationStats stats;
        const int rc = st->is.run_stream(s->sp, test != 0, writeout != 0, &stats);
        if (rc != SS_OK) {
            delete s;
            return rc;
        }
        if (before) *before = (int64_t)stats.original_num_edges;
        if (after) *after = (int64_t)stats.final_num_edges;
        *out = s;
        return SS_OK;
    } catch (const std::bad_alloc &) {
        return fail(SS_ERR_NOMEM, "host allocation failed");
    }
}

extern "C" int ssh_sparsifier_triggered(const ssh_sparsifier *s, int64_t *count) {
    if (!s || !count) return fail(SS_ERR_INVALID, "null argument");
    *count = s->sp.triggered;
    return SS_OK;
}
