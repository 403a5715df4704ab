// spec_spars.hpp -- C++ host layer mirroring the reference's sparsifier
// interface (src/sparsifier.rs, src/stream.rs) on top of the C ABI in
// spec_spars_hip.h.  Names, argument meanings and the order of operations
// follow the Rust code; panics and asserts become status codes (SS_*).
//
// The Laplacian lives on the host as CSC (like Sparsifier::current_laplacian);
// the sketch, the solve and the fused ER/sample kernel run on the GPU.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "spec_spars_hip.h"

namespace spec_spars {

// src/sparsifier.rs:146-210
struct SparsifierParameters {
    double epsilon = 0.5;
    int32_t beta_constant = 4;
    int32_t row_constant = 2;
    bool verbose = false;
    double jl_factor = 1.5;
    double jl_scaling_factor = 1.7320508075688772;   // sqrt(3)
    uint64_t sketch_seed = 1;
    uint64_t sampling_seed = 1;
    bool bench = false;
    bool sketch_uniform = true;
    std::string output_file;
    bool streaming = false;
    // solve settings (the reference hard-codes these in C++:
    // include/custom_cg.hpp:150-159, src/example.cc:99-100)
    uint64_t factor_seed = 0;
    int32_t ordering = SS_ORDER_ROUNDS;   // shallow solves (the natural order is a chain)
    int32_t precision = SS_PREC_FP64;
    int32_t stop_rule = SS_STOP_REF_ABS;
    double tol = 2e-7;
    int32_t max_iter = 400;
    // Device-resident pass: the Laplacian goes to the device once per pass
    // and the solution, probabilities, coins, sample and reweighted
    // Laplacian are computed there (only the kept edges and the new diagonal
    // come back, for the next host factor).  Results are the same bits as
    // the host-driven pass.
    bool resident = false;

    // SparsifierParameters::new_default (src/sparsifier.rs:182-209)
    static SparsifierParameters new_default(bool bench);
};

// src/sparsifier.rs:213-237
struct SparsificationStats {
    size_t original_num_edges = 0;
    size_t final_num_edges = 0;
    double sparsification_rate() const {
        return original_num_edges ? (double)final_num_edges / (double)original_num_edges : 0.0;
    }
};

// Square CSC matrix with int32 indices (sprs CsMatI<f64, i32>).
struct CscMatrix {
    int32_t n = 0;
    std::vector<int32_t> colptr{0};
    std::vector<int32_t> rowidx;
    std::vector<double> vals;
    size_t nnz() const { return rowidx.size(); }
    const double *get(int32_t row, int32_t col) const;
};

// src/sparsifier.rs:23-143
struct Triplet {
    int32_t num_nodes = 0;
    std::vector<int32_t> col_indices, row_indices;
    std::vector<double> diagonal;
    std::vector<double> values;

    explicit Triplet(int32_t n = 0) : num_nodes(n), diagonal((size_t)n, 0.0) {}
    // -value at (v1,v2) and (v2,v1); diagonal += value; self-loops skipped.
    int insert(int32_t v1, int32_t v2, double value);
    void process_diagonal();
    // triplets -> CSC, duplicates summed in insertion order
    CscMatrix to_csc() const;
    void delete_state();
};

// A + B with exact zeros dropped (sprs' add).
CscMatrix csc_add(const CscMatrix &a, const CscMatrix &b);

class Sparsifier {
public:
    // Sparsifier::new (src/sparsifier.rs:262-297); ctx runs the GPU work.
    Sparsifier(ss_context *ctx, int32_t num_nodes, const SparsifierParameters &p);

    int insert(int32_t v1, int32_t v2, double value);          // :357-376
    size_t num_edges() const;                                   // :346-353
    int form_laplacian(bool check);                             // :693-709
    int check_diagonal() const;                                 // :806-817
    // Solve + effective resistances + probabilities (:586-622).
    int get_probs(size_t length, const std::vector<double> &sketch_rowmajor, std::vector<double> &probs);
    // Coins and sampling on the GPU for given probabilities (:635-680).
    int sample_and_reweight(const std::vector<double> &probs, Triplet &reweightings);
    int apply_reweightings(Triplet &reweightings, bool report, SparsificationStats *stats);   // :682-691
    // The whole pass (:722-770).
    int sparsify(bool check, SparsificationStats *stats);
    // Passes the streaming trigger ran (insert: num_edges() > threshold,
    // :370-374).
    int64_t triggered = 0;
    // JL sketch of the incidence matrix of new_entries (:490-539), row-major n x jl_dim.
    int jl_sketch(std::vector<double> &Y) const;

    ss_context *ctx;
    int32_t num_nodes;
    Triplet new_entries;
    CscMatrix current_laplacian;
    size_t threshold;
    double epsilon;
    int32_t beta_constant, row_constant;
    double beta;
    bool verbose;
    double jl_factor;
    int32_t jl_dim;
    double jl_scaling_factor;
    uint64_t sketch_seed, sampling_seed;
    bool sketch_uniform, streaming;
    SparsifierParameters params;

private:
    // canonical upper-triangle edges of current_laplacian (:547-548)
    void upper_edges(std::vector<int32_t> &u, std::vector<int32_t> &v, std::vector<double> &lval) const;
    // the pass after the sketch and form_laplacian, on the device
    int sparsify_resident(const std::vector<double> &Y, SparsificationStats *stats);
    int reweight_from_mask(const std::vector<int32_t> &u, const std::vector<int32_t> &v,
                           const std::vector<double> &lval, const std::vector<double> &p,
                           const std::vector<uint8_t> &mask, Triplet &reweightings) const;
};

// A MatrixMarket matrix as CSC (rows sorted in each column).
struct MtxMatrix {
    int32_t rows = 0, cols = 0;
    std::vector<int32_t> colptr{0};
    std::vector<int32_t> rowidx;
    std::vector<double> vals;
};
// read_mtx (src/utils.rs:149-154): coordinate real / integer / pattern,
// general / symmetric (mirrored) / skew-symmetric; duplicates summed.
int read_mtx(const std::string &path, MtxMatrix &out);
// write_mtx (src/utils.rs:189-191): coordinate real general, storage order.
int write_mtx(const std::string &path, int32_t rows, int32_t cols, const int32_t *colptr, const int32_t *rowidx,
              const double *vals);

// src/stream.rs:8-64
struct InputStream {
    CscMatrix input_matrix;       // the file, diagonal zeroed
    int32_t num_nodes = 0;
    std::string dataset_name;
    int32_t diag_zeros = 0;       // diagonal positions the file does not store
    // InputStream::new (src/stream.rs:18-37)
    int open(const std::string &filename, const std::string &name);
    // run_stream (src/stream.rs:41-64) into a sparsifier built for
    // num_nodes vertices; writes s.params.output_file when writeout and not test
    int run_stream(Sparsifier &s, bool test, bool writeout, SparsificationStats *stats) const;
};

}  // namespace spec_spars
