// MatrixMarket input and output, and the stream driver that feeds a file to
// the sparsifier (src/utils.rs:149-154, :189-191; src/stream.rs:18-64).
//
// read_mtx restates sprs' read_matrix_market + to_csc: a coordinate file
// (real, integer or pattern values; general, symmetric or skew-symmetric
// storage) becomes a CSC matrix with rows sorted inside each column,
// symmetric storage mirrored (every off-diagonal (i, j) also at (j, i),
// negated for skew-symmetric), and duplicate entries summed in file order.
// Explicit zeros are kept.  Pattern files read every value as 1.0 (the
// reference's own pattern reader is marked broken, src/utils.rs:166).
//
// InputStream::new zeroes the stored diagonal (src/stream.rs:23-28);
// run_stream inserts every stored (row, col, value) with row < col in CSC
// order -- column ascending, row ascending -- then sparsifies
// (src/stream.rs:41-64), and optionally writes the sparsified Laplacian.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spec_spars.hpp"
#include "../../include/spec_spars_host.h"
#include "ss_internal.h"

namespace spec_spars {

using ss::fail;

namespace {

struct Reader {
    FILE *f = nullptr;
    std::vector<char> buf;
    size_t pos = 0, len = 0;
    bool eof = false;
    explicit Reader(FILE *fp) : f(fp), buf(1 << 20) {}
    bool fill() {
        if (eof) return false;
        len = fread(buf.data(), 1, buf.size(), f);
        pos = 0;
        if (len == 0) eof = true;
        return len > 0;
    }
    // next line without its newline; false at end of file
    bool line(std::string &out) {
        out.clear();
        for (;;) {
            if (pos >= len && !fill()) return !out.empty();
            char *s = buf.data() + pos;
            char *nl = (char *)memchr(s, '\n', len - pos);
            if (nl) {
                out.append(s, (size_t)(nl - s));
                pos = (size_t)(nl - buf.data()) + 1;
                if (!out.empty() && out.back() == '\r') out.pop_back();
                return true;
            }
            out.append(s, len - pos);
            pos = len;
        }
    }
};

std::string lower(std::string s) {
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

}  // namespace

int read_mtx(const std::string &path, MtxMatrix &out) {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) return fail(SS_ERR_INVALID, "read_mtx: cannot open %s (%s)", path.c_str(), std::strerror(errno));
    Reader rd(fp);
    std::string ln;
    auto done = [&](int rc) {
        std::fclose(fp);
        return rc;
    };
    if (!rd.line(ln)) return done(fail(SS_ERR_INVALID, "read_mtx: %s is empty", path.c_str()));
    char banner[64] = {0}, object[64] = {0}, format[64] = {0}, field[64] = {0}, symmetry[64] = {0};
    if (std::sscanf(ln.c_str(), "%63s %63s %63s %63s %63s", banner, object, format, field, symmetry) != 5 ||
        lower(banner) != "%%matrixmarket" || lower(object) != "matrix")
        return done(fail(SS_ERR_INVALID, "read_mtx: %s: not a MatrixMarket matrix header", path.c_str()));
    if (lower(format) != "coordinate")
        return done(fail(SS_ERR_INVALID, "read_mtx: %s: only the coordinate format is supported", path.c_str()));
    const std::string fld = lower(field), sym = lower(symmetry);
    const bool pattern = fld == "pattern";
    if (!pattern && fld != "real" && fld != "integer" && fld != "double")
        return done(fail(SS_ERR_INVALID, "read_mtx: %s: unsupported value field '%s'", path.c_str(), field));
    int mirror = 0;   // 1 symmetric, -1 skew-symmetric
    if (sym == "symmetric") mirror = 1;
    else if (sym == "skew-symmetric") mirror = -1;
    else if (sym != "general")
        return done(fail(SS_ERR_INVALID, "read_mtx: %s: unsupported symmetry '%s'", path.c_str(), symmetry));
    // size line (after comments and blank lines)
    long long rows = -1, cols = -1, nnz = -1;
    while (rd.line(ln)) {
        if (ln.empty() || ln[0] == '%') continue;
        if (std::sscanf(ln.c_str(), "%lld %lld %lld", &rows, &cols, &nnz) != 3)
            return done(fail(SS_ERR_INVALID, "read_mtx: %s: bad size line '%s'", path.c_str(), ln.c_str()));
        break;
    }
    if (rows < 0 || cols < 0 || nnz < 0 || rows > INT32_MAX || cols > INT32_MAX)
        return done(fail(SS_ERR_INVALID, "read_mtx: %s: bad or missing size line", path.c_str()));
    std::vector<int32_t> ti, tj;
    std::vector<double> tv;
    const size_t cap = (size_t)nnz * (mirror ? 2 : 1);
    ti.reserve(cap);
    tj.reserve(cap);
    tv.reserve(cap);
    long long seen = 0;
    while (seen < nnz && rd.line(ln)) {
        const char *s = ln.c_str();
        while (*s == ' ' || *s == '\t') s++;
        if (*s == 0 || *s == '%') continue;
        char *e1 = nullptr, *e2 = nullptr, *e3 = nullptr;
        const long long i = std::strtoll(s, &e1, 10);
        const long long j = std::strtoll(e1, &e2, 10);
        double v = 1.0;
        if (!pattern) {
            v = std::strtod(e2, &e3);
            if (e3 == e2) return done(fail(SS_ERR_INVALID, "read_mtx: %s: entry %lld has no value", path.c_str(), seen + 1));
        }
        if (e1 == s || e2 == e1 || i < 1 || j < 1 || i > rows || j > cols)
            return done(fail(SS_ERR_INVALID, "read_mtx: %s: bad entry %lld: '%s'", path.c_str(), seen + 1, ln.c_str()));
        ti.push_back((int32_t)(i - 1));
        tj.push_back((int32_t)(j - 1));
        tv.push_back(v);
        if (mirror && i != j) {
            ti.push_back((int32_t)(j - 1));
            tj.push_back((int32_t)(i - 1));
            tv.push_back(mirror > 0 ? v : -v);
        }
        seen++;
    }
    if (seen != nnz)
        return done(fail(SS_ERR_INVALID, "read_mtx: %s: %lld entries announced, %lld read", path.c_str(), nnz, seen));
    std::fclose(fp);
    // to CSC: stable sort by (col, row), duplicates summed in file order
    std::vector<int64_t> idx(ti.size());
    for (size_t t = 0; t < idx.size(); t++) idx[t] = (int64_t)t;
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        return tj[a] != tj[b] ? tj[a] < tj[b] : ti[a] < ti[b];
    });
    out.rows = (int32_t)rows;
    out.cols = (int32_t)cols;
    out.colptr.assign((size_t)cols + 1, 0);
    out.rowidx.clear();
    out.vals.clear();
    for (size_t q = 0; q < idx.size();) {
        const int32_t c = tj[idx[q]], r = ti[idx[q]];
        double s = tv[idx[q]];
        size_t e = q + 1;
        while (e < idx.size() && tj[idx[e]] == c && ti[idx[e]] == r) s += tv[idx[e++]];
        out.rowidx.push_back(r);
        out.vals.push_back(s);
        out.colptr[(size_t)c + 1]++;
        q = e;
    }
    for (int64_t c = 0; c < cols; c++) out.colptr[(size_t)c + 1] += out.colptr[(size_t)c];
    return SS_OK;
}

int write_mtx(const std::string &path, int32_t rows, int32_t cols, const int32_t *colptr, const int32_t *rowidx,
              const double *vals) {
    FILE *fp = std::fopen(path.c_str(), "wb");
    if (!fp) return fail(SS_ERR_INVALID, "write_mtx: cannot open %s (%s)", path.c_str(), std::strerror(errno));
    const int64_t nnz = colptr[cols];
    std::fprintf(fp, "%%%%MatrixMarket matrix coordinate real general\n%d %d %lld\n", rows, cols, (long long)nnz);
    // entries in storage order: column ascending, rows as stored; values
    // with 17 significant digits (every double reads back exactly)
    char line[96];
    for (int32_t c = 0; c < cols; c++)
        for (int32_t q = colptr[c]; q < colptr[c + 1]; q++) {
            const int len = std::snprintf(line, sizeof line, "%d %d %.17g\n", rowidx[q] + 1, c + 1, vals[q]);
            std::fwrite(line, 1, (size_t)len, fp);
        }
    const bool ok = std::ferror(fp) == 0;
    if (std::fclose(fp) != 0 || !ok) return fail(SS_ERR_INVALID, "write_mtx: writing %s failed", path.c_str());
    return SS_OK;
}

int InputStream::open(const std::string &filename, const std::string &name) {
    // src/stream.rs:18-37
    MtxMatrix m;
    int rc = read_mtx(filename, m);
    if (rc != SS_OK) return rc;
    if (m.rows != m.cols)
        return fail(SS_ERR_INVALID, "InputStream: %s is %d x %d, not square", filename.c_str(), m.rows, m.cols);
    diag_zeros = 0;
    for (int32_t c = 0; c < m.cols; c++) {
        bool found = false;
        for (int32_t q = m.colptr[c]; q < m.colptr[c + 1]; q++)
            if (m.rowidx[q] == c) {
                m.vals[q] = 0.0;
                found = true;
            }
        if (!found) diag_zeros++;
    }
    input_matrix.n = m.rows;
    input_matrix.colptr.swap(m.colptr);
    input_matrix.rowidx.swap(m.rowidx);
    input_matrix.vals.swap(m.vals);
    num_nodes = m.rows;
    dataset_name = name;
    return SS_OK;
}

int InputStream::run_stream(Sparsifier &s, bool test, bool writeout, SparsificationStats *stats) const {
    // src/stream.rs:41-64: the upper triangle in CSC order
    const CscMatrix &A = input_matrix;
    for (int32_t c = 0; c < A.n; c++)
        for (int32_t q = A.colptr[c]; q < A.colptr[c + 1]; q++)
            if (A.rowidx[q] < c) {
                int rc = s.insert(A.rowidx[q], c, A.vals[q]);
                if (rc != SS_OK) return rc;
            }
    int rc = s.sparsify(test, stats);
    if (rc != SS_OK) return rc;
    if (test) {
        if ((rc = s.check_diagonal()) != SS_OK) return rc;
    }
    if (!test && writeout) {
        const CscMatrix &H = s.current_laplacian;
        if (s.params.output_file.empty()) return fail(SS_ERR_INVALID, "run_stream: writeout without an output file");
        return write_mtx(s.params.output_file, H.n, H.n, H.colptr.data(), H.rowidx.data(), H.vals.data());
    }
    return SS_OK;
}

}  // namespace spec_spars

// ---------------------------------------------------------------------------
// C wrapper
// ---------------------------------------------------------------------------
#include "host_handles.h"

extern "C" {

int ssh_mtx_read(const char *path, ssh_matrix **out) {
    if (!path || !out) return ss::fail(SS_ERR_INVALID, "null argument");
    try {
        ssh_matrix *m = new ssh_matrix();
        const int rc = spec_spars::read_mtx(path, m->m);
        if (rc != SS_OK) {
            delete m;
            return rc;
        }
        *out = m;
        return SS_OK;
    } catch (const std::bad_alloc &) {
        return ss::fail(SS_ERR_NOMEM, "host allocation failed");
    }
}

int ssh_matrix_info(const ssh_matrix *m, int32_t *rows, int32_t *cols, int64_t *nnz) {
    if (!m) return ss::fail(SS_ERR_INVALID, "null matrix");
    if (rows) *rows = m->m.rows;
    if (cols) *cols = m->m.cols;
    if (nnz) *nnz = (int64_t)m->m.rowidx.size();
    return SS_OK;
}

int ssh_matrix_csc(const ssh_matrix *m, int32_t *colptr, int32_t *rowidx, double *vals) {
    if (!m) return ss::fail(SS_ERR_INVALID, "null matrix");
    if (colptr) std::memcpy(colptr, m->m.colptr.data(), sizeof(int32_t) * m->m.colptr.size());
    if (rowidx) std::memcpy(rowidx, m->m.rowidx.data(), sizeof(int32_t) * m->m.rowidx.size());
    if (vals) std::memcpy(vals, m->m.vals.data(), sizeof(double) * m->m.vals.size());
    return SS_OK;
}

int ssh_matrix_destroy(ssh_matrix *m) {
    delete m;
    return SS_OK;
}

int ssh_mtx_write(const char *path, int32_t rows, int32_t cols, const int32_t *colptr, const int32_t *rowidx,
                  const double *vals) {
    if (!path || !colptr || rows < 0 || cols < 0) return ss::fail(SS_ERR_INVALID, "bad argument");
    if (colptr[cols] > 0 && (!rowidx || !vals)) return ss::fail(SS_ERR_INVALID, "null entry arrays");
    return spec_spars::write_mtx(path, rows, cols, colptr, rowidx, vals);
}

int ssh_stream_open(const char *path, const char *dataset_name, ssh_stream **out) {
    if (!path || !out) return ss::fail(SS_ERR_INVALID, "null argument");
    try {
        ssh_stream *s = new ssh_stream();
        const int rc = s->is.open(path, dataset_name ? dataset_name : "");
        if (rc != SS_OK) {
            delete s;
            return rc;
        }
        *out = s;
        return SS_OK;
    } catch (const std::bad_alloc &) {
        return ss::fail(SS_ERR_NOMEM, "host allocation failed");
    }
}

int ssh_stream_info(const ssh_stream *s, int32_t *num_nodes, int64_t *stored, int32_t *diag_zeros) {
    if (!s) return ss::fail(SS_ERR_INVALID, "null stream");
    if (num_nodes) *num_nodes = s->is.num_nodes;
    if (stored) *stored = (int64_t)s->is.input_matrix.nnz();
    if (diag_zeros) *diag_zeros = s->is.diag_zeros;
    return SS_OK;
}

int ssh_stream_destroy(ssh_stream *s) {
    delete s;
    return SS_OK;
}

}  // extern "C"
