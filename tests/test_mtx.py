"""MatrixMarket input/output and the stream driver's host side
(src/utils.rs:149-191, src/stream.rs:18-64), checked against scipy's reader
on fixtures written here."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

from conftest import GOLDEN


def _scipy_csc(path):
    A = scipy.io.mmread(path)
    A = sp.csc_matrix(A) if not sp.issparse(A) else A.tocsc()
    A.sum_duplicates()
    A.sort_indices()
    return A


@pytest.mark.parametrize("name", ["sym_dup.mtx", "skew_pattern.mtx"])
def test_read_mtx_matches_scipy(name):
    """Symmetric storage mirrored, duplicates summed, pattern values 1.0."""
    import spec_spars_amd as ss
    path = os.path.join(GOLDEN, name)
    rows, cols, cp, ri, va = ss.read_mtx(path)
    A = _scipy_csc(path)
    assert (rows, cols) == A.shape
    # scipy drops nothing either: the explicit zero (6,5) stays stored
    assert np.array_equal(cp, A.indptr) and np.array_equal(ri, A.indices) and np.array_equal(va, A.data)


def test_read_mtx_symmetric_details():
    import spec_spars_amd as ss
    rows, cols, cp, ri, va = ss.read_mtx(os.path.join(GOLDEN, "sym_dup.mtx"))
    A = sp.csc_matrix((va, ri, cp), shape=(rows, cols)).toarray()
    assert np.array_equal(A, A.T)
    assert A[2, 0] == 0.75 and A[0, 2] == 0.75          # (3,1) twice: 0.25 + 0.5
    assert A[0, 0] == 4.5 and A[6, 6] == 9.0            # diagonal once, not mirrored
    assert len(va) == 2 * 9 + 2                         # 9 distinct off-diagonal pairs, 2 diagonals


def test_write_read_roundtrip(tmp_path):
    import spec_spars_amd as ss
    rng = np.random.default_rng(1)
    A = sp.random(40, 40, density=0.15, random_state=2, format="csc")
    A.data = rng.standard_normal(A.nnz) * 10.0 ** rng.uniform(-12, 12, A.nnz)
    A.sort_indices()
    p = tmp_path / "a.mtx"
    ss.write_mtx(p, A.indptr, A.indices, A.data)
    rows, cols, cp, ri, va = ss.read_mtx(p)
    assert (rows, cols) == (40, 40)
    assert np.array_equal(cp, A.indptr) and np.array_equal(ri, A.indices) and np.array_equal(va, A.data)
    B = _scipy_csc(p)
    assert np.array_equal(B.data, A.data)


def test_read_mtx_errors(tmp_path):
    import spec_spars_amd as ss
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n")
    with pytest.raises(ss.SSError):
        ss.read_mtx(bad)
    bad.write_text("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n")
    with pytest.raises(ss.SSError):
        ss.read_mtx(bad)                                # fewer entries than announced
    bad.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n")
    with pytest.raises(ss.SSError):
        ss.read_mtx(bad)                                # index out of range
    with pytest.raises(ss.SSError):
        ss.read_mtx(tmp_path / "missing.mtx")


def test_input_stream_host_side():
    """InputStream::new: square, diagonal zeroed (kept stored), the
    diagonal positions the file lacks counted."""
    import spec_spars_amd as ss
    s = ss.InputStream(os.path.join(GOLDEN, "sym_dup.mtx"), "sym_dup")
    assert s.num_nodes == 7 and s.stored == 20 and s.diag_zeros == 5
    with pytest.raises(ss.SSError):
        ss.InputStream(os.path.join(GOLDEN, "skew_pattern.mtx"))   # 4 x 5: not square
