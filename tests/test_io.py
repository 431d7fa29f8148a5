"""Matrix Market ingest and output (host only), against the reference's own
fixture files (T-mtl/matrix_market/, copied under tests/golden/matrix_market)
and the reference test's known answers (matrix_market_test.cpp), with SciPy's
independent reader as a cross-check of the symmetric / skew-symmetric /
pattern expansion."""
import os

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import mtl4gpu as g

HERE = os.path.dirname(os.path.abspath(__file__))
MM = os.path.join(HERE, "golden", "matrix_market")


def dense(S, I, V, nr, nc):
    return sp.csr_matrix((V, I, S), shape=(nr, nc)).toarray()


def scipy_dense(name):
    return np.asarray(scipy.io.mmread(os.path.join(MM, name)).todense())


def test_bcsstk01_real_symmetric():
    # matrix_market_test.cpp:54 reads it into compressed2D<double>
    S, I, V, nr, nc = g.mm_read(os.path.join(MM, "bcsstk01.mtx"))
    assert (nr, nc) == (48, 48)
    # 224 stored entries, mirrored off the diagonal
    lines = [l for l in open(os.path.join(MM, "bcsstk01.mtx")) if not l.startswith("%")][1:]
    diag = sum(1 for l in lines if l.split()[0] == l.split()[1])
    assert len(V) == 2 * 224 - diag
    for r in range(nr):
        assert np.all(np.diff(I[S[r]:S[r + 1]]) > 0)
    np.testing.assert_array_equal(dense(S, I, V, nr, nc), scipy_dense("bcsstk01.mtx"))


def test_pattern_files_have_zero_values():
    # which_value(pattern_type) == 0 (matrix_market.hpp:153)
    for name, sym in (("jgl009.mtx", False), ("bcspwr02.mtx", True)):
        S, I, V, nr, nc = g.mm_read(os.path.join(MM, name))
        assert np.all(V == 0)
        ref = scipy_dense(name) != 0
        pat = sp.csr_matrix((np.ones(len(V)), I, S), shape=(nr, nc)).toarray() != 0
        np.testing.assert_array_equal(pat, ref)
        if name == "jgl009.mtx":
            assert (nr, nc, len(V)) == (9, 9, 50)


def test_skew_symmetric_negates_the_mirror():
    S, I, V, nr, nc = g.mm_read(os.path.join(MM, "plskz362.mtx"))
    A = dense(S, I, V, nr, nc)
    np.testing.assert_array_equal(A, -A.T)
    np.testing.assert_array_equal(A, scipy_dense("plskz362.mtx"))


def test_dense_array_column_major():
    # matrix_market_test.cpp read_test: data(r, c) = 1 + 4 r + c from a
    # column-major array file
    S, I, V, nr, nc = g.mm_read(os.path.join(MM, "dense_read.mtx"))
    assert (nr, nc) == (2, 4)
    want = np.arange(1, 9, dtype=float).reshape(2, 4)
    np.testing.assert_array_equal(dense(S, I, V, nr, nc), want)
    assert len(V) == 8


def test_missing_file_is_file_not_found():
    with pytest.raises(g.M4GError) as e:
        g.mm_read(os.path.join(MM, "File_not_exist_test.mtx"))
    assert e.value.status == g.ERR_FILE


@pytest.mark.parametrize("text,status", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1.0 2.0\n", "format"),
    ("%%NotMarket matrix coordinate real general\n1 1 1\n1 1 1.0\n", "format"),
    ("%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1.0\n", "format"),
    ("%%MatrixMarket matrix coordinate real unknown\n1 1 1\n1 1 1.0\n", "format"),
    ("%%MatrixMarket matrix coordinate Real general\n1 1 1\n1 1 1.0\n", "format"),   # value tag is case-sensitive
])
def test_malformed_headers(tmp_path, text, status):
    p = tmp_path / "m.mtx"
    p.write_text(text)
    with pytest.raises(g.M4GError) as e:
        g.mm_read(str(p))
    assert e.value.status == g.ERR_FORMAT


def test_comments_case_and_duplicates(tmp_path):
    # comments and blank lines skipped; sparsity/symmetry case-insensitive;
    # of two entries at one position the later wins (update_store), also when
    # one of them is a mirror; integer values
    p = tmp_path / "m.mtx"
    p.write_text("%%MatrixMarket matrix Coordinate integer SYMMETRIC\n% a comment\n\n  % another\n"
                 "3 3 5\n1 1 4\n2 1 -1\n1 2 7\n3 3 2\n1 1 9\n")
    S, I, V, nr, nc = g.mm_read(str(p))
    A = dense(S, I, V, nr, nc)
    want = np.array([[9, 7, 0], [7, 0, 0], [0, 0, 2]], float)
    np.testing.assert_array_equal(A, want)


def test_write_then_read_round_trip(tmp_path):
    S, I, V, nr, nc = g.mm_read(os.path.join(MM, "bcsstk01.mtx"))
    p = tmp_path / "out.mtx"
    g.mm_write(str(p), S, I, V)
    lines = p.read_text().splitlines()
    # matrix_market_ostream's status and size lines (matrix_market.hpp:243, 325)
    assert lines[0] == "%%MatrixMarket  matrix coordinate real general"
    assert lines[1] == f"{nr} {nc} {len(V)}"
    r, c, v = lines[2].split()
    assert (int(r), int(c)) == (1, 1) and v == f"{V[0]:.16e}"
    S2, I2, V2, nr2, nc2 = g.mm_read(str(p))
    assert (nr2, nc2) == (nr, nc)
    np.testing.assert_array_equal(S2, S)
    np.testing.assert_array_equal(I2, I)
    np.testing.assert_array_equal(V2, V)     # 17 significant digits: exact for doubles
