"""The oracle's ell_matrix and sparse_banded restatements (layouts from the
inserters, products from mat_vec_mult.hpp:692-809) against the reference's
own tests and against the CSR product: both formats store the same entries
in the same per-row order (plus explicit zeros), so every product must equal
the CSR one bit for bit."""
import numpy as np
import pytest

import oracle_lib as o


def csr_from_entries(nrows, entries):
    """CSR of {(r, c): v}; rows sorted like compressed2D's inserter leaves them."""
    S, I, V = [0], [], []
    for r in range(nrows):
        row = sorted((c, v) for (rr, c), v in entries.items() if rr == r)
        I += [c for c, _ in row]
        V += [v for _, v in row]
        S.append(len(I))
    return o.CSR(np.array(S, np.int64), np.array(I, np.int64), np.array(V, np.float64))


def fill_matrix():
    # ell_matrix_test.cpp: fill_matrix
    e = {(0, 0): 2, (0, 1): 9, (1, 1): 1, (1, 2): 5, (1, 3): 5, (1, 4): 1, (2, 2): 6, (2, 3): 9,
         (3, 2): 2, (3, 3): 4, (4, 0): 7, (4, 4): 3}
    return csr_from_entries(5, {k: float(v) for k, v in e.items()})


def rectangle(nr, nc):
    # sparse_banded_matrix_test.cpp: rectangle_test, values 1, 2, ... in insertion order
    e, i = {}, 1
    for r in range(nr):
        if r < nc - 4:
            e[(r, r + 4)] = float(i); i += 1
        if r < nc:
            e[(r, r)] = float(i); i += 1
        if 2 <= r < nc + 2:
            e[(r, r - 2)] = float(i); i += 1
        if 4 <= r < nc + 4:
            e[(r, r - 4)] = float(i); i += 1
    return csr_from_entries(nr, e)


def lower_only():   # dr5: strictly lower bands only, first rows empty
    return csr_from_entries(5, {(2, 0): 1.0, (3, 1): 2.0, (4, 2): 3.0, (4, 0): 4.0})


def upper_only():   # dr6: strictly upper bands only, last rows empty
    return csr_from_entries(5, {(0, 2): 1.0, (1, 3): 2.0, (2, 4): 3.0, (0, 4): 4.0})


def test_ell_known_answers():
    A = fill_matrix()
    slots, stride, I, D = o.ell_layout(A)
    assert slots == 4 and stride == 32            # longest row: row 1; 5 rows rounded up to 32
    # row 0 = (0,2) (1,9) then two padding slots repeating column 1 with value 0
    assert [I[0 + k * stride] for k in range(4)] == [0, 1, 1, 1]
    assert [D[0 + k * stride] for k in range(4)] == [2, 9, 0, 0]
    # A[2][3] = 9, A[2][4] = 0 (ell_matrix_test.cpp:66-68)
    row2 = {I[2 + k * stride]: D[2 + k * stride] for k in range(slots) if D[2 + k * stride] != 0}
    assert row2.get(3) == 9.0 and 4 not in row2 and 0 not in row2
    x = np.arange(1.0, 6.0)                       # iota(x, 1)
    np.testing.assert_array_equal(o.ell_spmv(5, slots, stride, I, D, x), [20, 42, 54, 22, 22])


@pytest.mark.parametrize("make", [fill_matrix, lambda: o.laplacian_setup(3, 4), lambda: o.laplacian_setup(9, 7),
                                  lambda: rectangle(6, 11), lambda: rectangle(11, 6)])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ell_equals_csr(make, mode):
    A = make()
    ncols = int(A.indices.max()) + 1 if A.nnz else A.n
    slots, stride, I, D = o.ell_layout(A)
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, max(ncols, A.n))
    w0 = rng.uniform(-1, 1, A.n)
    ref = w0.copy()
    o.lib.orc_spmv(A.n, A.starts, A.indices, np.ascontiguousarray(A.data, np.float64), x, ref, mode)
    np.testing.assert_array_equal(o.ell_spmv(A.n, slots, stride, I, D, x, w0, mode), ref)


def test_ell_empty_rows_pad_with_column_zero():
    A = csr_from_entries(4, {(1, 2): 3.0, (1, 3): 1.0, (3, 0): 5.0})
    slots, stride, I, D = o.ell_layout(A)
    assert slots == 2
    assert I[0] == 0 and I[0 + stride] == 0 and D[0] == 0 and D[stride] == 0   # empty row 0
    assert I[3] == 0 and I[3 + stride] == 0 and D[3] == 5.0                  # row 3 padded with its column 0


def test_banded_laplacian_layout():
    # sparse_banded_matrix_test.cpp laplacian_test on 3 x 4: bands -4, -1, 0, 1, 4
    A = o.laplacian_setup(3, 4)
    B, D = o.banded_layout(A)
    np.testing.assert_array_equal(B, [-4, -1, 0, 1, 4])
    nb = len(B)
    assert D[0 * nb + 2] == 4.0 and D[0 * nb + 3] == -1.0 and D[0 * nb + 4] == -1.0   # A[0][0], [0][1], [0][4]
    assert D[1 * nb + 1] == -1.0 and D[4 * nb + 0] == -1.0                              # A[1][0], A[4][0]
    assert D[2 * nb + 4] == 0.0 or A.dense()[2, 6] == D[2 * nb + 4]


@pytest.mark.parametrize("make,ncols", [(lambda: rectangle(6, 11), 11), (lambda: rectangle(11, 6), 6),
                                        (lambda: rectangle(6, 5), 5), (lower_only, 5), (upper_only, 5),
                                        (lambda: o.laplacian_setup(3, 4), 12),
                                        (lambda: o.laplacian_setup(16, 9), 144)])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_banded_equals_csr(make, ncols, mode):
    # mat_vec_mult_test: v = A * w with w = 3 against the compressed2D product
    A = make()
    B, D = o.banded_layout(A)
    rng = np.random.default_rng(6)
    for x in (np.full(ncols, 3.0), rng.uniform(-1, 1, ncols)):
        w0 = rng.uniform(-1, 1, A.n)
        ref = w0.copy()
        o.lib.orc_spmv(A.n, A.starts, A.indices, np.ascontiguousarray(A.data, np.float64), x, ref, mode)
        got = o.banded_spmv(A.n, ncols, B, D, x, w0, mode)
        if mode != 0:
            # rows before the highest band starts and after the lowest ends are
            # not touched by += / -= (the CSR adds an empty row's 0 instead)
            lo, hi = B.min(), B.max()
            untouched = [r for r in range(A.n) if not (r + hi >= 0 and r + lo < ncols)]
            for r in untouched:
                assert got[r] == w0[r]
                ref[r] = w0[r]
        np.testing.assert_array_equal(got, ref)


def test_banded_rectangle_values():
    # rectangle_test(dr2) 6 x 11: row 0 holds (0,4)=1, (0,0)=2; row 2: (2,6)=5, (2,2)=6, (2,0)=7
    A = rectangle(6, 11)
    B, D = o.banded_layout(A)
    np.testing.assert_array_equal(B, [-4, -2, 0, 4])
    nb = len(B)
    assert list(D[0:nb]) == [0, 0, 2, 1]
    assert list(D[2 * nb:3 * nb]) == [0, 7, 6, 5]
    x = np.arange(11.0)
    np.testing.assert_array_equal(o.banded_spmv(6, 11, B, D, x)[:3], [1 * 4 + 2 * 0, 3 * 5 + 4 * 1, 5 * 6 + 6 * 2 + 7 * 0])


@pytest.mark.parametrize("m,n", [(2, 2), (2, 7), (4, 5), (9, 3), (13, 17), (30, 31)])
def test_poisson2d_dirichlet_operator(m, n):
    # matrix_free_1_test.cpp: the operator equals laplacian_setup's matrix times
    # v; on integer vectors every operation is exact, so the two agree bit for
    # bit although they add in different orders
    A = o.laplacian_setup(m, n)
    v = np.arange(m * n, dtype=np.float64)                   # iota(v)
    ref = o.spmv(A, v)
    np.testing.assert_array_equal(o.poisson2d_dirichlet_mult(m, n, v), ref)
    x = np.random.default_rng(m * 100 + n).uniform(-1, 1, m * n)
    np.testing.assert_allclose(o.poisson2d_dirichlet_mult(m, n, x), o.spmv(A, x), rtol=0, atol=1e-14)
    w0 = np.random.default_rng(1).uniform(-1, 1, m * n)
    np.testing.assert_array_equal(o.poisson2d_dirichlet_mult(m, n, v, w0, 1), w0 + ref)
    np.testing.assert_array_equal(o.poisson2d_dirichlet_mult(m, n, v, w0, 2), w0 - ref)
