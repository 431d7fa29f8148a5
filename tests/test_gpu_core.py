"""GPU parity of the HIP kernels against the oracle, through the C ABI.

SpMV rows are summed in the reference's stored order, so SpMV and the
factorisations are compared bit for bit; reductions use a different (tree)
order and are compared to 1e-12 relative.  The north-star bar for factors is
1e-12 relative; the sweeps and factors are expected to be bit-identical."""
import sys

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


MATS = {
    "lap64": lambda g: g.laplacian2d(64, 64),
    "lap_rect": lambda g: g.laplacian2d(37, 101),
    "p3d32": lambda g: g.poisson3d(32),
    "p3d_rect": lambda g: g.poisson3d(40, 23, 17),
    "vc5": lambda g: g.varcoef5(70, 55),
    "cd27": lambda g: g.convdiff27(14, 13, 12),
}


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("name", sorted(MATS))
def test_spmv_bitexact(g, name):
    say(f"spmv {name}")
    S, I, V = MATS[name](g)
    A = o.CSR(S, I, V)
    Ag = g.Matrix(S, I, V)
    x = g.rng_uniform(1, 1, A.n)
    w0 = g.rng_uniform(1, 2, A.n)
    xv, wv, bv, rv = g.Vector(data=x), g.Vector(data=w0), g.Vector(data=w0), g.Vector(A.n)
    g.spmv(Ag, xv, wv, g.ASSIGN)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x))
    wv.upload(w0)
    g.spmv(Ag, xv, wv, g.PLUS)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x, w0, 1))
    wv.upload(w0)
    g.spmv(Ag, xv, wv, g.MINUS)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x, w0, 2))
    g.residual(Ag, xv, bv, rv)
    np.testing.assert_array_equal(rv.numpy(), o.residual(A, x, w0))
    q, d = o.fused_spmv_dot(A, x)
    qv = g.Vector(A.n)
    dg = g.spmv_dot(Ag, xv, qv)
    np.testing.assert_array_equal(qv.numpy(), q)
    assert abs(dg - d) <= 1e-12 * abs(d)


def mixed_width_matrix(n=70000, seed=5):
    """Rows with near entries only (16-bit column deltas), far entries in every
    third slice of 64 rows (32-bit columns), deltas of exactly ±32767 (still
    16-bit) and +32768 (32-bit), empty rows and ragged lengths."""
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(n):
        if r % 97 == 5:
            rows.append([])
            continue
        c = {r}
        if r > 0:
            c.add(r - 1)
        if r + 1 < n:
            c.add(r + 1)
        if (r // 64) % 3 == 0:
            c.add((r + 40000) % n)
        if r % 5 == 0 and r - 32767 >= 0:
            c.add(r - 32767)
        if r % 7 == 0 and r + 32768 < n:
            c.add(r + 32768)
        if r % 11 == 0 and r + 32767 < n:
            c.add(r + 32767)
        rows.append(sorted(c))
    S = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int64)
    I = np.array([j for c in rows for j in c], np.int64)
    V = rng.standard_normal(len(I))
    return S, I, V


def test_spmv_mixed_index_widths(g):
    say("spmv mixed 16/32-bit slices")
    S, I, V = mixed_width_matrix()
    A = o.CSR(S, I, V)
    Ag = g.Matrix(S, I, V)
    x = g.rng_uniform(4, 1, A.n)
    w0 = g.rng_uniform(4, 2, A.n)
    xv, wv, rv = g.Vector(data=x), g.Vector(data=w0), g.Vector(A.n)
    g.spmv(Ag, xv, wv, g.ASSIGN)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x))
    wv.upload(w0)
    g.spmv(Ag, xv, wv, g.MINUS)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x, w0, 2))
    g.residual(Ag, xv, g.Vector(data=w0), rv)
    np.testing.assert_array_equal(rv.numpy(), o.residual(A, x, w0))
    q, d = o.fused_spmv_dot(A, x)
    qv = g.Vector(A.n)
    dg = g.spmv_dot(Ag, xv, qv)
    np.testing.assert_array_equal(qv.numpy(), q)
    assert abs(dg - d) <= 1e-12 * abs(d)


def banded_ragged_matrix(n, maxlen, empty, seed):
    """Columns within +-100 of the row (16-bit offsets), 0..maxlen entries per
    row, the rows in `empty` without entries."""
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(n):
        if r in empty:
            rows.append([])
            continue
        lo, hi = max(0, r - 100), min(n, r + 101)
        k = int(rng.integers(1, maxlen + 1))
        rows.append(sorted(rng.choice(np.arange(lo, hi), size=min(k, hi - lo), replace=False)))
    S = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int64)
    I = np.array([j for c in rows for j in c], np.int64)
    V = rng.standard_normal(len(I))
    return S, I, V


@pytest.mark.parametrize("case", ["w4", "w8", "w16", "empty"])
def test_spmv_pipelined_shapes(g, case):
    # the pipelined SpMV (all slices 16-bit, rows of <= 16 entries): whole
    # empty slices, an empty last slice, ragged rows, every mode and the fused dot
    say(f"spmv pipelined {case}")
    n = 5000
    if case == "empty":
        S, I, V = np.zeros(301, np.int64), np.zeros(0, np.int64), np.zeros(0)
        n = 300
    else:
        maxlen = {"w4": 4, "w8": 7, "w16": 13}[case]
        empty = set(range(64, 128)) | set(range(4992, 5000)) | {7, 300, 301}
        S, I, V = banded_ragged_matrix(n, maxlen, empty, seed=len(case))
    A = o.CSR(S, I, V)
    Ag = g.Matrix(S, I, V)
    x = g.rng_uniform(6, 1, n)
    w0 = g.rng_uniform(6, 2, n)
    xv, wv, rv = g.Vector(data=x), g.Vector(data=w0), g.Vector(n)
    g.spmv(Ag, xv, wv, g.ASSIGN)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x))
    for mode, m in ((g.PLUS, 1), (g.MINUS, 2)):
        wv.upload(w0)
        g.spmv(Ag, xv, wv, mode)
        np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x, w0, m))
    g.residual(Ag, xv, g.Vector(data=w0), rv)
    np.testing.assert_array_equal(rv.numpy(), o.residual(A, x, w0))
    q, d = o.fused_spmv_dot(A, x)
    qv = g.Vector(n)
    dg = g.spmv_dot(Ag, xv, qv)
    np.testing.assert_array_equal(qv.numpy(), q)
    assert abs(dg - d) <= 1e-12 * max(abs(d), 1e-300)


def test_spmv_more_rows_than_columns(g):
    say("spmv 300 x 40")
    rng = np.random.default_rng(9)
    m, n = 300, 40
    rows = [sorted(rng.choice(n, size=int(rng.integers(0, 6)), replace=False)) for _ in range(m)]
    S = np.concatenate([[0], np.cumsum([len(c) for c in rows])]).astype(np.int64)
    I = np.array([j for c in rows for j in c], np.int64)
    V = rng.standard_normal(len(I))
    Ag = g.Matrix(S, I, V, ncols=n)
    x = rng.standard_normal(n)
    w = g.Vector(m)
    g.spmv(Ag, g.Vector(data=x), w)
    ref = np.zeros(m)
    for r in range(m):   # sequential row sums in stored order, as the oracle
        acc = 0.0
        for e in range(S[r], S[r + 1]):
            acc += V[e] * x[I[e]]
        ref[r] = acc
    np.testing.assert_array_equal(w.numpy(), ref)


def test_matrix_vector_product_reference_stencil(g):
    # T-mtl/matrix_vector_product_test.cpp:56-89
    S, I, V = g.laplacian2d(5, 5)
    A = g.Matrix(S, I, V)
    v = o.CSR(S, I, V).dense()[12].copy()
    vv, w = g.Vector(data=v), g.Vector(25)
    g.spmv(A, vv, w)
    r = w.numpy()
    assert (r[12], r[13], r[14], r[15], r[17], r[18], r[22]) == (20, -8, 1, 0, -8, 2, 1)
    g.spmv(A, vv, w, g.PLUS)
    assert w.numpy()[12] == 40
    g.spmv(A, vv, w, g.MINUS)
    assert w.numpy()[18] == 2


@pytest.mark.parametrize("n", [1, 5, 1000, 123457])
def test_level1(g, n):
    say(f"level1 n={n}")
    x = g.rng_uniform(2, 1, n)
    y = g.rng_uniform(2, 2, n)
    xv, yv = g.Vector(data=x), g.Vector(data=y)
    assert abs(g.dot(xv, yv) - o.dot(x, y)) <= 1e-12 * np.sum(np.abs(x * y))
    assert abs(g.unary_dot(xv) - o.unary_dot(x)) <= 1e-12 * o.unary_dot(x)
    assert abs(g.two_norm(xv) - o.two_norm(x)) <= 1e-12 * o.two_norm(x)
    g.axpy(0.37, xv, yv)
    np.testing.assert_array_equal(yv.numpy(), y + 0.37 * x)
    yv.upload(y)
    g.xpay(xv, -1.3, yv)
    np.testing.assert_array_equal(yv.numpy(), x + (-1.3) * y)
    yv.upload(y)
    rr = g.axpy_unary_dot(0.7, xv, yv)
    r_ref, rr_ref = o.fused_axpy_unary_dot(0.7, x, y)
    np.testing.assert_array_equal(yv.numpy(), r_ref)
    assert abs(rr - rr_ref) <= 1e-12 * rr_ref
    yv.fill(2.5)
    assert np.all(yv.numpy() == 2.5)


def test_reductions_deterministic(g):
    x = g.rng_normal(4, 1, 3_000_001)
    xv = g.Vector(data=x)
    vals = {g.dot(xv, xv) for _ in range(5)} | {g.unary_dot(xv) for _ in range(3)}
    assert len(vals) == 1


def _ilu_storage(A, L, U):
    fac = np.zeros(A.nnz)
    for r in range(A.n):
        a = A.starts[r]
        nl = L.starts[r + 1] - L.starts[r]
        fac[a:a + nl] = L.data[L.starts[r]:L.starts[r + 1]]
        fac[a + nl:A.starts[r + 1]] = U.data[U.starts[r]:U.starts[r + 1]]
    return fac


@pytest.mark.parametrize("name", ["lap64", "lap_rect", "p3d32", "p3d_rect", "vc5"])
def test_ic0_factor_and_solve(g, name):
    say(f"ic0 {name}")
    S, I, V = MATS[name](g)
    A = o.CSR(S, I, V)
    st, U, L = o.ic0_factor(A)
    Ag = g.Matrix(S, I, V)
    P = g.PC(Ag, g.PC_IC0)
    Sg, Ig, Vg = P.factor(1)
    np.testing.assert_array_equal(Sg, U.starts)
    np.testing.assert_array_equal(Ig, U.indices)
    assert relerr(Vg, U.data) <= 1e-12
    exact = np.array_equal(Vg, U.data)
    Sl, Il, Vl = P.factor(0)
    np.testing.assert_array_equal(Il, L.indices)
    assert relerr(Vl, L.data) <= 1e-12
    x = g.rng_uniform(5, 1, A.n)
    ref = o.ic0_solve(L, U, x)
    xv, yv = g.Vector(data=x), g.Vector(A.n)
    for rep in range(3):                     # repeated applications reuse the exchange buffers
        P.solve(xv, yv)
        got = yv.numpy()
        if exact:
            np.testing.assert_array_equal(got, ref)
        assert relerr(got, ref) <= 1e-12
    P.adjoint_solve(xv, yv)                  # IC(0): adjoint_solve == solve (ic_0.hpp:97-103)
    assert relerr(yv.numpy(), ref) <= 1e-12
    say(f"  stats {P.stats()} factor bit-exact={exact}")


@pytest.mark.parametrize("name", ["lap64", "p3d_rect", "vc5", "cd27"])
def test_ilu0_factor_and_solves(g, name):
    say(f"ilu0 {name}")
    S, I, V = MATS[name](g)
    A = o.CSR(S, I, V)
    st, L, U = o.ilu0_factor(A)
    Ag = g.Matrix(S, I, V)
    P = g.PC(Ag, g.PC_ILU0)
    _, Il, Vl = P.factor(0)
    _, Iu, Vu = P.factor(1)
    np.testing.assert_array_equal(Il, L.indices)
    np.testing.assert_array_equal(Iu, U.indices)
    assert relerr(Vl, L.data) <= 1e-12 and relerr(Vu, U.data) <= 1e-12
    x = g.rng_uniform(6, 1, A.n)
    xv, yv = g.Vector(data=x), g.Vector(A.n)
    P.solve(xv, yv)
    assert relerr(yv.numpy(), o.ilu0_solve(L, U, x)) <= 1e-12
    P.adjoint_solve(xv, yv)
    assert relerr(yv.numpy(), o.ilu0_adjoint_solve(L, U, x)) <= 1e-12
    P.solve(xv, yv)
    assert relerr(yv.numpy(), o.ilu0_solve(L, U, x)) <= 1e-12


@pytest.mark.parametrize("n,extra,long_rows", [(240, 12, False), (300, 45, False), (300, 70, True)])
def test_ilu0_factor_general_patterns(g, n, extra, long_rows):
    # Random patterns as T-itl/ilu_0_test.cpp builds them: rows of up to 64
    # entries take the wavefront-per-row factorisation, longer ones the
    # thread-per-row one; both apply the same operations in the same order
    # as the oracle (ilu_0.hpp:63-78), so the factors match it exactly.
    from test_oracle import _rand_sparse
    say(f"ilu0 random n={n} extra={extra}")
    A, _ = _rand_sparse(n, seed=n + extra, extra=extra)
    assert (max(np.diff(A.starts)) > 64) == long_rows
    st, L, U = o.ilu0_factor(A)
    assert st == o.OK
    P = g.PC(g.Matrix(A.starts, A.indices, A.data), g.PC_ILU0)
    for which, ref in ((0, L), (1, U)):
        S, I, V = P.factor(which)
        np.testing.assert_array_equal(S, ref.starts)
        np.testing.assert_array_equal(I, ref.indices)
        np.testing.assert_array_equal(V, ref.data)
    x = g.rng_uniform(8, 1, A.n)
    xv, yv = g.Vector(data=x), g.Vector(A.n)
    P.solve(xv, yv)
    np.testing.assert_array_equal(yv.numpy(), o.ilu0_solve(L, U, x))


def test_ic0_golden_float_factors(g):
    # T-itl/ic_0_test.cpp:22-36 on the GPU: ic_0<compressed2D<double>, float>
    S, I, V = g.laplacian2d(3, 3)
    P = g.PC(g.Matrix(S, I, V), g.PC_IC0, factor_dtype=np.float32)
    xv, yv = g.Vector(data=np.arange(9.0)), g.Vector(9)
    P.solve(xv, yv)
    yc = np.array([1.03194, 1.60198, 1.45357, 2.5258, 3.55386, 3.16, 2.91787, 4.09338, 3.81335])
    assert np.linalg.norm(yv.numpy() - yc) < 1e-4
    A = o.CSR(S, I, V)
    _, U, L = o.ic0_factor(A)
    np.testing.assert_array_equal(yv.numpy(), o.ic0_solve(L, U, np.arange(9.0), factor_dtype=np.float32))


def test_float_factor_float_vector_sweeps(g):
    S, I, V = g.poisson3d(20)
    A = o.CSR(S, I, V)
    _, U, L = o.ic0_factor(A)
    P = g.PC(g.Matrix(S, I, V), g.PC_IC0, factor_dtype=np.float32)
    x = g.rng_uniform(7, 1, A.n)
    xv = g.Vector(data=x.astype(np.float32), dtype=np.float32)
    yv = g.Vector(A.n, dtype=np.float32)
    P.solve(xv, yv)
    ref, _ = g.analysis_pc_solve(g.PC_IC0, S, I, U.data, x.astype(np.float32), factor_dtype=np.float32,
                                 vector_dtype=np.float32)
    np.testing.assert_array_equal(yv.numpy(), ref)
    assert relerr(yv.numpy().astype(float), o.ic0_solve(L, U, x)) < 1e-5


def test_diagonal_pc(g):
    S, I, V = g.varcoef5(9)
    A = o.CSR(S, I, V)
    P = g.PC(g.Matrix(S, I, V), g.PC_DIAGONAL)
    x = g.rng_uniform(1, 3, A.n)
    xv, yv = g.Vector(data=x), g.Vector(A.n)
    P.solve(xv, yv)
    np.testing.assert_array_equal(yv.numpy(), (1.0 / np.diag(A.dense())) * x)


def test_errors(g):
    # non-square, missing diagonal, unsorted rows, size mismatch
    with pytest.raises(g.M4GError) as e:
        g.Matrix(np.array([0, 2, 3]), np.array([1, 0, 1]), np.ones(3))
    assert e.value.status == g.ERR_LOGIC
    M = g.Matrix(np.array([0, 1, 2]), np.array([1, 0]), np.ones(2))
    with pytest.raises(g.M4GError) as e:
        g.PC(M, g.PC_IC0)
    assert e.value.status == g.ERR_MISSING_DIAGONAL
    R = g.Matrix(np.array([0, 1, 2]), np.array([0, 1]), np.ones(2), ncols=3)
    with pytest.raises(g.M4GError) as e:
        g.PC(R, g.PC_ILU0)
    assert e.value.status == g.ERR_NOT_SQUARE
    with pytest.raises(g.M4GError) as e:
        g.dot(g.Vector(3), g.Vector(4))
    assert e.value.status == g.ERR_SIZE
