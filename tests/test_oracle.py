"""Pins the CPU oracle (oracle/mtl4_oracle.cpp) against the reference's own
known-answer tests and against what the mathematics fixes.  Each test names the
reference test (libs/numeric/{itl,mtl}/test/...) or the identity it checks."""
import numpy as np
import pytest

import oracle_lib as o


def dense_ic0_textbook(A):
    """dense_ic_0 of T-itl/ic_0_cg_test.cpp:17-40 (full textbook IC(0) on the
    dense upper triangle, stored with the true diagonal)."""
    U = np.triu(A).copy()
    n = len(U)
    for k in range(n):
        dia = U[k, k] = np.sqrt(U[k, k])
        for i in range(k + 1, n):
            d = U[k, i] = U[k, i] / dia
            for j in range(k + 1, i + 1):
                if U[j, i] != 0.0:
                    U[j, i] -= d * U[k, j]
    return U


def dense_ilu0_textbook(A):
    """dense_ilu_0 of T-itl/ilu_0_test.cpp:31-50."""
    LU = A.copy()
    n = len(LU)
    for i in range(1, n):
        for k in range(i):
            LU[i, k] /= LU[k, k]
            for j in range(k + 1, n):
                if LU[i, j] != 0:
                    LU[i, j] -= LU[i, k] * LU[k, j]
    return LU


# ---------------------------------------------------------------- generators
def test_laplacian_setup_stencil():
    # T-mtl/laplacian_setup_test.cpp: 4 on the diagonal, -1 to E/S/W/N
    A = o.laplacian_setup(3, 4).dense()
    assert A.shape == (12, 12)
    for r in range(12):
        i, j = divmod(r, 4)
        assert A[r, r] == 4
        for (di, dj) in ((0, 1), (1, 0), (0, -1), (-1, 0)):
            ii, jj = i + di, j + dj
            if 0 <= ii < 3 and 0 <= jj < 4:
                assert A[r, ii * 4 + jj] == -1
    assert o.laplacian_setup(64, 64).nnz == 20224   # the 64x64 5-point system


# ---------------------------------------------------------------- SpMV
def test_matrix_vector_product_stencil():
    # T-mtl/matrix_vector_product_test.cpp:56-89 on the 5x5 grid
    A = o.laplacian_setup(5, 5)
    v = A.dense()[12].copy()
    w = o.spmv(A, v)
    assert (w[12], w[13], w[14], w[15], w[17], w[18], w[22]) == (20, -8, 1, 0, -8, 2, 1)
    w = o.spmv(A, v, w, mode=1)
    assert w[12] == 40 and w[18] == 4
    w = o.spmv(A, v, w, mode=2)
    assert w[12] == 20 and w[18] == 2


@pytest.mark.parametrize("m,n", [(7, 9), (16, 16)])
def test_spmv_laplacian_eigenvectors(m, n):
    # A v = lambda v for v = sin(pi i p/(m+1)) sin(pi j q/(n+1)), lambda = 4 - 2cos - 2cos
    A = o.laplacian_setup(m, n)
    i = np.arange(m)[:, None] + 1
    j = np.arange(n)[None, :] + 1
    for p, q in ((1, 1), (2, 3), (m, n)):
        v = (np.sin(np.pi * i * p / (m + 1)) * np.sin(np.pi * j * q / (n + 1))).ravel()
        lam = 4 - 2 * np.cos(np.pi * p / (m + 1)) - 2 * np.cos(np.pi * q / (n + 1))
        np.testing.assert_allclose(o.spmv(A, v), lam * v, atol=1e-13)


def test_residual_is_b_minus_Ax():
    A = o.laplacian_setup(6, 5)
    rng = np.random.default_rng(1)
    x, b = rng.standard_normal(30), rng.standard_normal(30)
    np.testing.assert_array_equal(o.residual(A, x, b), b - o.spmv(A, x))


# ---------------------------------------------------------------- reductions
def test_dot_known_answer():
    # T-mtl/dot_test.cpp: dot(1..9, 1..9) = 285
    u = np.arange(1, 10, dtype=float)
    assert o.dot(u, u) == 285.0


def test_unary_dot_and_two_norm():
    # T-mtl/unary_dot_test.cpp: unary_dot(u) == dot(u, u) == two_norm(u)^2
    u = np.arange(1, 12, dtype=float)
    assert o.unary_dot(u) == o.dot(u, u) == 506.0
    assert abs(o.two_norm(u) ** 2 - 506.0) < 1e-12


def test_reduction_orders():
    # dot<4> puts element i into partial i%4 and the tail into partial 0
    # (dot.hpp:136-147): a cancellation pattern that only that order gets exactly right
    a = np.array([1e16, 1.0, -1e16, 1.0, 3.0], dtype=float)
    b = np.ones(5)
    # ((1e16 + 3) + 1) + ((-1e16) + 1)
    t0 = 1e16 + 3.0
    exp = (t0 + 1.0) + (-1e16 + 1.0)
    assert o.dot(a, b) == exp
    # fused lazy_dot: ((t0 + t1) + t2) + t3 (dot_index_evaluator.hpp:30)
    A = o.CSR(np.arange(6), np.arange(5), np.ones(5))
    _, d = o.fused_spmv_dot(A, a)
    sq = a * a
    assert d == ((sq[0] + sq[4] + sq[1]) + sq[2]) + sq[3]


def test_fuse_test_values():
    # T-mtl/fuse_test.cpp:36-49: B = 2 I (6x6), v = 1 -> d = 12; r = 6, q = 2, alpha = 7.8 -> rho = 552.96
    B = o.CSR(np.arange(7), np.arange(6), np.full(6, 2.0))
    w, d = o.fused_spmv_dot(B, np.ones(6))
    assert d == 12.0 and np.all(w == 2.0)
    r, rho = o.fused_axpy_unary_dot(7.8, np.full(6, 2.0), np.full(6, 6.0))
    assert abs(rho - 552.96) < 1e-9 and np.allclose(r, -9.6)


# ---------------------------------------------------------------- IC(0)
def test_ic0_golden_solve():
    # T-itl/ic_0_test.cpp:22-36: ic_0<compressed2D<double>, float> on the 3x3 grid, solve(P, iota)
    A = o.laplacian_setup(3, 3)
    st, U, L = o.ic0_factor(A)
    assert st == o.OK
    y = o.ic0_solve(L, U, np.arange(9.0), factor_dtype=np.float32)
    yc = np.array([1.03194, 1.60198, 1.45357, 2.5258, 3.55386, 3.16, 2.91787, 4.09338, 3.81335])
    assert np.linalg.norm(y - yc) < 1e-4          # reference bound 1e-3; goldens carry 6 digits


@pytest.mark.parametrize("m,n", [(3, 3), (4, 5), (8, 8)])
def test_ic0_equals_dense_textbook_on_5pt(m, n):
    # T-itl/ic_0_cg_test.cpp:17-40 compares U[2][3] and 1/U[3][3]; on 5-point
    # grids (no triangles) every entry must agree
    A = o.laplacian_setup(m, n)
    st, U, L = o.ic0_factor(A)
    Ud = U.dense()
    Ut = dense_ic0_textbook(A.dense())
    off = ~np.eye(A.n, dtype=bool) & (Ud != 0)
    np.testing.assert_allclose(Ud[off], Ut[off], rtol=1e-14)
    np.testing.assert_allclose(np.diag(Ud), 1.0 / np.diag(Ut), rtol=1e-14)


def test_ic0_diagonal_recurrence_closed_form():
    # D_i = a_ii - sum_{k in N-(i)} a_ki^2 / D_k for 5-point stencils
    m = n = 6
    A = o.laplacian_setup(m, n)
    _, U, _ = o.ic0_factor(A)
    Ad = A.dense()
    D = np.zeros(A.n)
    for i in range(A.n):
        D[i] = Ad[i, i] - sum(Ad[k, i] ** 2 / D[k] for k in range(i) if Ad[k, i] != 0)
    np.testing.assert_allclose(np.diag(U.dense()), 1 / np.sqrt(D), rtol=1e-14)


def test_ic0_pattern_invariant():
    # U^T U == A on the pattern of A (triangle-free graph); L is exactly trans(U)
    A = o.laplacian_setup(7, 5)
    _, U, L = o.ic0_factor(A)
    Ud = U.dense().copy()
    np.fill_diagonal(Ud, 1 / np.diag(Ud))          # true diagonal
    M = Ud.T @ Ud
    Ad = A.dense()
    mask = Ad != 0
    np.testing.assert_allclose(M[mask], Ad[mask], rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(L.dense(), U.dense().T)


def test_ic0_trisolve_reference_matrix():
    # the 5x5 matrix of T-mtl/inverse_trisolve_test.cpp:20-64: upper part
    # {7, 8, 2, 4, 9} with (1,3)=2, (1,4)=3, diagonal inverted; recover x = 1..5
    Ufull = np.diag([7.0, 8, 2, 4, 9])
    Ufull[1, 3], Ufull[1, 4] = 2, 3
    S, I, D = [0], [], []
    for r in range(5):
        for c in range(r, 5):
            if Ufull[r, c] != 0:
                I.append(c)
                D.append(1 / Ufull[r, c] if r == c else Ufull[r, c])
        S.append(len(I))
    U = o.CSR(np.array(S), np.array(I), np.array(D))
    L = o.csr_transpose(U)
    x = np.arange(1.0, 6.0)
    b = Ufull.T @ (Ufull @ x)
    np.testing.assert_allclose(o.ic0_solve(L, U, b), x, rtol=1e-14)


def test_ic0_missing_diagonal():
    A = o.CSR(np.array([0, 1, 2]), np.array([1, 0]), np.array([1.0, 1.0]))
    st, _, _ = o.ic0_factor(A)
    assert st == o.ERR_MISSING_DIAGONAL


# ---------------------------------------------------------------- ILU(0)
def _rand_sparse(n, seed, extra=2):
    rng = np.random.default_rng(seed)
    A = np.eye(n)
    for _ in range(extra * n):       # like T-itl/ilu_0_test.cpp:16-28 (A = 1; A[r][c] -= 1; A[r][r] += 1)
        r, c = rng.integers(n), rng.integers(n)
        A[r, c] -= 1
        A[r, r] += 1
    S, I, V = [0], [], []
    for r in range(n):
        for c in range(n):
            if A[r, c] != 0 or r == c:
                I.append(c)
                V.append(A[r, c])
        S.append(len(I))
    return o.CSR(np.array(S), np.array(I), np.array(V)), A


@pytest.mark.parametrize("case", ["grid3", "rand3", "rand12", "rand30"])
def test_ilu0_equals_dense_textbook(case):
    # T-itl/ilu_0_test.cpp:31-50 and ilu_0_bicgstab_test.cpp: sparse ILU(0) == dense textbook ILU(0) on the pattern
    if case == "grid3":
        A = o.laplacian_setup(3, 3)
        Ad = A.dense()
    else:
        A, Ad = _rand_sparse(int(case[4:]), seed=len(case) + int(case[4:]))
    st, L, U = o.ilu0_factor(A)
    assert st == o.OK
    LUt = dense_ilu0_textbook(Ad)
    Ld, Ud = L.dense(), U.dense()
    mask = (Ad != 0)
    lo = np.tril(mask, -1)
    np.testing.assert_allclose(Ld[lo], LUt[lo], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(np.diag(Ud), 1 / np.diag(LUt), rtol=1e-12)
    up = np.triu(mask, 1)
    np.testing.assert_allclose(Ud[up], LUt[up], rtol=1e-12, atol=1e-14)


def test_ilu0_pattern_invariant_and_solves():
    # (I + L) U == A on A's pattern; solve and adjoint_solve invert (I+L)U and its transpose
    A, Ad = _rand_sparse(20, seed=5)
    _, L, U = o.ilu0_factor(A)
    Lm = L.dense() + np.eye(A.n)
    Um = U.dense().copy()
    np.fill_diagonal(Um, 1 / np.diag(Um))
    P = Lm @ Um
    mask = Ad != 0
    np.testing.assert_allclose(P[mask], Ad[mask], rtol=1e-12, atol=1e-12)
    x = np.linspace(-1, 2, A.n)
    np.testing.assert_allclose(o.ilu0_solve(L, U, P @ x), x, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(o.ilu0_adjoint_solve(L, U, P.T @ x), x, rtol=1e-10, atol=1e-10)


def test_ilu0_is_exact_lu_without_fill():
    # tridiagonal: ILU(0) == LU, so the preconditioned CG converges in one iteration
    n = 12
    S, I, V = [0], [], []
    for r in range(n):
        for c in (r - 1, r, r + 1):
            if 0 <= c < n:
                I.append(c)
                V.append(2.0 if c == r else -1.0)
        S.append(len(I))
    A = o.CSR(np.array(S), np.array(I), np.array(V))
    b = o.spmv(A, np.ones(n))
    res = o.solve("cg", A, b, pc=o.PC_ILU0, max_iter=n, rtol=1e-12)
    assert res["status"] == o.OK and res["iterations"] == 1


# ---------------------------------------------------------------- solvers
def test_ilu0_cg_iteration_pin():
    # T-itl/ilu_0_cg_test.cpp:22-33: CG + ILU(0) on the 3x3 grid reaches 1e-6 in <= 4 iterations
    A = o.laplacian_setup(3, 3)
    b = o.spmv(A, np.ones(9))
    res = o.solve("cg", A, b, pc=o.PC_ILU0, max_iter=9, rtol=1e-6)
    assert res["status"] == o.OK and res["iterations"] <= 4


def test_absolute_termination_criterion():
    # T-itl/absolut_termination_criterion_test.cpp:21-39: diagonal PC, rtol 0, atol 1e-6
    A = o.laplacian_setup(4, 4)
    b = o.spmv(A, np.ones(16))
    res = o.solve("cg", A, b, pc=o.PC_DIAGONAL, max_iter=500, rtol=0.0, atol=1e-6)
    assert res["status"] == o.OK
    assert o.two_norm(o.residual(A, res["x"], b)) <= 1e-6


@pytest.mark.parametrize("method,pc", [("cg", o.PC_IC0), ("cg", o.PC_IDENTITY), ("bicgstab", o.PC_ILU0),
                                       ("bicgstab", o.PC_DIAGONAL)])
def test_krylov_krylov_test_grid(method, pc):
    # T-itl/ic_0_krylov_test.cpp:20-66: 4x4 grid, 1e-4 within 150 iterations
    A = o.laplacian_setup(4, 4)
    x_true = np.linspace(-1, 1, 16)
    b = o.spmv(A, x_true)
    res = o.solve(method, A, b, pc=pc, max_iter=150, rtol=1e-4)
    assert res["status"] == o.OK
    assert o.two_norm(o.residual(A, res["x"], b)) <= 1e-4 * o.two_norm(b) * 10


def test_cg_exact_arithmetic_bound_and_history():
    # CG converges in at most n iterations; history[0] = |b|, recurrence residual decreases to tol
    A = o.laplacian_setup(5, 5)
    b = o.spmv(A, np.linspace(0, 1, 25))
    res = o.solve("cg", A, b, pc=o.PC_IDENTITY, max_iter=100, rtol=1e-12)
    assert res["status"] == o.OK and res["iterations"] <= 25
    assert res["history"][0] == pytest.approx(o.two_norm(b))
    assert res["history"][-1] <= 1e-12 * o.two_norm(b)


def test_max_iterations_error():
    A = o.laplacian_setup(8, 8)
    b = o.spmv(A, np.ones(64))
    res = o.solve("cg", A, b, pc=o.PC_IDENTITY, max_iter=3, rtol=1e-12)
    assert res["status"] == o.MAX_ITER and res["iterations"] == 3


def test_bicgstab_orthogonality_breakdown():
    # r = b - A x = 0 with tolerances that no residual meets: the loop is
    # entered and rho_1 = dot(rtilde, r) = 0 throws (bicgstab.hpp:46)
    A = o.laplacian_setup(3, 3)
    res = o.solve("bicgstab", A, np.zeros(9), pc=o.PC_IDENTITY, max_iter=10, rtol=-1.0, norm_r0=1.0, atol=-1.0)
    assert res["status"] == o.ERR_ORTHOGONALITY


def test_oracle_plant_a_mistake_is_detected():
    # the golden vector must reject a factor with a dropped Schur term
    A = o.laplacian_setup(3, 3)
    _, U, L = o.ic0_factor(A)
    bad = o.CSR(U.starts, U.indices, U.data.copy())
    bad.data[bad.starts[4]] *= 1.01           # perturb one inverted diagonal
    Lb = o.csr_transpose(bad)
    y = o.ic0_solve(Lb, bad, np.arange(9.0), factor_dtype=np.float32)
    yc = np.array([1.03194, 1.60198, 1.45357, 2.5258, 3.55386, 3.16, 2.91787, 4.09338, 3.81335])
    assert np.linalg.norm(y - yc) > 1e-3


def test_openmp_baseline_matches_the_oracle_solve():
    # the OpenMP timing baseline (oracle/cpu_openmp.cpp) runs the same IC(0)-CG:
    # iteration count within one of the oracle's (its dots are summed in
    # per-thread partials), residual at the target
    S, I, V = o.laplacian_setup(24, 24).starts, o.laplacian_setup(24, 24).indices, o.laplacian_setup(24, 24).data
    A = o.CSR(S, I, V)
    st, U, L = o.ic0_factor(A)
    b = np.linspace(-1.0, 1.0, A.n)
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=500, rtol=1e-10)
    x, its = o.omp_ic0_pcg(A, U, L, b, 1e-10, 500, 4)
    assert abs(its - ref["iterations"]) <= 1
    assert o.two_norm(o.residual(A, x, b)) <= 1e-10 * o.two_norm(b) * 1.0001
    np.testing.assert_allclose(x, ref["x"], rtol=0, atol=1e-8 * np.max(np.abs(ref["x"])))


def test_more_krylov_restatements_converge():
    # cgs / bicg / tfqmr (oracle restatements of cgs.hpp, bicg.hpp, tfqmr.hpp):
    # on an SPD system without preconditioner BiCG is CG in exact arithmetic,
    # so its iteration count equals CG's up to one; all three reach the tolerance
    A = o.laplacian_setup(20, 20)
    b = np.linspace(-1.0, 1.0, A.n)
    cg = o.solve("cg", A, b, pc=0, max_iter=2000, rtol=1e-10)
    for m in ("cgs", "bicg", "tfqmr"):
        for pc in (0, o.PC_IC0, o.PC_ILU0):
            r = o.solve(m, A, b, pc=pc, max_iter=2000, rtol=1e-10)
            assert r["status"] == 0, (m, pc)
            assert o.two_norm(o.residual(A, r["x"], b)) <= 20 * 1e-10 * o.two_norm(b), (m, pc)
    bicg = o.solve("bicg", A, b, pc=0, max_iter=2000, rtol=1e-10)
    assert abs(bicg["iterations"] - cg["iterations"]) <= 1
    # the oracle's restatement keeps MTL4's rho_1 = dot(z_tilde, z) (bicg.hpp:42): with IC(0) it is
    # not PCG's r.z, so BiCG needs more iterations than PCG here
    assert o.solve("bicg", A, b, pc=o.PC_IC0, max_iter=2000, rtol=1e-10)["iterations"] > \
        o.solve("cg", A, b, pc=o.PC_IC0, max_iter=2000, rtol=1e-10)["iterations"]


def test_more_krylov_finite_termination():
    # exact arithmetic fixes these counts: with a minimal polynomial of degree m
    # BiCG, CGS and TFQMR reach the exact solution after m steps (barring
    # breakdown).  diag(1,2,5) blocks: m = 3; A = I + N with N^2 = 0 (nonsymmetric,
    # so BiCG's shadow sequence needs the true A^T): m = 2
    d = np.repeat([1.0, 2.0, 5.0], 10)
    D = o.CSR(np.arange(31, dtype=np.int64), np.arange(30, dtype=np.int64), d)
    n, h = 40, 20
    rng = np.random.default_rng(2)
    S, I, V = [0], [], []
    for i in range(n):
        row = [(i, 1.0)]
        if i < h:
            row += [(int(j), float(rng.uniform(-1, 1))) for j in rng.choice(np.arange(h, n), 3, replace=False)]
        row.sort()
        I += [c for c, _ in row]
        V += [v for _, v in row]
        S.append(len(I))
    N = o.CSR(np.array(S, np.int64), np.array(I, np.int64), np.array(V))
    for A, m in ((D, 3), (N, 2)):
        b = rng.uniform(-1, 1, A.n)
        for meth in ("cgs", "bicg", "tfqmr"):
            r = o.solve(meth, A, b, pc=0, max_iter=50, rtol=1e-12)
            assert r["status"] == 0 and r["iterations"] == m, (meth, m, r["iterations"])
            assert o.two_norm(o.residual(A, r["x"], b)) <= 1e-12 * o.two_norm(b) * 10
    # ic_0_krylov_test.cpp: 4x4-grid Laplacian, x0 = 0.01, b = A * ones,
    # cyclic_iteration(b, 150, 1e-4): bicg (I), cgs (ILU), tfqmr (ILU, I) all return 0
    A = o.laplacian_setup(4, 4)
    b = o.spmv(A, np.ones(16))
    for meth, pc in (("bicg", o.PC_IDENTITY), ("cgs", o.PC_ILU0), ("tfqmr", o.PC_ILU0)):
        r = o.solve(meth, A, b, pc=pc, max_iter=150, rtol=1e-4, x0=np.full(16, 0.01))
        assert r["status"] == 0, meth


def _nilpotent_plus_identity(n=40, h=20, seed=2):
    rng = np.random.default_rng(seed)
    S, I, V = [0], [], []
    for i in range(n):
        row = [(i, 1.0)]
        if i < h:
            row += [(int(j), float(rng.uniform(-1, 1))) for j in rng.choice(np.arange(h, n), 3, replace=False)]
        row.sort()
        I += [c for c, _ in row]
        V += [v for _, v in row]
        S.append(len(I))
    return o.CSR(np.array(S, np.int64), np.array(I, np.int64), np.array(V))


def test_qmr_idr_finite_termination():
    # QMR's Lanczos process ends after m steps for a minimal polynomial of
    # degree m: the next Lanczos vector is 0, so rho = 0 passes the test and x
    # is exact.  MTL4's IDR(s) (s steps per cycle) ends within m + 1 steps here.
    d = np.repeat([1.0, 2.0, 5.0], 10)
    D = o.CSR(np.arange(31, dtype=np.int64), np.arange(30, dtype=np.int64), d)
    N = _nilpotent_plus_identity()
    rng = np.random.default_rng(5)
    for A, m in ((D, 3), (N, 2)):
        b = rng.uniform(-1, 1, A.n)
        r = o.solve("qmr", A, b, pc=0, max_iter=40, rtol=1e-12)
        assert r["status"] == 0 and r["iterations"] == m
        assert o.two_norm(o.residual(A, r["x"], b)) <= 1e-12 * o.two_norm(b)
    b = rng.uniform(-1, 1, N.n)
    r = o.solve("idr_s", N, b, pc=0, max_iter=40, rtol=1e-12, param=2)
    assert r["status"] == 0 and r["iterations"] <= 3
    assert o.two_norm(o.residual(N, r["x"], b)) <= 1e-12 * o.two_norm(b)


def test_gmres_known_answers():
    # gmres_test.cpp: 2x2-grid Laplacian, b = 1, x = 5, one GMRES(1) step with
    # basic_iteration(b, 1, 1e-8): r0 is an eigenvector, so the step is exact
    A = o.laplacian_setup(2, 2)
    b = np.ones(4)
    r = o.solve("gmres", A, b, x0=np.full(4, 5.0), pc=0, max_iter=1, rtol=1e-8, param=1)
    assert o.two_norm(o.residual(A, r["x"], b)) <= 5e-5
    # one GMRES(1) cycle is the minimal-residual step x + ((r.Ar)/(Ar.Ar)) r
    N = _nilpotent_plus_identity()
    rng = np.random.default_rng(7)
    b = rng.uniform(-1, 1, N.n)
    x0 = rng.uniform(-1, 1, N.n)
    r0 = o.residual(N, x0, b)
    Ar = o.spmv(N, r0)
    expect = x0 + (np.dot(r0, Ar) / np.dot(Ar, Ar)) * r0
    r = o.solve("gmres", N, b, x0=x0, pc=0, max_iter=1, rtol=1e-14, param=1)
    assert r["iterations"] == 1
    np.testing.assert_allclose(r["x"], expect, rtol=1e-13, atol=1e-14)
    # the restarted driver stops at the first cycle whose true residual passes
    r = o.solve("gmres", N, b, pc=0, max_iter=200, rtol=1e-10, param=1)
    assert r["status"] == 0 and r["resid"] <= 1e-10 * o.two_norm(b)
    assert abs(o.two_norm(o.residual(N, r["x"], b)) - r["resid"]) <= 1e-12 * o.two_norm(b)


def _csr_from_dense(M):
    S, I, V = [0], [], []
    for row in M:
        for j, v in enumerate(row):
            if v != 0:
                I.append(j)
                V.append(float(v))
        S.append(len(I))
    return o.CSR(np.array(S), np.array(I), np.array(V))


@pytest.mark.parametrize("method", ["gmres", "gmres_full"])
def test_gmres_worked_swap_case(method):
    # Worked by hand through gmres.hpp:59-116 for A = [[0, 1], [1, 0]],
    # b = e1, x0 = 0 (identity preconditioners, atol 0), restart 2:
    #   k = 0: v0 = e1, A v0 = e2, H00 = 0, H10 = 1, v1 = e2; the rotation
    #          c0 = 0, s0 = -1 makes H00 = 1, H10 = 0 and g = [0, -1].
    #   k = 1: A v1 = e1, H01 = 1, H11 = 0, H21 = 0 (breakdown: v2 = 0).
    #          The loop over i < k builds its rotation from column k - 1,
    #          (H00, H10) = (1, 0): c = 1, s = 0 (givens.hpp:43-44), so column
    #          1 is NOT rotated by (c0, s0) (gmres.hpp:84-85); nu = 0.
    #   lu_solve(H[0:2][0:2] = [[1, 1], [0, 0]], [0, -1]) does not throw: lu()
    #   tests the pivots of rows 0 .. n-2 only (lu.hpp:76-84), and the
    #   backward substitution divides by the zero pivot: y = [+inf, -inf].
    #   x += V y multiplies 0 * inf in both entries: x = [NaN, NaN], and
    #   terminate(b - A x) records a NaN residual.
    # gmres_full's terminate() returns the error code, 0 while i < max_iter
    # (basic_iteration.hpp:43-47, 80); the restarted driver keeps cycling on
    # NaN until max_iter (status 1).
    # Textbook GMRES (stored rotations applied to column 1) gives
    # H = [[1, 0], [0, 1]], g = [0, 1] and the exact x = e2 in two steps; a
    # pivot test on the last row would fall back to one vector and x = 0.
    A = _csr_from_dense([[0, 1], [1, 0]])
    b = np.array([1.0, 0.0])
    for max_iter in (2, 6):
        r = o.solve(method, A, b, pc=0, max_iter=max_iter, rtol=1e-10, param=2)
        assert np.isnan(r["x"]).all() and np.isnan(r["resid"])
        if method == "gmres_full":
            assert r["iterations"] == 2 and r["status"] == (o.MAX_ITER if max_iter == 2 else o.OK)
        else:
            assert r["iterations"] == max_iter and r["status"] == o.MAX_ITER
    # restart 1: g = [0, -1] after the first rotation, so the loop that drops
    # trailing directions with |g[k-1]| <= atol (gmres.hpp:95) leaves none and
    # the cycle fails with code 2 ("did not find any direction", :107-108),
    # x untouched
    r = o.solve("gmres", A, b, pc=0, max_iter=3, rtol=1e-10, param=1)
    np.testing.assert_array_equal(r["x"], [0.0, 0.0])
    assert r["status"] == 2 and r["iterations"] == 1 and r["resid"] == 1.0


@pytest.mark.parametrize("restart", [2, 3, 5])
def test_gmres_restart_cycle_invariants(restart):
    # Each cycle of itl::gmres ends with iter.terminate(b - A x)
    # (gmres.hpp:115-116): the residual it reports after c cycles is the true
    # residual of the x it returns (identity preconditioners), bit for bit.
    # And a cycle only adds a combination of its Arnoldi vectors:
    # x_c - x_{c-1} lies in span{r, A r, ..., A^(m-1) r} with r = b - A x_{c-1}.
    A = o.CSR(*_varcoef_like(9))
    rng = np.random.default_rng(11)
    b = rng.uniform(-1, 1, A.n)
    x_prev = np.zeros(A.n)
    for cycle in range(1, 5):
        r = o.solve("gmres", A, b, x0=x_prev, pc=0, max_iter=restart, rtol=1e-300, param=restart)
        assert r["iterations"] == restart and r["status"] == o.MAX_ITER
        assert r["resid"] == o.two_norm(o.residual(A, r["x"], b))
        rr = o.residual(A, x_prev, b)
        K = [rr / np.linalg.norm(rr)]
        for _ in range(restart - 1):
            v = o.spmv(A, K[-1])
            K.append(v / np.linalg.norm(v))
        K = np.stack(K, axis=1)
        d = r["x"] - x_prev
        coef, *_ = np.linalg.lstsq(K, d, rcond=None)
        assert np.linalg.norm(K @ coef - d) <= 1e-10 * max(np.linalg.norm(d), 1e-300), cycle
        x_prev = r["x"]


def _varcoef_like(m):
    # nonsymmetric 5-point operator on an m x m grid (convection-like
    # off-diagonals), diagonal dominant
    S, I, V = [0], [], []
    for r in range(m * m):
        i, j = r % m, r // m
        for c, v in ((r - m, -1.3), (r - 1, -0.7), (r, 4.5), (r + 1, -1.1), (r + m, -0.9)):
            if c < 0 or c >= m * m:
                continue
            if (c == r - 1 and i == 0) or (c == r + 1 and i == m - 1):
                continue
            I.append(c)
            V.append(v)
        S.append(len(I))
    return np.array(S), np.array(I), np.array(V)


@pytest.mark.parametrize("method,param", [("bicgstab_2", None), ("bicgstab_ell", 1), ("bicgstab_ell", 3),
                                          ("idr_s", 2), ("idr_s", 4)])
def test_ell_idr_residual_consistency(method, param):
    # x and the recurrence residual must agree (identity PC, so r_hat[0] is
    # b - A x).  BiCGStab(l) stopped by the iteration limit leaves at its first
    # inner test of iteration k, before that iteration touches x
    # (bicgstab_ell.hpp:87-99), so ||b - A x|| is the norm tested at the end of
    # iteration k-1; IDR(s) updates x and r before each test.  A dropped or
    # mis-signed term in the MR or IDR updates breaks this.
    A = o.laplacian_setup(12, 12)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, A.n)
    for k in (2, 3, 5):
        r = o.solve(method, A, b, pc=0, max_iter=k, rtol=1e-30, param=param)
        assert r["status"] == 1 and r["iterations"] == k
        true = o.two_norm(o.residual(A, r["x"], b))
        tested = r["resid"] if method == "idr_s" else r["history"][k - 1]
        assert abs(true - tested) <= 1e-10 * o.two_norm(b), (k, true, tested)
