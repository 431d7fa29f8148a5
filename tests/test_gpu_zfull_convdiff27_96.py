"""The mixed-precision workload at its full size: the 27-point
convection-diffusion system on a 96^3 grid (n = 884,736, nnz = 23,393,656;
kappa ~ lognormal(0,1) per cell, velocity (1, 0.5, 0.25) at cell Peclet 0.5,
seed 3), b = A x_true with x_true ~ U[-1,1] (seed 3).

Checked against the oracle on the same inputs:
  * the SpMV bit for bit; the ILU(0) factors L (strict lower) and U
    (inverted diagonal) to 1e-12 relative (ilu_0.hpp:45-82), and one
    application of the FP64 factors (solve and adjoint_solve) and of the FP32
    factors (ilu.hpp:83-115; FP32 factors as ilu_0.hpp:80-81 store them), bit
    for bit whenever the factors are;
  * plain FP64 ILU(0)-BiCGStab to 1e-10: iteration count within 2 % of the
    oracle's, residuals at the tolerance, x close to the oracle's
    (bicgstab.hpp:27-78);
  * the FP32-inner (1e-4) / FP64-refined (1e-10) solve: the true FP64
    residual at 1e-10, and its error against the oracle's FP64 solution."""
import time

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu

N = 96
RTOL = 1e-10


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


@pytest.fixture(scope="module")
def case(g):
    say(f"convdiff27 {N}^3: generating")
    S, I, V = g.convdiff27(N)
    A = o.CSR(S, I, V)
    assert A.n == N ** 3 and A.nnz == 23_393_656
    x_true = g.rng_uniform(3, 1, A.n)
    b = o.spmv(A, x_true)
    st, L, U = o.ilu0_factor(A)
    assert st == o.OK
    return dict(S=S, I=I, V=V, A=A, Ag=g.Matrix(S, I, V), b=b, L=L, U=U)


@pytest.fixture(scope="module")
def oracle_fp64(case):
    t0 = time.perf_counter()
    ref = o.solve("bicgstab", case["A"], case["b"], pc=o.PC_ILU0, max_iter=2000, rtol=RTOL)
    say(f"  oracle ILU(0)-BiCGStab: {ref['iterations']} iterations in {time.perf_counter() - t0:.1f} s, "
        f"status {ref['status']}")
    assert ref["status"] == o.OK
    return ref


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_spmv_and_ilu0_96(g, case):
    say("spmv + ilu0 factor + applications 96^3 27-point")
    A, Ag = case["A"], case["Ag"]
    x = g.rng_uniform(9, 1, A.n)
    xv, wv = g.Vector(data=x), g.Vector(A.n)
    g.spmv(Ag, xv, wv, g.ASSIGN)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x))
    P = g.PC(Ag, g.PC_ILU0)
    exact = True
    for which, ref in ((0, case["L"]), (1, case["U"])):
        S, I, V = P.factor(which)
        np.testing.assert_array_equal(S, ref.starts)
        np.testing.assert_array_equal(I, ref.indices)
        assert relerr(V, ref.data) <= 1e-12
        exact = exact and np.array_equal(V, ref.data)
    say(f"  ilu0 factors bit-exact: {exact}")

    def check(got, want):
        if exact:
            np.testing.assert_array_equal(got, want)
        assert relerr(got, want) <= 1e-12

    yv = g.Vector(A.n)
    P.solve(xv, yv)
    check(yv.numpy(), o.ilu0_solve(case["L"], case["U"], x))
    P.adjoint_solve(xv, yv)
    check(yv.numpy(), o.ilu0_adjoint_solve(case["L"], case["U"], x))
    P32 = g.PC(Ag, g.PC_ILU0, factor_dtype=np.float32)
    P32.solve(xv, yv)
    want32 = o.ilu0_solve(case["L"], case["U"], x, factor_dtype=np.float32)
    if exact:
        np.testing.assert_array_equal(yv.numpy(), want32)
    assert relerr(yv.numpy(), want32) <= 1e-6


def test_ilu0_bicgstab_fp64_96(g, case, oracle_fp64):
    A, b = case["A"], case["b"]
    nb = o.two_norm(b)
    P = g.PC(case["Ag"], g.PC_ILU0)
    x = g.Vector(A.n)
    x.fill(0.0)
    it, _ = g.make_iteration(2000, RTOL, 0.0, nb)
    st = g.bicgstab(case["Ag"], x, g.Vector(data=b), P, it)
    xs = x.numpy()
    true_res = o.two_norm(o.residual(A, xs, b))
    err = float(np.max(np.abs(xs - oracle_fp64["x"])))
    say(f"  gpu ILU(0)-BiCGStab: {it.iterations} iterations (oracle {oracle_fp64['iterations']}), status {st}, "
        f"true relres {true_res / nb:.2e}, max|x - x_oracle| {err:.2e}")
    assert st == 0
    slack = max(1, int(np.ceil(0.02 * oracle_fp64["iterations"])))
    assert abs(it.iterations - oracle_fp64["iterations"]) <= slack
    assert it.resid <= RTOL * nb
    assert true_res <= 10 * RTOL * nb
    assert err <= 1e-7


def test_refined_solve_96(g, case, oracle_fp64):
    A, b = case["A"], case["b"]
    nb = o.two_norm(b)
    P32 = g.PC(case["Ag"], g.PC_ILU0, factor_dtype=np.float32)
    x = g.Vector(A.n)
    x.fill(0.0)
    it, _ = g.make_iteration(50, RTOL, 0.0, nb)
    st, info = g.refined_solve(g.METHOD_BICGSTAB, case["Ag"], x, g.Vector(data=b), P32, it, 1e-4)
    xs = x.numpy()
    relres = o.two_norm(o.residual(A, xs, b)) / nb
    err = float(np.max(np.abs(xs - oracle_fp64["x"])) / np.max(np.abs(oracle_fp64["x"])))
    say(f"  refined: {info.outer_iterations} outer, {info.inner_iterations} inner iterations, status {st}, "
        f"true FP64 relres {relres:.2e}, error vs the oracle's FP64 x {err:.2e}")
    assert st == 0
    assert relres <= RTOL * 1.01
    # the refined and the FP64 solutions both satisfy the 1e-10 residual; their
    # difference is bounded by the conditioning times that residual
    assert err <= 1e-7
    assert 2 <= info.outer_iterations <= 6
