"""The headline workload at its full size: 3D 7-point Poisson on a 128^3 grid
(n = 2,097,152, nnz = 14,581,760), FP64, b = A x_true with x_true ~ U[-1,1]
(seed 1), IC(0)-PCG to a 1e-8 relative residual.

At this size the library takes code paths that smaller grids do not reach:
the pipelined SELL SpMV over 32,768 slices with 16-bit column offsets, and the
IC(0) sweeps compiled into 256 tiles of 128x8x8 rows with imports that cross
many tiles.  Checked against the oracle on the same inputs:
  * SpMV (=, +=, -=, b - A x) and the fused SpMV + dot bit for bit
    (mat_vec_mult.hpp:619-689, fused_expr.hpp:42-91);
  * the IC(0) factor U and L = trans(U) bit for bit (ic_0.hpp:135-186, :57);
  * one preconditioner application y = U^-1 L^-1 x bit for bit
    (ic_0.hpp:68-87, lower_trisolve.hpp:191-209, upper_trisolve.hpp:159-176);
  * the PCG iteration count within 2 % of the oracle's complete solve, the
    recurrence and true residuals at 1e-8, and x close to the oracle's
    (cg.hpp:75-112)."""
import time

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu

N = 128
RTOL = 1e-8


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


@pytest.fixture(scope="module")
def case(g):
    say(f"poisson3d {N}^3: generating")
    S, I, V = g.poisson3d(N)
    A = o.CSR(S, I, V)
    assert A.n == N ** 3 and A.nnz == 14_581_760
    Ag = g.Matrix(S, I, V)
    x_true = g.rng_uniform(1, 1, A.n)
    b = o.spmv(A, x_true)
    return dict(S=S, I=I, V=V, A=A, Ag=Ag, x_true=x_true, b=b)


@pytest.fixture(scope="module")
def factors(case):
    t0 = time.perf_counter()
    st, U, L = o.ic0_factor(case["A"])
    assert st == o.OK
    say(f"  oracle IC(0) factor {time.perf_counter() - t0:.2f} s, {U.nnz} entries")
    return U, L


def test_spmv_bitexact_128(g, case):
    say("spmv 128^3 (=, +=, -=, b - Ax, fused dot)")
    A, Ag = case["A"], case["Ag"]
    x = g.rng_uniform(1, 7, A.n)
    w0 = g.rng_uniform(1, 8, A.n)
    xv, wv, rv = g.Vector(data=x), g.Vector(data=w0), g.Vector(A.n)
    g.spmv(Ag, xv, wv, g.ASSIGN)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x))
    wv.upload(w0)
    g.spmv(Ag, xv, wv, g.PLUS)
    np.testing.assert_array_equal(wv.numpy(), o.spmv(A, x, w0, 1))
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


def test_ic0_factor_and_apply_bitexact_128(g, case, factors):
    say("ic0 factor + one application 128^3")
    U, L = factors
    P = g.PC(case["Ag"], g.PC_IC0)
    # L = trans(U) with the inverted diagonal last in each row (which=0), U (which=1)
    for which, ref in ((0, L), (1, U)):
        S, I, V = P.factor(which)
        np.testing.assert_array_equal(S, ref.starts)
        np.testing.assert_array_equal(I, ref.indices)
        np.testing.assert_array_equal(V, ref.data)
    for stream in (3, 4):
        x = g.rng_uniform(5, stream, case["A"].n)
        want = o.ic0_solve(L, U, x)
        xv, yv = g.Vector(data=x), g.Vector(case["A"].n)
        P.solve(xv, yv)
        np.testing.assert_array_equal(yv.numpy(), want)
    # adjoint_solve is solve for IC(0) (ic_0.hpp:89-103)
    yv2 = g.Vector(case["A"].n)
    P.adjoint_solve(xv, yv2)
    np.testing.assert_array_equal(yv2.numpy(), want)


def test_ic0_pcg_iterations_128(g, case):
    A, b = case["A"], case["b"]
    nb = o.two_norm(b)
    t0 = time.perf_counter()
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=1000, rtol=RTOL)
    say(f"ic0-pcg 128^3: oracle {ref['iterations']} iterations in {time.perf_counter() - t0:.2f} s "
        f"(factorisation included), status {ref['status']}")
    assert ref["status"] == o.OK
    P = g.PC(case["Ag"], g.PC_IC0)
    x = g.Vector(A.n)
    x.fill(0.0)
    bv = g.Vector(data=b)
    it, hist = g.make_iteration(1000, RTOL, 0.0, nb)
    st = g.cg(case["Ag"], x, bv, P, it)
    xs = x.numpy()
    true_res = o.two_norm(o.residual(A, xs, b))
    err = float(np.max(np.abs(xs - ref["x"])))
    say(f"  gpu {it.iterations} iterations, status {st}, recurrence resid/|b| {it.resid / nb:.3e}, "
        f"true {true_res / nb:.3e}, max|x - x_oracle| {err:.2e}")
    assert st == 0
    slack = max(1, int(np.ceil(0.02 * ref["iterations"])))
    assert abs(it.iterations - ref["iterations"]) <= slack
    assert it.resid <= RTOL * nb
    assert true_res <= 2 * RTOL * nb
    # both iterates sit within the tolerance of x_true; they agree far closer
    assert err <= 1e-6
    n = min(len(ref["history"]), it.iterations + 1, 40)
    np.testing.assert_allclose(hist[:n], ref["history"][:n], rtol=1e-6)
