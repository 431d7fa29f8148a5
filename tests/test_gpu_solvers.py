"""GPU CG / BiCGStab against the oracle: iteration counts within 2 % of the
oracle's, final recurrence and true residuals at the requested tolerance, and
solutions close to the oracle's (reductions run in a different order, so the
iterates agree to rounding, not bit for bit)."""
import sys

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


def run(g, method, S, I, V, b, pc, rtol, max_iter=2000, atol=0.0, factor_dtype=np.float64, norm_r0=None):
    A = g.Matrix(S, I, V)
    P = g.PC(A, pc, factor_dtype=factor_dtype)
    x = g.Vector(len(b))
    x.fill(0.0)
    bv = g.Vector(data=b)
    it, hist = g.make_iteration(max_iter, rtol, atol, o.two_norm(b) if norm_r0 is None else norm_r0)
    f = g.cg if method == "cg" else g.bicgstab
    st = f(A, x, bv, P, it)
    return dict(status=st, iterations=it.iterations, resid=it.resid, x=x.numpy(),
                history=hist[: it.iterations + 1])


def compare(g, method, gen, pc, rtol, seed=1, max_iter=2000, strict=False):
    """strict: iteration counts within 2 % (the bar for the benchmark
    configurations); otherwise within max(2, 2 %), since unpreconditioned
    BiCGStab reacts to the different summation order of the reductions."""
    S, I, V = gen
    A = o.CSR(S, I, V)
    x_true = g.rng_uniform(seed, 1, A.n)
    b = o.spmv(A, x_true)
    ref = o.solve(method, A, b, pc=pc, max_iter=max_iter, rtol=rtol)
    got = run(g, method, S, I, V, b, pc, rtol, max_iter)
    say(f"  {method} pc={pc} n={A.n}: gpu {got['iterations']} it (status {got['status']}), "
        f"oracle {ref['iterations']} it (status {ref['status']})")
    assert got["status"] == ref["status"] == o.OK
    slack = int(np.ceil(0.02 * ref["iterations"]))
    assert abs(got["iterations"] - ref["iterations"]) <= (max(1, slack) if strict else max(2, slack))
    nb = o.two_norm(b)
    assert got["resid"] <= rtol * nb
    true_res = o.two_norm(o.residual(A, got["x"], b))
    assert true_res <= 10 * rtol * nb
    assert np.max(np.abs(got["x"] - ref["x"])) <= 100 * rtol * np.max(np.abs(x_true)) * 10
    n = min(len(got["history"]), len(ref["history"]), 20)
    np.testing.assert_allclose(got["history"][:n], ref["history"][:n], rtol=1e-6)
    return got, ref


def test_laplacian64_ic0_pcg(g):
    # 2D 5-point 64x64, IC(0)-PCG to 1e-10
    say("64x64 laplacian ic0 pcg")
    compare(g, "cg", g.laplacian2d(64, 64), o.PC_IC0, 1e-10, strict=True)


@pytest.mark.parametrize("pc", [o.PC_IDENTITY, o.PC_DIAGONAL, o.PC_IC0, o.PC_ILU0])
def test_cg_pcs(g, pc):
    say(f"cg pc={pc}")
    compare(g, "cg", g.varcoef5(40, 33), pc, 1e-9)


@pytest.mark.parametrize("pc", [o.PC_IDENTITY, o.PC_DIAGONAL, o.PC_ILU0])
def test_bicgstab_convdiff(g, pc):
    say(f"bicgstab pc={pc}")
    compare(g, "bicgstab", g.convdiff27(16), pc, 1e-10)


def test_pcg_3d_poisson(g):
    say("pcg p3d 40")
    compare(g, "cg", g.poisson3d(40), o.PC_IC0, 1e-8, strict=True)


def test_ilu0_cg_iteration_pin(g):
    # T-itl/ilu_0_cg_test.cpp: 3x3 grid, CG + ILU(0) to 1e-6 in <= 4 iterations
    S, I, V = g.laplacian2d(3, 3)
    b = o.spmv(o.CSR(S, I, V), np.ones(9))
    got = run(g, "cg", S, I, V, b, o.PC_ILU0, 1e-6, max_iter=9)
    assert got["status"] == 0 and got["iterations"] <= 4


def test_absolute_termination(g):
    # T-itl/absolut_termination_criterion_test.cpp: rtol 0, atol 1e-6, diagonal PC
    S, I, V = g.laplacian2d(4, 4)
    A = o.CSR(S, I, V)
    b = o.spmv(A, np.ones(16))
    got = run(g, "cg", S, I, V, b, o.PC_DIAGONAL, 0.0, max_iter=500, atol=1e-6)
    assert got["status"] == 0
    assert o.two_norm(o.residual(A, got["x"], b)) <= 1e-6


def test_max_iter_and_counter(g):
    S, I, V = g.laplacian2d(16, 16)
    b = o.spmv(o.CSR(S, I, V), np.ones(256))
    got = run(g, "cg", S, I, V, b, o.PC_IDENTITY, 1e-14, max_iter=5)
    assert got["status"] == g.MAX_ITER and got["iterations"] == 5
    ref = o.solve("cg", o.CSR(S, I, V), b, pc=o.PC_IDENTITY, max_iter=5, rtol=1e-14)
    assert ref["status"] == o.MAX_ITER and ref["iterations"] == 5


def test_bicgstab_breakdown(g):
    S, I, V = g.laplacian2d(3, 3)
    A = g.Matrix(S, I, V)
    P = g.PC(A, g.PC_IDENTITY)
    x, b = g.Vector(9), g.Vector(9)
    x.fill(0)
    b.fill(0)
    it, _ = g.make_iteration(10, -1.0, -1.0, 1.0)
    with pytest.raises(g.M4GError) as e:
        g.bicgstab(A, x, b, P, it)
    assert e.value.status == g.ERR_ORTHOGONALITY


def test_warm_restart(g):
    # x is in/out: a second call resumes from the current iterate after
    # restart() (i = 0, basic_iteration.hpp:114), as base_solver::operator()
    # does (base_solver.hpp:43-55).  The error member is only ever set
    # (basic_iteration.hpp:45), so it still reports 1 after converging.
    S, I, V = g.laplacian2d(20, 20)
    Ah = o.CSR(S, I, V)
    A = g.Matrix(S, I, V)
    P = g.PC(A, g.PC_IC0)
    bh = o.spmv(Ah, np.ones(400))
    b = g.Vector(data=bh)
    x = g.Vector(400)
    x.fill(0)
    full = o.solve("cg", Ah, bh, pc=o.PC_IC0, max_iter=1000, rtol=1e-10)["iterations"]
    it, _ = g.make_iteration(3, 1e-10, 0.0, o.two_norm(bh))
    assert g.cg(A, x, b, P, it) == g.MAX_ITER and it.iterations == 3
    it.max_iter, it.iterations = 1000, 0
    assert g.cg(A, x, b, P, it) == g.MAX_ITER
    assert it.resid <= 1e-10 * it.norm_r0 and 0 < it.iterations < full
    assert o.two_norm(o.residual(Ah, x.numpy(), bh)) <= 1e-9 * o.two_norm(bh)
