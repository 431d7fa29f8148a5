"""The one-workgroup IC(0)-PCG of small 2D grid systems (k_cg_ic0_grid):
systems of at most 4096 rows on at most 64 grid lines whose IC(0) factor has
the 5-point pattern run whole batches of CG iterations in one workgroup, the
sweeps on one wavefront with a DPP lane shift per step.  Checked against the
oracle (cg.hpp:75-112 with ic_0.hpp:135-186) and against the multi-kernel
path (M4G_GRID_CG=0) on the same inputs:
  * one PC application's worth of arithmetic: the final x agrees with the
    oracle's to the tolerance, iteration counts within one, the residual
    histories to rounding (only the reductions are summed in another order);
  * the iteration protocol: status, max_iter stop, absolute tolerance, a
    nonzero start vector and a caller-given norm_r0;
  * systems the path does not take (more than 64 lines, a 7-point pattern)
    still solve through the multi-kernel path.
The multi-kernel CG's scalars are checked the same way: its folded
finalizes against separate finalize kernels (bit for bit), and rho taken
from the last sweep's per-tile partials against a separate dot kernel (to
rounding, and the oracle's iteration count).
"""
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


def solve(g, S, I, V, b, rtol, max_iter=5000, atol=0.0, x0=None, norm_r0=None):
    A = g.Matrix(S, I, V)
    P = g.PC(A, g.PC_IC0)
    x = g.Vector(data=np.zeros(len(b)) if x0 is None else x0)
    it, hist = g.make_iteration(max_iter, rtol, atol, o.two_norm(b) if norm_r0 is None else norm_r0)
    st = g.cg(A, x, g.Vector(data=b), P, it)
    return dict(status=st, iterations=it.iterations, resid=it.resid, x=x.numpy(), history=hist[: it.iterations + 1])


CASES = {
    "lap64": lambda g: g.laplacian2d(64, 64),        # the reference's 64 x 64 case
    "lap_rect": lambda g: g.laplacian2d(50, 37),     # 50 lines of 37 cells
    "vc5": lambda g: g.varcoef5(70, 58, seed=6),      # variable coefficients, 58 lines of 70
    "lap_two": lambda g: g.laplacian2d(2, 500),      # two lines of 500 cells
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_grid_cg_matches_oracle_and_multikernel(g, case, monkeypatch):
    S, I, V = CASES[case](g)
    A = o.CSR(S, I, V)
    x_true = g.rng_uniform(4, 1, A.n)
    b = o.spmv(A, x_true)
    rtol = 1e-10
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=5000, rtol=rtol)
    got = solve(g, S, I, V, b, rtol)
    monkeypatch.setenv("M4G_GRID_CG", "0")
    multi = solve(g, S, I, V, b, rtol)
    monkeypatch.delenv("M4G_GRID_CG")
    say(f"  {case} n={A.n}: grid {got['iterations']} it, multi-kernel {multi['iterations']}, "
        f"oracle {ref['iterations']}")
    assert got["status"] == multi["status"] == ref["status"] == o.OK
    assert abs(got["iterations"] - ref["iterations"]) <= 1
    assert abs(got["iterations"] - multi["iterations"]) <= 1
    nb = o.two_norm(b)
    assert got["resid"] <= rtol * nb
    assert o.two_norm(o.residual(A, got["x"], b)) <= 10 * rtol * nb
    scale = max(1.0, np.max(np.abs(ref["x"])))
    assert np.max(np.abs(got["x"] - ref["x"])) <= 1e-8 * scale
    assert np.max(np.abs(got["x"] - multi["x"])) <= 1e-8 * scale
    k = min(len(got["history"]), len(ref["history"]), 30)
    np.testing.assert_allclose(got["history"][:k], ref["history"][:k], rtol=1e-8)
    np.testing.assert_allclose(got["history"][:k], multi["history"][:k], rtol=1e-8)


def test_grid_cg_iteration_protocol(g):
    S, I, V = g.laplacian2d(40, 40)
    A = o.CSR(S, I, V)
    b = o.spmv(A, g.rng_uniform(5, 1, A.n))
    nb = o.two_norm(b)
    # max_iter stop: status 1 after exactly max_iter iterations
    r = solve(g, S, I, V, b, 1e-14, max_iter=7)
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=7, rtol=1e-14)
    assert r["status"] == ref["status"] == o.MAX_ITER
    assert r["iterations"] == ref["iterations"] == 7
    np.testing.assert_allclose(r["x"], ref["x"], rtol=0, atol=1e-10 * np.max(np.abs(ref["x"])))
    # absolute tolerance with norm_r0 = 0 (basic_iteration.hpp:90-91)
    r = solve(g, S, I, V, b, 0.0, atol=1e-6 * nb, norm_r0=0.0)
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=5000, rtol=0.0, atol=1e-6 * nb, norm_r0=0.0)
    assert r["status"] == ref["status"] == o.OK
    assert abs(r["iterations"] - ref["iterations"]) <= 1
    # a nonzero start vector
    x0 = g.rng_uniform(6, 1, A.n)
    r = solve(g, S, I, V, b, 1e-10, x0=x0)
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=5000, rtol=1e-10, x0=x0)
    assert r["status"] == o.OK and abs(r["iterations"] - ref["iterations"]) <= 1
    assert np.max(np.abs(r["x"] - ref["x"])) <= 1e-8 * np.max(np.abs(ref["x"]))
    # converged before the first iteration: x unchanged
    xs = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=5000, rtol=1e-12)["x"]
    r = solve(g, S, I, V, b, 1e-3, x0=xs)
    assert r["status"] == o.OK and r["iterations"] == 0
    np.testing.assert_array_equal(r["x"], xs)


@pytest.mark.parametrize("case", ["lines_100", "p3d"])
def test_systems_outside_the_grid_path(g, case):
    S, I, V = g.laplacian2d(100, 30) if case == "lines_100" else g.poisson3d(12)
    A = o.CSR(S, I, V)
    b = o.spmv(A, g.rng_uniform(7, 1, A.n))
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=5000, rtol=1e-10)
    r = solve(g, S, I, V, b, 1e-10)
    assert r["status"] == ref["status"] == o.OK
    assert abs(r["iterations"] - ref["iterations"]) <= 1


@pytest.mark.parametrize("pc", ["ic0", "diag", "none"])
def test_folded_finalizes_are_bit_identical(g, pc, monkeypatch):
    # below 256 reduction workgroups the CG forms rho and alpha inside their
    # consumer kernels; the scalars, hence x and the history, must be those
    # of the separate finalize kernels bit for bit (M4G_CG_FOLD=0)
    S, I, V = g.varcoef5(200, 150, seed=9)   # n = 30,000: 30 reduction workgroups
    A = o.CSR(S, I, V)
    b = o.spmv(A, g.rng_uniform(8, 1, A.n))
    kind = {"ic0": g.PC_IC0, "diag": g.PC_DIAGONAL, "none": g.PC_IDENTITY}[pc]

    def run():
        Ag = g.Matrix(S, I, V)
        P = g.PC(Ag, kind)
        x = g.Vector(A.n)
        x.fill(0.0)
        it, hist = g.make_iteration(3000, 1e-10, 0.0, o.two_norm(b))
        st = g.cg(Ag, x, g.Vector(data=b), P, it)
        return st, it.iterations, x.numpy(), hist[: it.iterations + 1].copy()

    folded = run()
    monkeypatch.setenv("M4G_CG_FOLD", "0")
    separate = run()
    say(f"  {pc}: {folded[1]} iterations folded, {separate[1]} separate")
    assert folded[0] == separate[0] == 0 and folded[1] == separate[1]
    np.testing.assert_array_equal(folded[2], separate[2])
    np.testing.assert_array_equal(folded[3], separate[3])


@pytest.mark.parametrize("case", ["p3d_ic0", "vc5_ilu0"])
def test_rho_from_the_last_sweep(g, case, monkeypatch):
    # with an IC(0)/ILU(0) PC the CG takes rho = r.z from the partials the
    # last sweep stores per tile; a separate dot kernel (M4G_CG_SWEEP_DOT=0)
    # sums the same products in another order, so the two solves agree to
    # rounding: same iteration count (or one apart), x and history close
    if case == "p3d_ic0":
        S, I, V = g.poisson3d(40)          # 64,000 rows: several tiles per sweep
        kind = g.PC_IC0
    else:
        S, I, V = g.varcoef5(300, 220, seed=3)
        kind = g.PC_ILU0
    A = o.CSR(S, I, V)
    b = o.spmv(A, g.rng_uniform(9, 1, A.n))

    def run():
        Ag = g.Matrix(S, I, V)
        P = g.PC(Ag, kind)
        x = g.Vector(A.n)
        x.fill(0.0)
        it, hist = g.make_iteration(3000, 1e-10, 0.0, o.two_norm(b))
        st = g.cg(Ag, x, g.Vector(data=b), P, it)
        tiles = P.stats()["tiles_upper"]
        return st, it.iterations, x.numpy(), hist[: it.iterations + 1].copy(), tiles

    fused = run()
    monkeypatch.setenv("M4G_CG_SWEEP_DOT", "0")
    separate = run()
    say(f"  {case}: {fused[1]} iterations with the sweep's partials ({fused[4]} tiles), {separate[1]} with a dot kernel")
    assert fused[4] > 1
    assert fused[0] == separate[0] == 0 and abs(fused[1] - separate[1]) <= 1
    scale = max(1.0, np.max(np.abs(separate[2])))
    assert np.max(np.abs(fused[2] - separate[2])) <= 1e-8 * scale
    k = min(len(fused[3]), len(separate[3]), 40)
    np.testing.assert_allclose(fused[3][:k], separate[3][:k], rtol=1e-7)
    ref = o.solve("cg", A, b, pc=o.PC_IC0 if kind == g.PC_IC0 else o.PC_ILU0, max_iter=3000, rtol=1e-10)
    assert abs(fused[1] - ref["iterations"]) <= 1
