"""The line-pipelined sweep schedule of 2D 5-point grids (selected with
M4G_LINE_SWEEP=1) executed on the CPU with the kernel's operation order
(m4g_analysis_pc_solve): bit-identical to the oracle's triangular solves for
IC(0), ILU(0) and its adjoint, and to the tiled schedule with FP32 factors."""
import numpy as np
import pytest

import mtl4gpu as g
import oracle_lib as o
from test_host import _ilu_storage

CASES = {
    # laplacian2d(m, n): m grid lines of n points
    "lap 64x64": lambda: g.laplacian2d(64, 64),
    "lap 101x200": lambda: g.laplacian2d(200, 101),
    "lap 280x7": lambda: g.laplacian2d(7, 280),
    "vc5 70x55": lambda: g.varcoef5(70, 55),
    "vc5 40x129": lambda: g.varcoef5(40, 129),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("kind", [g.PC_IC0, g.PC_ILU0])
def test_line_schedule_bitexact(monkeypatch, case, kind):
    monkeypatch.setenv("M4G_LINE_SWEEP", "1")
    S, I, V = CASES[case]()
    A = o.CSR(S, I, V)
    rhs = g.rng_uniform(3, 7, A.n)
    if kind == g.PC_IC0:
        st, U, L = o.ic0_factor(A)
        ref, fac = o.ic0_solve(L, U, rhs), U.data
    else:
        st, L, U = o.ilu0_factor(A)
        ref, fac = o.ilu0_solve(L, U, rhs), _ilu_storage(A, L, U)
    assert st == o.OK
    out, stats = g.analysis_pc_solve(kind, S, I, fac, rhs)
    assert stats["line_bands"] == (stats["grid_ny"] + 63) // 64
    np.testing.assert_array_equal(out, ref)
    if kind == g.PC_ILU0:
        outa, sa = g.analysis_pc_solve(kind, S, I, fac, rhs, adjoint=True)
        assert sa["line_bands"] > 0
        np.testing.assert_array_equal(outa, o.ilu0_adjoint_solve(L, U, rhs))


@pytest.mark.parametrize("vdtype", [np.float64, np.float32])
def test_line_schedule_float_factors(monkeypatch, vdtype):
    S, I, V = g.varcoef5(130, 70)
    A = o.CSR(S, I, V)
    _, L, U = o.ilu0_factor(A)
    fac = _ilu_storage(A, L, U)
    rhs = g.rng_uniform(3, 8, A.n).astype(vdtype)
    tiled, st = g.analysis_pc_solve(g.PC_ILU0, S, I, fac, rhs, factor_dtype=np.float32, vector_dtype=vdtype)
    assert st["line_bands"] == 0
    monkeypatch.setenv("M4G_LINE_SWEEP", "1")
    line, sl = g.analysis_pc_solve(g.PC_ILU0, S, I, fac, rhs, factor_dtype=np.float32, vector_dtype=vdtype)
    assert sl["line_bands"] == 2
    np.testing.assert_array_equal(line, tiled)


def test_line_schedule_gate(monkeypatch):
    # short lines and 3D grids keep the tiled schedule
    monkeypatch.setenv("M4G_LINE_SWEEP", "1")
    for S, I, V in (g.laplacian2d(9, 7), g.poisson3d(10)):
        A = o.CSR(S, I, V)
        _, U, L = o.ic0_factor(A)
        out, st = g.analysis_pc_solve(g.PC_IC0, S, I, U.data, g.rng_uniform(3, 9, A.n))
        assert st["line_bands"] == 0
