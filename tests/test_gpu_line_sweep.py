"""The line-pipelined sweeps of 2D 5-point grids (one band of 64 grid lines
per workgroup, lane l owning line 64k + l; selected with M4G_LINE_SWEEP=1,
the tiled sweeps being the faster default): every IC(0) / ILU(0) application,
the ILU(0) adjoint and the FP32-factor variants are bit-identical to the
tiled sweeps, and in FP64 to the oracle's triangular solves; grids without
that structure keep the tiled form."""
import os

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


def make_pc(g, A, kind, tiled, **kw):
    old = os.environ.get("M4G_LINE_SWEEP")
    os.environ["M4G_LINE_SWEEP"] = "0" if tiled else "1"
    try:
        return g.PC(A, kind, **kw)
    finally:
        if old is None:
            del os.environ["M4G_LINE_SWEEP"]
        else:
            os.environ["M4G_LINE_SWEEP"] = old


GRIDS = {
    # laplacian2d(m, n): m grid lines of n points (row stride n)
    "lap 64x64": lambda g: g.laplacian2d(64, 64),
    "lap 101x200": lambda g: g.laplacian2d(200, 101),
    "lap 280x7": lambda g: g.laplacian2d(7, 280),
    "lap 1000x3": lambda g: g.laplacian2d(3, 1000),
    "vc5 70x55": lambda g: g.varcoef5(70, 55),
    "vc5 256x256": lambda g: g.varcoef5(256, 256),
    "vc5 40x129": lambda g: g.varcoef5(40, 129),
}


def apply(g, P, x, adjoint=False, dtype=np.float64):
    xv, yv = g.Vector(data=x.astype(dtype), dtype=dtype), g.Vector(len(x), dtype=dtype)
    (P.adjoint_solve if adjoint else P.solve)(xv, yv)
    return yv.numpy()


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_ic0_line_equals_tiled_and_oracle(g, name):
    print(name, flush=True)
    S, I, V = GRIDS[name](g)
    A = g.Matrix(S, I, V)
    Pl, Pt = make_pc(g, A, g.PC_IC0, False), make_pc(g, A, g.PC_IC0, True)
    st = Pl.stats()
    assert st["line_bands"] == (st["grid_ny"] + 63) // 64 > 0 and Pt.stats()["line_bands"] == 0
    assert g.PC(A, g.PC_IC0).stats()["line_bands"] == 0 or os.environ.get("M4G_LINE_SWEEP") == "1"
    Ao = o.CSR(S, I, V)
    _, U, L = o.ic0_factor(Ao)
    exact = np.array_equal(Pl.factor(1)[2], U.data)
    x = g.rng_uniform(9, 1, len(S) - 1)
    ref = o.ic0_solve(L, U, x)
    for rep in range(3):           # repeated applications reuse the exchange buffers
        yl = apply(g, Pl, x)
        np.testing.assert_array_equal(yl, apply(g, Pt, x))
        if exact:
            np.testing.assert_array_equal(yl, ref)
        assert np.max(np.abs(yl - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("name", ["lap 64x64", "vc5 70x55", "vc5 40x129"])
def test_ilu0_line_equals_tiled(g, name):
    print(name, flush=True)
    S, I, V = GRIDS[name](g)
    A = g.Matrix(S, I, V)
    Pl, Pt = make_pc(g, A, g.PC_ILU0, False), make_pc(g, A, g.PC_ILU0, True)
    assert Pl.stats()["line_bands"] > 0
    _, L, U = o.ilu0_factor(o.CSR(S, I, V))
    x = g.rng_uniform(10, 1, len(S) - 1)
    for adjoint in (False, True, False):
        yl = apply(g, Pl, x, adjoint)
        np.testing.assert_array_equal(yl, apply(g, Pt, x, adjoint))
        ref = o.ilu0_adjoint_solve(L, U, x) if adjoint else o.ilu0_solve(L, U, x)
        assert np.max(np.abs(yl - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("vdtype", [np.float64, np.float32])
def test_float_factors_line_equals_tiled(g, vdtype):
    S, I, V = g.varcoef5(130, 70)
    A = g.Matrix(S, I, V)
    Pl = make_pc(g, A, g.PC_ILU0, False, factor_dtype=np.float32)
    Pt = make_pc(g, A, g.PC_ILU0, True, factor_dtype=np.float32)
    assert Pl.stats()["line_bands"] == 2
    x = g.rng_uniform(11, 1, len(S) - 1)
    for adjoint in (False, True):
        np.testing.assert_array_equal(apply(g, Pl, x, adjoint, vdtype), apply(g, Pt, x, adjoint, vdtype))


def test_pcg_line_equals_tiled(g):
    # a whole IC(0)-PCG solve: the same iterations and the same solution bits
    S, I, V = g.varcoef5(256, 256)
    A = g.Matrix(S, I, V)
    n = len(S) - 1
    b = g.Vector(data=g.rng_uniform(12, 1, n))
    nb = g.two_norm(b)
    out = []
    for tiled in (False, True):
        P = make_pc(g, A, g.PC_IC0, tiled)
        x = g.Vector(n)
        x.fill(0.0)
        it, _ = g.make_iteration(5000, 1e-10, 0.0, nb, history=False)
        assert g.cg(A, x, b, P, it) == 0
        out.append((it.iterations, x.numpy()))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_structure_gate(g):
    # short lines, 3D grids and 9-point couplings keep the tiled sweeps
    for S, I, V in (g.laplacian2d(9, 7), g.poisson3d(16), g.convdiff27(12, 11, 10)):
        assert make_pc(g, g.Matrix(S, I, V), g.PC_ILU0, False).stats()["line_bands"] == 0
