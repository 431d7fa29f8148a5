"""CPU tests of the product's host logic: generators, grid inference, the
sweep schedule compiler (emulated on the CPU with the kernel's operation
order and compared bit for bit with the oracle) and the C ABI exports."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as o
import mtl4gpu as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(g.LIB_PATH)
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(m4g_\w+)\s*\(", text))
    assert len(names) > 40
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_laplacian2d_matches_reference_generator():
    S, I, V = g.laplacian2d(5, 7)
    A = o.laplacian_setup(5, 7)
    np.testing.assert_array_equal(S, A.starts)
    np.testing.assert_array_equal(I, A.indices)
    np.testing.assert_array_equal(V, A.data)


def test_poisson3d_counts_and_stencil():
    S, I, V = g.poisson3d(128, 2, 3)
    n = 128 * 2 * 3
    assert len(S) == n + 1
    S, I, V = g.poisson3d(9)
    A = o.CSR(S, I, V)
    assert A.nnz == 7 * 729 - 6 * 81
    d = A.dense()
    assert np.allclose(d, d.T) and np.all(np.diag(d) == 6)
    assert np.all(np.abs(d.sum(1)) <= 6) and d.sum() > 0
    # row sums: 6 - #neighbours
    nb = np.array([(S[r + 1] - S[r] - 1) for r in range(729)])
    np.testing.assert_array_equal(d.sum(1), 6 - nb)


def test_poisson3d_128_nnz():
    # only the count (generation of 14.6M entries takes a moment)
    nnz = g._f["m4g_gen_poisson3d"](128, 128, 128, None, None, None)
    assert nnz == 14581760


def test_convdiff27_properties():
    S, I, V = g.convdiff27(8, seed=3, pe=0.5)
    A = o.CSR(S, I, V)
    assert A.nnz == (3 * 8 - 2) ** 3
    d = A.dense()
    assert not np.allclose(d, d.T)              # convection makes it nonsymmetric
    off = d - np.diag(np.diag(d))
    assert np.all(off <= 0)                     # cell Peclet < 2: M-matrix sign pattern
    assert np.all(np.diag(d) > 0)
    # symmetric part is the diffusion operator: row sums >= 0 (Dirichlet boundary)
    sym = 0.5 * (d + d.T)
    assert np.all(sym.sum(1) > -1e-12)
    # diffusion-only part (pe = 0) is symmetric
    S0, I0, V0 = g.convdiff27(8, seed=3, pe=0.0)
    d0 = o.CSR(S0, I0, V0).dense()
    assert np.allclose(d0, d0.T)
    assert g._f["m4g_gen_convdiff27"](96, 96, 96, 3, 0.5, None, None, None) == 23393656


def test_varcoef5_spd():
    S, I, V = g.varcoef5(12, seed=4)
    d = o.CSR(S, I, V).dense()
    assert np.allclose(d, d.T)
    assert np.all(np.linalg.eigvalsh(d) > 0)


def test_edge_laplacian_assembly():
    nx, ny = 6, 5
    e = g.edge_laplacian(nx, ny)
    assert e["n_edges"] == (nx - 1) * ny + nx * (ny - 1) + 2 * (nx + ny)
    theta = g.rng_lognormal(7, 9, e["n_edges"])
    e2 = g.edge_laplacian(nx, ny, theta)
    # values = M theta through the assembly map
    vals = np.zeros(len(e2["I"]))
    for p in range(len(vals)):
        for q in range(e2["aptr"][p], e2["aptr"][p + 1]):
            vals[p] += e2["acoef"][q] * theta[e2["aidx"][q]]
    np.testing.assert_allclose(vals, e2["V"], rtol=1e-15)
    # A = sum theta_e (e_i - e_j)(e_i - e_j)^T + boundary
    A = np.zeros((nx * ny, nx * ny))
    for k in range(e2["n_edges"]):
        i, j = e2["ei"][k], e2["ej"][k]
        if j < 0:
            A[i, i] += theta[k]
        else:
            A[i, i] += theta[k]; A[j, j] += theta[k]; A[i, j] -= theta[k]; A[j, i] -= theta[k]
    np.testing.assert_allclose(o.CSR(e2["S"], e2["I"], e2["V"]).dense(), A, rtol=1e-14)


def test_rng_reproducible_and_distributed():
    a = g.rng_uniform(11, 1, 100000)
    b = g.rng_uniform(11, 1, 10, offset=500)
    np.testing.assert_array_equal(a[500:510], b)
    assert -1 <= a.min() and a.max() < 1 and abs(a.mean()) < 0.01
    z = g.rng_normal(11, 2, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    k = g.rng_int(5, 3, 10000, 32, 256)
    assert k.min() >= 32 and k.max() <= 256


@pytest.mark.parametrize("gen,expect", [
    (lambda: g.laplacian2d(64, 64), (64, 64, 1)),
    (lambda: g.poisson3d(20, 12, 7), (20, 12, 7)),
    (lambda: g.convdiff27(10, 9, 8), (10, 9, 8)),
    (lambda: g.varcoef5(33, 17), (33, 17, 1)),
])
def test_grid_inference(gen, expect):
    S, I, V = gen()
    assert g.analysis_grid(S, I) == expect


def test_grid_inference_rejects_random_pattern():
    rng = np.random.default_rng(0)
    n = 200
    S, I = [0], []
    for r in range(n):
        cols = sorted(set(rng.integers(0, n, 6).tolist()) | {r})
        I += cols
        S.append(len(I))
    assert g.analysis_grid(np.array(S), np.array(I)) == (0, 0, 0)


def _ilu_storage(A, L, U):
    fac = np.zeros(A.nnz)
    for r in range(A.n):
        a = A.starts[r]
        nl = L.starts[r + 1] - L.starts[r]
        fac[a:a + nl] = L.data[L.starts[r]:L.starts[r + 1]]
        fac[a + nl:A.starts[r + 1]] = U.data[U.starts[r]:U.starts[r + 1]]
    return fac


CASES = {
    "lap3": lambda: g.laplacian2d(3, 3),
    "lap64": lambda: g.laplacian2d(64, 64),
    "lap_rect": lambda: g.laplacian2d(40, 97),
    "p3d": lambda: g.poisson3d(18, 11, 9),
    "vc5": lambda: g.varcoef5(45, 31),
    "cd27": lambda: g.convdiff27(11, 10, 9),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("kind", [g.PC_IC0, g.PC_ILU0])
@pytest.mark.parametrize("tile_rows", [0, 96])
def test_sweep_schedule_emulation_bitexact(case, kind, tile_rows):
    S, I, V = CASES[case]()
    A = o.CSR(S, I, V)
    if kind == g.PC_IC0 and case == "cd27":
        pytest.skip("IC(0) needs a symmetric matrix")
    rhs = g.rng_uniform(3, 5, A.n)
    if kind == g.PC_IC0:
        st, U, L = o.ic0_factor(A)
        ref = o.ic0_solve(L, U, rhs)
        fac = U.data
    else:
        st, L, U = o.ilu0_factor(A)
        ref = o.ilu0_solve(L, U, rhs)
        fac = _ilu_storage(A, L, U)
    assert st == o.OK
    out, stats = g.analysis_pc_solve(kind, S, I, fac, rhs, max_tile_rows=tile_rows)
    np.testing.assert_array_equal(out, ref)
    if kind == g.PC_ILU0:
        outa, _ = g.analysis_pc_solve(kind, S, I, fac, rhs, adjoint=True, max_tile_rows=tile_rows)
        np.testing.assert_array_equal(outa, o.ilu0_adjoint_solve(L, U, rhs))
    if tile_rows and A.n > tile_rows:
        assert stats["tiles_lower"] > 1


def test_sweep_schedule_float_factors_bitexact():
    # ic_0<Matrix, float>: float factors, double accumulation (ic_0.hpp:122-130)
    S, I, V = g.laplacian2d(30, 30)
    A = o.CSR(S, I, V)
    _, U, L = o.ic0_factor(A)
    rhs = g.rng_uniform(3, 6, A.n)
    out, _ = g.analysis_pc_solve(g.PC_IC0, S, I, U.data, rhs, factor_dtype=np.float32, max_tile_rows=128)
    np.testing.assert_array_equal(out, o.ic0_solve(L, U, rhs, factor_dtype=np.float32))


def test_sweep_schedule_general_pattern():
    # a pattern that is not a grid: row-block tiles
    from test_oracle import _rand_sparse
    A, Ad = _rand_sparse(300, seed=9, extra=3)
    _, L, U = o.ilu0_factor(A)
    fac = _ilu_storage(A, L, U)
    rhs = g.rng_uniform(1, 2, A.n)
    for cap in (0, 64):
        out, st = g.analysis_pc_solve(g.PC_ILU0, A.starts, A.indices, fac, rhs, max_tile_rows=cap)
        assert st["grid_nx"] == 0
        np.testing.assert_array_equal(out, o.ilu0_solve(L, U, rhs))


@pytest.mark.parametrize("width", [3, 4, 8])
def test_sweep_schedule_split_rows(width, monkeypatch):
    # rows longer than the piece width run as several pieces that carry the
    # running value through LDS; the operation order, hence the result, is
    # unchanged
    monkeypatch.setenv("M4G_SWEEP_WIDTH", str(width))
    S, I, V = g.convdiff27(9, 8, 7)
    A = o.CSR(S, I, V)
    _, L, U = o.ilu0_factor(A)
    fac = _ilu_storage(A, L, U)
    rhs = g.rng_uniform(5, 1, A.n)
    for cap in (0, 100):
        out, _ = g.analysis_pc_solve(g.PC_ILU0, S, I, fac, rhs, max_tile_rows=cap)
        np.testing.assert_array_equal(out, o.ilu0_solve(L, U, rhs))
        outa, _ = g.analysis_pc_solve(g.PC_ILU0, S, I, fac, rhs, adjoint=True, max_tile_rows=cap)
        np.testing.assert_array_equal(outa, o.ilu0_adjoint_solve(L, U, rhs))


def test_sweep_schedule_long_rows():
    # rows of more than 16 entries (the widest piece) on a general pattern
    from test_oracle import _rand_sparse
    A, _ = _rand_sparse(200, seed=4, extra=24)
    _, L, U = o.ilu0_factor(A)
    assert max(np.diff(L.starts)) > 16
    fac = _ilu_storage(A, L, U)
    rhs = g.rng_uniform(2, 2, A.n)
    for cap in (0, 64):
        out, _ = g.analysis_pc_solve(g.PC_ILU0, A.starts, A.indices, fac, rhs, max_tile_rows=cap)
        np.testing.assert_array_equal(out, o.ilu0_solve(L, U, rhs))
