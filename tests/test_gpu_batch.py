"""Batched IC(0)-PCG (one factorisation launch, one persistent solve launch)
against the oracle, system by system and right-hand side by right-hand side:
factors bit for bit, iteration counts within 2 % (at least 1) of the
oracle's itl::cg + ic_0 on the same system and rhs, residuals at the target,
solutions close to the oracle's."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


def make_systems(g, sides, seed0=11):
    systems = [g.varcoef5(s, s, seed=seed0 + k) for k, s in enumerate(sides)]
    return systems


def interleave(vecs_per_sys, nrhs):
    # entry (system s, row i, rhs j) at (row0(s) + i) * nrhs + j
    return np.concatenate([np.stack(v, axis=1).reshape(-1) for v in vecs_per_sys])


@pytest.mark.parametrize("block,rpi", [(512, None), (256, 2), (1024, 1)])
def test_batch_matches_single_solves(g, block, rpi):
    sides = [7, 16, 5, 23, 12, 31, 9, 18]
    nrhs = 4
    print(f"batch block={block} rhs_per_item={rpi}", flush=True)
    systems = make_systems(g, sides)
    B = g.Batch(systems)
    B.factor_ic0()
    # factors: each system's upper(A) storage, inverted diagonal first (ic_0.hpp:148-185)
    fac = B.factor_values()
    off = 0
    for S, I, V in systems:
        A = o.CSR(S, I, V)
        st, U, _ = o.ic0_factor(A)
        assert st == o.OK
        np.testing.assert_array_equal(fac[off:off + len(U.data)], U.data)
        off += len(U.data)
    assert off == len(fac)
    # right-hand sides
    bs = []
    for k, (S, I, V) in enumerate(systems):
        n = len(S) - 1
        bs.append([g.rng_normal(21 + k, j, n) for j in range(nrhs)])
    bh = interleave(bs, nrhs)
    x = g.Vector(len(bh))
    x.fill(0.0)
    b = g.Vector(data=bh)
    rtol = 1e-8
    rpi_arr = None if rpi is None else [rpi] * len(sides)
    its, res, err = B.cg_ic0(nrhs, x, b, 1000, rtol, rhs_per_item=rpi_arr, block=block)
    xs = x.numpy()
    for k, (S, I, V) in enumerate(systems):
        A = o.CSR(S, I, V)
        n = A.n
        r0 = int(B.row0[k])
        for j in range(nrhs):
            ref = o.solve("cg", A, bs[k][j], pc=o.PC_IC0, max_iter=1000, rtol=rtol)
            assert err[k, j] == 0 and ref["status"] == 0
            slack = max(1, int(np.ceil(0.02 * ref["iterations"])))
            assert abs(its[k, j] - ref["iterations"]) <= slack, (k, j, its[k, j], ref["iterations"])
            nb = o.two_norm(bs[k][j])
            assert res[k, j] <= rtol * nb
            xk = xs[(r0 * nrhs + j):((r0 + n) * nrhs):nrhs]
            assert o.two_norm(o.residual(A, xk, bs[k][j])) <= 10 * rtol * nb
            assert np.max(np.abs(xk - ref["x"])) <= 1e-6 * max(1.0, np.max(np.abs(ref["x"])))


def test_batch_max_iter_and_zero_rhs(g):
    systems = make_systems(g, [6, 10])
    B = g.Batch(systems)
    B.factor_ic0()
    nrhs = 2
    bs = [[g.rng_normal(3, 0, 36), np.zeros(36)], [g.rng_normal(3, 1, 100), g.rng_normal(3, 2, 100)]]
    bh = interleave(bs, nrhs)
    x = g.Vector(len(bh))
    x.fill(0.0)
    b = g.Vector(data=bh)
    its, res, err = B.cg_ic0(nrhs, x, b, 2, 1e-14)
    # norm_r0 == 0: converged at once by the atol rule (basic_iteration.hpp:90-91)
    assert err[0, 1] == 0 and its[0, 1] == 0
    for k, j in [(0, 0), (1, 0), (1, 1)]:
        assert err[k, j] == g.MAX_ITER and its[k, j] == 2
        A = o.CSR(*systems[k])
        ref = o.solve("cg", A, bs[k][j], pc=o.PC_IC0, max_iter=2, rtol=1e-14)
        assert ref["status"] == o.MAX_ITER and ref["iterations"] == 2


def test_batch_missing_diagonal(g):
    S, I, V = g.varcoef5(4, 4)
    I2 = np.array(I).copy()
    V2 = np.array(V).copy()
    # drop the diagonal of row 5 by moving it off the diagonal is not sorted; use a pattern without it
    rows = np.repeat(np.arange(16), np.diff(S))
    keep = ~((rows == 5) & (I2 == 5))
    S2 = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=16))])
    with pytest.raises(g.M4GError) as e:
        g.Batch([(S, I, V), (S2, I2[keep], V2[keep])])
    assert e.value.status == g.ERR_MISSING_DIAGONAL
