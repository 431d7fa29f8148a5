"""The committed 8^3 convection-diffusion fixture (tests/golden/
convdiff27_8_seed3.npz, written by tests/golden/make_convdiff27_8.py from a
plain-Python statement of the stencil).  The library's generator must
reproduce it bit for bit, the oracle's SpMV must reproduce its b = A x_true,
and the matrix has the properties the stencil promises."""
import os

import numpy as np
import pytest

import oracle_lib as o
import mtl4gpu as g

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "convdiff27_8_seed3.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIXTURE) as f:
        return {k: f[k] for k in f.files}


def test_generator_reproduces_fixture(fx):
    S, I, V = g.convdiff27(8, seed=3, pe=0.5)
    np.testing.assert_array_equal(S, fx["starts"])
    np.testing.assert_array_equal(I, fx["indices"])
    np.testing.assert_array_equal(V, fx["values"])
    np.testing.assert_array_equal(g.rng_lognormal(3, 2, 512), fx["kappa"])
    np.testing.assert_array_equal(g.rng_uniform(3, 1, 512), fx["x_true"])


def test_oracle_spmv_reproduces_fixture_rhs(fx):
    A = o.CSR(fx["starts"], fx["indices"].astype(np.int64), fx["values"])
    np.testing.assert_array_equal(o.spmv(A, fx["x_true"]), fx["b"])


def test_fixture_properties(fx):
    S, I, V = fx["starts"], fx["indices"].astype(np.int64), fx["values"]
    n = 512
    assert len(S) == n + 1 and len(I) == (3 * 8 - 2) ** 3
    A = o.CSR(S, I, V)
    d = A.dense()
    for r in range(n):                                   # columns ascending, diagonal stored
        cols = I[S[r]:S[r + 1]]
        assert np.all(np.diff(cols) > 0) and r in cols
    off = d - np.diag(np.diag(d))
    assert np.all(off <= 0) and np.all(np.diag(d) > 0)   # M-matrix sign pattern (cell Peclet < 2)
    assert not np.allclose(d, d.T)
    # the symmetric part is the diffusion operator (convection is skew): its
    # interior rows sum to zero, its boundary rows keep the Dirichlet terms
    sym = 0.5 * (d + d.T)
    ii = np.arange(n)
    i, j, k = ii % 8, (ii // 8) % 8, ii // 64
    interior = (i > 0) & (i < 7) & (j > 0) & (j < 7) & (k > 0) & (k < 7)
    assert np.allclose(sym.sum(1)[interior], 0.0, atol=1e-12)
    assert np.all(sym.sum(1)[~interior] > 1e-3)


def test_oracle_bicgstab_ilu0_recovers_x_true(fx):
    A = o.CSR(fx["starts"], fx["indices"].astype(np.int64), fx["values"])
    ref = o.solve("bicgstab", A, fx["b"], pc=o.PC_ILU0, max_iter=500, rtol=1e-10)
    assert ref["status"] == o.OK
    assert np.max(np.abs(ref["x"] - fx["x_true"])) <= 1e-8
