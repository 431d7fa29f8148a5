"""A real matrix from the reference's fixtures on the GPU path: bcsstk01
(48x48 stiffness matrix, T-mtl/matrix_market/bcsstk01.mtx) read through the
Matrix Market reader, IC(0)-factorised and solved with itl::cg.  Its graph has
triangles, so the IC(0) factor exercises the diagonal-only Schur update on a
pattern where it differs from a complete-graph reading (ic_0.hpp:153-185);
the GPU factor, iteration count and solution are compared with the oracle."""
import os

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


@pytest.mark.parametrize("rtol", [1e-8, 1e-12])
def test_bcsstk01_ic0_cg(g, rtol):
    print(f"bcsstk01 ic0-cg rtol={rtol}", flush=True)
    S, I, V, n, _ = g.mm_read(os.path.join(HERE, "golden", "matrix_market", "bcsstk01.mtx"))
    A = o.CSR(S, I, V)
    st, U, _ = o.ic0_factor(A)
    assert st == o.OK
    Ag = g.Matrix(S, I, V)
    P = g.PC(Ag, g.PC_IC0)
    Sg, Ig, Vg = P.factor(1)
    np.testing.assert_array_equal(Sg, U.starts)
    np.testing.assert_array_equal(Ig, U.indices)
    np.testing.assert_array_equal(Vg, U.data)
    b = g.rng_uniform(5, 1, n)
    ref = o.solve("cg", A, b, pc=o.PC_IC0, max_iter=1000, rtol=rtol)
    x = g.Vector(n)
    x.fill(0.0)
    it, _ = g.make_iteration(1000, rtol, 0.0, o.two_norm(b), history=False)
    assert g.cg(Ag, x, g.Vector(data=b), P, it) == 0
    assert abs(it.iterations - ref["iterations"]) <= max(1, int(np.ceil(0.02 * ref["iterations"])))
    nb = o.two_norm(b)
    assert o.two_norm(o.residual(A, x.numpy(), b)) <= 10 * rtol * nb
    xs = x.numpy()
    assert np.max(np.abs(xs - ref["x"])) <= 1e-6 * np.max(np.abs(ref["x"]))
