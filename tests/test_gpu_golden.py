"""The GPU path on the committed 8^3 convection-diffusion fixture
(tests/golden/convdiff27_8_seed3.npz): SpMV reproduces the fixture's b bit for
bit; ILU(0) factors and one application match the oracle bit for bit; the
FP64 ILU(0)-BiCGStab and the FP32-inner / FP64-refined solves recover x_true."""
import os

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "convdiff27_8_seed3.npz")


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


@pytest.fixture(scope="module")
def fx(g):
    with np.load(FIXTURE) as f:
        d = {k: f[k] for k in f.files}
    d["indices"] = d["indices"].astype(np.int64)
    d["A"] = o.CSR(d["starts"], d["indices"], d["values"])
    d["Ag"] = g.Matrix(d["starts"], d["indices"], d["values"])
    return d


def test_golden_spmv(g, fx):
    print("golden convdiff27 8^3: spmv", flush=True)
    w = g.Vector(512)
    g.spmv(fx["Ag"], g.Vector(data=fx["x_true"]), w, g.ASSIGN)
    np.testing.assert_array_equal(w.numpy(), fx["b"])


def test_golden_ilu0_factor_and_apply(g, fx):
    print("golden convdiff27 8^3: ilu0", flush=True)
    st, L, U = o.ilu0_factor(fx["A"])
    assert st == o.OK
    P = g.PC(fx["Ag"], g.PC_ILU0)
    for which, ref in ((0, L), (1, U)):
        S, I, V = P.factor(which)
        np.testing.assert_array_equal(S, ref.starts)
        np.testing.assert_array_equal(I, ref.indices)
        np.testing.assert_array_equal(V, ref.data)
    y = g.Vector(512)
    P.solve(g.Vector(data=fx["b"]), y)
    np.testing.assert_array_equal(y.numpy(), o.ilu0_solve(L, U, fx["b"]))
    P.adjoint_solve(g.Vector(data=fx["b"]), y)
    np.testing.assert_array_equal(y.numpy(), o.ilu0_adjoint_solve(L, U, fx["b"]))


def test_golden_bicgstab_fp64(g, fx):
    print("golden convdiff27 8^3: bicgstab", flush=True)
    b = fx["b"]
    nb = o.two_norm(b)
    ref = o.solve("bicgstab", fx["A"], b, pc=o.PC_ILU0, max_iter=500, rtol=1e-10)
    x = g.Vector(512)
    x.fill(0.0)
    it, _ = g.make_iteration(500, 1e-10, 0.0, nb)
    st = g.bicgstab(fx["Ag"], x, g.Vector(data=b), g.PC(fx["Ag"], g.PC_ILU0), it)
    assert st == 0
    # reductions in another order than the oracle's: at most one iteration apart
    assert abs(it.iterations - ref["iterations"]) <= 1
    assert it.resid <= 1e-10 * nb
    assert np.max(np.abs(x.numpy() - fx["x_true"])) <= 1e-8


def test_golden_refined(g, fx):
    print("golden convdiff27 8^3: refined", flush=True)
    b = fx["b"]
    nb = o.two_norm(b)
    x = g.Vector(512)
    x.fill(0.0)
    it, _ = g.make_iteration(50, 1e-10, 0.0, nb)
    P32 = g.PC(fx["Ag"], g.PC_ILU0, factor_dtype=np.float32)
    st, info = g.refined_solve(g.METHOD_BICGSTAB, fx["Ag"], x, g.Vector(data=b), P32, it, 1e-4)
    assert st == 0
    xs = x.numpy()
    assert o.two_norm(o.residual(fx["A"], xs, b)) <= 1e-10 * nb * 1.01
    assert np.max(np.abs(xs - fx["x_true"])) <= 1e-8
