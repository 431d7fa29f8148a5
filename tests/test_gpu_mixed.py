"""FP32-inner / FP64-refined solves: the refined solution must reach the FP64
residual target, and its error against the FP64 oracle solution is bounded by
the conditioning (stated per case)."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


@pytest.mark.parametrize("case", ["cd27_bicgstab_ilu0", "p3d_cg_ic0", "vc5_cg_ic0"])
def test_refined_solve(g, case):
    print(f"mixed {case}", flush=True)
    if case == "cd27_bicgstab_ilu0":
        S, I, V = g.convdiff27(20)
        method, pc, opc, tol = g.METHOD_BICGSTAB, g.PC_ILU0, o.PC_ILU0, 1e-10
    elif case == "p3d_cg_ic0":
        S, I, V = g.poisson3d(30)
        method, pc, opc, tol = g.METHOD_CG, g.PC_IC0, o.PC_IC0, 1e-10
    else:
        S, I, V = g.varcoef5(60, 50)
        method, pc, opc, tol = g.METHOD_CG, g.PC_IC0, o.PC_IC0, 1e-9
    A = o.CSR(S, I, V)
    x_true = g.rng_uniform(3, 1, A.n)
    b = o.spmv(A, x_true)
    ref = o.solve("bicgstab" if method == g.METHOD_BICGSTAB else "cg", A, b, pc=opc, max_iter=5000, rtol=tol)
    assert ref["status"] == 0
    Ag = g.Matrix(S, I, V)
    P32 = g.PC(Ag, pc, factor_dtype=np.float32)
    x = g.Vector(A.n)
    x.fill(0.0)
    bv = g.Vector(data=b)
    it, hist = g.make_iteration(50, tol, 0.0, o.two_norm(b))
    st, info = g.refined_solve(method, Ag, x, bv, P32, it, 1e-4)
    xs = x.numpy()
    relres = o.two_norm(o.residual(A, xs, b)) / o.two_norm(b)
    err = np.max(np.abs(xs - ref["x"])) / np.max(np.abs(ref["x"]))
    print(f"  outer {info.outer_iterations} inner {info.inner_iterations} relres {relres:.2e} "
          f"err vs fp64 oracle {err:.2e} (oracle {ref['iterations']} it)", flush=True)
    assert st == 0
    assert relres <= tol * 1.01
    assert err < 1e-6
    assert 1 <= info.outer_iterations <= 6
