"""itl::cgs, itl::bicg and itl::tfqmr on the GPU against the oracle's
restatements (oracle/mtl4_oracle.cpp: orc_cgs, orc_bicg, orc_tfqmr): same
iteration count within 2 % (at least 1; dots are reduced in a different
order), true residual at the target, solution close to the oracle's."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


CASES = [
    # (matrix, preconditioners)
    ("lap", lambda g: g.laplacian2d(32, 32), ("none", "diag", "ic0")),
    ("vc5", lambda g: g.varcoef5(30, 27), ("ic0", "ilu0")),
    ("cd27", lambda g: g.convdiff27(10, 9, 8), ("none", "ilu0")),
]
PCS = {"none": ("PC_IDENTITY", 0), "diag": ("PC_DIAGONAL", 1), "ic0": ("PC_IC0", 2), "ilu0": ("PC_ILU0", 3)}


@pytest.mark.parametrize("method", ["cgs", "bicg", "tfqmr"])
@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_solver_matches_oracle(g, method, case):
    name, make, pcs = next(c for c in CASES if c[0] == case)
    S, I, V = make(g)
    A = o.CSR(S, I, V)
    n = A.n
    b = g.rng_uniform(8, 1, n)
    nb = o.two_norm(b)
    rtol = 1e-9
    Ag = g.Matrix(S, I, V)
    for pc in pcs:
        print(f"{method} {name} {pc}", flush=True)
        gname, okind = PCS[pc]
        ref = o.solve(method, A, b, pc=okind, max_iter=600, rtol=rtol, history=False)
        P = g.PC(Ag, getattr(g, gname))
        x = g.Vector(n)
        x.fill(0.0)
        it, _ = g.make_iteration(600, rtol, 0.0, nb, history=False)
        fn = {"cgs": g.cgs, "bicg": g.bicg, "tfqmr": g.tfqmr}[method]
        st = fn(Ag, x, g.Vector(data=b), P, it)
        if ref["status"] != 0:
            # MTL4's bicg takes rho_1 = dot(z_tilde, z) (bicg.hpp:42), which with a
            # preconditioner is not PBiCG's z_tilde.r and can stall: both sides must
            # then run out of iterations alike
            assert ref["status"] == g.MAX_ITER and st == g.MAX_ITER and it.iterations == ref["iterations"]
            continue
        assert st == 0, (method, name, pc, st)
        slack = max(1, int(np.ceil(0.02 * ref["iterations"])))
        assert abs(it.iterations - ref["iterations"]) <= slack, (method, name, pc, it.iterations, ref["iterations"])
        xs = x.numpy()
        assert o.two_norm(o.residual(A, xs, b)) <= 20 * rtol * nb
        assert np.max(np.abs(xs - ref["x"])) <= 1e-6 * max(1.0, np.max(np.abs(ref["x"])))


def test_zero_rhs_and_max_iter(g):
    S, I, V = g.laplacian2d(12, 12)
    Ag = g.Matrix(S, I, V)
    P = g.PC(Ag, g.PC_IC0)
    n = len(S) - 1
    for fn in (g.cgs, g.bicg, g.tfqmr):
        x = g.Vector(n)
        x.fill(0.0)
        it, _ = g.make_iteration(100, 1e-10, 0.0, 0.0, history=False)   # norm_r0 = 0: converged at once
        assert fn(Ag, x, g.Vector(data=np.zeros(n)), P, it) == 0 and it.iterations == 0
        it, _ = g.make_iteration(2, 1e-14, 0.0, 1.0, history=False)
        x.fill(0.0)
        assert fn(Ag, x, g.Vector(data=np.ones(n)), P, it) == g.MAX_ITER and it.iterations == 2
