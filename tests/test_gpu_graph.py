"""Iteration batches replayed as a hipGraph against eager launches.

CG and BiCGStab run their iterations in batches; with M4G_KRYLOV_GRAPH=1 a
batch from the second on is captured once and replayed (by default every
batch is launched eagerly).  The replays do the same work as the eager launches, in the same
order, so the two solves are bit-identical: same status and iteration count,
the same residual history (every replay's iterations are in it, with the
host's own copies and synchronisations between the replays) and the same x.
The counts are also held to the oracle's (cg.hpp:75-112, bicgstab.hpp:27-78),
within 2 % of the range the oracle itself spans over rounding-level changes
of b.
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


CASES = [
    # (name, matrix, method, PC, iteration count held to the oracle's)
    ("p3d40_ic0_cg", lambda g: g.poisson3d(40), "cg", "PC_IC0", True),
    ("vc5_ilu0_cg", lambda g: g.varcoef5(120, 90, seed=4), "cg", "PC_ILU0", True),
    ("lap2d_diag_cg", lambda g: g.laplacian2d(150, 140), "cg", "PC_DIAGONAL", True),
    ("lap2d_none_cg", lambda g: g.laplacian2d(150, 140), "cg", "PC_IDENTITY", True),
    ("cd27_ilu0_bicgstab", lambda g: g.convdiff27(40, 36, 32), "bicgstab", "PC_ILU0", True),
    # unpreconditioned BiCGStab on this nonsymmetric system moves by tens of
    # iterations under 1e-15 changes of b (the oracle: 172-201), so only the
    # graph/eager identity and the true residual are checked
    ("cd27_none_bicgstab", lambda g: g.convdiff27(40, 36, 32), "bicgstab", "PC_IDENTITY", False),
]
OKIND = {"PC_IC0": o.PC_IC0, "PC_ILU0": o.PC_ILU0, "PC_DIAGONAL": 1, "PC_IDENTITY": 0}


@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_graph_batches_equal_eager(g, case, monkeypatch):
    name, make, method, pc, held = next(c for c in CASES if c[0] == case)
    S, I, V = make(g)
    A = o.CSR(S, I, V)
    b = o.spmv(A, g.rng_uniform(11, 1, A.n))
    nb = o.two_norm(b)
    rtol = 1e-10 if method == "cg" else 1e-11

    def run():
        Ag = g.Matrix(S, I, V)
        P = g.PC(Ag, getattr(g, pc))
        x = g.Vector(A.n)
        x.fill(0.0)
        it, hist = g.make_iteration(4000, rtol, 0.0, nb)
        st = getattr(g, method)(Ag, x, g.Vector(data=b), P, it)
        return st, it.iterations, x.numpy(), hist[: it.iterations + 1].copy()

    monkeypatch.setenv("M4G_KRYLOV_GRAPH", "0")
    eager = run()
    monkeypatch.setenv("M4G_KRYLOV_GRAPH", "1")
    graph = run()
    say(f"  {case}: {graph[1]} iterations replayed, {eager[1]} eager")
    assert eager[1] > 16        # several batches: the graph was captured and replayed
    assert graph[0] == eager[0] == 0 and graph[1] == eager[1]
    assert np.array_equal(graph[3], eager[3])
    assert np.array_equal(graph[2], eager[2])
    assert o.two_norm(o.residual(A, graph[2], b)) <= 20 * rtol * nb
    if not held:
        return
    # the oracle on b and on b * (1 + 1e-15 noise)
    spread = []
    for k in range(4):
        bb = b if k == 0 else b * (1.0 + 1e-15 * np.random.default_rng(k).standard_normal(A.n))
        ref = o.solve(method, A, bb, pc=OKIND[pc], max_iter=4000, rtol=rtol, history=False)
        assert ref["status"] == 0
        spread.append(ref["iterations"])
    slack = max(1, int(np.ceil(0.02 * max(spread))))
    assert min(spread) - slack <= graph[1] <= max(spread) + slack, (graph[1], spread)
