"""itl::cgs, bicg, tfqmr, qmr, gmres, bicgstab_2, bicgstab_ell and idr_s on the
GPU against the oracle's restatements (oracle/mtl4_oracle.cpp).  cgs, bicg and
tfqmr: same status, iteration count within 2 % (at least 1; dots are reduced
in a different order), solution close to the oracle's.  The others follow the
reference's own control flow (QMR tests the Lanczos vector's norm,
BiCGStab(l) leaves before its last x update, GMRES rotates H its own way,
IDR(s) takes s steps per cycle), whose outcome moves with rounding: they are
held to the spread of the oracle over rounding-level changes of b."""
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


def test_adjoint_products_follow_new_values(g):
    """bicg and qmr keep trans(A)'s pattern with A and gather its values at
    each solve: after set_values both follow the oracle on the new system (and
    on the first one again once its values are put back).  bicg converges
    within the oracle's own iteration range over 1e-14 changes of b (66-70
    here: the residual wanders near the tolerance) +- 2 %; MTL4's qmr tests
    the Lanczos vector's norm, which
    here stays above 1e-8 * ||r0||, so it runs its 60 iterations on both
    sides and x is compared (the oracle's own spread is about 1e-8)."""
    S, I, V = g.convdiff27(9, 8, 7)
    S2, I2 = np.asarray(S), np.asarray(I)
    rows = np.repeat(np.arange(len(S2) - 1), np.diff(S2))
    V1 = np.asarray(V, dtype=np.float64)
    # a nonsymmetric change of the off-diagonal entries: trans(A) changes too
    V2 = np.where(rows != I2, V1 * (1.0 + 0.3 * np.sin(rows + 2.0 * I2)), V1)
    n = len(S2) - 1
    b = g.rng_uniform(5, 1, n)
    nb = o.two_norm(b)
    rtol = 1e-8
    Ag = g.Matrix(S, I, V1)
    P = g.PC(Ag, g.PC_IDENTITY)
    for vals in (V1, V2, V1):
        Ag.set_values(vals)
        A = o.CSR(S, I, vals)
        for method, max_iter in (("bicg", 400), ("qmr", 60)):
            print(f"{method} after set_values", flush=True)
            ref = o.solve(method, A, b, pc=0, max_iter=max_iter, rtol=rtol, history=False)
            x = g.Vector(n)
            x.fill(0.0)
            it, _ = g.make_iteration(max_iter, rtol, 0.0, nb, history=False)
            bv = g.Vector(data=b)
            st = g.bicg(Ag, x, bv, P, it) if method == "bicg" else g.qmr(Ag, x, bv, P, it)
            xs = x.numpy()
            if method == "bicg":
                assert st == 0 and ref["status"] == 0, (st, ref["status"])
                spread = [ref["iterations"]]
                for k in range(1, 12):
                    bb = b * (1.0 + 1e-14 * np.random.default_rng(k).standard_normal(n))
                    spread.append(o.solve(method, A, bb, pc=0, max_iter=max_iter, rtol=rtol, history=False)["iterations"])
                slack = max(1, int(np.ceil(0.02 * max(spread))))
                assert min(spread) - slack <= it.iterations <= max(spread) + slack, (it.iterations, spread)
                assert o.two_norm(o.residual(A, xs, b)) <= 20 * rtol * nb
            else:
                assert st == g.MAX_ITER and ref["status"] == g.MAX_ITER and it.iterations == max_iter
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


MORE = [
    # (method, param, right preconditioner?)
    ("qmr", None, False),
    ("qmr", None, True),
    ("gmres", 1, False),
    ("gmres", 8, False),
    ("gmres", 4, True),
    ("gmres_full", None, False),
    ("bicgstab_2", None, False),
    ("bicgstab_ell", 1, False),
    ("bicgstab_ell", 4, False),
    ("bicgstab_ell", 3, True),
    ("idr_s", 1, False),
    ("idr_s", 4, False),
]


def _run_gpu(g, method, param, Ag, x, bv, P, R, it):
    if method == "qmr":
        return g.qmr(Ag, x, bv, P, it, R=R)
    if method == "gmres":
        return g.gmres(Ag, x, bv, P, it, restart=param, R=R)
    if method == "gmres_full":
        return g.gmres_full(Ag, x, bv, P, it, R=R)
    if method == "bicgstab_2":
        return g.bicgstab_2(Ag, x, bv, P, it)
    if method == "bicgstab_ell":
        return g.bicgstab_ell(Ag, x, bv, P, it, l=param, R=R)
    if method == "idr_s":
        return g.idr_s(Ag, x, bv, it, s=param, L=P, R=R)
    raise ValueError(method)


def _oracle_spread(A, b, method, okind, right, param, max_iter, rtol, runs=6):
    """The oracle on b and on b * (1 + 1e-15 * noise) (runs - 1 draws): how
    far the reference's own outcome moves under rounding-level changes."""
    out = []
    for k in range(runs):
        bb = b if k == 0 else b * (1.0 + 1e-15 * np.random.default_rng(k).standard_normal(A.n))
        o.lib.orc_set_random_counter(0)
        out.append(o.solve(method, A, bb, pc=okind, max_iter=max_iter, rtol=rtol, history=False,
                           right=okind if right else None, param=param))
    return out


@pytest.mark.parametrize("method,param,right", MORE, ids=[f"{m}-{p}-{'R' if r else 'L'}" for m, p, r in MORE])
@pytest.mark.parametrize("case", [c[0] for c in CASES])
def test_more_solvers_match_oracle(g, method, param, right, case):
    # Iteration counts of BiCGStab(l), IDR(s) and MTL4's GMRES move by several
    # steps when b changes by 1e-15 (e.g. IDR(4) on vc5: 208-259), so no
    # differently ordered dot can hold a fixed 2 % to one oracle run.  The bar
    # is the oracle's own spread over six such runs: status among theirs,
    # iterations within [min, max] +- 2 %, and x within 1e-6 (or 20x their x
    # spread) where that spread is below 1e-3; beyond that (GMRES(8),
    # gmres_full, IDR(1)) x is chaotic and not compared.
    name, make, pcs = next(c for c in CASES if c[0] == case)
    S, I, V = make(g)
    A = o.CSR(S, I, V)
    n = A.n
    b = g.rng_uniform(9, 1, n)
    nb = o.two_norm(b)
    rtol = 1e-9
    max_iter = 80 if method in ("qmr", "gmres", "gmres_full") else 400
    Ag = g.Matrix(S, I, V)
    for pc in pcs:
        print(f"{method}({param}) {name} {pc} right={right}", flush=True)
        gname, okind = PCS[pc]
        runs = _oracle_spread(A, b, method, okind, right, param, max_iter, rtol)
        ref = runs[0]
        P = g.PC(Ag, getattr(g, gname))
        x = g.Vector(n)
        x.fill(0.0)
        it, _ = g.make_iteration(max_iter, rtol, 0.0, nb, history=False)
        g.random_seed_counter(0)
        st = _run_gpu(g, method, param, Ag, x, g.Vector(data=b), P, P if right else None, it)
        xs = x.numpy()
        if st == g.OK and method in ("gmres", "gmres_full", "idr_s"):
            # these test the true residual b - A x before they stop
            assert o.two_norm(o.residual(A, xs, b)) <= 20 * rtol * nb
        scale = max(1.0, np.max(np.abs(ref["x"])))
        spread = max(np.max(np.abs(r["x"] - ref["x"])) for r in runs[1:])
        if not (np.all(np.isfinite(ref["x"])) and spread <= 1e-3 * scale):
            # chaotic in the reference itself (e.g. IDR(1) on the Laplacian
            # converges for about 1 in 5 such perturbations): the GPU's outcome
            # must be one the method can produce, stated honestly
            print(f"  chaotic: the oracle's x moves by {spread:.1e}; GPU status {st}, {it.iterations} its")
            assert st in (g.OK, g.MAX_ITER, g.BREAKDOWN)
            if st == g.MAX_ITER:
                assert it.iterations == max_iter
            continue
        to_status = lambda e: {0: g.OK, 1: g.MAX_ITER}.get(e, g.BREAKDOWN)
        assert st in {to_status(r["status"]) for r in runs}, (st, [r["status"] for r in runs])
        if st == g.BREAKDOWN:
            assert it.error in {r["status"] for r in runs}
        its = [r["iterations"] for r in runs]
        slack = max(1, int(np.ceil(0.02 * max(its))))
        assert min(its) - slack <= it.iterations <= max(its) + slack, (it.iterations, its)
        tol = max(1e-6 * scale, 20 * spread)
        assert np.max(np.abs(xs - ref["x"])) <= tol, (np.max(np.abs(xs - ref["x"])), tol)


def test_idr_s_seed_counter(g):
    # random(P) advances mtl::seed<double>::counter once per column
    S, I, V = g.laplacian2d(10, 10)
    Ag = g.Matrix(S, I, V)
    n = len(S) - 1
    b = np.ones(n)
    g.random_seed_counter(5)
    x = g.Vector(n)
    x.fill(0.0)
    it, _ = g.make_iteration(300, 1e-8, 0.0, o.two_norm(b), history=False)
    assert g.idr_s(Ag, x, g.Vector(data=b), it, s=4) == 0
    assert g.random_seed_counter() == 9


def test_more_solvers_argument_errors(g):
    S, I, V = g.laplacian2d(6, 6)
    Ag = g.Matrix(S, I, V)
    n = len(S) - 1
    x, bv = g.Vector(n), g.Vector(data=np.ones(n))
    it, _ = g.make_iteration(10, 1e-8, 0.0, 1.0, history=False)
    with pytest.raises(g.M4GError) as e:
        g.gmres(Ag, x, bv, None, it, restart=0)
    assert e.value.status == g.ERR_INVALID
    with pytest.raises(g.M4GError) as e:
        g.bicgstab_ell(Ag, x, bv, None, it, l=0)
    assert e.value.status == g.ERR_INVALID


def test_interrupt_flag_stops_host_driven_solvers(g):
    # interruptible_iteration: a raised flag finishes the loop at the next
    # residual test; the C ABI reports M4G_ERR_INTERRUPTED
    import ctypes
    S, I, V = g.laplacian2d(40, 40)
    Ag = g.Matrix(S, I, V)
    n = len(S) - 1
    flag = ctypes.c_int(1)
    for fn in (lambda x, it: g.gmres(Ag, x, g.Vector(data=np.ones(n)), None, it, restart=4),
               lambda x, it: g.qmr(Ag, x, g.Vector(data=np.ones(n)), None, it),
               lambda x, it: g.idr_s(Ag, x, g.Vector(data=np.ones(n)), it, s=2)):
        x = g.Vector(n)
        x.fill(0.0)
        it, _ = g.make_iteration(500, 1e-30, 0.0, 1.0, history=False)
        it.interrupt = ctypes.pointer(flag)
        with pytest.raises(g.M4GError) as e:
            fn(x, it)
        assert e.value.status == g.ERR_INTERRUPTED
        assert it.iterations <= 2


@pytest.mark.parametrize("method,restart", [("gmres", 2), ("gmres_full", None), ("gmres", 1)])
def test_gmres_worked_swap_case(g, method, restart):
    # The case worked by hand in tests/test_oracle.py (A = [[0, 1], [1, 0]],
    # b = e1, x0 = 0): the reference's rotation sequence leaves H singular,
    # lu_solve divides by the unchecked last pivot and x becomes NaN
    # (restart 2, gmres_full); restart 1 fails with code 2, x untouched.  The
    # GPU path must give the oracle's outcome, not textbook GMRES's exact e2.
    S, I, V = np.array([0, 1, 2]), np.array([1, 0]), np.array([1.0, 1.0])
    A = o.CSR(S, I, V)
    b = np.array([1.0, 0.0])
    ref = o.solve(method, A, b, pc=0, max_iter=6, rtol=1e-10, history=False, param=restart)
    Ag = g.Matrix(S, I, V)
    P = g.PC(Ag, g.PC_IDENTITY)
    x = g.Vector(2)
    x.fill(0.0)
    it, _ = g.make_iteration(6, 1e-10, 0.0, 1.0, history=False)
    try:
        st = _run_gpu(g, method, restart, Ag, x, g.Vector(data=b), P, None, it)
    except g.M4GError as e:
        st = e.status
    xs = x.numpy()
    print(f"{method}({restart}): gpu status {st} it {it.iterations} x {xs}; oracle status {ref['status']} "
          f"it {ref['iterations']} x {ref['x']}", flush=True)
    assert it.iterations == ref["iterations"]
    if np.isnan(ref["x"]).any():
        assert np.isnan(xs).all() and st == ref["status"]
    else:
        np.testing.assert_array_equal(xs, ref["x"])
        assert (st == g.BREAKDOWN) if ref["status"] == 2 else (st == ref["status"])
