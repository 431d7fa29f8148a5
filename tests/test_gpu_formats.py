"""ell_matrix and sparse_banded on the GPU (formats.hip) against the oracle's
restatements: the downloaded layouts equal the inserters' layouts entry for
entry, and every product (=, +=, -=) is bit-identical to the oracle's loops,
from the reference's small test matrices up to the 128^3 7-point Laplacian."""
import numpy as np
import pytest

import oracle_lib as o
from test_formats import fill_matrix, rectangle, lower_only, upper_only, csr_from_entries

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


def random_csr(nr, nc, seed, empty_every=7, max_row=9):
    rng = np.random.default_rng(seed)
    e = {}
    for r in range(nr):
        if r % empty_every == 3:
            continue
        for c in rng.choice(nc, size=int(rng.integers(1, max_row + 1)), replace=False):
            e[(r, int(c))] = float(rng.uniform(-1, 1))
    return csr_from_entries(nr, e)


def cases_ell(g):
    S, I, V = g.poisson3d(40, 30, 20)
    yield "fill_matrix", fill_matrix(), 5
    yield "laplacian 9x7", o.laplacian_setup(9, 7), 63
    yield "random 300x170 ragged", random_csr(300, 170, 1), 170
    yield "rectangle 11x6", rectangle(11, 6), 6
    yield "poisson3d 40x30x20", o.CSR(S, I, V), 24000


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ell_matches_oracle(g, mode):
    rng = np.random.default_rng(2)
    for name, A, ncols in cases_ell(g):
        print(name, flush=True)
        E = g.Ell(A.starts, A.indices, A.data, ncols=ncols)
        slots, stride, I, D = o.ell_layout(A)
        assert (E.slots, E.stride, E.nnz) == (slots, stride, A.nnz)
        gi, gd = E.download()
        np.testing.assert_array_equal(gi, I)
        np.testing.assert_array_equal(gd, D)
        x = rng.uniform(-1, 1, ncols)
        w0 = rng.uniform(-1, 1, A.n)
        ref = o.ell_spmv(A.n, slots, stride, I, D, x, w0, mode)
        w = g.Vector(data=w0)
        E.spmv(g.Vector(data=x), w, mode)
        np.testing.assert_array_equal(w.numpy(), ref)


def cases_banded(g):
    yield "rectangle 6x11", rectangle(6, 11), 11
    yield "rectangle 11x6", rectangle(11, 6), 6
    yield "rectangle 6x5", rectangle(6, 5), 5
    yield "lower only", lower_only(), 5
    yield "upper only", upper_only(), 5
    yield "laplacian 16x9", o.laplacian_setup(16, 9), 144
    rng = np.random.default_rng(5)
    e = {}
    for r in range(50):
        for d in (-7, -3, 0, 2, 5):
            if 0 <= r + d < 40 and rng.uniform() < 0.8:
                e[(r, r + d)] = float(rng.uniform(-1, 1))
    yield "random bands 50x40", csr_from_entries(50, e), 40


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_banded_matches_oracle(g, mode):
    rng = np.random.default_rng(3)
    for name, A, ncols in cases_banded(g):
        print(name, flush=True)
        Bm = g.Banded(A.starts, A.indices, A.data, ncols=ncols)
        B, D = o.banded_layout(A)
        assert Bm.nbands == len(B)
        gb, gd = Bm.download()
        np.testing.assert_array_equal(gb, B)
        np.testing.assert_array_equal(gd, D)
        x = rng.uniform(-1, 1, ncols)
        w0 = rng.uniform(-1, 1, A.n)
        ref = o.banded_spmv(A.n, ncols, B, D, x, w0, mode)
        w = g.Vector(data=w0)
        Bm.spmv(g.Vector(data=x), w, mode)
        np.testing.assert_array_equal(w.numpy(), ref)
        # the reference layout constructor gives the same matrix
        B2 = g.Banded(nrows=A.n, ncols=ncols, bands=B, data=D)
        w2 = g.Vector(data=w0)
        B2.spmv(g.Vector(data=x), w2, mode)
        np.testing.assert_array_equal(w2.numpy(), ref)


def test_formats_at_benchmark_size(g):
    # 128^3 7-point Laplacian: ELL and banded products equal the oracle's and the CSR product
    S, I, V = g.poisson3d(128, 128, 128)
    A = o.CSR(S, I, V)
    n = A.n
    x = g.rng_uniform(11, 1, n)
    csr = np.zeros(n)
    o.lib.orc_spmv(n, A.starts, A.indices, np.ascontiguousarray(A.data, np.float64), x, csr, 0)
    E = g.Ell(S, I, V)
    w = g.Vector(n)
    E.spmv(g.Vector(data=x), w, g.ASSIGN)
    np.testing.assert_array_equal(w.numpy(), csr)
    E.close()
    Bm = g.Banded(S, I, V)
    assert Bm.nbands == 7
    B, D = o.banded_layout(A)
    ref = o.banded_spmv(n, n, B, D, x)
    w.fill(0.0)
    Bm.spmv(g.Vector(data=x), w, g.ASSIGN)
    got = w.numpy()
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, csr)


def test_format_errors(g):
    with pytest.raises(g.M4GError) as e:
        g.Ell([0, 2], [1, 0], [1.0, 2.0], ncols=2)          # unsorted row
    assert e.value.status == g.ERR_LOGIC
    with pytest.raises(g.M4GError) as e:
        g.Banded(nrows=3, ncols=3, bands=[1, 0], data=np.zeros(6))
    assert e.value.status == g.ERR_LOGIC
    E = g.Ell([0, 1, 2], [0, 1], [1.0, 2.0])
    with pytest.raises(g.M4GError) as e:
        E.spmv(g.Vector(data=np.ones(3)), g.Vector(2))
    assert e.value.status == g.ERR_SIZE
    # an empty banded matrix: = clears, += leaves w alone
    Z = g.Banded(nrows=3, ncols=3, bands=[], data=np.zeros(0))
    w = g.Vector(data=np.array([1.0, 2.0, 3.0]))
    Z.spmv(g.Vector(data=np.ones(3)), w, g.PLUS)
    np.testing.assert_array_equal(w.numpy(), [1, 2, 3])
    Z.spmv(g.Vector(data=np.ones(3)), w, g.ASSIGN)
    np.testing.assert_array_equal(w.numpy(), [0, 0, 0])


@pytest.mark.parametrize("m,n", [(2, 2), (2, 9), (7, 3), (33, 65), (512, 384)])
def test_poisson2d_dirichlet_operator(g, m, n):
    # the matrix-free operator (poisson2D_dirichlet::mult): every mode bit-identical to the oracle's loops
    rng = np.random.default_rng(m + n)
    x = rng.uniform(-1, 1, m * n)
    w0 = rng.uniform(-1, 1, m * n)
    for mode in (g.ASSIGN, g.PLUS, g.MINUS):
        ref = o.poisson2d_dirichlet_mult(m, n, x, w0, mode)
        w = g.Vector(data=w0)
        g.poisson2d_dirichlet_mult(m, n, g.Vector(data=x), w, mode)
        np.testing.assert_array_equal(w.numpy(), ref)
    with pytest.raises(g.M4GError):
        g.poisson2d_dirichlet_mult(1, n, g.Vector(n), g.Vector(n))
