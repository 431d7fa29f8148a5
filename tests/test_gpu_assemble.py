"""Assembly on the device (m4g_matrix_assemble) against the oracle's inserter:
the same pattern and bit-identical values for every updater, with and without
a starting matrix, from host and from device arrays, up to the 128^3 7-point
Laplacian assembled from edge contributions."""
import numpy as np
import pytest

import oracle_lib as o
from test_assemble import run_steps, entry

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


def as_csr(M):
    S, I, V = M.download()
    return o.CSR(S, I, V)


def test_inserter_test_sequence(g):
    dev = run_steps(lambda r, c, v, u, A: g.assemble(8, 6, r, c, v, u, A))
    ref = run_steps(lambda r, c, v, u, A: o.assemble(8, r, c, v, u, A))
    for D, R in zip(dev, ref):
        C = as_csr(D)
        np.testing.assert_array_equal(C.starts, R.starts)
        np.testing.assert_array_equal(C.indices, R.indices)
        np.testing.assert_array_equal(C.data, R.data)
    C = as_csr(dev[2])
    assert entry(C, 4, 2) == 54 and entry(C, 7, 2) == 43 and entry(as_csr(dev[3]), 2, 2) == 28


@pytest.mark.parametrize("updater", [0, 1, 2, 3])
def test_random_triplets_match_oracle(g, updater):
    rng = np.random.default_rng(10 + updater)
    nr, nc, m = 500, 300, 20000          # about 13 values per position: long duplicate runs
    rows = rng.integers(0, nr, m)
    cols = rng.integers(0, nc, m)
    vals = rng.uniform(0.5, 1.5, m) * rng.choice([-1.0, 1.0], m)
    ref = o.assemble(nr, rows, cols, vals, updater)
    C = as_csr(g.assemble(nr, nc, rows, cols, vals, updater))
    np.testing.assert_array_equal(C.starts, ref.starts)
    np.testing.assert_array_equal(C.indices, ref.indices)
    np.testing.assert_array_equal(C.data, ref.data)
    # starting from a matrix: its entries are present before the first insertion
    base = o.assemble(nr, rows[:3000], cols[:3000], vals[:3000], 0)
    E = g.Matrix(base.starts, base.indices, base.data, ncols=nc)
    ref2 = o.assemble(nr, rows[3000:], cols[3000:], vals[3000:], updater, existing=base)
    C2 = as_csr(g.assemble(nr, nc, rows[3000:], cols[3000:], vals[3000:], updater, existing=E))
    np.testing.assert_array_equal(C2.starts, ref2.starts)
    np.testing.assert_array_equal(C2.indices, ref2.indices)
    np.testing.assert_array_equal(C2.data, ref2.data)


def edge_triplets(nx, ny, nz):
    """The 7-point Laplacian as a sum of edge contributions (+1 on both
    diagonals, -1 on the two couplings of every grid edge), plus 6 - degree on
    each diagonal for the Dirichlet boundary, in element order."""
    idx = np.arange(nx * ny * nz, dtype=np.int64).reshape(nz, ny, nx)
    rows, cols, vals = [], [], []
    for a, b in ((idx[:, :, :-1], idx[:, :, 1:]), (idx[:, :-1, :], idx[:, 1:, :]), (idx[:-1, :, :], idx[1:, :, :])):
        a, b = a.ravel(), b.ravel()
        e = np.stack([a, b, a, b], 1).ravel()
        f = np.stack([a, b, b, a], 1).ravel()
        rows.append(e)
        cols.append(f)
        vals.append(np.tile([1.0, 1.0, -1.0, -1.0], len(a)))
    n = nx * ny * nz
    all_r = np.concatenate(rows)
    all_c = np.concatenate(cols)
    all_v = np.concatenate(vals)
    diag = all_r == all_c
    deg = np.bincount(all_r[diag], minlength=n).astype(float)
    rows.append(np.arange(n, dtype=np.int64))
    cols.append(np.arange(n, dtype=np.int64))
    vals.append(6.0 - deg)
    return np.concatenate([all_r, rows[-1]]), np.concatenate([all_c, cols[-1]]), np.concatenate([all_v, vals[-1]])


def test_laplacian_from_edges_device_arrays(g):
    import torch
    nx = ny = nz = 128
    r, c, v = edge_triplets(nx, ny, nz)
    n = nx * ny * nz
    dev = [torch.from_numpy(a).to("cuda") for a in (r, c, v)]
    M = g.assemble(n, n, dev[0], dev[1], dev[2], g.UPDATE_PLUS)
    S, I, V = g.poisson3d(nx, ny, nz)
    C = as_csr(M)
    np.testing.assert_array_equal(C.starts, S)
    np.testing.assert_array_equal(C.indices, I)
    np.testing.assert_array_equal(C.data, V)


def test_assemble_errors(g):
    with pytest.raises(g.M4GError) as e:
        g.assemble(3, 3, [0, 3], [0, 0], [1.0, 2.0])
    assert e.value.status == g.ERR_INVALID
    Z = g.assemble(4, 4, [], [], [])
    assert Z.nnz == 0
    np.testing.assert_array_equal(as_csr(Z).starts, np.zeros(5))
