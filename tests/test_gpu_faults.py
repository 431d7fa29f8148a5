"""Failure handling forced from the host.

A sweep's bounded waits give up after M4G_SPIN_LIMIT polls (read at every
launch).  With a limit of one poll, the tiles of a multi-tile sweep give up
waiting for values their upstream tiles have not exported yet, every wave
still runs to its end, and the call reports M4G_ERR_TIMEOUT.  The fault flag
is cleared when it is reported, so later calls with the normal limit give the
exact results again.  Also: the torch binding rejects tensors of a type the
kernels do not read."""
import os

import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


def test_sweep_timeout_is_reported_then_cleared(g, monkeypatch):
    S, I, V = g.poisson3d(64)
    A = o.CSR(S, I, V)
    st, U, L = o.ic0_factor(A)
    Ag = g.Matrix(S, I, V)
    P = g.PC(Ag, g.PC_IC0)
    tiles = P.stats()["tiles_lower"]
    print(f"64^3 IC(0): {tiles} tiles per sweep", flush=True)
    assert tiles >= 8
    v = g.rng_uniform(8, 1, A.n)
    ref = o.ic0_solve(L, U, v)
    xv, yv = g.Vector(data=v), g.Vector(A.n)
    P.solve(xv, yv)
    np.testing.assert_array_equal(yv.numpy(), ref)

    monkeypatch.setenv("M4G_SPIN_LIMIT", "1")
    with pytest.raises(g.M4GError) as e:
        P.solve(xv, yv)
    assert e.value.status == g.ERR_TIMEOUT
    b = g.Vector(data=o.spmv(A, v))
    x = g.Vector(A.n)
    x.fill(0.0)
    it, _ = g.make_iteration(300, 1e-8, 0.0, o.two_norm(o.spmv(A, v)))
    with pytest.raises(g.M4GError) as e:
        g.cg(Ag, x, b, P, it)
    assert e.value.status == g.ERR_TIMEOUT
    print("timeouts reported with a one-poll limit", flush=True)

    monkeypatch.delenv("M4G_SPIN_LIMIT")
    for _ in range(2):
        P.solve(xv, yv)
        np.testing.assert_array_equal(yv.numpy(), ref)
    bh = o.spmv(A, v)
    ref_cg = o.solve("cg", A, bh, pc=o.PC_IC0, max_iter=300, rtol=1e-8, history=False)
    x.fill(0.0)
    it, _ = g.make_iteration(300, 1e-8, 0.0, o.two_norm(bh))
    assert g.cg(Ag, x, b, P, it) == 0
    assert abs(it.iterations - ref_cg["iterations"]) <= max(1, int(np.ceil(0.02 * ref_cg["iterations"])))
    print(f"after the limit is restored: exact application, CG {it.iterations} it "
          f"(oracle {ref_cg['iterations']})", flush=True)


def test_vector_rejects_unsupported_tensor_types(g):
    import torch
    for dt in (torch.float16, torch.bfloat16, torch.int32, torch.int64):
        t = torch.zeros(16, dtype=dt, device="cuda")
        with pytest.raises(TypeError):
            g.Vector(tensor=t)
    for dt in (torch.float32, torch.float64):
        t = torch.arange(16, dtype=dt, device="cuda")
        v = g.Vector(tensor=t)
        np.testing.assert_array_equal(v.numpy(), np.arange(16))
