"""The batched workload at its full size: 512 SPD 2D 5-point
variable-coefficient diffusion systems, side ~ U{32..256} (seed 7), per-cell
kappa ~ lognormal(0,1) (seed 1000 + k), 8 right-hand sides ~ N(0,1) per
system (seed 2000 + k), IC(0)-PCG to 1e-8 in one persistent launch.

Every (system, rhs) pair runs itl::cg + pc::ic_0 with basic_iteration's test
(cg.hpp:75-112, ic_0.hpp:135-186, basic_iteration.hpp:83-99).  Checked:
  * every pair converged, its recurrence residual at 1e-8 and its true
    residual ||b - A x|| (the oracle's product) within 10x of that;
  * the smallest, median and largest system of each third of the batch by
    size (the autotuner's size classes), first and last right-hand side:
    iteration counts within 2 % of the oracle's single solve and x close to
    the oracle's;
  * both with the autotuned grouping / block size and with 8 right-hand
    sides per work item in 1024-thread workgroups."""
import numpy as np
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu

NSYS, NRHS, RTOL = 512, 8, 1e-8


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def g():
    import mtl4gpu
    return mtl4gpu


@pytest.fixture(scope="module")
def batch(g):
    say(f"batch {NSYS}x{NRHS}: generating")
    sides = g.rng_int(7, 1, NSYS, 32, 256)
    systems = [g.varcoef5(int(sd), int(sd), seed=1000 + k) for k, sd in enumerate(sides)]
    sizes = np.array([len(S) - 1 for S, _, _ in systems])
    bs = [[g.rng_normal(2000 + k, j, int(sizes[k])) for j in range(NRHS)] for k in range(NSYS)]
    bh = np.concatenate([np.stack(v, axis=1).reshape(-1) for v in bs])
    B = g.Batch(systems)
    B.factor_ic0()
    b = g.Vector(data=bh)
    order = np.argsort(sizes, kind="stable")
    sample = []
    for c in range(3):
        cls = order[c * NSYS // 3:(c + 1) * NSYS // 3]
        sample += [int(cls[0]), int(cls[len(cls) // 2]), int(cls[-1])]
    return dict(systems=systems, sizes=sizes, bs=bs, B=B, b=b, sample=sample)


@pytest.fixture(scope="module")
def oracle_runs(batch):
    runs = {}
    for k in batch["sample"]:
        A = o.CSR(*batch["systems"][k])
        for j in (0, NRHS - 1):
            runs[k, j] = o.solve("cg", A, batch["bs"][k][j], pc=o.PC_IC0, max_iter=10000, rtol=RTOL, history=False)
            assert runs[k, j]["status"] == o.OK
    return runs


def check_batch(g, batch, oracle_runs, rpi, block):
    B, sizes, bs = batch["B"], batch["sizes"], batch["bs"]
    x = g.Vector(int(B.n_total) * NRHS)
    x.fill(0.0)
    its, res, err = B.cg_ic0(NRHS, x, batch["b"], 10000, RTOL, rhs_per_item=rpi, block=block)
    xs = x.numpy()
    say(f"  block {block}: iterations min/median/max {its.min()}/{int(np.median(its))}/{its.max()}, "
        f"errors {int((err != 0).sum())}")
    assert (err == 0).all()
    worst = 0.0
    for k in range(NSYS):
        A = o.CSR(*batch["systems"][k])
        r0 = int(B.row0[k])
        for j in range(NRHS):
            nb = o.two_norm(bs[k][j])
            assert res[k, j] <= RTOL * nb, (k, j)
            xk = xs[r0 * NRHS + j:(r0 + A.n) * NRHS:NRHS]
            tr = o.two_norm(o.residual(A, xk, bs[k][j])) / nb
            worst = max(worst, tr)
            assert tr <= 10 * RTOL, (k, j, tr)
            if (k, j) in oracle_runs:
                ref = oracle_runs[k, j]
                slack = max(1, int(np.ceil(0.02 * ref["iterations"])))
                assert abs(int(its[k, j]) - ref["iterations"]) <= slack, (k, j, its[k, j], ref["iterations"])
                assert np.max(np.abs(xk - ref["x"])) <= 1e-6 * max(1.0, np.max(np.abs(ref["x"])))
    say(f"  worst true relres {worst:.2e}; sampled sides "
        f"{[int(np.sqrt(sizes[k])) for k in batch['sample']]} match the oracle")


def test_batch_autotuned_512x8(g, batch, oracle_runs):
    rpi, block, _ = batch["B"].autotune(NRHS, batch["b"])
    say(f"batch autotuned: block {block}, rhs per item {sorted(set(int(r) for r in rpi))}")
    check_batch(g, batch, oracle_runs, rpi, block)


def test_batch_eight_rhs_per_item_512x8(g, batch, oracle_runs):
    say("batch 8 rhs per item, block 1024")
    check_batch(g, batch, oracle_runs, [8] * NSYS, 1024)
