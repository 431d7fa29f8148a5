"""Differentiation of x = A(theta)^{-1} b through the C ABI (torch.autograd
test binding): the hand-written adjoint is checked against central finite
differences (step 1e-6) on 64 random theta entries and against a CPU
reference gradient (scipy direct solves + the closed form
dL/dtheta_e = -(lambda_i - lambda_j)(x_i - x_j), boundary edges -lambda_i x_i)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    import mtl4gpu as g
    import mtl4gpu_torch as gt
    return torch, g, gt


def cpu_reference_gradient(e, theta, b, x_target):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = len(e["S"]) - 1
    vals = np.zeros(len(e["I"]))
    rows = np.repeat(np.arange(len(vals)), np.diff(e["aptr"]))
    np.add.at(vals, rows, e["acoef"] * theta[e["aidx"]])
    A = sp.csr_matrix((vals, e["I"], e["S"]), shape=(n, n)).tocsc()
    lu = spla.splu(A)
    x = lu.solve(b)
    lam = lu.solve(x - x_target, trans="T")
    ei, ej = e["ei"], e["ej"]
    interior = ej >= 0
    grad = np.empty(len(theta))
    grad[interior] = -(lam[ei[interior]] - lam[ej[interior]]) * (x[ei[interior]] - x[ej[interior]])
    grad[~interior] = -lam[ei[~interior]] * x[ei[~interior]]
    return x, grad


def test_gradcheck_small(env):
    torch, g, gt = env
    sysm = gt.EdgeSystem(6, 5, rtol=1e-14)
    rng = np.random.default_rng(0)
    theta = torch.tensor(np.exp(rng.standard_normal(sysm.n_edges)), device="cuda", requires_grad=True)
    b = torch.tensor(rng.standard_normal(sysm.n), device="cuda", requires_grad=True)
    assert torch.autograd.gradcheck(lambda th, bb: gt.solve(th, bb, sysm), (theta, b), eps=1e-6, atol=1e-7,
                                    rtol=1e-5)


@pytest.mark.parametrize("side", [64, 256])
def test_edge_diffusion_gradient(env, side):
    torch, g, gt = env
    print(f"edge diffusion side={side}", flush=True)
    sysm = gt.EdgeSystem(side, side, rtol=1e-10)
    ne, n = sysm.n_edges, sysm.n
    theta_h = g.rng_lognormal(5, 1, ne)
    b_h = g.rng_normal(5, 2, n)
    xt_h = g.rng_normal(5, 3, n)
    theta = torch.tensor(theta_h, device="cuda", requires_grad=True)
    b = torch.tensor(b_h, device="cuda")
    x_target = torch.tensor(xt_h, device="cuda")
    x = gt.solve(theta, b, sysm)
    loss = 0.5 * torch.sum((x - x_target) ** 2)
    loss.backward()
    grad = theta.grad.cpu().numpy()
    x_ref, grad_ref = cpu_reference_gradient(sysm.e, theta_h, b_h, xt_h)
    err_x = np.max(np.abs(x.detach().cpu().numpy() - x_ref)) / np.max(np.abs(x_ref))
    err_g = np.max(np.abs(grad - grad_ref)) / np.max(np.abs(grad_ref))
    print(f"  iterations {sysm.last_iterations}, |x - x_cpu| rel {err_x:.2e}, |grad - grad_cpu| rel {err_g:.2e}",
          flush=True)
    assert err_x < 1e-7 and err_g < 1e-6

    # central finite differences on 64 random entries, forward solves tightened to 1e-13
    fd_sys = gt.EdgeSystem(side, side, rtol=1e-13)
    rng = np.random.default_rng(11)
    idx = rng.choice(ne, 64, replace=False)
    h = 1e-6

    def L(th):
        with torch.no_grad():
            xx = gt.solve(torch.tensor(th, device="cuda"), b, fd_sys)
            return 0.5 * float(torch.sum((xx - x_target) ** 2))

    fd = np.empty(64)
    for k, e in enumerate(idx):
        tp, tm = theta_h.copy(), theta_h.copy()
        tp[e] += h
        tm[e] -= h
        fd[k] = (L(tp) - L(tm)) / (2 * h)
    scale = np.max(np.abs(grad[idx]))
    err_fd = np.max(np.abs(fd - grad[idx])) / scale
    print(f"  finite differences on 64 entries: max rel err {err_fd:.2e} (scale {scale:.3e})", flush=True)
    assert err_fd < 1e-4


def test_backward_after_another_forward(env):
    # the backward of the first solve must use the first theta's factors even
    # though the same system was refactorised for a second theta in between
    torch, g, gt = env
    sysm = gt.EdgeSystem(24, 20, rtol=1e-12)
    rng = np.random.default_rng(7)
    th1 = np.exp(rng.standard_normal(sysm.n_edges))
    th2 = np.exp(rng.standard_normal(sysm.n_edges))
    b_h, xt_h = rng.standard_normal(sysm.n), rng.standard_normal(sysm.n)
    b, xt = torch.tensor(b_h, device="cuda"), torch.tensor(xt_h, device="cuda")
    theta1 = torch.tensor(th1, device="cuda", requires_grad=True)
    theta2 = torch.tensor(th2, device="cuda", requires_grad=True)
    loss1 = 0.5 * torch.sum((gt.solve(theta1, b, sysm) - xt) ** 2)
    loss2 = 0.5 * torch.sum((gt.solve(theta2, b, sysm) - xt) ** 2)
    loss1.backward()
    loss2.backward()
    for th_h, th in ((th1, theta1), (th2, theta2)):
        _, ref = cpu_reference_gradient(sysm.e, th_h, b_h, xt_h)
        got = th.grad.cpu().numpy()
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-8
