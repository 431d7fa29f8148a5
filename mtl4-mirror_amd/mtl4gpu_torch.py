"""torch.autograd binding of x = A(theta)^{-1} b (test binding of the C ABI).

The operator is A(theta) = sum_e theta_e (e_i - e_j)(e_i - e_j)^T plus
boundary terms theta_b e_i e_i^T on a 2D grid (m4g_gen_edge_laplacian), i.e.
its stored values are M theta for a fixed sparse assembly map M.  The forward
pass assembles the values on the GPU (SpMV with M), refactorises IC(0) and
runs IC(0)-PCG; the backward pass solves the adjoint system A^T lambda = dL/dx
with the same kernels (A is symmetric), forms dL/dA_ij = -lambda_i x_j on the
pattern (m4g_pattern_gradient) and maps it to theta with M^T (SpMV).  The
backward reuses the forward's values and IC(0) factors unless the same
EdgeSystem was given another theta in between.
All device work goes through libmtl4gpu; torch only holds the memory.
"""
import numpy as np
import torch

import mtl4gpu as g


class EdgeSystem:
    """Device state of one A(theta) pattern: matrix, assembly maps, IC(0)."""

    def __init__(self, nx, ny, rtol=1e-10, max_iter=20000):
        e = g.edge_laplacian(nx, ny)
        self.nx, self.ny, self.n, self.n_edges = nx, ny, nx * ny, e["n_edges"]
        self.rtol, self.max_iter = rtol, max_iter
        self.S, self.I = e["S"], e["I"]
        nnz = len(e["I"])
        self.nnz = nnz
        # M (nnz x n_edges): values = M theta; Mt (n_edges x nnz): dtheta = M^T G
        rows = np.repeat(np.arange(nnz), np.diff(e["aptr"]))
        self.M = self._csr(rows, e["aidx"], e["acoef"], nnz, self.n_edges)
        self.Mt = self._csr(e["aidx"], rows, e["acoef"], self.n_edges, nnz)
        g.set_stream(torch.cuda.current_stream().cuda_stream)
        theta0 = np.ones(self.n_edges)
        self.A = g.Matrix(self.S, self.I, self._assemble_host(e, theta0))
        self.P = g.PC(self.A, g.PC_IC0)
        self.e = e
        self.epoch = 0              # counts set_theta calls: which theta the factors belong to

    @staticmethod
    def _csr(r, c, v, m, n):
        order = np.lexsort((c, r))
        r, c, v = np.asarray(r)[order], np.asarray(c)[order], np.asarray(v, float)[order]
        S = np.zeros(m + 1, np.int64)
        np.add.at(S, r + 1, 1)
        S = np.cumsum(S)
        return g.Matrix(S, c, v, ncols=n)

    def _assemble_host(self, e, theta):
        vals = np.zeros(len(e["I"]))
        rows = np.repeat(np.arange(len(vals)), np.diff(e["aptr"]))
        np.add.at(vals, rows, e["acoef"] * theta[e["aidx"]])
        return vals

    def set_theta(self, theta):
        """Assembles A's values = M theta on the GPU and refactorises IC(0)."""
        vals = torch.empty(self.nnz, dtype=torch.float64, device=theta.device)
        g.spmv(self.M, g.Vector(tensor=theta.detach().contiguous()), g.Vector(tensor=vals))
        g.synchronize()
        self.A.set_values(vals)
        self.P.refactor(self.A)
        self.epoch += 1

    def solve(self, rhs):
        """x = A^{-1} rhs with IC(0)-PCG to rtol (x0 = 0)."""
        x = torch.zeros_like(rhs)
        xv, bv = g.Vector(tensor=x), g.Vector(tensor=rhs.contiguous())
        nb = g.two_norm(bv)
        it, _ = g.make_iteration(self.max_iter, self.rtol, 0.0, nb, history=False)
        st = g.cg(self.A, xv, bv, self.P, it)
        if st != 0:
            raise RuntimeError(f"IC(0)-PCG did not converge (status {st}, {it.iterations} iterations)")
        self.last_iterations = it.iterations
        return x


class SparseSolve(torch.autograd.Function):
    """x = A(theta)^{-1} b with a hand-written adjoint."""

    @staticmethod
    def forward(ctx, theta, b, system):
        g.set_stream(torch.cuda.current_stream().cuda_stream)   # order against torch's work
        system.set_theta(theta)
        x = system.solve(b)
        ctx.system = system
        ctx.epoch, ctx.theta_version = system.epoch, theta._version
        ctx.save_for_backward(theta, x)
        return x

    @staticmethod
    def backward(ctx, gx):
        system = ctx.system
        theta, x = ctx.saved_tensors
        g.set_stream(torch.cuda.current_stream().cuda_stream)
        if system.epoch != ctx.epoch or theta._version != ctx.theta_version:
            system.set_theta(theta)                  # another theta was factorised since the forward
        lam = system.solve(gx.contiguous())          # A^T lambda = dL/dx (A symmetric)
        grad_vals = torch.empty(system.nnz, dtype=torch.float64, device=x.device)
        g.pattern_gradient(system.A, g.Vector(tensor=lam), g.Vector(tensor=x), g.Vector(tensor=grad_vals))
        dtheta = torch.empty_like(theta)
        g.spmv(system.Mt, g.Vector(tensor=grad_vals), g.Vector(tensor=dtheta))
        g.synchronize()
        return dtheta, lam, None


def solve(theta, b, system):
    return SparseSolve.apply(theta, b, system)
