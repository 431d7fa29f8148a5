// itl::cgs, itl::bicg and itl::tfqmr on the path's kernels (FP64 vectors).
//
// Unlike cg / bicgstab (krylov.hip), whose scalars stay on the device, these
// solvers run host-driven: each dot or norm is one deterministic two-stage
// reduction read back to the host, where the scalar recurrences and
// basic_iteration's test (basic_iteration.hpp:52-93) run.  Every vector
// update is one kernel that evaluates the reference's expression element by
// element in its operation order (e.g. p = u + beta*(q + beta*p)), so the
// iterates differ from the CPU only through the reduction order of dots.
#include "m4g_internal.hpp"
#include "m4g_device.hpp"
#include "pc.hpp"
#include "krylov.hpp"

#include <cmath>
#include <memory>

namespace m4g {

// y = a + s * b  (b may alias y)
__global__ __launch_bounds__(256) void k_lin1(int64_t n, const double* __restrict__ a, double s, const double* b,
                                              double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = a[i] + s * b[i];
}

// y = a + s * (b + s * c)  (c may alias y)
__global__ __launch_bounds__(256) void k_lin2(int64_t n, const double* __restrict__ a, double s,
                                              const double* __restrict__ b, const double* c, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = a[i] + s * (b[i] + s * c[i]);
}

// y = a - s * b
__global__ __launch_bounds__(256) void k_lin_minus(int64_t n, const double* __restrict__ a, double s,
                                                   const double* __restrict__ b, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = a[i] - s * b[i];
}

namespace {

int lin_grid(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
}

// The vector algebra of the host-driven solvers on one stream.
struct Ops {
    int64_t n;
    hipStream_t s;
    double* red;   // one double of scratch for reductions

    void lin1(const double* a, double sc, const double* b, double* y) const   // y = a + sc*b
    {
        if (n) hipLaunchKernelGGL(k_lin1, dim3(lin_grid(n)), dim3(256), 0, s, n, a, sc, b, y);
    }
    void lin2(const double* a, double sc, const double* b, const double* c, double* y) const
    {
        if (n) hipLaunchKernelGGL(k_lin2, dim3(lin_grid(n)), dim3(256), 0, s, n, a, sc, b, c, y);
    }
    void minus(const double* a, double sc, const double* b, double* y) const   // y = a - sc*b
    {
        if (n) hipLaunchKernelGGL(k_lin_minus, dim3(lin_grid(n)), dim3(256), 0, s, n, a, sc, b, y);
    }
    void copy(const double* a, double* y) const
    {
        if (n && a != y) M4G_HIP(hipMemcpyAsync(y, a, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    double dot(const double* a, const double* b) const
    {
        launch_dot(n, a, b, red, s);
        return read_scalar(red);
    }
    double norm(const double* a) const
    {
        launch_sumsq(n, a, red, s);
        return std::sqrt(read_scalar(red));
    }
};

// basic_iteration::finished(r) (basic_iteration.hpp:52-65, 85-93)
bool finished(m4g_iteration* it, double r)
{
    it->resid = r;
    if (it->history && it->iterations < it->history_capacity) it->history[it->iterations] = r;
    const bool conv = it->norm_r0 == 0 ? r <= it->atol : (r <= it->rtol * it->norm_r0 || r <= it->atol);
    if (conv) return true;
    if (it->iterations >= it->max_iter) {
        it->error = M4G_MAX_ITER;
        return true;
    }
    return false;
}

void apply_pc(m4g_pc_s* P, const double* x, double* y, bool adjoint, const Ops& o)
{
    if (!P || P->kind == M4G_PC_IDENTITY) o.copy(x, y);
    else pc_apply(*P, x, y, M4G_F64, adjoint, o.s);
}

int sweep_fault(m4g_pc_s* P)
{
    return P && P->kind >= M4G_PC_IC0 ? check_sweep_fault(P->lower) | check_sweep_fault(P->upper) : 0;
}

// status of a finished solve: the iteration's error, breakdowns (error 2) as
// M4G_BREAKDOWN, a failed sweep wait as M4G_ERR_TIMEOUT
int status_of(m4g_iteration* it, int fault)
{
    if (fault) return M4G_ERR_TIMEOUT;
    if (it->error == 2) return M4G_BREAKDOWN;
    return it->error;
}

struct Vecs {
    std::vector<DevBuf<double>> v;
    Vecs(int count, int64_t n)
    {
        for (int k = 0; k < count; k++) v.emplace_back(size_t(std::max<int64_t>(n, 1)));
    }
    double* operator[](int k) { return v[size_t(k)].ptr; }
};

}  // namespace

// itl::cgs (cgs.hpp:24-62)
static int cgs_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* P, m4g_iteration* it)
{
    const int64_t n = A.nrows;
    DevBuf<double> red(1);
    Ops o{n, current_stream(), red.ptr};
    Vecs V(9, n);
    double *p = V[0], *phat = V[1], *q = V[2], *qhat = V[3], *vhat = V[4], *u = V[5], *uhat = V[6], *r = V[7],
           *rtilde = V[8];
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
    launch_spmv(A, x, r, M4G_ASSIGN, b, o.s);   // r = b - A*x
    o.copy(r, rtilde);
    while (!finished(it, o.norm(r))) {
        ++it->iterations;
        rho_1 = o.dot(rtilde, r);
        if (rho_1 == 0.0) it->error = 2;   // iter.fail(2, "cgs breakdown"); the loop goes on
        if (it->iterations <= 1) {
            o.copy(r, u);
            o.copy(r, p);
        } else {
            beta = rho_1 / rho_2;
            o.lin1(r, beta, q, u);        // u = r + beta*q
            o.lin2(u, beta, q, p, p);     // p = u + beta*(q + beta*p)
        }
        apply_pc(P, p, phat, false, o);
        launch_spmv(A, phat, vhat, M4G_ASSIGN, nullptr, o.s);   // vhat = A * solve(M, p)
        alpha = rho_1 / o.dot(rtilde, vhat);
        o.minus(u, alpha, vhat, q);       // q = u - alpha*vhat
        o.lin1(u, 1.0, q, u);             // u += q
        apply_pc(P, u, uhat, false, o);
        o.lin1(x, alpha, uhat, x);        // x += alpha*uhat
        launch_spmv(A, uhat, qhat, M4G_ASSIGN, nullptr, o.s);
        o.minus(r, alpha, qhat, r);       // r -= alpha*qhat
        rho_2 = rho_1;
    }
    M4G_HIP(hipStreamSynchronize(o.s));
    return status_of(it, sweep_fault(P));
}

// itl::bicg (bicg.hpp:27-68); adjoint(A) * v runs as A^T (built once, rows in
// A's row order, so each entry sums in the order of the transposed traversal)
static int bicg_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* P, m4g_iteration* it)
{
    const int64_t n = A.nrows;
    DevBuf<double> red(1);
    Ops o{n, current_stream(), red.ptr};
    m4g_matrix At = nullptr;
    M4G_REQUIRE(m4g_matrix_transpose(&A, &At) == M4G_OK, M4G_ERR_HIP, m4g_last_error());
    std::unique_ptr<m4g_matrix_s, int (*)(m4g_matrix)> hold(At, m4g_matrix_destroy);
    Vecs V(8, n);
    double *r = V[0], *z = V[1], *p = V[2], *q = V[3], *r_t = V[4], *z_t = V[5], *p_t = V[6], *q_t = V[7];
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
    launch_spmv(A, x, r, M4G_ASSIGN, b, o.s);
    o.copy(r, r_t);
    while (!finished(it, o.norm(r))) {
        ++it->iterations;
        apply_pc(P, r, z, false, o);
        apply_pc(P, r_t, z_t, true, o);
        rho_1 = o.dot(z_t, z);
        if (rho_1 == 0.0) {   // return iter.fail(2, "bicg breakdown")
            it->error = 2;
            break;
        }
        if (it->iterations <= 1) {
            o.copy(z, p);
            o.copy(z_t, p_t);
        } else {
            beta = rho_1 / rho_2;
            o.lin1(z, beta, p, p);        // p = z + beta*p
            o.lin1(z_t, beta, p_t, p_t);  // p_tilde = z_tilde + conj(beta)*p_tilde
        }
        launch_spmv(A, p, q, M4G_ASSIGN, nullptr, o.s);
        launch_spmv(*At, p_t, q_t, M4G_ASSIGN, nullptr, o.s);
        alpha = rho_1 / o.dot(p_t, q);
        o.lin1(x, alpha, p, x);
        o.minus(r, alpha, q, r);
        o.minus(r_t, alpha, q_t, r_t);
        rho_2 = rho_1;
    }
    M4G_HIP(hipStreamSynchronize(o.s));
    return status_of(it, sweep_fault(P));
}

// itl::tfqmr (tfqmr.hpp:32-99) with left preconditioner L and right R (null: identity)
static int tfqmr_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* L, m4g_pc_s* R, m4g_iteration* it)
{
    const int64_t n = A.nrows;
    M4G_REQUIRE(n > 0, M4G_ERR_LOGIC, "empty rhs vector");
    DevBuf<double> red(1);
    Ops o{n, current_stream(), red.ptr};
    Vecs V(11, n);
    double *rt = V[0], *r = V[1], *u1 = V[2], *u2 = V[3], *y1 = V[4], *y2 = V[5], *w = V[6], *d = V[7], *v = V[8],
           *tmp = V[9];
    launch_fill(n, 0.0, d, M4G_F64, o.s);
    double theta = 0, eta = 0, tau, rho, rhon, sigma, alpha, beta, c;
    apply_pc(R, x, tmp, false, o);
    launch_spmv(A, tmp, rt, M4G_ASSIGN, b, o.s);   // rt = b - A * solve(R, x)
    apply_pc(L, rt, r, false, o);
    auto done = [&](int st) {
        M4G_HIP(hipStreamSynchronize(o.s));
        return st;
    };
    if (finished(it, o.norm(rt))) return done(status_of(it, sweep_fault(L) | sweep_fault(R)));
    o.copy(r, y1);
    o.copy(r, w);
    apply_pc(R, y1, tmp, false, o);
    launch_spmv(A, tmp, rt, M4G_ASSIGN, nullptr, o.s);
    apply_pc(L, rt, u1, false, o);
    o.copy(u1, v);
    tau = o.norm(r);
    rho = tau * tau;
    while (!finished(it, tau)) {
        ++it->iterations;
        sigma = o.dot(r, v);
        if (sigma == 0.0) {   // return iter.fail(1, "tfgmr breakdown, sigma=0 #1")
            it->error = 1;
            return done(status_of(it, sweep_fault(L) | sweep_fault(R)));
        }
        alpha = rho / sigma;
        for (int j = 1; j < 3; j++) {
            if (j == 1) {
                o.minus(w, alpha, u1, w);                          // w -= alpha*u1
                o.lin1(y1, theta * theta * eta / alpha, d, d);     // d = y1 + (theta^2 eta / alpha) d
            } else {
                o.minus(y1, alpha, v, y2);                         // y2 = y1 - alpha*v
                apply_pc(R, y2, tmp, false, o);
                launch_spmv(A, tmp, rt, M4G_ASSIGN, nullptr, o.s);
                apply_pc(L, rt, u2, false, o);
                o.minus(w, alpha, u2, w);
                o.lin1(y2, theta * theta * eta / alpha, d, d);
            }
            theta = o.norm(w) / tau;
            c = 1.0 / std::sqrt(1.0 + theta * theta);
            tau *= theta * c;
            eta = c * c * alpha;
            o.lin1(x, eta, d, x);                                  // x += eta*d
        }
        if (rho == 0.0) {     // return iter.fail(1, "tfgmr breakdown, rho=0 #2")
            it->error = 1;
            return done(status_of(it, sweep_fault(L) | sweep_fault(R)));
        }
        rhon = o.dot(r, w);
        beta = rhon / rho;
        rho = rhon;
        o.lin1(w, beta, y2, y1);                                   // y1 = w + beta*y2
        apply_pc(R, y1, tmp, false, o);
        launch_spmv(A, tmp, rt, M4G_ASSIGN, nullptr, o.s);
        apply_pc(L, rt, u1, false, o);
        o.lin2(u1, beta, u2, v, v);                                // v = u1 + beta*(u2 + beta*v)
        // rt = A * x - b (tfqmr.hpp:94) is not used afterwards
    }
    apply_pc(R, x, tmp, false, o);                                 // shift back: x = solve(R, x)
    o.copy(tmp, x);
    return done(status_of(it, sweep_fault(L) | sweep_fault(R)));
}

}  // namespace m4g

using namespace m4g;

namespace {
int check_args(const char* who, m4g_matrix A, m4g_vector x, m4g_vector b, m4g_iteration* it)
{
    M4G_REQUIRE(A && x && b && it, M4G_ERR_INVALID, std::string(who) + ": null argument");
    M4G_REQUIRE(A->nrows == A->ncols, M4G_ERR_NOT_SQUARE, "Matrix is not square");
    M4G_REQUIRE(x->type == M4G_F64 && b->type == M4G_F64, M4G_ERR_INVALID, std::string(who) + ": double vectors");
    M4G_REQUIRE(x->n == A->ncols && b->n == A->nrows, M4G_ERR_SIZE, std::string(who) + ": incompatible size");
    return M4G_OK;
}
}  // namespace

extern "C" {

int m4g_cgs(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_cgs", A, x, b, it);
        return cgs_t(*A, (double*)x->ptr, (const double*)b->ptr, P, it);
    });
}

int m4g_bicg(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_bicg", A, x, b, it);
        return bicg_t(*A, (double*)x->ptr, (const double*)b->ptr, P, it);
    });
}

int m4g_tfqmr(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_tfqmr", A, x, b, it);
        return tfqmr_t(*A, (double*)x->ptr, (const double*)b->ptr, L, R, it);
    });
}

}  // extern "C"
