// itl::cgs, bicg, tfqmr, qmr, gmres, bicgstab_2, bicgstab_ell and idr_s on
// the path's kernels (FP64 vectors).
//
// Unlike cg / bicgstab (krylov.hip), whose scalars stay on the device, these
// solvers run host-driven: each dot or norm is one deterministic two-stage
// reduction whose last kernel posts the value into pinned host memory, where
// the host spins for it and runs the scalar recurrences and basic_iteration's
// test (basic_iteration.hpp:52-93).  Every vector
// update is one kernel that evaluates the reference's expression element by
// element in its operation order (e.g. p = u + beta*(q + beta*p)), so the
// iterates differ from the CPU only through the reduction order of dots.
#include "m4g_internal.hpp"
#include "m4g_device.hpp"
#include "pc.hpp"
#include "krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <memory>
#include <vector>

namespace m4g {

// The update kernels below evaluate one expression per element; any operand
// may be the output (each element is read before it is written).

// y = a + s * b
__global__ __launch_bounds__(256) void k_lin1(int64_t n, const double* a, double s, const double* b, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = a[i] + s * b[i];
}

// y = a + s * (b + s * c)
__global__ __launch_bounds__(256) void k_lin2(int64_t n, const double* a, double s, const double* b, const double* c,
                                              double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = a[i] + s * (b[i] + s * c[i]);
}

// y = a - s * b
__global__ __launch_bounds__(256) void k_lin_minus(int64_t n, const double* a, double s, const double* b, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = a[i] - s * b[i];
}

// y = s * a
__global__ __launch_bounds__(256) void k_scale(int64_t n, double s, const double* a, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = s * a[i];
}

// y = a / s
__global__ __launch_bounds__(256) void k_div(int64_t n, const double* a, double s, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = a[i] / s;
}

// y = s * a + t * b
__global__ __launch_bounds__(256) void k_lin_sum(int64_t n, double s, const double* a, double t, const double* b,
                                                 double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = s * a[i] + t * b[i];
}

// y = s * a - b
__global__ __launch_bounds__(256) void k_scale_minus(int64_t n, double s, const double* a, const double* b,
                                                     double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        y[i] = s * a[i] - b[i];
}

// Linear combination of up to kComboMax vectors, summed in list order:
// y = (((init ? y : 0) + v_0 c_0) + v_1 c_1) + ...  (multi_vector * vector,
// mat_vec_mult.hpp:359-366).  y must not be one of the v_j.
constexpr int kComboMax = 16;
struct ComboArgs {
    const double* v[kComboMax];
    double c[kComboMax];
    int count;
    int init;
};
__global__ __launch_bounds__(256) void k_combo(int64_t n, ComboArgs a, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
        double acc = a.init ? y[i] : 0.0;
        for (int j = 0; j < a.count; j++) acc += a.v[j][i] * a.c[j];
        y[i] = acc;
    }
}

namespace {

int lin_grid(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
}

// The vector algebra of the host-driven solvers on one stream.  Dots and
// norms come back through a mailbox: the reduction's last kernel posts the
// value into pinned host memory and the host spins on it.
struct Ops {
    int64_t n;
    hipStream_t s;
    double* red;   // one double of scratch for reductions
    std::unique_ptr<Mailbox> mb;

    Ops(int64_t n_, hipStream_t s_, double* red_) : n(n_), s(s_), red(red_), mb(std::make_unique<Mailbox>(s_)) {}

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
        const unsigned q = mb->arm();
        launch_dot_post(n, a, b, red, mb->box(), q, s);
        return mb->wait(q)[0];
    }
    double norm(const double* a) const
    {
        const unsigned q = mb->arm();
        launch_sumsq_post(n, a, red, mb->box(), q, s);
        return std::sqrt(mb->wait(q)[0]);
    }
    void scale(double sc, const double* a, double* y) const   // y = sc*a
    {
        if (n) hipLaunchKernelGGL(k_scale, dim3(lin_grid(n)), dim3(256), 0, s, n, sc, a, y);
    }
    void div(const double* a, double sc, double* y) const     // y = a / sc
    {
        if (n) hipLaunchKernelGGL(k_div, dim3(lin_grid(n)), dim3(256), 0, s, n, a, sc, y);
    }
    void lin_sum(double sa, const double* a, double sb, const double* b, double* y) const   // y = sa*a + sb*b
    {
        if (n) hipLaunchKernelGGL(k_lin_sum, dim3(lin_grid(n)), dim3(256), 0, s, n, sa, a, sb, b, y);
    }
    void scale_minus(double sa, const double* a, const double* b, double* y) const   // y = sa*a - b
    {
        if (n) hipLaunchKernelGGL(k_scale_minus, dim3(lin_grid(n)), dim3(256), 0, s, n, sa, a, b, y);
    }
    // y = sum_j v[j] * c[j] in order j = 0, 1, ... (y not among v)
    void combo(const std::vector<const double*>& v, const std::vector<double>& c, double* y) const
    {
        if (!n) return;
        if (v.empty()) {
            launch_fill(n, 0.0, y, M4G_F64, s);
            return;
        }
        for (size_t j0 = 0; j0 < v.size(); j0 += kComboMax) {
            ComboArgs a{};
            a.count = int(std::min<size_t>(kComboMax, v.size() - j0));
            a.init = j0 > 0;
            for (int j = 0; j < a.count; j++) {
                a.v[j] = v[j0 + j];
                a.c[j] = c[j0 + j];
            }
            hipLaunchKernelGGL(k_combo, dim3(lin_grid(n)), dim3(256), 0, s, n, a, y);
        }
    }
    // w[i] = dot(v[i], a) for every i, with one read-back (trans(P) * a)
    std::vector<double> dots(const std::vector<const double*>& v, const double* a, double* scratch) const
    {
        std::vector<double> w(v.size());
        if (v.empty()) return w;
        for (size_t i = 0; i < v.size(); i++) launch_dot(n, v[i], a, scratch + i, s);
        for (size_t i0 = 0; i0 < v.size(); i0 += kPostMax) {
            const int count = int(std::min<size_t>(kPostMax, v.size() - i0));
            const unsigned q = mb->arm();
            launch_post(scratch + i0, count, mb->box(), q, s);
            const double* got = mb->wait(q);
            std::copy(got, got + count, w.begin() + std::ptrdiff_t(i0));
        }
        return w;
    }
};

// basic_iteration::finished(r) (basic_iteration.hpp:52-65, 85-93); a raised
// interrupt flag finishes too (interruptible_iteration.hpp:37-40)
bool finished(m4g_iteration* it, double r)
{
    it->resid = r;
    if (it->history && it->iterations < it->history_capacity) it->history[it->iterations] = r;
    const bool conv = it->norm_r0 == 0 ? r <= it->atol : (r <= it->rtol * it->norm_r0 || r <= it->atol);
    if (conv) return true;
    if (it->interrupt && *it->interrupt) {
        it->error = M4G_ERR_INTERRUPTED;
        return true;
    }
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
    return P ? check_pc_faults(*P) : 0;
}

// status of a finished solve: the iteration's error, codes set by
// iter.fail(k >= 2, ...) as M4G_BREAKDOWN, a failed sweep wait as M4G_ERR_TIMEOUT
int status_of(m4g_iteration* it, int fault)
{
    if (fault) return M4G_ERR_TIMEOUT;
    if (it->error == M4G_ERR_INTERRUPTED) return M4G_ERR_INTERRUPTED;
    if (it->error >= 2) return M4G_BREAKDOWN;
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

// itl::bicg (bicg.hpp:27-68); adjoint(A) * v runs as A^T (pattern kept with A,
// rows in A's row order, so each entry sums in the order of the transposed
// traversal; values gathered from A at the start of each solve)
static int bicg_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* P, m4g_iteration* it)
{
    const int64_t n = A.nrows;
    DevBuf<double> red(1);
    Ops o{n, current_stream(), red.ptr};
    const m4g_matrix_s& At = transpose_of(A, o.s);
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
        launch_spmv(At, p_t, q_t, M4G_ASSIGN, nullptr, o.s);
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

// ---- QMR, GMRES, BiCGStab(l), IDR(s) -----------------------------------------

// itl::qmr (qmr.hpp:29-109) with left L and right R (null: identity);
// trans(A) * q runs on A's kept transpose.  The loop tests rho = ||y||,
// the norm of the preconditioned Lanczos vector, as the reference does;
// breakdowns record their codes 1..6.
static int qmr_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* L, m4g_pc_s* R, m4g_iteration* it)
{
    const int64_t n = A.nrows;
    M4G_REQUIRE(n > 0, M4G_ERR_LOGIC, "empty rhs vector");
    DevBuf<double> red(1);
    Ops o{n, current_stream(), red.ptr};
    const m4g_matrix_s& At = transpose_of(A, o.s);
    Vecs V(14, n);
    double *r = V[0], *v_tld = V[1], *y = V[2], *w_tld = V[3], *z = V[4], *v = V[5], *w = V[6], *y_tld = V[7],
           *z_tld = V[8], *p = V[9], *q = V[10], *p_tld = V[11], *d = V[12], *sv = V[13];
    auto done = [&] {
        M4G_HIP(hipStreamSynchronize(o.s));
        return status_of(it, sweep_fault(L) | sweep_fault(R));
    };
    auto fail = [&](int code) {
        it->error = code;
        return done();
    };
    double beta, gamma = 1.0, gamma_1, delta, eta = -1.0, ep = 1.0, rho_1, theta = 0.0, theta_1;
    launch_spmv(A, x, r, M4G_ASSIGN, b, o.s);   // r = b - A*x
    o.copy(r, v_tld);
    apply_pc(L, v_tld, y, false, o);
    o.copy(r, w_tld);
    apply_pc(R, w_tld, z, true, o);
    if (finished(it, o.norm(r))) return done();
    double rho = o.norm(y), xi = o.norm(z);
    while (!finished(it, rho)) {
        ++it->iterations;
        if (rho == 0.0) return fail(1);   // "qmr breakdown #1, rho=0"
        if (xi == 0.0) return fail(2);    // "qmr breakdown #2, xi=0"
        o.div(v_tld, rho, v);
        o.div(y, rho, y);
        o.div(w_tld, xi, w);
        o.div(z, xi, z);
        delta = o.dot(z, y);
        if (delta == 0.0) return fail(3);
        apply_pc(R, y, y_tld, false, o);
        apply_pc(L, z, z_tld, true, o);
        if (it->iterations <= 1) {
            o.copy(y_tld, p);
            o.copy(z_tld, q);
        } else {
            o.minus(y_tld, (xi * delta) / ep, p, p);    // p = y_tld - ((xi*delta)/ep) * p
            o.minus(z_tld, (rho * delta) / ep, q, q);
        }
        launch_spmv(A, p, p_tld, M4G_ASSIGN, nullptr, o.s);
        ep = o.dot(q, p_tld);
        if (ep == 0.0) return fail(4);
        beta = ep / delta;
        if (beta == 0.0) return fail(5);
        o.minus(p_tld, beta, v, v_tld);
        apply_pc(L, v_tld, y, false, o);
        rho_1 = rho;
        rho = o.norm(y);
        launch_spmv(At, q, w_tld, M4G_ASSIGN, nullptr, o.s);
        o.minus(w_tld, beta, w, w_tld);               // w_tld = trans(A)*q - beta*w
        apply_pc(R, w_tld, z, true, o);
        xi = o.norm(z);
        gamma_1 = gamma;
        theta_1 = theta;
        theta = rho / (gamma_1 * beta);
        gamma = 1.0 / std::sqrt(1.0 + theta * theta);
        if (gamma == 0.0) return fail(6);
        eta = -eta * rho_1 * gamma * gamma / (beta * gamma_1 * gamma_1);
        if (it->iterations <= 1) {
            o.scale(eta, p, d);
            o.scale(eta, p_tld, sv);
        } else {
            const double f = theta_1 * theta_1 * gamma * gamma;
            o.lin_sum(eta, p, f, d, d);                 // d = eta*p + f*d
            o.lin_sum(eta, p_tld, f, sv, sv);
        }
        o.lin1(x, 1.0, d, x);                           // x += d
        o.minus(r, 1.0, sv, r);                         // r -= s
    }
    return done();
}

namespace {

// lu_solve(A, b) (mtl/operation/lu.hpp:132-140): partial pivoting on the first
// largest |a| of each column (lu.hpp:63-88), a zero pivot among the first
// m-1 is matrix_singular (false), then unit-lower and upper triangular
// solves; A is m x m row-major.
bool dense_lu_solve(int m, std::vector<double> A, const double* b, double* y)
{
    std::vector<int> P(static_cast<size_t>(m));
    for (int i = 0; i < m; i++) P[size_t(i)] = i;
    auto a = [&](int r, int c) -> double& { return A[size_t(r) * size_t(m) + size_t(c)]; };
    for (int i = 0; i + 1 < m; i++) {
        int rmax = i;
        for (int r = i + 1; r < m; r++)
            if (std::fabs(a(r, i)) > std::fabs(a(rmax, i))) rmax = r;
        if (rmax != i) {
            for (int c = 0; c < m; c++) std::swap(a(i, c), a(rmax, c));
            std::swap(P[size_t(i)], P[size_t(rmax)]);
        }
        if (std::fabs(a(i, i)) <= 0.0) return false;
        for (int r = i + 1; r < m; r++) a(r, i) /= a(i, i);
        for (int r = i + 1; r < m; r++)
            for (int c = i + 1; c < m; c++) a(r, c) -= a(r, i) * a(i, c);
    }
    std::vector<double> w(static_cast<size_t>(m));
    for (int r = 0; r < m; r++) {
        double t = b[P[size_t(r)]];
        for (int c = 0; c < r; c++) t -= a(r, c) * w[size_t(c)];
        w[size_t(r)] = t;
    }
    for (int r = m - 1; r >= 0; r--) {
        double t = w[size_t(r)];
        for (int c = r + 1; c < m; c++) t -= a(r, c) * y[c];
        y[r] = t / a(r, r);
    }
    return true;
}

// The Hessenberg matrix of gmres_full with the reference's rotations:
// mat::givens (mtl/operation/givens.hpp:40-84), G = [c s; -s c] applied as
// rows k, k+1 <- trans(G) * rows, then columns k, k+1 <- columns * G.
struct Hessenberg {
    int rows, cols;
    std::vector<double> h;
    Hessenberg(int r, int c) : rows(r), cols(c), h(size_t(r) * size_t(std::max(c, 1)), 0.0) {}
    double& operator()(int i, int j) { return h[size_t(i) * size_t(cols) + size_t(j)]; }
    static void rotation(double a, double b, double& c, double& s)
    {
        if (b == 0.0) {
            c = 1.0;
            s = 0.0;
        } else if (std::fabs(b) > std::fabs(a)) {
            const double t = -a / b;
            s = 1.0 / std::sqrt(1.0 + t * t);
            c = s * t;
        } else {
            const double t = -b / a;
            c = 1.0 / std::sqrt(1.0 + t * t);
            s = c * t;
        }
    }
    void trafo(int k, double c, double s)
    {
        for (int j = 0; j < cols; j++) {
            const double h0 = (*this)(k, j), h1 = (*this)(k + 1, j);
            (*this)(k, j) = c * h0 + (-s) * h1;
            (*this)(k + 1, j) = s * h0 + c * h1;
        }
        for (int i = 0; i < rows; i++) {
            const double h0 = (*this)(i, k), h1 = (*this)(i, k + 1);
            (*this)(i, k) = h0 * c + h1 * (-s);
            (*this)(i, k + 1) = h0 * s + h1 * c;
        }
    }
};

}  // namespace

// itl::gmres_full (gmres.hpp:37-119) on the iteration `it`: kmax = min(n,
// max_iter - iterations) Arnoldi steps (modified Gram-Schmidt, one
// re-orthogonalisation), the reference's rotation sequence on H, then
// H y = g on the largest leading block lu_solve accepts, x += solve(R, V y)
// and the final test on the true residual (iter.terminate(b - A*x)).
static int gmres_full_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* L, m4g_pc_s* R, m4g_iteration* it)
{
    const int64_t n = A.nrows;
    M4G_REQUIRE(n > 0, M4G_ERR_LOGIC, "empty rhs vector");
    const int64_t room = int64_t(it->max_iter) - int64_t(it->iterations);
    const int64_t kmax64 = room < 0 ? n : std::min<int64_t>(n, room);
    M4G_REQUIRE(kmax64 < (int64_t(1) << 20), M4G_ERR_INVALID, "gmres: too many Krylov vectors");
    const int kmax = int(kmax64);
    DevBuf<double> red(1);
    Ops o{n, current_stream(), red.ptr};
    Vecs W(4, n);
    double *r0 = W[0], *r = W[1], *va0 = W[2], *tmp = W[3];
    Vecs V(kmax + 1, n);
    Hessenberg H(kmax + 1, kmax);
    std::vector<double> sn(size_t(kmax) + 1, 0.0), cs(size_t(kmax) + 1, 0.0), g(size_t(kmax) + 1, 0.0),
        y(size_t(std::max(kmax, 1)), 0.0);
    auto done = [&] {
        M4G_HIP(hipStreamSynchronize(o.s));
        return status_of(it, sweep_fault(L) | sweep_fault(R));
    };
    launch_spmv(A, x, r0, M4G_ASSIGN, b, o.s);
    apply_pc(L, r0, r, false, o);
    double rho = g[0] = o.norm(r);
    if (finished(it, rho)) return done();
    o.div(r, rho, V[0]);
    int k;
    for (k = 0; k < kmax; ++k, ++it->iterations) {
        apply_pc(R, V[k], tmp, false, o);
        launch_spmv(A, tmp, va0, M4G_ASSIGN, nullptr, o.s);
        apply_pc(L, va0, V[k + 1], false, o);
        for (int j = 0; j < k + 1; j++) {
            H(j, k) = o.dot(V[j], V[k + 1]);
            o.minus(V[k + 1], H(j, k), V[j], V[k + 1]);
        }
        H(k + 1, k) = o.norm(V[k + 1]);
        for (int j = 0; j < k + 1; j++) {   // re-orthogonalise
            const double hr = o.dot(V[k + 1], V[j]);
            H(j, k) += hr;
            o.minus(V[k + 1], hr, V[j], V[k + 1]);
        }
        H(k + 1, k) = o.norm(V[k + 1]);
        if (H(k + 1, k) != 0.0) o.scale(1. / H(k + 1, k), V[k + 1], V[k + 1]);
        for (int i = 0; i < k; i++) {       // "k Given's rotations" as the reference applies them
            double c, s;
            Hessenberg::rotation(H(i, k - 1), H(i + 1, k - 1), c, s);
            H.trafo(i, c, s);
        }
        const double nu = std::sqrt(H(k, k) * H(k, k) + H(k + 1, k) * H(k + 1, k));
        if (nu != 0.0) {
            cs[size_t(k)] = H(k, k) / nu;
            sn[size_t(k)] = -H(k + 1, k) / nu;
            H(k, k) = cs[size_t(k)] * H(k, k) - sn[size_t(k)] * H(k + 1, k);
            H(k + 1, k) = 0;
            const double w1 = cs[size_t(k)] * g[size_t(k)] - sn[size_t(k)] * g[size_t(k) + 1],
                         w2 = sn[size_t(k)] * g[size_t(k)] + cs[size_t(k)] * g[size_t(k) + 1];
            g[size_t(k)] = w1;
            g[size_t(k) + 1] = w2;
        }
        rho = std::fabs(g[size_t(k) + 1]);
    }
    while (k > 0 && std::fabs(g[size_t(k) - 1]) <= it->atol) k--;
    int m = k;
    for (; m > 0; --m) {
        std::vector<double> Hm(size_t(m) * size_t(m));
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) Hm[size_t(i) * size_t(m) + size_t(j)] = H(i, j);
        if (dense_lu_solve(m, std::move(Hm), g.data(), y.data())) break;
    }
    if (m == 0) {   // "GMRES did not find any direction to correct x"
        it->error = 2;
        return done();
    }
    std::vector<const double*> vs(static_cast<size_t>(m));
    for (int j = 0; j < m; j++) vs[size_t(j)] = V[j];
    o.combo(vs, std::vector<double>(y.begin(), y.begin() + m), va0);   // V * y
    apply_pc(R, va0, tmp, false, o);
    o.lin1(x, 1.0, tmp, x);
    launch_spmv(A, x, r, M4G_ASSIGN, b, o.s);
    finished(it, o.norm(r));   // iter.terminate(r)
    return done();
}

// itl::gmres with restart (gmres.hpp:122-137): gmres_full on a copy of the
// iteration allowed `restart` more steps, then update_progress
// (basic_iteration.hpp:145-155), until the iteration is finished.
static int gmres_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* L, m4g_pc_s* R, m4g_iteration* it,
                   int64_t restart)
{
    M4G_REQUIRE(restart >= 1, M4G_ERR_INVALID, "gmres: restart must be at least 1");
    for (;;) {
        m4g_iteration inner = *it;
        inner.max_iter = int(std::min<int64_t>(int64_t(it->iterations) + restart, it->max_iter));
        const int st = gmres_full_t(A, x, b, L, R, &inner);
        if (st == M4G_ERR_TIMEOUT) return st;
        it->iterations = inner.iterations;
        it->resid = inner.resid;
        if (inner.error > 1) {
            it->error = inner.error;
            break;
        }
        if (finished(it, it->resid)) break;
    }
    return status_of(it, 0);
}

// itl::bicgstab_ell (bicgstab_ell.hpp:37-135) with left L and right R (null:
// identity); itl::bicgstab_2 (bicgstab_2.hpp:35-130) is l = 2 without R.
// As in the reference, the inner test checks r_hat[j] and leaves before that
// step's x += alpha * u_hat[0]; the MR part orthogonalises r_hat[1..l] in
// place (orthogonalize_factors, orth.hpp:61-85); omega == 0 is error 3.
static int bicgstab_ell_t(m4g_matrix_s& A, double* x, const double* b, m4g_pc_s* L, m4g_pc_s* R,
                          m4g_iteration* it, int64_t l64)
{
    const int64_t n = A.nrows;
    M4G_REQUIRE(n > 0, M4G_ERR_LOGIC, "empty rhs vector");
    M4G_REQUIRE(l64 >= 1 && l64 <= 64, M4G_ERR_INVALID, "bicgstab_ell: l must be in [1, 64]");
    const int l = int(l64);
    DevBuf<double> red(1);
    Ops o{n, current_stream(), red.ptr};
    Vecs W(4, n), RH(l + 1, n), UH(l + 1, n);
    double *x0 = W[0], *y = W[1], *tmp = W[2], *r0_tilde = W[3];
    auto done = [&] {
        M4G_HIP(hipStreamSynchronize(o.s));
        return status_of(it, sweep_fault(L) | sweep_fault(R));
    };
    auto finish_x = [&] {   // x = solve(R, x); x += x0
        apply_pc(R, x, tmp, false, o);
        o.lin1(tmp, 1.0, x0, x);
    };
    launch_fill(n, 0.0, x0, M4G_F64, o.s);
    o.copy(b, RH[0]);
    if (o.norm(x) != 0.0) {   // shift the problem
        launch_spmv(A, x, tmp, M4G_ASSIGN, nullptr, o.s);
        o.minus(RH[0], 1.0, tmp, RH[0]);
        o.copy(x, x0);
        launch_fill(n, 0.0, x, M4G_F64, o.s);
    }
    o.div(RH[0], o.norm(RH[0]), r0_tilde);
    apply_pc(L, RH[0], y, false, o);
    o.copy(y, RH[0]);
    launch_fill(n, 0.0, UH[0], M4G_F64, o.s);
    double rho_0 = 1.0, rho_1 = 0.0, alpha = 0.0, Gamma = 0.0, beta = 0.0, omega = 1.0;
    const size_t L1 = size_t(l) + 1;
    std::vector<double> tau(size_t(l) * size_t(l)), gamma(L1), gamma_a(L1), gamma_aa(L1);
    auto T = [&](int i, int j) -> double& { return tau[size_t(i - 1) * size_t(l) + size_t(j - 1)]; };
    while (!finished(it, o.norm(RH[0]))) {
        ++it->iterations;
        rho_0 = -omega * rho_0;
        for (int j = 0; j < l; ++j) {
            rho_1 = o.dot(r0_tilde, RH[j]);
            beta = alpha * rho_1 / rho_0;
            rho_0 = rho_1;
            for (int i = 0; i <= j; ++i) o.minus(RH[i], beta, UH[i], UH[i]);   // u_hat[i] = r_hat[i] - beta*u_hat[i]
            apply_pc(R, UH[j], tmp, false, o);
            launch_spmv(A, tmp, y, M4G_ASSIGN, nullptr, o.s);
            apply_pc(L, y, UH[j + 1], false, o);
            Gamma = o.dot(r0_tilde, UH[j + 1]);
            alpha = rho_0 / Gamma;
            for (int i = 0; i <= j; ++i) o.minus(RH[i], alpha, UH[i + 1], RH[i]);
            if (finished(it, o.norm(RH[j]))) {
                finish_x();
                return done();
            }
            apply_pc(R, RH[j], RH[j + 1], false, o);
            launch_spmv(A, RH[j + 1], y, M4G_ASSIGN, nullptr, o.s);
            apply_pc(L, y, RH[j + 1], false, o);
            o.lin1(x, alpha, UH[0], x);
        }
        // MR part: orthogonalize_factors(r_hat[1..l])
        std::fill(tau.begin(), tau.end(), 0.0);
        T(1, 1) = o.dot(RH[1], RH[1]);
        for (int j = 2; j <= l; j++) {
            for (int i = 1; i < j; i++) {
                const double t = o.dot(RH[i], RH[j]) / T(i, i);
                T(i, j) = t;
                o.minus(RH[j], t, RH[i], RH[j]);
            }
            T(j, j) = o.dot(RH[j], RH[j]);
        }
        for (int j = 1; j <= l; ++j) gamma_a[size_t(j)] = o.dot(RH[j], RH[0]) / T(j, j);
        gamma[size_t(l)] = gamma_a[size_t(l)];
        omega = gamma[size_t(l)];
        if (omega == 0.0) {   // "bicg breakdown #2"
            it->error = 3;
            return done();
        }
        for (int j = l - 1; j > 0; --j) {
            double sum = 0.0;
            for (int i = j + 1; i <= l; ++i) sum += T(j, i) * gamma[size_t(i)];
            gamma[size_t(j)] = gamma_a[size_t(j)] - sum;
        }
        for (int j = 1; j < l; ++j) {   // strict_upper(tau[1:l][1:l]) * gamma[2:] + gamma[2:]
            double sum = 0.0;
            for (int i = j + 1; i < l; ++i) sum += T(j, i) * gamma[size_t(i) + 1];
            gamma_aa[size_t(j)] = sum + gamma[size_t(j) + 1];
        }
        o.lin1(x, gamma[1], RH[0], x);
        o.minus(RH[0], gamma_a[size_t(l)], RH[l], RH[0]);
        o.minus(UH[0], gamma[size_t(l)], UH[l], UH[0]);
        for (int j = 1; j < l; ++j) {
            o.minus(UH[0], gamma[size_t(j)], UH[j], UH[0]);
            o.lin1(x, gamma_aa[size_t(j)], RH[j], x);
            o.minus(RH[0], gamma_a[size_t(j)], RH[j], RH[0]);
        }
    }
    finish_x();
    return done();
}

// mtl::seed<double>::counter (mtl/operation/random.hpp:29-38)
static int g_seed_counter = 0;

// itl::idr_s (idr_s.hpp:35-103); the preconditioners are ignored as in the
// reference.  P = random(P) (column j: srand(17 * ++counter), rand()), its
// column 0 replaced by r, then orth(P) (orth.hpp:35-54); each cycle takes s
// steps.  A singular M in lu_solve is M4G_ERR_SINGULAR (mtl::matrix_singular).
static int idr_s_t(m4g_matrix_s& A, double* x, const double* b, m4g_iteration* it, int64_t s64)
{
    const int64_t n = A.nrows;
    M4G_REQUIRE(n > 0, M4G_ERR_LOGIC, "empty rhs vector");
    M4G_REQUIRE(s64 <= 64, M4G_ERR_INVALID, "idr_s: s must be at most 64");
    const int s = int(std::max<int64_t>(s64, 1));
    DevBuf<double> red(size_t(s) + 1);
    Ops o{n, current_stream(), red.ptr};
    Vecs W(5, n), dR(s, n), dX(s, n), P(s, n);
    double *v = W[0], *t = W[1], *q = W[2], *r = W[3], *tmp = W[4];
    auto done = [&] {
        M4G_HIP(hipStreamSynchronize(o.s));
        return status_of(it, 0);
    };
    std::vector<const double*> Pp(static_cast<size_t>(s)), dRp(static_cast<size_t>(s)), dXp(static_cast<size_t>(s));
    for (int j = 0; j < s; j++) {
        launch_fill(n, 0.0, dR[j], M4G_F64, o.s);
        launch_fill(n, 0.0, dX[j], M4G_F64, o.s);
        Pp[size_t(j)] = P[j];
        dRp[size_t(j)] = dR[j];
        dXp[size_t(j)] = dX[j];
    }
    std::vector<double> m(static_cast<size_t>(s)), c(static_cast<size_t>(s)), M(size_t(s) * size_t(s), 0.0);
    launch_spmv(A, x, r, M4G_ASSIGN, b, o.s);
    {
        std::vector<double> host(static_cast<size_t>(n));
        for (int j = 0; j < s; j++) {
            srand(unsigned(17 * ++g_seed_counter));
            for (int64_t e = 0; e < n; e++) host[size_t(e)] = double(rand());
            M4G_HIP(hipMemcpyAsync(P[j], host.data(), size_t(n) * sizeof(double), hipMemcpyHostToDevice, o.s));
            M4G_HIP(hipStreamSynchronize(o.s));   // host buffer reused
        }
    }
    o.copy(r, P[0]);
    for (int j = 0; j < s; j++) {   // orth(P)
        for (int i = 0; i < j; i++) o.minus(P[j], o.dot(P[i], P[j]), P[i], P[j]);
        o.div(P[j], o.norm(P[j]), P[j]);
    }
    double omega = 0.0;
    for (int k = 0; k < s; k++) {
        launch_spmv(A, r, v, M4G_ASSIGN, nullptr, o.s);
        omega = o.dot(v, r) / o.dot(v, v);
        o.scale(omega, r, dX[k]);
        o.scale(-omega, v, dR[k]);
        o.lin1(x, 1.0, dX[k], x);
        o.lin1(r, 1.0, dR[k], r);
        ++it->iterations;
        if (finished(it, o.norm(r))) return done();
        const std::vector<double> col = o.dots(Pp, dR[k], red.ptr);
        for (int i = 0; i < s; i++) M[size_t(i) * size_t(s) + size_t(k)] = col[size_t(i)];
    }
    int oldest = 0;
    m = o.dots(Pp, r, red.ptr);
    while (!finished(it, o.norm(r))) {
        for (int k = 0; k < s; k++) {
            M4G_REQUIRE(dense_lu_solve(s, M, m.data(), c.data()), M4G_ERR_SINGULAR, "matrix_singular in idr_s");
            std::vector<double> mc(static_cast<size_t>(s));
            for (int j = 0; j < s; j++) mc[size_t(j)] = -c[size_t(j)];
            o.combo(dRp, mc, q);                          // q = dR * -c
            o.lin1(r, 1.0, q, v);                         // v = r + q
            o.combo(dXp, c, tmp);                         // dX * c
            if (k == 0) {
                launch_spmv(A, v, t, M4G_ASSIGN, nullptr, o.s);
                omega = o.dot(t, v) / o.dot(t, t);
                o.minus(q, omega, t, dR[oldest]);         // dR[oldest] = q - omega*t
                o.scale_minus(omega, v, tmp, dX[oldest]); // dX[oldest] = omega*v - dX*c
            } else {
                o.scale_minus(omega, v, tmp, dX[oldest]);
                o.scale(-1.0, dX[oldest], tmp);
                launch_spmv(A, tmp, dR[oldest], M4G_ASSIGN, nullptr, o.s);   // A * -dX[oldest]
            }
            o.lin1(r, 1.0, dR[oldest], r);
            o.lin1(x, 1.0, dX[oldest], x);
            ++it->iterations;
            if (finished(it, o.norm(r))) return done();
            const std::vector<double> dm = o.dots(Pp, dR[oldest], red.ptr);
            for (int i = 0; i < s; i++) {
                M[size_t(i) * size_t(s) + size_t(oldest)] = dm[size_t(i)];
                m[size_t(i)] += dm[size_t(i)];
            }
            oldest = (oldest + 1) % s;
        }
    }
    return done();
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
        TraceRange tr(kTraceCgs);
        return cgs_t(*A, (double*)x->ptr, (const double*)b->ptr, P, it);
    });
}

int m4g_bicg(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_bicg", A, x, b, it);
        TraceRange tr(kTraceBicg);
        return bicg_t(*A, (double*)x->ptr, (const double*)b->ptr, P, it);
    });
}

int m4g_tfqmr(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_tfqmr", A, x, b, it);
        TraceRange tr(kTraceTfqmr);
        return tfqmr_t(*A, (double*)x->ptr, (const double*)b->ptr, L, R, it);
    });
}

int m4g_qmr(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_qmr", A, x, b, it);
        TraceRange tr(kTraceQmr);
        return qmr_t(*A, (double*)x->ptr, (const double*)b->ptr, L, R, it);
    });
}

int m4g_gmres_full(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_gmres_full", A, x, b, it);
        TraceRange tr(kTraceGmresFull);
        return gmres_full_t(*A, (double*)x->ptr, (const double*)b->ptr, L, R, it);
    });
}

int m4g_gmres(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it, int64_t restart)
{
    return guard([&] {
        check_args("m4g_gmres", A, x, b, it);
        TraceRange tr(kTraceGmres);
        return gmres_t(*A, (double*)x->ptr, (const double*)b->ptr, L, R, it, restart);
    });
}

int m4g_bicgstab_2(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_iteration* it)
{
    return guard([&] {
        check_args("m4g_bicgstab_2", A, x, b, it);
        TraceRange tr(kTraceBicgstab2);
        return bicgstab_ell_t(*A, (double*)x->ptr, (const double*)b->ptr, L, nullptr, it, 2);
    });
}

int m4g_bicgstab_ell(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it, int64_t l)
{
    return guard([&] {
        check_args("m4g_bicgstab_ell", A, x, b, it);
        TraceRange tr(kTraceBicgstabEll);
        return bicgstab_ell_t(*A, (double*)x->ptr, (const double*)b->ptr, L, R, it, l);
    });
}

int m4g_idr_s(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it, int64_t s)
{
    (void)L;
    (void)R;
    return guard([&] {
        check_args("m4g_idr_s", A, x, b, it);
        TraceRange tr(kTraceIdrS);
        return idr_s_t(*A, (double*)x->ptr, (const double*)b->ptr, it, s);
    });
}

int m4g_random_seed_counter(void) { return g_seed_counter; }

void m4g_set_random_seed_counter(int counter) { g_seed_counter = counter; }

}  // extern "C"
