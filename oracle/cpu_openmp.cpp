// CPU timing baseline only (test infrastructure, like the rest of oracle/):
// IC(0)-preconditioned CG as MTL4 runs it when built with MTL_WITH_OPENMP —
// the row loops of the SpMV, the dots and the vector updates split over
// OpenMP threads, the IC(0) factorisation and both triangular sweeps
// sequential, no loop fusion.  The algorithm is itl::cg (cg.hpp:75-112) with
// basic_iteration's test (basic_iteration.hpp:85-93); the sweeps are
// lower_trisolve with the inverted diagonal last (lower_trisolve.hpp:191-209)
// on L = trans(U) and upper_trisolve with it first (upper_trisolve.hpp:
// 159-176) on U, as the oracle's factors hold them.  Dots are summed in
// per-thread partials, so iterates are not bit-identical to the oracle's.
#include <cmath>
#include <cstdint>
#include <vector>

#include <omp.h>

typedef int64_t idx_t;

extern "C" int ompc_ic0_pcg(idx_t n, const idx_t* S, const idx_t* I, const double* V, const idx_t* US,
                            const idx_t* UI, const double* UD, const idx_t* LS, const idx_t* LI, const double* LD,
                            const double* b, double* x, double rtol, int max_iter, int threads, int* iterations)
{
    omp_set_num_threads(threads);
    std::vector<double> r(n), z(n), p(n), q(n), y(n);
    auto spmv = [&](const double* v, double* w) {
#pragma omp parallel for schedule(static)
        for (idx_t i = 0; i < n; i++) {
            double s = 0;
            for (idx_t e = S[i]; e < S[i + 1]; e++) s += V[e] * v[I[e]];
            w[i] = s;
        }
    };
    auto dot = [&](const double* a, const double* c) {
        double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
        for (idx_t i = 0; i < n; i++) s += a[i] * c[i];
        return s;
    };
    spmv(x, q.data());
#pragma omp parallel for schedule(static)
    for (idx_t i = 0; i < n; i++) r[i] = b[i] - q[i];
    const double norm_r0 = std::sqrt(dot(b, b));
    double rho_1 = 0;
    int it = 0;
    for (;;) {
        const double res = std::sqrt(dot(r.data(), r.data()));
        if (res <= rtol * norm_r0 || it >= max_iter) break;
        ++it;
        // z = solve(P, r): sequential sweeps
        for (idx_t i = 0; i < n; i++) {
            double acc = r[i];
            const idx_t last = LS[i + 1] - 1;
            for (idx_t e = LS[i]; e < last; e++) acc -= LD[e] * y[LI[e]];
            y[i] = acc * LD[last];
        }
        for (idx_t i = n - 1; i >= 0; i--) {
            double acc = y[i];
            for (idx_t e = US[i] + 1; e < US[i + 1]; e++) acc -= UD[e] * z[UI[e]];
            z[i] = acc * UD[US[i]];
        }
        const double rho = dot(r.data(), z.data());
        if (it == 1) {
#pragma omp parallel for schedule(static)
            for (idx_t i = 0; i < n; i++) p[i] = z[i];
        } else {
            const double beta = rho / rho_1;
#pragma omp parallel for schedule(static)
            for (idx_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        }
        spmv(p.data(), q.data());
        const double alpha = rho / dot(p.data(), q.data());
#pragma omp parallel for schedule(static)
        for (idx_t i = 0; i < n; i++) x[i] += alpha * p[i];
        rho_1 = rho;
#pragma omp parallel for schedule(static)
        for (idx_t i = 0; i < n; i++) r[i] -= alpha * q[i];
    }
    *iterations = it;
    return 0;
}
