// Stopping on an absolute tolerance alone: with rtol = 0 the only way for a
// cyclic_iteration to report convergence is resid <= atol.  Known answer of
// the reference's absolut_termination_criterion_test (T-itl/, line 39): CG
// with pc::diagonal on the 4x4-grid Laplacian, atol = 1e-6, from x = 0 with
// b = A * ones, leaves a true residual below atol.  Also checked on a 9x7
// grid, where the reference has no test, as the same rule must hold there.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cstdio>

namespace {

// true residual of CG + diagonal PC stopped by atol only
double residual_after_atol_stop(int nx, int ny, double atol, int& iterations)
{
    using Mat = mtl::compressed2D<double>;
    const int n = nx * ny;
    Mat K(n, n);
    laplacian_setup(K, nx, ny);
    mtl::dense_vector<double> ones(n, 1.0), rhs(n), sol(n, 0.0);
    rhs = K * ones;
    itl::pc::diagonal<Mat> jacobi(K);
    itl::cyclic_iteration<double> ctl(rhs, 1000, 0.0, atol, 1000);
    const int code = itl::cg(K, sol, rhs, jacobi, ctl);
    if (code != 0) std::printf("cg returned %d\n", code);
    iterations = ctl.iterations();
    mtl::dense_vector<double> res(rhs - K * sol);
    return two_norm(res);
}

}  // namespace

int main()
{
    const double atol = 1.0e-6;
    int failures = 0;
    const int grids[2][2] = {{4, 4}, {9, 7}};
    for (const auto& g : grids) {
        int its = 0;
        const double rn = residual_after_atol_stop(g[0], g[1], atol, its);
        std::printf("grid %dx%d: %d iterations, |b - A x| = %.3e (atol %.1e)\n", g[0], g[1], its, rn, atol);
        if (!(rn <= atol)) failures++;
    }
    if (failures) {
        std::printf("absolute tolerance not honoured on %d grid(s)\n", failures);
        return 1;
    }
    return 0;
}
