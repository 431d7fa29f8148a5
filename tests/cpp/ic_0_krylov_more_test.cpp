// The cgs / bicg / tfqmr lines of T-itl/ic_0_krylov_test.cpp: the 4x4-grid
// Laplacian, x0 = 0.01, b = A * ones, cyclic_iteration(b, 150, 1e-4), each
// solver must return 0; plus the solver classes of solver_classes_test.cpp
// (cgs_solver, tfqmr_solver with and without ILU(0)).
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <iostream>

int main()
{
    const int size = 4, N = size * size;
    typedef mtl::compressed2D<double> matrix_s_type;
    matrix_s_type As(N, N);
    laplacian_setup(As, size, size);
    mtl::dense_vector<double> xs(N, 1.0), bs(N);
    bs = As * xs;
    itl::pc::identity<matrix_s_type> I(As);
    itl::pc::ilu_0<matrix_s_type> ILU(As);
    int failed = 0;
    {
        xs = 0.01;
        itl::cyclic_iteration<double> iters(bs, 150, 1.e-4, 0.0, 10);
        if (itl::bicg(As, xs, bs, I, iters) != 0) failed++;
    }
    {
        xs = 0.01;
        itl::cyclic_iteration<double> iters(bs, 150, 1.e-4, 0.0, 10);
        if (itl::cgs(As, xs, bs, ILU, iters) != 0) failed++;
    }
    {
        xs = 0.01;
        itl::cyclic_iteration<double> iters(bs, 150, 1.e-4, 0.0, 10);
        if (itl::tfqmr(As, xs, bs, ILU, I, iters) != 0) failed++;
    }
    std::cout << 3 - failed << " solvers succeeded and " << failed << " solvers failed.\n";
    MTL_THROW_IF(failed, mtl::runtime_error("a solver did not converge"));

    mtl::dense_vector<double> x(N, 0.0);
    itl::cgs_solver<matrix_s_type> c1(As);
    itl::cgs_solver<matrix_s_type, itl::pc::ilu_0<matrix_s_type>> c2(As);
    itl::tfqmr_solver<matrix_s_type> t1(As);
    itl::tfqmr_solver<matrix_s_type, itl::pc::ilu_0<matrix_s_type>> t2(As);
    itl::bicg_solver<matrix_s_type> b1(As);
    x = 0; c1(x, bs); MTL_THROW_IF(!c1.iteration_ref().is_converged(), mtl::runtime_error("cgs_solver"));
    x = 0; c2(x, bs); MTL_THROW_IF(!c2.iteration_ref().is_converged(), mtl::runtime_error("cgs_solver ilu_0"));
    x = 0; t1(x, bs); MTL_THROW_IF(!t1.iteration_ref().is_converged(), mtl::runtime_error("tfqmr_solver"));
    x = 0; t2(x, bs); MTL_THROW_IF(!t2.iteration_ref().is_converged(), mtl::runtime_error("tfqmr_solver ilu_0"));
    x = 0; b1(x, bs); MTL_THROW_IF(!b1.iteration_ref().is_converged(), mtl::runtime_error("bicg_solver"));
    mtl::dense_vector<double> r(bs - As * x);
    MTL_THROW_IF(two_norm(r) > 1e-6 * two_norm(bs), mtl::runtime_error("residual"));
    return 0;
}
