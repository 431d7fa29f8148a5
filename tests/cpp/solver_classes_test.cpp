// Same use as T-itl/solver_classes_test.cpp: the solver classes
// cg_solver / bicgstab_solver with operator()(x, b) and iteration_ref().
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>

int main()
{
    const int size = 3, N = size * size;
    typedef mtl::compressed2D<double> Matrix;
    Matrix A;
    laplacian_setup(A, size, size);
    mtl::dense_vector<double> x(N, 1.0), b(N);
    b = A * x;

    itl::cg_solver<Matrix, itl::pc::ic_0<Matrix>> s1(A);
    x = 0;
    s1(x, b);
    if (s1.iteration_ref().iterations() > 9) std::cerr << "Too many iterations\n";
    MTL_THROW_IF(!s1.iteration_ref().is_converged(), mtl::runtime_error("cg_solver did not converge"));

    itl::bicgstab_solver<Matrix, itl::pc::ilu_0<Matrix>> s2(A);
    x = 0;
    s2(x, b);
    MTL_THROW_IF(!s2.iteration_ref().is_converged(), mtl::runtime_error("bicgstab_solver did not converge"));
    mtl::dense_vector<double> r(b - A * x);
    MTL_THROW_IF(two_norm(r) > 1e-8 * two_norm(b), mtl::runtime_error("residual too large"));
    return 0;
}
