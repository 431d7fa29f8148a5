// Same check as T-itl/absolut_termination_criterion_test.cpp: diagonal PC,
// rtol = 0, atol = 1e-6; the true residual must be below atol.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>

int main()
{
    const int size = 4, N = size * size;
    const double abs = 1.e-6;
    typedef mtl::compressed2D<double> matrix_type;
    mtl::compressed2D<double> A(N, N);
    laplacian_setup(A, size, size);
    itl::pc::diagonal<matrix_type> P(A);
    mtl::dense_vector<double> x(N, 1.0), b(N);
    b = A * x;
    x = 0;
    itl::cyclic_iteration<double> iter(b, 500, 0, abs, 5);
    cg(A, x, b, P, iter);
    mtl::dense_vector<double> r(b - A * x);
    std::cout << "Norm r = " << two_norm(r) << "\n";
    if (two_norm(r) > abs) throw "Residuum not reduced enough.";
    return 0;
}
