// Same check as T-itl/ilu_0_cg_test.cpp: CG + ILU(0) on the 3x3 grid reaches
// 1e-6 within 4 iterations.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>

int main()
{
    const int size = 3, N = size * size;
    typedef mtl::compressed2D<double> Matrix;
    Matrix A;
    laplacian_setup(A, size, size);
    itl::pc::ilu_0<Matrix> P(A);
    mtl::dense_vector<double> x(N, 1.0), b(N);
    b = A * x;
    x = 0;
    itl::cyclic_iteration<double> iter(b, N, 1.e-6, 0.0, 1);
    cg(A, x, b, P, iter);
    MTL_THROW_IF(size == 3 && iter.iterations() > 4, mtl::runtime_error("Too many iterations in cg with ILU(0)"));
    return 0;
}
