// Same checks as T-itl/ic_0_test.cpp: ic_0<compressed2D<double>, float> on the
// 3x3-grid Laplacian; solve(P, iota) must equal the reference's golden vector.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>

int main()
{
    const int size = 3, N = size * size;
    typedef mtl::compressed2D<double> matrix_type;
    mtl::compressed2D<double> A(N, N);
    laplacian_setup(A, size, size);

    itl::pc::ic_0<matrix_type, float> P(A);
    mtl::dense_vector<double> x(N), y, yc(N);
    iota(x);
    yc = 1.03194, 1.60198, 1.45357, 2.5258, 3.55386, 3.16, 2.91787, 4.09338, 3.81335;

    y = solve(P, x);
    std::cout << y << '\n';
    MTL_THROW_IF(two_norm(mtl::dense_vector<double>(y - yc)) > 0.0001,
                 mtl::logic_error("IC(0) doesn't yield expected result"));
    return 0;
}
