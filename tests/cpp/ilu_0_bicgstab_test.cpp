// Same checks as T-itl/ilu_0_bicgstab_test.cpp: sparse ILU(0) against a dense
// textbook ILU(0) ([3][2] and 1/U[3][3]), then BiCGStab + ILU(0) to 1e-6.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <vector>

int main()
{
    const int size = 3, N = size * size;
    typedef mtl::compressed2D<double> matrix_type;
    mtl::compressed2D<double> A(N, N);
    laplacian_setup(A, size, size);
    itl::pc::ilu_0<matrix_type> P(A);

    std::vector<std::vector<double>> LU(N, std::vector<double>(N));
    for (int r = 0; r < N; r++)
        for (int c = 0; c < N; c++) LU[r][c] = A[r][c];
    for (int i = 1; i < N; i++)
        for (int k = 0; k < i; k++) {
            LU[i][k] /= LU[k][k];
            for (int j = k + 1; j < N; j++)
                if (LU[i][j] != 0) LU[i][j] -= LU[i][k] * LU[k][j];
        }
    matrix_type L = P.get_L(), U = P.get_U();
    if (std::abs(LU[3][2] - L[3][2]) > 0.001) throw "Wrong value in L for sparse ILU(0) factorization";
    if (std::abs(LU[3][3] - 1. / U[3][3]) > 0.001) throw "Wrong value in U for sparse ILU(0) factorization";

    mtl::dense_vector<double> x(N, 1.0), b(N);
    b = A * x;
    x = 0;
    itl::cyclic_iteration<double> iter(b, N, 1.e-6, 0.0, 5);
    int err = bicgstab(A, x, b, P, iter);
    mtl::dense_vector<double> r(b - A * x);
    MTL_THROW_IF(err != 0 || two_norm(r) > 1e-5 * two_norm(b), mtl::runtime_error("BiCGStab + ILU(0) failed"));
    return 0;
}
