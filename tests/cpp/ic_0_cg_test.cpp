// Same checks as T-itl/ic_0_cg_test.cpp: sparse IC(0) against a dense
// textbook IC(0) (U[2][3] and 1/U[3][3]), then CG + IC(0) to 1e-6.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <vector>

int main()
{
    const int size = 3, N = size * size;
    typedef mtl::compressed2D<double> matrix_type;
    mtl::compressed2D<double> A(N, N);
    laplacian_setup(A, size, size);

    itl::pc::ic_0<matrix_type, float> P(A);
    // dense textbook IC(0) on the upper triangle (ic_0_cg_test.cpp:17-40)
    std::vector<std::vector<double>> U(N, std::vector<double>(N, 0.0));
    for (int r = 0; r < N; r++)
        for (int c = r; c < N; c++) U[r][c] = A[r][c];
    for (int k = 0; k < N; k++) {
        double dia = U[k][k] = std::sqrt(U[k][k]);
        for (int i = k + 1; i < N; i++) {
            double d = U[k][i] /= dia;
            for (int j = k + 1; j <= i; j++)
                if (U[j][i] != 0.0) U[j][i] -= d * U[k][j];
        }
    }
    auto Us = P.get_U();
    if (std::abs(U[2][3] - Us[2][3]) > 0.001) throw "Wrong value in L for sparse IC(0) factorization";
    if (std::abs(U[3][3] - 1. / Us[3][3]) > 0.001) throw "Wrong value in U for sparse IC(0) factorization";

    mtl::dense_vector<double> x(N, 1.0), b(N);
    b = A * x;
    x = 0;
    itl::cyclic_iteration<double> iter(b, N, 1.e-6, 0.0, 1);
    int err = cg(A, x, b, P, iter);
    mtl::dense_vector<double> r(b - A * x);
    MTL_THROW_IF(err != 0 || two_norm(r) > 1e-6 * two_norm(b) * 10, mtl::runtime_error("CG + IC(0) failed"));
    return 0;
}
