// ILU(0) of the 3x4-grid Laplacian against a dense incomplete LU written here
// (IKJ order, updates restricted to A's pattern; the same check the
// reference's ilu_0_bicgstab_test makes on two entries, T-itl/, lines 18-48).
// Every entry of get_L() (strict lower, unit diagonal implied) and get_U()
// (diagonal stored inverted, ilu_0.hpp:79-81) is compared.  Then BiCGStab +
// ILU(0) from x = 0 with b = A * ones reaches 1e-6.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cmath>
#include <vector>
#include "check.hpp"

int main()
{
    using Mat = mtl::compressed2D<double>;
    Checks c;
    const int nx = 3, ny = 4, n = nx * ny;
    Mat K(n, n);
    laplacian_setup(K, nx, ny);

    std::vector<std::vector<double>> LU(n, std::vector<double>(n));
    std::vector<std::vector<bool>> in(n, std::vector<bool>(n));
    for (int r = 0; r < n; r++)
        for (int s = 0; s < n; s++) {
            LU[r][s] = K[r][s];
            in[r][s] = LU[r][s] != 0.0;
        }
    for (int i = 1; i < n; i++)
        for (int k = 0; k < i; k++) {
            if (!in[i][k]) continue;
            LU[i][k] /= LU[k][k];
            for (int j = k + 1; j < n; j++)
                if (in[i][j]) LU[i][j] -= LU[i][k] * LU[k][j];
        }

    itl::pc::ilu_0<Mat> P(K);
    const Mat L = P.get_L(), U = P.get_U();
    for (int r = 0; r < n; r++)
        for (int s = 0; s < n; s++) {
            if (s < r) c.near(L[r][s], LU[r][s], 1e-14 * std::fabs(LU[r][s]), "L entry");
            else if (s == r) c.near(U[r][s], 1.0 / LU[r][r], 1e-14 / std::fabs(LU[r][r]), "inverted U diagonal");
            else c.near(U[r][s], LU[r][s], 1e-14 * std::fabs(LU[r][s]), "U entry");
        }

    mtl::dense_vector<double> ones(n, 1.0), rhs(n), sol(n, 0.0);
    rhs = K * ones;
    itl::cyclic_iteration<double> ctl(rhs, n, 1.0e-6, 0.0, n);
    const int code = itl::bicgstab(K, sol, rhs, P, ctl);
    mtl::dense_vector<double> res(rhs - K * sol);
    std::printf("BiCGStab + ILU(0): code %d, %d iterations, true relres %.2e\n", code, ctl.iterations(),
                two_norm(res) / two_norm(rhs));
    c.expect(code == 0, "BiCGStab + ILU(0) converges");
    c.expect(two_norm(res) <= 1.0e-5 * two_norm(rhs), "true residual near the tolerance");
    return c.status();
}
