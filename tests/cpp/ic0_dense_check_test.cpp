// IC(0) of the 4x3-grid Laplacian against a dense incomplete Cholesky written
// here (updates restricted to the pattern of upper(A); on this triangle-free
// grid graph every formulation of IC(0) agrees, cf. the dense check of the
// reference's ic_0_cg_test, T-itl/, lines 17-40).  Every entry of get_U() is
// compared, the diagonal stored inverted (ic_0.hpp:160-161), for double and
// for float factor storage.  Then CG + IC(0) from x = 0 with b = A * ones
// reaches 1e-6 (the reference test's tolerance) and its true residual agrees.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cmath>
#include <vector>
#include "check.hpp"

namespace {

using Mat = mtl::compressed2D<double>;
using Dense = std::vector<std::vector<double>>;

Dense dense_ic0(const Mat& K, int n)
{
    Dense U(n, std::vector<double>(n, 0.0));
    std::vector<std::vector<bool>> in(n, std::vector<bool>(n, false));
    for (int r = 0; r < n; r++)
        for (int s = r; s < n; s++) {
            U[r][s] = K[r][s];
            in[r][s] = K[r][s] != 0.0;
        }
    for (int k = 0; k < n; k++) {
        U[k][k] = std::sqrt(U[k][k]);
        for (int s = k + 1; s < n; s++)
            if (in[k][s]) U[k][s] /= U[k][k];
        for (int r = k + 1; r < n; r++) {
            if (!in[k][r]) continue;
            for (int s = r; s < n; s++)
                if (in[k][s] && in[r][s]) U[r][s] -= U[k][r] * U[k][s];
        }
    }
    return U;
}

template <typename Value>
void compare_factor(Checks& c, const Mat& K, int n, double tol)
{
    itl::pc::ic_0<Mat, Value> P(K);
    const Dense D = dense_ic0(K, n);
    const auto U = P.get_U();
    for (int r = 0; r < n; r++)
        for (int s = r; s < n; s++) {
            const double want = r == s ? 1.0 / D[r][r] : D[r][s];
            c.near(double(U[r][s]), want, tol * std::fabs(want) + (want == 0.0 ? 0.0 : 1e-300), "U entry");
        }
    std::printf("factor (%s storage) compared on %d x %d entries\n", sizeof(Value) == 8 ? "double" : "float", n, n);
}

}  // namespace

int main()
{
    Checks c;
    const int nx = 4, ny = 3, n = nx * ny;
    Mat K(n, n);
    laplacian_setup(K, nx, ny);
    compare_factor<double>(c, K, n, 1e-14);
    compare_factor<float>(c, K, n, 1e-6);

    itl::pc::ic_0<Mat> P(K);
    mtl::dense_vector<double> ones(n, 1.0), rhs(n), sol(n, 0.0);
    rhs = K * ones;
    itl::cyclic_iteration<double> ctl(rhs, n, 1.0e-6, 0.0, n);
    const int code = itl::cg(K, sol, rhs, P, ctl);
    mtl::dense_vector<double> res(rhs - K * sol);
    std::printf("CG + IC(0): code %d, %d iterations, true relres %.2e\n", code, ctl.iterations(),
                two_norm(res) / two_norm(rhs));
    c.expect(code == 0, "CG + IC(0) converges within n iterations");
    c.expect(two_norm(res) <= 2.0e-6 * two_norm(rhs), "true residual at the tolerance");
    return c.status();
}
