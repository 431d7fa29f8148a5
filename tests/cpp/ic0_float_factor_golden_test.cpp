// IC(0) with factors stored in float (ic_0<compressed2D<double>, float>)
// applied to v = (0, 1, ..., 8) on the 3x3-grid Laplacian.  The expected
// values are the reference's golden vector (T-itl/ic_0_test.cpp:29), given
// there to 6 significant digits, hence the 1e-4 bound on the 2-norm of the
// difference.  IC(0) is symmetric, so adjoint_solve must give the same vector.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cmath>
#include <cstdio>

int main()
{
    using Mat = mtl::compressed2D<double>;
    const int side = 3, n = side * side;
    Mat K(n, n);
    laplacian_setup(K, side, side);
    itl::pc::ic_0<Mat, float> prec(K);

    const double golden[n] = {1.03194, 1.60198, 1.45357, 2.5258, 3.55386, 3.16, 2.91787, 4.09338, 3.81335};
    mtl::dense_vector<double> v(n), forward(n), backward(n);
    iota(v);
    forward = solve(prec, v);
    backward = adjoint_solve(prec, v);

    double err2 = 0.0, sym = 0.0;
    for (int i = 0; i < n; i++) {
        const double d = forward[i] - golden[i];
        err2 += d * d;
        sym = std::fmax(sym, std::fabs(forward[i] - backward[i]));
        std::printf("%d: %.6f (expected %.5f)\n", i, forward[i], golden[i]);
    }
    const double err = std::sqrt(err2);
    std::printf("|y - golden| = %.3e, max |solve - adjoint_solve| = %.3e\n", err, sym);
    if (err > 1.0e-4 || sym != 0.0) {
        std::printf("IC(0) application with float factors differs from the golden vector\n");
        return 1;
    }
    return 0;
}
