// Operators without a stored matrix (the operator protocol of
// mat_cvec_multiplier.hpp:29-88 and the library's poisson2D_dirichlet,
// poisson2D_dirichlet.hpp:24-102).
//  * A user type whose operator* returns the product vector, written here as
//    a loop over grid cells and their in-grid neighbours, and
//    mtl::mat::poisson2D_dirichlet, must both give the Laplacian's product
//    (the reference's matrix_free_1_test / matrix_free_2_test compare the same
//    two with laplacian_setup); the multiplier's += and -= as well.
//  * itl::cg with the identity on poisson2D_dirichlet reaches 1e-11 and
//    x = ones (examples/matrix_free_cg.cpp); cg on the user type takes the
//    same number of iterations.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/itl/itl.hpp>
#include <cmath>
#include "check.hpp"

namespace {

// 5-point Laplacian on an ny x nx grid, Dirichlet boundary, row = y * nx + x
struct grid_operator {
    grid_operator(int ny, int nx) : ny(ny), nx(nx) {}
    template <typename Vector>
    Vector operator*(const Vector& v) const
    {
        Vector w(ny * nx);
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                const int r = y * nx + x;
                double acc = 4.0 * v[r];
                if (y > 0) acc -= v[r - nx];
                if (x > 0) acc -= v[r - 1];
                if (x + 1 < nx) acc -= v[r + 1];
                if (y + 1 < ny) acc -= v[r + nx];
                w[r] = acc;
            }
        return w;
    }
    int ny, nx;
};

}  // namespace

namespace mtl { namespace ashape {
template <> struct ashape_aux<grid_operator> { typedef nonscal type; };
}}  // namespace mtl::ashape

int main()
{
    using Vec = mtl::dense_vector<double>;
    Checks c;
    {
        const int ny = 4, nx = 5, n = ny * nx;
        mtl::compressed2D<double> K;
        laplacian_setup(K, ny, nx);
        Vec v(n);
        for (int i = 0; i < n; i++) v[i] = 0.5 * i - 3.0;
        Vec ref(K * v);
        grid_operator G(ny, nx);
        Vec wg;
        wg = G * v;
        mtl::mat::poisson2D_dirichlet P(ny, nx);
        Vec wp(n);
        wp = P * v;
        double dg = 0.0, dp = 0.0;
        for (int i = 0; i < n; i++) {
            dg = std::fmax(dg, std::fabs(wg[i] - ref[i]));
            dp = std::fmax(dp, std::fabs(wp[i] - ref[i]));
        }
        c.expect(dg == 0.0, "user operator's product equals the Laplacian's");
        c.expect(dp == 0.0, "poisson2D_dirichlet's product equals the Laplacian's");
        wp += P * v;
        for (int i = 0; i < n; i++) c.near(wp[i], 2 * ref[i], 0.0, "w += P * v");
        wp -= P * v;
        for (int i = 0; i < n; i++) c.near(wp[i], ref[i], 0.0, "w -= P * v");
    }
    {
        const int side = 10, n = side * side;
        mtl::mat::poisson2D_dirichlet P(side, side);
        itl::pc::identity<mtl::mat::poisson2D_dirichlet> none(P);
        Vec ones(n, 1.0), b(n), x(n, 0.0);
        b = P * ones;
        itl::cyclic_iteration<double> ctl(b, 100, 1.e-11, 0.0, 100);
        const int code = itl::cg(P, x, b, none, ctl);
        double err = 0.0;
        for (int i = 0; i < n; i++) err = std::fmax(err, std::fabs(x[i] - 1.0));
        std::printf("matrix-free cg: code %d, %d iterations, max |x - 1| = %.1e\n", code, ctl.iterations(), err);
        c.expect(code == 0 && err <= 1e-9, "matrix-free cg solves to x = ones");
        grid_operator G(side, side);
        Vec y(n, 0.0);
        itl::basic_iteration<double> ctl2(b, 100, 1.e-11);
        c.expect(itl::cg(G, y, b, ctl2) == 0, "cg on the user operator converges");
        c.expect(ctl2.iterations() == ctl.iterations(), "same iteration count on both operators");
    }
    return c.status();
}
