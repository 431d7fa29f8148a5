// w = A * v with v = row 12 (the centre) of the 5x5-grid Laplacian, i.e. the
// centre row of A*A: the squared 5-point stencil 20 / -8 / 2 / 1 (the known
// answer of the reference's matrix_vector_product_test, T-mtl/, lines 56-89),
// checked on all 25 entries, then w += A * v doubles it and w -= A * v
// restores it.  Products of integers are exact, so equality is required.
#include <boost/numeric/mtl/mtl.hpp>
#include <cstdlib>
#include "check.hpp"

namespace {

// entry (dx, dy) of the squared 5-point stencil
double squared_stencil(int dx, int dy)
{
    const int ax = std::abs(dx), ay = std::abs(dy);
    if (ax == 0 && ay == 0) return 20.0;
    if (ax + ay == 1) return -8.0;
    if (ax == 1 && ay == 1) return 2.0;
    if (ax + ay == 2) return 1.0;   // two steps along one axis
    return 0.0;
}

}  // namespace

int main()
{
    Checks c;
    const int side = 5, centre = 12;
    mtl::compressed2D<double> K;
    laplacian_setup(K, side, side);
    mtl::dense_vector<double> v(side * side), w(side * side);
    for (int j = 0; j < side * side; j++) v[j] = K[centre][j];
    w = K * v;
    for (int k = 0; k < side * side; k++)
        c.near(w[k], squared_stencil(k % side - centre % side, k / side - centre / side), 0.0, "A * v");
    w += K * v;
    for (int k = 0; k < side * side; k++)
        c.near(w[k], 2 * squared_stencil(k % side - centre % side, k / side - centre / side), 0.0, "w += A * v");
    w -= K * v;
    for (int k = 0; k < side * side; k++)
        c.near(w[k], squared_stencil(k % side - centre % side, k / side - centre / side), 0.0, "w -= A * v");
    return c.status();
}
