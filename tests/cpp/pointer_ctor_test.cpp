// compressed2D(m, n, nnz, starts, indices, data) copies three caller arrays
// (compressed2D.hpp:355-364): later writes to the caller's arrays must not
// reach the matrix.  Checked for std::size_t (the default size_type) and
// unsigned indices on a 4x4 matrix with an empty row and an off-diagonal-only
// row; element access and the product are compared with the arrays.
#include <boost/numeric/mtl/mtl.hpp>
#include <cstddef>
#include "check.hpp"

namespace {

template <typename Index>
void run(Checks& c)
{
    // rows: {0: 3 at 0, 1 at 3}, {1: -2 at 1}, {2: empty}, {3: 5 at 0, 7 at 2}
    Index starts[] = {0, 2, 3, 3, 5};
    Index cols[] = {0, 3, 1, 0, 2};
    double vals[] = {3.0, 1.0, -2.0, 5.0, 7.0};
    mtl::compressed2D<double> M(4, 4, 5, starts, cols, vals);
    vals[0] = -100.0;
    cols[4] = 3;
    c.expect(num_rows(M) == 4 && num_cols(M) == 4 && M.nnz() == 5, "dimensions");
    c.near(M[0][0], 3.0, 0.0, "M[0][0] kept its value");
    c.near(M[3][2], 7.0, 0.0, "M[3][2] kept its column");
    c.near(M[3][3], 0.0, 0.0, "M[3][3] absent");
    c.near(M[2][1], 0.0, 0.0, "empty row");
    mtl::dense_vector<double> x(4), y(4);
    for (int i = 0; i < 4; i++) x[i] = 1.0 + i;
    y = M * x;
    const double expect[] = {3.0 * 1 + 1.0 * 4, -2.0 * 2, 0.0, 5.0 * 1 + 7.0 * 3};
    for (int i = 0; i < 4; i++) c.near(y[i], expect[i], 0.0, "M * x");
}

}  // namespace

int main()
{
    Checks c;
    run<std::size_t>(c);
    run<unsigned>(c);
    return c.status();
}
