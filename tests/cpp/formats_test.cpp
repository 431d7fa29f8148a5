// ell_matrix and sparse_banded through the drop-in headers.
//  * ell_matrix: a 5 x 5 matrix given as triplets (the matrix and the
//    product with x = 1..5, [20, 42, 54, 22, 22], are the known answers of
//    the reference's ell_matrix_test, T-mtl/), its element access, slot
//    layout (4 slots, stride 32 = rows rounded up to 32, ell_matrix.hpp:127-177)
//    and product; the 3 x 4-grid Laplacian as ell_matrix.
//  * sparse_banded: the 3 x 4-grid Laplacian (5 bands, 60 entries), band
//    patterns on square, wide and tall matrices (offsets -4, -2, 0, +4) and
//    lower-only / upper-only matrices; every product equals the product of
//    the same entries in a compressed2D, and compressed2D <-> sparse_banded
//    conversions keep the entries.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/mtl/matrix/ell_matrix.hpp>
#include <boost/numeric/mtl/matrix/sparse_banded.hpp>
#include <cmath>
#include <vector>
#include "check.hpp"

namespace {

struct Entry {
    unsigned r, c;
    double v;
};

const std::vector<Entry> five_by_five = {{0, 0, 2}, {0, 1, 9}, {1, 1, 1}, {1, 2, 5}, {1, 3, 5}, {1, 4, 1},
                                         {2, 2, 6}, {2, 3, 9}, {3, 2, 2}, {3, 3, 4}, {4, 0, 7}, {4, 4, 3}};

template <typename Matrix>
void insert_all(Matrix& M, const std::vector<Entry>& es)
{
    mtl::mat::inserter<Matrix> ins(M, 4);
    for (const Entry& e : es) ins[e.r][e.c] << e.v;
}

// entries on the given diagonals (offset = column - row) of an m x n matrix
std::vector<Entry> band_entries(unsigned m, unsigned n, const std::vector<int>& offsets)
{
    std::vector<Entry> es;
    for (unsigned r = 0; r < m; r++)
        for (int off : offsets) {
            const long c = long(r) + off;
            if (c >= 0 && c < long(n)) es.push_back({r, unsigned(c), 1.0 + 0.25 * r + 0.5 * (off + 4)});
        }
    return es;
}

// product of M with x = 1, 2, ... against the compressed2D of the same entries
template <typename Matrix>
void same_product(Checks& c, const Matrix& M, unsigned m, unsigned n, const std::vector<Entry>& es, const char* what)
{
    mtl::compressed2D<double> C(m, n);
    insert_all(C, es);
    mtl::dense_vector<double> x(n), y(m), z(m);
    for (unsigned i = 0; i < n; i++) x[i] = 1.0 + i;
    y = M * x;
    z = C * x;
    double d = 0.0;
    for (unsigned i = 0; i < m; i++) d = std::fmax(d, std::fabs(y[i] - z[i]));
    c.expect(d == 0.0, what);
}

}  // namespace

int main()
{
    Checks c;
    {
        mtl::mat::ell_matrix<double> E(5, 5);
        insert_all(E, five_by_five);
        c.near(E[2][3], 9.0, 0.0, "ell E[2][3]");
        c.near(E[2][4], 0.0, 0.0, "ell E[2][4] absent");
        c.near(E[4][0], 7.0, 0.0, "ell E[4][0]");
        c.expect(E.nnz() == five_by_five.size(), "ell nnz");
        c.expect(E.slots() == 4 && E.stride() == 32, "ell slot layout");
        mtl::dense_vector<double> x(5), y(5);
        for (int i = 0; i < 5; i++) x[i] = i + 1;
        y = E * x;
        const double want[] = {20, 42, 54, 22, 22};
        for (int i = 0; i < 5; i++) c.near(y[i], want[i], 0.0, "ell product");
        same_product(c, E, 5, 5, five_by_five, "ell product equals the compressed2D product");
        mtl::mat::ell_matrix<double> L;
        laplacian_setup(L, 3, 4);
        c.expect(L.num_rows() == 12, "ell Laplacian size");
        c.near(L[0][0], 4.0, 0.0, "ell Laplacian diagonal");
        c.near(L[0][1], -1.0, 0.0, "ell Laplacian neighbour");
    }
    {
        mtl::mat::sparse_banded<double> L;
        laplacian_setup(L, 3, 4);
        c.expect(L.ref_bands().size() == 5 && L.nnz() == 60, "banded Laplacian: 5 bands, 60 entries");
        c.near(L[0][0], 4.0, 0.0, "banded diagonal");
        c.near(L[1][0], -1.0, 0.0, "banded west");
        c.near(L[0][4], -1.0, 0.0, "banded south");
        c.near(L[0][2], 0.0, 0.0, "banded absent");
        mtl::compressed2D<double> C;
        laplacian_setup(C, 3, 4);
        mtl::mat::sparse_banded<double> B;
        B = C;
        mtl::compressed2D<double> back;
        back = B;
        c.expect(back.nnz() == C.nnz() && back.data() == C.data(), "compressed2D -> banded -> compressed2D");
        const std::vector<int> offs = {-4, -2, 0, 4};
        const unsigned shapes[3][2] = {{6, 11}, {11, 6}, {6, 5}};
        for (const auto& sh : shapes) {
            mtl::mat::sparse_banded<double> M(sh[0], sh[1]);
            const auto es = band_entries(sh[0], sh[1], offs);
            insert_all(M, es);
            same_product(c, M, sh[0], sh[1], es, "banded product on a rectangle");
        }
        const std::vector<Entry> lower = {{2, 0, 1}, {3, 1, 2}, {4, 2, 3}, {4, 0, 4}};
        const std::vector<Entry> upper = {{0, 2, 1}, {1, 3, 2}, {2, 4, 3}, {0, 4, 4}};
        for (const auto* es : {&lower, &upper}) {
            mtl::mat::sparse_banded<double> M(5, 5);
            insert_all(M, *es);
            same_product(c, M, 5, 5, *es, "banded product, one-sided bands");
        }
    }
    return c.status();
}
