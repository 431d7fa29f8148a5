// Same checks as T-mtl/compressed2D_pointer_constructor_test.cpp: a
// compressed2D built from raw starts / indices / data copies its input.
#include <boost/numeric/mtl/mtl.hpp>
#include <cstddef>

int main()
{
    std::size_t starts[] = {0, 2, 4, 6};
    std::size_t indices[] = {0, 1, 0, 1, 1, 2};
    double data[] = {1, 2, 3, 4, 5, 6};
    mtl::compressed2D<double> A(3, 3, 6, starts, indices, data);
    data[0] = 100;   // no aliasing
    MTL_THROW_IF(A[0][0] != 1 || A[1][0] != 3 || A[2][2] != 6 || A[0][2] != 0, mtl::runtime_error("wrong entries"));
    mtl::dense_vector<double> x(3), y(3);
    x = 1.0, 2.0, 3.0;
    y = A * x;
    MTL_THROW_IF(y[0] != 5 || y[1] != 11 || y[2] != 28, mtl::runtime_error("wrong product"));
    MTL_THROW_IF(A.nnz() != 6 || num_rows(A) != 3, mtl::runtime_error("wrong sizes"));
    return 0;
}
