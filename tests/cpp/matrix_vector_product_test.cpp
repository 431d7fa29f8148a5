// Same checks as T-mtl/matrix_vector_product_test.cpp:56-89: the 5x5-grid
// Laplacian times its column 12 gives the stencil 20/-8/2/1 exactly; += and -=.
#include <boost/numeric/mtl/mtl.hpp>

int main()
{
    mtl::compressed2D<double> A;
    laplacian_setup(A, 5, 5);
    mtl::dense_vector<double> v(25), w(25);
    for (unsigned i = 0; i < num_cols(A); i++) v[i] = A[12][i];
    w = A * v;
    MTL_THROW_IF(w[12] != 20.0, mtl::runtime_error("wrong diagonal"));
    MTL_THROW_IF(w[13] != -8.0, mtl::runtime_error("wrong east neighbor"));
    MTL_THROW_IF(w[14] != 1.0, mtl::runtime_error("wrong east east neighbor"));
    MTL_THROW_IF(w[15] != 0.0, mtl::runtime_error("wrong zero-element"));
    MTL_THROW_IF(w[17] != -8.0, mtl::runtime_error("wrong south neighbor"));
    MTL_THROW_IF(w[18] != 2.0, mtl::runtime_error("wrong south east neighbor"));
    MTL_THROW_IF(w[22] != 1.0, mtl::runtime_error("wrong south south neighbor"));
    w += A * v;
    MTL_THROW_IF(w[12] != 40.0 || w[18] != 4.0, mtl::runtime_error("wrong +="));
    w -= A * v;
    MTL_THROW_IF(w[12] != 20.0 || w[18] != 2.0, mtl::runtime_error("wrong -="));
    return 0;
}
