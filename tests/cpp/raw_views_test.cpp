// compressed2D's raw views (compressed2D.hpp:609-625): reading them gives the
// CSR arrays; writing a value through address_data() changes the product
// (the device copy is rebuilt on next use).
#include <boost/numeric/mtl/mtl.hpp>
#include <iostream>

int main()
{
    mtl::compressed2D<double> A;
    laplacian_setup(A, 3, 3);
    const mtl::compressed2D<double>& C = A;
    MTL_THROW_IF(C.address_major()[0] != 0 || C.address_major()[9] != int64_t(A.nnz()),
                 mtl::runtime_error("address_major"));
    MTL_THROW_IF(C.address_minor()[0] != 0 || C.address_data()[0] != 4.0, mtl::runtime_error("address_minor/data"));
    MTL_THROW_IF(C.ref_major().size() != 10 || C.ref_minor().size() != A.nnz(), mtl::runtime_error("ref_major/minor"));
    mtl::dense_vector<double> x(9, 1.0), y(9);
    y = A * x;
    MTL_THROW_IF(y[0] != 2.0, mtl::runtime_error("A * x before"));
    A.address_data()[0] = 10.0;   // A[0][0] = 10
    y = A * x;
    MTL_THROW_IF(y[0] != 8.0 || y[1] != 1.0, mtl::runtime_error("A * x after writing through address_data"));
    std::cout << "raw views ok\n";
    return 0;
}
