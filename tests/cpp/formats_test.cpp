// T-mtl/ell_matrix_test.cpp and T-mtl/sparse_banded_matrix_test.cpp through
// the drop-in headers: inserter-built ell_matrix and sparse_banded matrices,
// element access, and products against the compressed2D product.
#include <boost/numeric/mtl/mtl.hpp>
#include <boost/numeric/mtl/matrix/ell_matrix.hpp>
#include <boost/numeric/mtl/matrix/sparse_banded.hpp>
#include <cmath>
#include <iostream>

template <typename Matrix>
void fill_matrix(Matrix& A)
{
    mtl::mat::inserter<Matrix> ins(A, 3);
    ins[0][0] << 2;
    ins[0][1] << 9;
    ins[1][1] << 1;
    ins[1][2] << 5;
    ins[1][3] << 5;
    ins[1][4] << 1;
    ins[2][2] << 6;
    ins[2][3] << 9;
    ins[3][2] << 2;
    ins[3][3] << 4;
    ins[4][0] << 7;
    ins[4][4] << 3;
}

template <typename Matrix>
void rectangle_test(Matrix& A)
{
    mtl::mat::inserter<Matrix> ins(A);
    int i = 1;
    unsigned nc = num_cols(A);
    for (unsigned r = 0; r < num_rows(A); r++) {
        if (r < nc - 4) ins(r, r + 4) << i++;
        if (r < nc) ins(r, r) << i++;
        if (r >= 2 && r < nc + 2) ins(r, r - 2) << i++;
        if (r >= 4 && r < nc + 4) ins[r][r - 4] << i++;
    }
}

template <typename Matrix>
void mat_vec_mult_test(const Matrix& A)
{
    mtl::dense_vector<double> v(num_rows(A)), w(num_cols(A), 3.0), v2(num_rows(A));
    v = A * w;
    mtl::compressed2D<double> B(A);
    v2 = B * w;
    v2 -= v;
    MTL_THROW_IF(two_norm(v2) > 0.001, mtl::runtime_error("wrong result for sparse banded times vector"));
}

int main()
{
    {   // ell_matrix_test.cpp
        typedef mtl::mat::ell_matrix<double> matrix_type;
        matrix_type A(5, 5);
        fill_matrix(A);
        MTL_THROW_IF(A[2][3] != 9.0, mtl::runtime_error("A[2][3]"));
        MTL_THROW_IF(A[2][4] != 0.0, mtl::runtime_error("A[2][4]"));
        MTL_THROW_IF(A[2][0] != 0.0, mtl::runtime_error("A[2][0]"));
        mtl::compressed2D<double> B(5, 5);
        fill_matrix(B);
        MTL_THROW_IF(A.nnz() != B.nnz(), mtl::runtime_error("nnz"));
        MTL_THROW_IF(A.slots() != 4 || A.stride() != 32, mtl::runtime_error("slot layout"));
        mtl::dense_vector<double> res(5), res2(5), x(5);
        for (int i = 0; i < 5; i++) x[i] = i + 1;   // iota(x, 1)
        res2 = B * x;
        res = A * x;
        MTL_THROW_IF(res[0] != 20 || res[1] != 42 || res[2] != 54 || res[3] != 22 || res[4] != 22,
                     mtl::runtime_error("A * x"));
        res2 -= res;
        MTL_THROW_IF(two_norm(res2) > 0.001, mtl::runtime_error("ell product"));
        matrix_type C;
        laplacian_setup(C, 3, 4);
        MTL_THROW_IF(C.num_rows() != 12 || C[0][0] != 4.0 || C[0][1] != -1.0, mtl::runtime_error("laplacian ell"));
    }
    {   // sparse_banded_matrix_test.cpp
        mtl::mat::sparse_banded<double> dr, dr2(6, 11), dr3(11, 6), dr4(6, 5);
        laplacian_setup(dr, 3, 4);
        MTL_THROW_IF(dr[0][0] != 4.0 || dr[0][1] != -1.0 || dr[0][4] != -1.0 || dr[0][2] != 0.0 ||
                         dr[1][0] != -1.0 || dr[4][0] != -1.0 || dr[2][0] != 0.0,
                     mtl::runtime_error("laplacian banded"));
        MTL_THROW_IF(dr.ref_bands().size() != 5 || dr.nnz() != 60, mtl::runtime_error("bands"));
        rectangle_test(dr2);
        rectangle_test(dr3);
        rectangle_test(dr4);
        mtl::compressed2D<double> C;
        laplacian_setup(C, 3, 4);
        mtl::mat::sparse_banded<double> D;
        D = C;
        mtl::compressed2D<double> E;
        E = D;
        MTL_THROW_IF(E.nnz() != C.nnz(), mtl::runtime_error("round trip"));
        mtl::mat::sparse_banded<double> dr5(5, 5), dr6(5, 5);
        {
            mtl::mat::inserter<mtl::mat::sparse_banded<double>> ins5(dr5), ins6(dr6);
            ins5[2][0] << 1; ins5[3][1] << 2; ins5[4][2] << 3; ins5[4][0] << 4;
            ins6[0][2] << 1; ins6[1][3] << 2; ins6[2][4] << 3; ins6[0][4] << 4;
        }
        mat_vec_mult_test(dr2);
        mat_vec_mult_test(dr3);
        mat_vec_mult_test(dr4);
        mat_vec_mult_test(dr5);
        mat_vec_mult_test(dr6);
    }
    std::cout << "formats ok\n";
    return 0;
}
