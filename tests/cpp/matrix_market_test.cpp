// The compressed2D<double> checks of T-mtl/matrix_market_test.cpp: files of
// the reference's matrix_market/ directory read with matrix_market_istream,
// through the io::matrix_market constructor and assignment, the
// file_not_found failure, and the dense (array) file read of read_test; plus
// a write / read round trip with matrix_market_ostream.
#include <boost/numeric/mtl/mtl.hpp>
#include <cstdio>
#include <iostream>
#include <string>
#include <unistd.h>

static std::string fixture(const char* argv0, const char* name)
{
    // tests/cpp/bin/<program> -> tests/golden/matrix_market/<name>
    std::string p(argv0);
    const std::size_t slash = p.rfind('/');
    const std::string dir = slash == std::string::npos ? std::string(".") : p.substr(0, slash);
    return dir + "/../../golden/matrix_market/" + name;
}

int main(int, char* argv[])
{
    using namespace mtl;
    compressed2D<double> A;
    io::matrix_market_istream(fixture(argv[0], "bcsstk01.mtx")) >> A;   // real symmetric
    MTL_THROW_IF(num_rows(A) != 48 || num_cols(A) != 48, runtime_error("bcsstk01 size"));
    MTL_THROW_IF(A[4][0] != 1.0e6 || A[0][4] != 1.0e6 || A[0][0] != 2.8322685185200e+06,
                 runtime_error("bcsstk01 entries"));

    compressed2D<double> B(io::matrix_market(fixture(argv[0], "jgl009.mtx"))), C;   // general pattern
    C = io::matrix_market(fixture(argv[0], "jgl009.mtx"));
    MTL_THROW_IF(num_rows(B) != 9 || B.nnz() != 50 || C.nnz() != 50, runtime_error("jgl009 size"));
    for (double v : B.data()) MTL_THROW_IF(v != 0.0, runtime_error("pattern entries must be 0"));

    bool caught = false;   // failure_test
    try {
        A = io::matrix_market("File_not_exist_test.mtx");
    } catch (const file_not_found& e) {
        std::cerr << "Successfully caught exception for inexistant file. Error message is:\n" << e.what();
        caught = true;
    }
    MTL_THROW_IF(!caught, runtime_error("No exception thrown for inexistant file."));

    compressed2D<double> D;   // read_test: data(r, c) = 1 + 4 r + c, array file
    io::matrix_market_istream(fixture(argv[0], "dense_read.mtx")) >> D;
    for (std::size_t r = 0; r < 2; r++)
        for (std::size_t c = 0; c < 4; c++)
            MTL_THROW_IF(D[r][c] != double(1 + 4 * r + c), runtime_error("could not read the dense matrix"));

    const std::string out = "/tmp/m4g_matrix_market_test_" + std::to_string(long(getpid())) + ".mtx";
    io::matrix_market_ostream(out) << A;
    compressed2D<double> E{io::matrix_market(out)};
    std::remove(out.c_str());
    MTL_THROW_IF(E.ref_major() != A.ref_major() || E.ref_minor() != A.ref_minor() || E.data() != A.data(),
                 runtime_error("round trip"));

    // the matrix works on the GPU path: A * x against the host entries
    dense_vector<double> x(48, 1.0), y(48);
    y = A * x;
    double s0 = 0;
    for (std::size_t c = 0; c < 48; c++) s0 += A[0][c];
    MTL_THROW_IF(y[0] != s0, runtime_error("product"));
    std::cout << "matrix_market_test passed\n";
    return 0;
}
