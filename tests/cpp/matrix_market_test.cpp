// Matrix Market input and output through the drop-in (matrix_market.hpp:44-335),
// on the reference's fixture files kept under tests/golden/matrix_market/:
//  * bcsstk01 (real symmetric, 48 x 48): size and entries stated in the file
//    (entry (1,1) = 2.8322685185200e+06, (5,1) mirrored to (1,5) = 1.0e+06);
//  * jgl009 (pattern general, 9 x 9, 50 entries): pattern entries read as 0,
//    through the io::matrix_market constructor and by assignment;
//  * dense_read (array format, 2 x 4): entry (r, c) = 1 + 4 r + c;
//  * a missing file throws mtl::file_not_found;
//  * bcsstk01 written with matrix_market_ostream and read back is unchanged,
//    and its product with ones equals its row sums.
#include <boost/numeric/mtl/mtl.hpp>
#include <cstdio>
#include <string>
#include <unistd.h>
#include "check.hpp"

namespace {

// the fixtures sit at tests/golden/matrix_market/, the program at tests/cpp/bin/
std::string fixture_path(const char* argv0, const char* file)
{
    std::string self(argv0);
    const std::size_t cut = self.find_last_of('/');
    const std::string bin = cut == std::string::npos ? std::string(".") : self.substr(0, cut);
    return bin + "/../../golden/matrix_market/" + file;
}

}  // namespace

int main(int, char** argv)
{
    using mtl::compressed2D;
    Checks c;
    compressed2D<double> S;
    mtl::io::matrix_market_istream(fixture_path(argv[0], "bcsstk01.mtx")) >> S;
    c.expect(num_rows(S) == 48 && num_cols(S) == 48, "bcsstk01 is 48 x 48");
    c.near(S[0][0], 2.8322685185200e+06, 0.0, "bcsstk01 (1,1)");
    c.near(S[4][0], 1.0e6, 0.0, "bcsstk01 (5,1)");
    c.near(S[0][4], 1.0e6, 0.0, "bcsstk01 (1,5) mirrored");

    compressed2D<double> P1(mtl::io::matrix_market(fixture_path(argv[0], "jgl009.mtx"))), P2;
    P2 = mtl::io::matrix_market(fixture_path(argv[0], "jgl009.mtx"));
    c.expect(num_rows(P1) == 9 && P1.nnz() == 50 && P2.nnz() == 50, "jgl009 is 9 x 9 with 50 entries");
    bool zeros = true;
    for (double v : P1.data()) zeros = zeros && v == 0.0;
    c.expect(zeros, "pattern entries read as 0");

    compressed2D<double> D;
    mtl::io::matrix_market_istream(fixture_path(argv[0], "dense_read.mtx")) >> D;
    for (int r = 0; r < 2; r++)
        for (int col = 0; col < 4; col++) c.near(D[r][col], 1.0 + 4 * r + col, 0.0, "dense_read entry");

    bool thrown = false;
    try {
        compressed2D<double> X;
        X = mtl::io::matrix_market("no_such_matrix_file.mtx");
    } catch (const mtl::file_not_found& e) {
        std::printf("missing file reported: %s\n", e.what());
        thrown = true;
    }
    c.expect(thrown, "a missing file throws file_not_found");

    const std::string tmp = "/tmp/m4g_mm_round_trip_" + std::to_string(long(getpid())) + ".mtx";
    mtl::io::matrix_market_ostream(tmp) << S;
    compressed2D<double> R{mtl::io::matrix_market(tmp)};
    std::remove(tmp.c_str());
    c.expect(R.ref_major() == S.ref_major() && R.ref_minor() == S.ref_minor() && R.data() == S.data(),
             "write + read round trip is exact");

    mtl::dense_vector<double> ones(48, 1.0), y(48);
    y = S * ones;
    for (int r = 0; r < 48; r++) {
        double sum = 0.0;
        for (int col = 0; col < 48; col++) sum += S[r][col];
        c.near(y[r], sum, 1e-9 * std::fabs(sum) + 1e-9, "product with ones = row sum");
    }
    return c.status();
}
