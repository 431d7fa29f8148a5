// Matrix Market ingest and output for compressed2D<double>: the reading and
// writing of mtl::io::matrix_market_istream / matrix_market_ostream
// (mtl/io/matrix_market.hpp:44-217, 219-335) restated on host CSR arrays.
//
// Reading (matrix_market_istream::read, :166-217):
//  * header tokens `%%MatrixMarket matrix <sparsity> <value> <symmetry>`;
//    sparsity and symmetry compared case-insensitively, the value tag
//    case-sensitively ("real", "integer", "complex", "pattern");
//  * comment lines start with '%' (leading blanks skipped);
//  * coordinate: nrows ncols nnz, then nnz entries "r c [value]" (1-based);
//    array: nrows ncols, then all nrows*ncols values column by column;
//  * each entry is inserted at (r, c) and, when r != c, mirrored at (c, r)
//    for symmetric (same value), skew-symmetric (negated) and hermitian
//    (conjugated: the same for real values) files (:128-141);
//  * pattern entries have value 0 (which_value(pattern_type), :153);
//  * insertion overwrites (the inserter's default update_store): of several
//    entries at one position the last one read is kept;
//  * the result is compressed2D: rows with ascending columns, every inserted
//    position stored (zeros included).
// Writing (write_sparse_matrix, :241-256): "%%MatrixMarket  matrix coordinate
// real general", "nrows ncols nnz", then every stored entry row by row as
// "r c v" (1-based) with v in scientific notation, digits10 + 1 = 16 digits
// after the point.
#include "m4g_internal.hpp"
#include "../../include/mtl4gpu_io.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace m4g {
namespace {

struct Entry {
    int64_t r, c, seq;
    double v;
};

std::string lower(std::string s)
{
    for (char& ch : s) ch = char(std::tolower(static_cast<unsigned char>(ch)));
    return s;
}

struct MMData {
    int64_t nrows = 0, ncols = 0;
    std::vector<int64_t> starts, indices;
    std::vector<double> values;
};

MMData read_mm(const char* path)
{
    std::ifstream in(path);
    M4G_REQUIRE(in.good(), M4G_ERR_FILE,
                std::string("matrix_market_istream: Error in input stream!\nProbably file ") + path + " not found.\n");
    std::string marker, type, sparsity, value, symmetry;
    in >> marker >> type >> sparsity >> value >> symmetry;
    M4G_REQUIRE(marker == "%%MatrixMarket", M4G_ERR_FORMAT, "File not in Matrix Market format");
    M4G_REQUIRE(type == "matrix", M4G_ERR_FORMAT, "Try to read matrix from non-matrix file");
    symmetry = lower(symmetry);
    sparsity = lower(sparsity);
    enum { GENERAL, SYMMETRIC, SKEW, HERMITIAN } sym;
    if (symmetry == "general") sym = GENERAL;
    else if (symmetry == "symmetric") sym = SYMMETRIC;
    else if (symmetry == "skew-symmetric") sym = SKEW;
    else if (symmetry == "hermitian") sym = HERMITIAN;
    else throw Error(M4G_ERR_FORMAT, "Search string not found");
    M4G_REQUIRE(sparsity == "coordinate" || sparsity == "array", M4G_ERR_FORMAT, "Search string not found");
    const bool coordinate = sparsity == "coordinate";
    // comment lines: a '%' as the first non-blank character
    for (;;) {
        char first = 0;
        if (!(in >> first)) break;
        if (first == '%') {
            std::string rest;
            std::getline(in, rest);
        } else {
            in.putback(first);
            break;
        }
    }
    MMData m;
    int64_t nnz = 0;
    in >> m.nrows >> m.ncols;
    if (coordinate) in >> nnz;
    M4G_REQUIRE(in.good() && m.nrows >= 0 && m.ncols >= 0 && nnz >= 0, M4G_ERR_FORMAT,
                "matrix_market: bad size line");
    enum { REAL, INTEGER, PATTERN } vf;
    if (value == "real") vf = REAL;
    else if (value == "integer") vf = INTEGER;
    else if (value == "pattern") vf = PATTERN;
    else if (value == "complex") throw Error(M4G_ERR_FORMAT, "Cannot convert complex value in real\n");
    else throw Error(M4G_ERR_FORMAT, "Unknown tag for matrix value type in file");
    auto read_value = [&]() -> double {
        if (vf == PATTERN) return 0.0;
        if (vf == INTEGER) {
            long v = 0;
            in >> v;
            return double(v);
        }
        double v = 0;
        in >> v;
        return v;
    };
    std::vector<Entry> es;
    int64_t seq = 0;
    auto insert = [&](int64_t r, int64_t c, double v) {
        M4G_REQUIRE(r >= 0 && r < m.nrows && c >= 0 && c < m.ncols, M4G_ERR_FORMAT,
                    "matrix_market: entry outside the matrix");
        es.push_back({r, c, seq++, v});
        if (r != c) {
            if (sym == SYMMETRIC || sym == HERMITIAN) es.push_back({c, r, seq++, v});
            else if (sym == SKEW) es.push_back({c, r, seq++, -v});
        }
    };
    if (coordinate) {
        es.reserve(size_t(sym == GENERAL ? nnz : 2 * nnz));
        for (int64_t i = 0; i < nnz; i++) {
            int64_t r = 0, c = 0;
            in >> r >> c;
            if (!in) break;   // as the reference: a short file ends the entries
            const double v = read_value();
            insert(r - 1, c - 1, v);
        }
    } else {
        for (int64_t c = 0; c < m.ncols; c++)
            for (int64_t r = 0; r < m.nrows; r++) insert(r, c, read_value());
    }
    // update_store: the last entry at a position wins
    std::sort(es.begin(), es.end(), [](const Entry& a, const Entry& b) {
        return a.r != b.r ? a.r < b.r : a.c != b.c ? a.c < b.c : a.seq < b.seq;
    });
    m.starts.assign(size_t(m.nrows) + 1, 0);
    for (size_t k = 0; k < es.size(); k++) {
        if (k + 1 < es.size() && es[k + 1].r == es[k].r && es[k + 1].c == es[k].c) continue;
        m.indices.push_back(es[k].c);
        m.values.push_back(es[k].v);
        m.starts[size_t(es[k].r) + 1]++;
    }
    for (int64_t r = 0; r < m.nrows; r++) m.starts[size_t(r) + 1] += m.starts[size_t(r)];
    return m;
}

}  // namespace
}  // namespace m4g

using namespace m4g;

extern "C" {

int m4g_mm_read(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nnz, int64_t* starts, int64_t* indices,
                double* values)
{
    return guard([&] {
        M4G_REQUIRE(path && nrows && ncols && nnz, M4G_ERR_INVALID, "m4g_mm_read: null argument");
        MMData m = read_mm(path);
        const int64_t cap = *nnz;
        const bool fill = starts != nullptr;
        *nrows = m.nrows;
        *ncols = m.ncols;
        *nnz = int64_t(m.values.size());
        if (fill) {
            M4G_REQUIRE(cap >= int64_t(m.values.size()) && (m.values.empty() || (indices && values)), M4G_ERR_SIZE,
                        "m4g_mm_read: arrays too small");
            std::copy(m.starts.begin(), m.starts.end(), starts);
            std::copy(m.indices.begin(), m.indices.end(), indices);
            std::copy(m.values.begin(), m.values.end(), values);
        }
        return M4G_OK;
    });
}

int m4g_mm_write(const char* path, int64_t nrows, int64_t ncols, const int64_t* starts, const int64_t* indices,
                 const double* values)
{
    return guard([&] {
        M4G_REQUIRE(path && starts && nrows >= 0 && ncols >= 0, M4G_ERR_INVALID, "m4g_mm_write: bad argument");
        const int64_t nnz = starts[nrows];
        M4G_REQUIRE(nnz == 0 || (indices && values), M4G_ERR_INVALID, "m4g_mm_write: null arrays");
        FILE* f = std::fopen(path, "w");
        M4G_REQUIRE(f, M4G_ERR_FILE, std::string("m4g_mm_write: cannot open ") + path);
        std::fprintf(f, "%%%%MatrixMarket  matrix coordinate real general\n");
        std::fprintf(f, "%lld %lld %lld\n", (long long)nrows, (long long)ncols, (long long)nnz);
        for (int64_t r = 0; r < nrows; r++)
            for (int64_t e = starts[r]; e < starts[r + 1]; e++)
                std::fprintf(f, "%lld %lld %.16e\n", (long long)(r + 1), (long long)(indices[e] + 1), values[e]);
        const bool ok = std::fclose(f) == 0;
        M4G_REQUIRE(ok, M4G_ERR_FILE, std::string("m4g_mm_write: cannot write ") + path);
        return M4G_OK;
    });
}

}  // extern "C"
