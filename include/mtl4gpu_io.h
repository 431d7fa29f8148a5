/* Matrix Market files for compressed2D<double> (host only; no GPU needed).
 *
 * Reference: mtl::io::matrix_market_istream and matrix_market_ostream
 * (boost/numeric/mtl/io/matrix_market.hpp:44-217, 219-335).
 */
#ifndef MTL4GPU_IO_H
#define MTL4GPU_IO_H

#include <stdint.h>

#include "mtl4gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Reads the file at `path` as `matrix_market_istream(path) >> A` fills a
 * compressed2D<double> (matrix_market.hpp:166-217): coordinate or array
 * files; real, integer or pattern values (pattern entries are 0; complex
 * files are an error for real matrices, :155); general, symmetric,
 * skew-symmetric or hermitian (entries off the diagonal mirrored, negated
 * for skew-symmetric, :128-141); of several entries at one position the last
 * one read is kept (update_store inserter).  Output is CSR with ascending
 * columns in each row, int64 indices.
 * Call once with starts == NULL: *nrows, *ncols, *nnz receive the sizes.
 * Call again with starts [nrows+1], indices and values [*nnz] (*nnz = their
 * capacity on entry) to receive the matrix.
 * Returns M4G_ERR_FILE if the file cannot be opened (mtl::file_not_found),
 * M4G_ERR_FORMAT for a malformed file (mtl::runtime_error). */
int m4g_mm_read(const char* path, int64_t* nrows, int64_t* ncols, int64_t* nnz, int64_t* starts, int64_t* indices,
                double* values);

/* Writes CSR arrays as `matrix_market_ostream(path) << A` writes a
 * compressed2D<double> (matrix_market.hpp:235-256, 319-325): header
 * "%%MatrixMarket  matrix coordinate real general", "nrows ncols nnz", then
 * every stored entry row by row, 1-based, value with 16 digits after the
 * point in scientific notation. */
int m4g_mm_write(const char* path, int64_t nrows, int64_t ncols, const int64_t* starts, const int64_t* indices,
                 const double* values);

#ifdef __cplusplus
}
#endif

#endif /* MTL4GPU_IO_H */
