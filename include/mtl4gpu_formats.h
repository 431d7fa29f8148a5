/* Alternative sparse formats on the device: mtl::mat::ell_matrix and
 * mtl::mat::sparse_banded, with their matrix-vector products.
 *
 * Reference: mtl/matrix/ell_matrix.hpp:31-177 (layout and inserter),
 * mtl/matrix/sparse_banded.hpp:38-258 (layout and inserter), and their
 * products in mtl/operation/mat_vec_mult.hpp:692-809.  Values are double; `mode` is
 * M4G_ASSIGN (w = A v), M4G_PLUS (w += A v) or M4G_MINUS (w -= A v), the
 * Assign::first_update of the reference.
 */
#ifndef MTL4GPU_FORMATS_H
#define MTL4GPU_FORMATS_H

#include "mtl4gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ell_matrix --------------------------------------------------------- */
typedef struct m4g_ell_s* m4g_ell;

/* ell_matrix_inserter (ell_matrix.hpp:127-177) on a CSR with sorted rows:
 * slots = longest row, stride = num_rows rounded up to 32, entry i of row r at
 * r + i * stride; a row's padding repeats its last column index (0 for an
 * empty row) with value 0.  Column indices are kept as 32-bit on the device. */
int m4g_ell_create(m4g_ell* E, int64_t nrows, int64_t ncols, const int64_t* starts, const int64_t* indices,
                   const double* values);
int m4g_ell_destroy(m4g_ell E);
/* num_rows, num_cols, slots(), stride(), nnz() (the CSR's entry count) */
int m4g_ell_shape(m4g_ell E, int64_t* nrows, int64_t* ncols, int64_t* slots, int64_t* stride, int64_t* nnz);
/* ref_minor() and ref_data(): stride * slots entries each */
int m4g_ell_download(m4g_ell E, int64_t* indices, double* data);
/* smat_cvec_mult(ell_matrix) (mat_vec_mult.hpp:692-704): per row
 * s = sum over all slots of data * v[index] in slot order, from 0 */
int m4g_ell_spmv(m4g_ell E, m4g_vector v, m4g_vector w, int mode);

/* ---- sparse_banded ------------------------------------------------------ */
typedef struct m4g_banded_s* m4g_banded;

/* From the reference layout: `bands` strictly ascending diagonal offsets
 * (column - row), `data` row-major, data[r * nbands + b] = A[r][r + bands[b]]
 * (sparse_banded_inserter, sparse_banded.hpp:168-258). */
int m4g_banded_create(m4g_banded* B, int64_t nrows, int64_t ncols, int64_t nbands, const int64_t* bands,
                      const double* data);
/* The inserter fed with a CSR's entries: bands = the distinct offsets of the
 * stored entries, ascending; positions of a band without a stored entry hold 0. */
int m4g_banded_from_csr(m4g_banded* B, int64_t nrows, int64_t ncols, const int64_t* starts, const int64_t* indices,
                        const double* values);
int m4g_banded_destroy(m4g_banded B);
int m4g_banded_shape(m4g_banded B, int64_t* nrows, int64_t* ncols, int64_t* nbands);
/* ref_bands() and the reference layout of the values (nrows * nbands) */
int m4g_banded_download(m4g_banded B, int64_t* bands, double* data);
/* smat_cvec_mult(sparse_banded) (mat_vec_mult.hpp:709-809): per row the
 * bands with 0 <= r + band < ncols in ascending order, summed from 0; rows
 * with no such band are set to 0 (assign) or left alone (+=, -=). */
int m4g_banded_spmv(m4g_banded B, m4g_vector v, m4g_vector w, int mode);

/* ---- matrix-free operator ------------------------------------------------ */
/* mtl::mat::poisson2D_dirichlet(m, n).mult(v, w, Assign)
 * (mtl/matrix/poisson2D_dirichlet.hpp:27-80): the 5-point Laplacian on an
 * m x n grid (m, n >= 2) with Dirichlet borders, no matrix stored; each entry
 * in the reference's operand order; mode as above (Assign::apply). */
int m4g_poisson2d_dirichlet_mult(int64_t m, int64_t n, m4g_vector v, m4g_vector w, int mode);

#ifdef __cplusplus
}
#endif

#endif /* MTL4GPU_FORMATS_H */
