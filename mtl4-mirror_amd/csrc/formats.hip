// mtl::mat::ell_matrix and mtl::mat::sparse_banded on the device, with their
// matrix-vector products (mtl/operation/mat_vec_mult.hpp:692-809).
//
// ELL keeps the reference's slot layout, which is already the coalesced one:
// entry i of row r sits at r + i * stride, so the 64 lanes of a wavefront
// (64 consecutive rows) read 64 consecutive values and indices per slot.
// sparse_banded is stored band-major on the device ([band][row]; the
// reference is row-major [row][band]) so that each band is one contiguous
// stream and x is read at r + offset, also contiguous across a wavefront.
// Both products are HBM-bound streams with one thread per row; each row sums
// its entries from 0 in the reference's order, so results are bit-identical
// to the CPU loops (no FMA contraction).
#include "m4g_internal.hpp"
#include "../../include/mtl4gpu_formats.h"

#include <algorithm>
#include <memory>
#include <vector>

struct m4g_ell_s {
    int64_t nrows = 0, ncols = 0, slots = 0, stride = 0, nnz = 0;
    m4g::DevBuf<int32_t> idx;   // stride * slots
    m4g::DevBuf<double> val;
};

struct m4g_banded_s {
    int64_t nrows = 0, ncols = 0, nbands = 0;
    m4g::DevBuf<int64_t> bands;   // ascending offsets
    m4g::DevBuf<double> val;      // band-major: val[b * nrows + r]
};

namespace m4g {

template <int MODE>
__device__ __forceinline__ void first_update(double* w, int64_t r, double s)
{
    if (MODE == M4G_ASSIGN) w[r] = s;
    else if (MODE == M4G_PLUS) w[r] = w[r] + s;
    else w[r] = w[r] - s;
}

// smat_cvec_mult(ell_matrix) (mat_vec_mult.hpp:692-704): s = 0; for each slot
// s += data[k] * v[indices[k]]; first_update(w[r], s).
template <int MODE>
__global__ __launch_bounds__(256) void k_ell_spmv(int64_t nrows, int64_t stride, int slots,
                                                  const int32_t* __restrict__ idx, const double* __restrict__ val,
                                                  const double* __restrict__ x, double* __restrict__ w)
{
    for (int64_t r = blockIdx.x * int64_t(256) + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * 256) {
        double s = 0.0;
        int64_t k = r;
#pragma unroll 8
        for (int i = 0; i < slots; ++i, k += stride) s += val[k] * x[idx[k]];
        first_update<MODE>(w, r, s);
    }
}

// smat_cvec_mult(sparse_banded) (mat_vec_mult.hpp:709-809): row r sums the
// bands b (ascending) with begin_rows[b] <= r < end_rows[b], i.e.
// 0 <= r + bands[b] < ncols, from 0.  The reference's segment loop updates
// the rows from the first row of the highest band to the end of the lowest
// band; the rows before and after are set to 0 by an assignment and left
// alone by += and -=.
template <int MODE>
__global__ __launch_bounds__(256) void k_banded_spmv(int64_t nrows, int64_t ncols, int nbands,
                                                     const int64_t* __restrict__ bands,
                                                     const double* __restrict__ val, const double* __restrict__ x,
                                                     double* __restrict__ w)
{
    const int64_t lowest = bands[0], highest = bands[nbands - 1];
    for (int64_t r = blockIdx.x * int64_t(256) + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * 256) {
        double s = 0.0;
        if (r + lowest >= 0 && r + highest < ncols) {
            // interior row: every band is inside the matrix
            for (int b = 0; b < nbands; ++b) s += val[b * nrows + r] * x[r + bands[b]];
        } else {
            for (int b = 0; b < nbands; ++b) {
                const int64_t c = r + bands[b];
                if (c >= 0 && c < ncols) s += val[b * nrows + r] * x[c];
            }
        }
        if (MODE == M4G_ASSIGN || (r + highest >= 0 && r + lowest < ncols)) first_update<MODE>(w, r, s);
    }
}

// poisson2D_dirichlet::mult (mtl/matrix/poisson2D_dirichlet.hpp:27-80): the
// 5-point operator on an m x n grid without a stored matrix; each w[k] is the
// reference's expression for its region (inner, border, corner) with the
// same operand order, then Assign::apply.
template <int MODE>
__global__ __launch_bounds__(256) void k_poisson2d_dirichlet(int64_t m, int64_t n, const double* __restrict__ v,
                                                             double* __restrict__ w)
{
    const int64_t total = m * n;
    for (int64_t k = blockIdx.x * int64_t(256) + threadIdx.x; k < total; k += int64_t(gridDim.x) * 256) {
        const int64_t i = k / n, j = k - i * n;
        const bool top = i == 0, bottom = i == m - 1, left = j == 0, right = j == n - 1;
        double x;
        if (top && left) x = 4 * v[0] - v[1] - v[n];
        else if (top && right) x = 4 * v[n - 1] - v[n - 2] - v[2 * n - 1];
        else if (bottom && left) x = 4 * v[k] - v[k - n] - v[k + 1];
        else if (bottom && right) x = 4 * v[k] - v[k - 1] - v[k - n];
        else if (top) x = 4 * v[k] - v[k + n] - v[k - 1] - v[k + 1];
        else if (bottom) x = 4 * v[k] - v[k - n] - v[k - 1] - v[k + 1];
        else if (left) x = 4 * v[k] - v[k - n] - v[k + n] - v[k + 1];
        else if (right) x = 4 * v[k] - v[k - n] - v[k + n] - v[k - 1];
        else x = 4 * v[k] - v[k - n] - v[k + n] - v[k - 1] - v[k + 1];
        first_update<MODE>(w, k, x);
    }
}

__global__ __launch_bounds__(256) void k_zero_rows(int64_t n, double* __restrict__ w)
{
    for (int64_t r = blockIdx.x * int64_t(256) + threadIdx.x; r < n; r += int64_t(gridDim.x) * 256) w[r] = 0.0;
}

namespace {

int rows_grid(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)));
}

void check_vectors(int64_t nrows, int64_t ncols, m4g_vector v, m4g_vector w, int mode, const char* who)
{
    M4G_REQUIRE(v && w, M4G_ERR_INVALID, std::string(who) + ": null vector");
    M4G_REQUIRE(v->type == M4G_F64 && w->type == M4G_F64, M4G_ERR_INVALID, std::string(who) + ": double vectors");
    M4G_REQUIRE(v->n == ncols && w->n == nrows, M4G_ERR_SIZE, std::string(who) + ": incompatible size");
    M4G_REQUIRE(v->ptr != w->ptr || nrows == 0, M4G_ERR_INVALID, std::string(who) + ": v and w alias");
    M4G_REQUIRE(mode == M4G_ASSIGN || mode == M4G_PLUS || mode == M4G_MINUS, M4G_ERR_INVALID,
                std::string(who) + ": unknown mode");
}

void check_csr(int64_t nrows, int64_t ncols, const int64_t* S, const int64_t* I, const char* who)
{
    M4G_REQUIRE(nrows >= 0 && ncols >= 0, M4G_ERR_INVALID, std::string(who) + ": negative dimension");
    M4G_REQUIRE(ncols < (int64_t(1) << 31), M4G_ERR_INVALID, std::string(who) + ": more than 2^31 columns");
    M4G_REQUIRE(S && (S[nrows] == 0 || I), M4G_ERR_INVALID, std::string(who) + ": null array");
    M4G_REQUIRE(S[0] == 0, M4G_ERR_LOGIC, std::string(who) + ": starts[0] != 0");
    for (int64_t r = 0; r < nrows; r++) {
        M4G_REQUIRE(S[r + 1] >= S[r], M4G_ERR_LOGIC, std::string(who) + ": starts not ascending");
        for (int64_t j = S[r]; j < S[r + 1]; j++) {
            M4G_REQUIRE(I[j] >= 0 && I[j] < ncols, M4G_ERR_LOGIC, std::string(who) + ": column out of range");
            M4G_REQUIRE(j == S[r] || I[j] > I[j - 1], M4G_ERR_LOGIC, std::string(who) + ": rows must be sorted");
        }
    }
}

}  // namespace
}  // namespace m4g

using namespace m4g;

extern "C" {

int m4g_ell_create(m4g_ell* E, int64_t nrows, int64_t ncols, const int64_t* S, const int64_t* I, const double* V)
{
    return guard([&] {
        M4G_REQUIRE(E, M4G_ERR_INVALID, "m4g_ell_create: null handle");
        check_csr(nrows, ncols, S, I, "m4g_ell_create");
        M4G_REQUIRE(S[nrows] == 0 || V, M4G_ERR_INVALID, "m4g_ell_create: null values");
        auto e = std::make_unique<m4g_ell_s>();
        e->nrows = nrows;
        e->ncols = ncols;
        e->nnz = S[nrows];
        e->stride = (nrows + 31) / 32 * 32;   // ell_matrix::set_stride (ell_matrix.hpp:43-44), alignment 32
        for (int64_t r = 0; r < nrows; r++) e->slots = std::max(e->slots, S[r + 1] - S[r]);
        const size_t total = size_t(e->stride) * size_t(e->slots);
        std::vector<int32_t> idx(total, 0);
        std::vector<double> val(total, 0.0);
        for (int64_t r = 0; r < nrows; r++) {
            int64_t k = r, patch = 0;
            for (int64_t j = S[r]; j < S[r + 1]; j++, k += e->stride) {
                patch = I[j];
                idx[size_t(k)] = int32_t(I[j]);
                val[size_t(k)] = V[j];
            }
            for (int64_t j = S[r + 1] - S[r]; j < e->slots; j++, k += e->stride) idx[size_t(k)] = int32_t(patch);
        }
        e->idx.reset(total);
        e->val.reset(total);
        e->idx.upload(idx.data(), total);
        e->val.upload(val.data(), total);
        M4G_HIP(hipStreamSynchronize(current_stream()));
        *E = e.release();
        return M4G_OK;
    });
}

int m4g_ell_destroy(m4g_ell E)
{
    return guard([&] {
        delete E;
        return M4G_OK;
    });
}

int m4g_ell_shape(m4g_ell E, int64_t* nrows, int64_t* ncols, int64_t* slots, int64_t* stride, int64_t* nnz)
{
    return guard([&] {
        M4G_REQUIRE(E, M4G_ERR_INVALID, "m4g_ell_shape: null handle");
        if (nrows) *nrows = E->nrows;
        if (ncols) *ncols = E->ncols;
        if (slots) *slots = E->slots;
        if (stride) *stride = E->stride;
        if (nnz) *nnz = E->nnz;
        return M4G_OK;
    });
}

int m4g_ell_download(m4g_ell E, int64_t* indices, double* data)
{
    return guard([&] {
        M4G_REQUIRE(E, M4G_ERR_INVALID, "m4g_ell_download: null handle");
        const size_t total = size_t(E->stride) * size_t(E->slots);
        if (indices) {
            std::vector<int32_t> idx(total);
            E->idx.download(idx.data(), total);
            for (size_t k = 0; k < total; k++) indices[k] = idx[k];
        }
        if (data) E->val.download(data, total);
        return M4G_OK;
    });
}

int m4g_ell_spmv(m4g_ell E, m4g_vector v, m4g_vector w, int mode)
{
    return guard([&] {
        M4G_REQUIRE(E, M4G_ERR_INVALID, "m4g_ell_spmv: null handle");
        check_vectors(E->nrows, E->ncols, v, w, mode, "m4g_ell_spmv");
        if (E->nrows == 0) return M4G_OK;
        const dim3 grid(rows_grid(E->nrows)), block(256);
        const double* x = (const double*)v->ptr;
        double* y = (double*)w->ptr;
        hipStream_t s = current_stream();
        const int slots = int(E->slots);
        if (mode == M4G_ASSIGN)
            hipLaunchKernelGGL(k_ell_spmv<M4G_ASSIGN>, grid, block, 0, s, E->nrows, E->stride, slots, E->idx.ptr,
                               E->val.ptr, x, y);
        else if (mode == M4G_PLUS)
            hipLaunchKernelGGL(k_ell_spmv<M4G_PLUS>, grid, block, 0, s, E->nrows, E->stride, slots, E->idx.ptr,
                               E->val.ptr, x, y);
        else
            hipLaunchKernelGGL(k_ell_spmv<M4G_MINUS>, grid, block, 0, s, E->nrows, E->stride, slots, E->idx.ptr,
                               E->val.ptr, x, y);
        M4G_HIP(hipGetLastError());
        return M4G_OK;
    });
}

int m4g_banded_create(m4g_banded* B, int64_t nrows, int64_t ncols, int64_t nbands, const int64_t* bands,
                      const double* data)
{
    return guard([&] {
        M4G_REQUIRE(B, M4G_ERR_INVALID, "m4g_banded_create: null handle");
        M4G_REQUIRE(nrows >= 0 && ncols >= 0 && nbands >= 0, M4G_ERR_INVALID, "m4g_banded_create: negative size");
        M4G_REQUIRE(nbands == 0 || (bands && data), M4G_ERR_INVALID, "m4g_banded_create: null array");
        M4G_REQUIRE(nbands <= (int64_t(1) << 20), M4G_ERR_INVALID, "m4g_banded_create: too many bands");
        for (int64_t b = 1; b < nbands; b++)
            M4G_REQUIRE(bands[b] > bands[b - 1], M4G_ERR_LOGIC, "m4g_banded_create: bands must ascend strictly");
        auto m = std::make_unique<m4g_banded_s>();
        m->nrows = nrows;
        m->ncols = ncols;
        m->nbands = nbands;
        const size_t total = size_t(nrows) * size_t(nbands);
        std::vector<double> bm(total);
        for (int64_t r = 0; r < nrows; r++)
            for (int64_t b = 0; b < nbands; b++) bm[size_t(b) * size_t(nrows) + size_t(r)] = data[r * nbands + b];
        m->bands.reset(size_t(nbands));
        m->val.reset(total);
        m->bands.upload(bands, size_t(nbands));
        m->val.upload(bm.data(), total);
        M4G_HIP(hipStreamSynchronize(current_stream()));
        *B = m.release();
        return M4G_OK;
    });
}

int m4g_banded_from_csr(m4g_banded* B, int64_t nrows, int64_t ncols, const int64_t* S, const int64_t* I,
                        const double* V)
{
    return guard([&] {
        M4G_REQUIRE(B, M4G_ERR_INVALID, "m4g_banded_from_csr: null handle");
        check_csr(nrows, ncols, S, I, "m4g_banded_from_csr");
        M4G_REQUIRE(S[nrows] == 0 || V, M4G_ERR_INVALID, "m4g_banded_from_csr: null values");
        std::vector<int64_t> offs;
        for (int64_t r = 0; r < nrows; r++)
            for (int64_t j = S[r]; j < S[r + 1]; j++) offs.push_back(I[j] - r);
        std::sort(offs.begin(), offs.end());
        offs.erase(std::unique(offs.begin(), offs.end()), offs.end());
        const int64_t nb = int64_t(offs.size());
        std::vector<double> data(size_t(nrows) * size_t(nb), 0.0);
        for (int64_t r = 0; r < nrows; r++)
            for (int64_t j = S[r]; j < S[r + 1]; j++) {
                const int64_t b = std::lower_bound(offs.begin(), offs.end(), I[j] - r) - offs.begin();
                data[size_t(r) * size_t(nb) + size_t(b)] = V[j];
            }
        const int st = m4g_banded_create(B, nrows, ncols, nb, offs.data(), data.data());
        M4G_REQUIRE(st == M4G_OK, st, m4g_last_error());
        return M4G_OK;
    });
}

int m4g_banded_destroy(m4g_banded B)
{
    return guard([&] {
        delete B;
        return M4G_OK;
    });
}

int m4g_banded_shape(m4g_banded B, int64_t* nrows, int64_t* ncols, int64_t* nbands)
{
    return guard([&] {
        M4G_REQUIRE(B, M4G_ERR_INVALID, "m4g_banded_shape: null handle");
        if (nrows) *nrows = B->nrows;
        if (ncols) *ncols = B->ncols;
        if (nbands) *nbands = B->nbands;
        return M4G_OK;
    });
}

int m4g_banded_download(m4g_banded B, int64_t* bands, double* data)
{
    return guard([&] {
        M4G_REQUIRE(B, M4G_ERR_INVALID, "m4g_banded_download: null handle");
        if (bands) B->bands.download(bands, size_t(B->nbands));
        if (data) {
            const size_t total = size_t(B->nrows) * size_t(B->nbands);
            std::vector<double> bm(total);
            B->val.download(bm.data(), total);
            for (int64_t r = 0; r < B->nrows; r++)
                for (int64_t b = 0; b < B->nbands; b++)
                    data[r * B->nbands + b] = bm[size_t(b) * size_t(B->nrows) + size_t(r)];
        }
        return M4G_OK;
    });
}

int m4g_banded_spmv(m4g_banded B, m4g_vector v, m4g_vector w, int mode)
{
    return guard([&] {
        M4G_REQUIRE(B, M4G_ERR_INVALID, "m4g_banded_spmv: null handle");
        check_vectors(B->nrows, B->ncols, v, w, mode, "m4g_banded_spmv");
        if (B->nrows == 0) return M4G_OK;
        const dim3 grid(rows_grid(B->nrows)), block(256);
        const double* x = (const double*)v->ptr;
        double* y = (double*)w->ptr;
        hipStream_t s = current_stream();
        const int nb = int(B->nbands);
        if (nb == 0) {   // no bands: an assignment clears w (mat_vec_mult.hpp:725-728)
            if (mode == M4G_ASSIGN) hipLaunchKernelGGL(k_zero_rows, grid, block, 0, s, B->nrows, y);
            M4G_HIP(hipGetLastError());
            return M4G_OK;
        }
        if (mode == M4G_ASSIGN)
            hipLaunchKernelGGL(k_banded_spmv<M4G_ASSIGN>, grid, block, 0, s, B->nrows, B->ncols, nb, B->bands.ptr,
                               B->val.ptr, x, y);
        else if (mode == M4G_PLUS)
            hipLaunchKernelGGL(k_banded_spmv<M4G_PLUS>, grid, block, 0, s, B->nrows, B->ncols, nb, B->bands.ptr,
                               B->val.ptr, x, y);
        else
            hipLaunchKernelGGL(k_banded_spmv<M4G_MINUS>, grid, block, 0, s, B->nrows, B->ncols, nb, B->bands.ptr,
                               B->val.ptr, x, y);
        M4G_HIP(hipGetLastError());
        return M4G_OK;
    });
}

int m4g_poisson2d_dirichlet_mult(int64_t m, int64_t n, m4g_vector v, m4g_vector w, int mode)
{
    return guard([&] {
        M4G_REQUIRE(m >= 2 && n >= 2, M4G_ERR_INVALID, "m4g_poisson2d_dirichlet_mult: grid must be at least 2 x 2");
        check_vectors(m * n, m * n, v, w, mode, "m4g_poisson2d_dirichlet_mult");
        const dim3 grid(rows_grid(m * n)), block(256);
        const double* x = (const double*)v->ptr;
        double* y = (double*)w->ptr;
        hipStream_t s = current_stream();
        if (mode == M4G_ASSIGN)
            hipLaunchKernelGGL(k_poisson2d_dirichlet<M4G_ASSIGN>, grid, block, 0, s, m, n, x, y);
        else if (mode == M4G_PLUS)
            hipLaunchKernelGGL(k_poisson2d_dirichlet<M4G_PLUS>, grid, block, 0, s, m, n, x, y);
        else
            hipLaunchKernelGGL(k_poisson2d_dirichlet<M4G_MINUS>, grid, block, 0, s, m, n, x, y);
        M4G_HIP(hipGetLastError());
        return M4G_OK;
    });
}

}  // extern "C"
