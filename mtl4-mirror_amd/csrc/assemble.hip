// Assembly of a compressed2D on the device from (row, col, value) triplets
// with the inserter's semantics (mtl/matrix/compressed2D.hpp:662-1151):
// entries of the same position are folded in insertion order with the
// updater (mtl/operation/update.hpp:20-115): the first one stores
// updater.init(v), each later one applies the updater; entries already in a
// given matrix count as present before the first insertion.  Rows come out
// sorted by column.
//
// Steps, all on the device: 64-bit keys row * ncols + col with the insertion
// index as payload; one stable radix sort (rocPRIM, least significant digit
// first, on only the key bits in use), so equal keys keep insertion order;
// segment heads and an inclusive scan give each distinct position its slot;
// one thread per position folds its segment sequentially (the same operation
// order as the CPU inserter, so values are bit-identical); per-row counts and
// an exclusive scan give the row starts.  The result is handed to the same
// SELL build as a host-created matrix.
#include "m4g_internal.hpp"

#include <rocprim/rocprim.hpp>

#include <memory>
#include <vector>

namespace m4g {
namespace {

__global__ __launch_bounds__(256) void k_assemble_keys(int64_t count, int64_t offset, const int64_t* __restrict__ rows,
                                                       const int64_t* __restrict__ cols, int64_t nrows, int64_t ncols,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ order,
                                                       int* __restrict__ bad)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < count; i += int64_t(gridDim.x) * 256) {
        const int64_t r = rows[i], c = cols[i];
        if (r < 0 || r >= nrows || c < 0 || c >= ncols) {
            atomicOr(bad, 1);
            keys[offset + i] = 0;
        } else {
            keys[offset + i] = uint64_t(r) * uint64_t(ncols) + uint64_t(c);
        }
        order[offset + i] = uint32_t(offset + i);
    }
}

// the present entries of an existing matrix (device CSR, int32 indices)
__global__ __launch_bounds__(256) void k_existing_keys(int64_t nrows, int64_t ncols, const int32_t* __restrict__ starts,
                                                       const int32_t* __restrict__ indices, uint64_t* __restrict__ keys,
                                                       uint32_t* __restrict__ order)
{
    for (int64_t r = blockIdx.x * int64_t(256) + threadIdx.x; r < nrows; r += int64_t(gridDim.x) * 256)
        for (int32_t j = starts[r]; j < starts[r + 1]; j++) {
            keys[j] = uint64_t(r) * uint64_t(ncols) + uint64_t(indices[j]);
            order[j] = uint32_t(j);
        }
}

__global__ __launch_bounds__(256) void k_heads(int64_t n, const uint64_t* __restrict__ keys, uint32_t* __restrict__ heads)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        heads[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_segment_starts(int64_t n, const uint32_t* __restrict__ heads,
                                                        const uint32_t* __restrict__ pos, int64_t* __restrict__ seg)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        if (heads[i]) seg[pos[i] - 1] = i;
}

// Folds each position's entries in insertion order: entries with index below
// `existing` are the matrix's present values (at most one per position, and
// first after the stable sort); the others are new values.
__global__ __launch_bounds__(256) void k_fold(int64_t nunique, int64_t n, int64_t existing, int64_t ncols, int updater,
                                              const int64_t* __restrict__ seg, const uint64_t* __restrict__ keys,
                                              const uint32_t* __restrict__ order, const double* __restrict__ old_vals,
                                              const double* __restrict__ new_vals, double* __restrict__ out_vals,
                                              int32_t* __restrict__ out_cols, int32_t* __restrict__ row_counts)
{
    for (int64_t u = blockIdx.x * int64_t(256) + threadIdx.x; u < nunique; u += int64_t(gridDim.x) * 256) {
        const int64_t i0 = seg[u], i1 = u + 1 < nunique ? seg[u + 1] : n;
        const int64_t first = order[i0];
        double x;
        if (first < existing) {
            x = old_vals[first];
        } else {
            const double v = new_vals[first - existing];
            x = updater == M4G_UPDATE_MINUS ? -v : v;   // updater.init(v)
        }
        for (int64_t i = i0 + 1; i < i1; i++) {
            const double v = new_vals[order[i] - existing];
            if (updater == M4G_UPDATE_STORE) x = v;
            else if (updater == M4G_UPDATE_PLUS) x += v;
            else if (updater == M4G_UPDATE_MINUS) x -= v;
            else x *= v;
        }
        const uint64_t key = keys[i0];
        out_vals[u] = x;
        out_cols[u] = int32_t(key % uint64_t(ncols));
        atomicAdd(&row_counts[key / uint64_t(ncols)], 1);
    }
}

int grid_for(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)));
}

unsigned bits_for(uint64_t maxkey)
{
    unsigned b = 1;
    while (b < 64 && (maxkey >> b) != 0) b++;
    return b;
}

}  // namespace
}  // namespace m4g

using namespace m4g;

extern "C" int m4g_matrix_assemble(m4g_matrix* out, int64_t nrows, int64_t ncols, int64_t count, const int64_t* rows,
                                   const int64_t* cols, const double* values, int on_device, int updater,
                                   m4g_matrix existing)
{
    return guard([&] {
        M4G_REQUIRE(out, M4G_ERR_INVALID, "m4g_matrix_assemble: null output");
        M4G_REQUIRE(nrows >= 0 && ncols >= 0 && count >= 0, M4G_ERR_INVALID, "m4g_matrix_assemble: negative size");
        M4G_REQUIRE(count == 0 || (rows && cols && values), M4G_ERR_INVALID, "m4g_matrix_assemble: null array");
        M4G_REQUIRE(updater >= M4G_UPDATE_STORE && updater <= M4G_UPDATE_TIMES, M4G_ERR_INVALID,
                    "m4g_matrix_assemble: unknown updater");
        M4G_REQUIRE(nrows < (int64_t(1) << 31) && ncols < (int64_t(1) << 31), M4G_ERR_INVALID,
                    "m4g_matrix_assemble: dimensions exceed int32 device indices");
        const int64_t n_old = existing ? existing->nnz : 0;
        if (existing)
            M4G_REQUIRE(existing->nrows == nrows && existing->ncols == ncols, M4G_ERR_SIZE,
                        "m4g_matrix_assemble: existing matrix has other dimensions");
        const int64_t n = n_old + count;
        M4G_REQUIRE(n < (int64_t(1) << 31), M4G_ERR_INVALID, "m4g_matrix_assemble: more than 2^31 entries");
        hipStream_t s = current_stream();

        // the triplets on the device
        DevBuf<int64_t> drows, dcols;
        DevBuf<double> dvals;
        const int64_t *R = rows, *C = cols;
        const double* V = values;
        if (!on_device && count) {
            drows.reset(size_t(count));
            dcols.reset(size_t(count));
            dvals.reset(size_t(count));
            drows.upload(rows, size_t(count));
            dcols.upload(cols, size_t(count));
            dvals.upload(values, size_t(count));
            R = drows.ptr;
            C = dcols.ptr;
            V = dvals.ptr;
        }

        DevBuf<uint64_t> keys(size_t(std::max<int64_t>(n, 1))), keys_s(size_t(std::max<int64_t>(n, 1)));
        DevBuf<uint32_t> order(size_t(std::max<int64_t>(n, 1))), order_s(size_t(std::max<int64_t>(n, 1)));
        DevBuf<int> bad(1);
        M4G_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int), s));
        if (n_old)
            hipLaunchKernelGGL(k_existing_keys, dim3(grid_for(nrows)), dim3(256), 0, s, nrows, ncols,
                               existing->starts.ptr, existing->indices.ptr, keys.ptr, order.ptr);
        if (count)
            hipLaunchKernelGGL(k_assemble_keys, dim3(grid_for(count)), dim3(256), 0, s, count, n_old, R, C, nrows,
                               ncols, keys.ptr, order.ptr, bad.ptr);
        M4G_HIP(hipGetLastError());
        int hbad = 0;
        M4G_HIP(hipMemcpyAsync(&hbad, bad.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        M4G_HIP(hipStreamSynchronize(s));
        M4G_REQUIRE(!hbad, M4G_ERR_INVALID, "m4g_matrix_assemble: index out of range");

        auto A = std::make_unique<m4g_matrix_s>();
        A->nrows = nrows;
        A->ncols = ncols;
        DevBuf<int32_t> counts(static_cast<size_t>(nrows) + 1);
        M4G_HIP(hipMemsetAsync(counts.ptr, 0, (size_t(nrows) + 1) * sizeof(int32_t), s));
        int64_t nunique = 0;
        DevBuf<double> ovals;
        DevBuf<int32_t> ocols;
        if (n) {
            const unsigned end_bit = bits_for(uint64_t(std::max<int64_t>(nrows, 1)) * uint64_t(std::max<int64_t>(ncols, 1)) - 1);
            size_t tmp_bytes = 0;
            M4G_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.ptr, keys_s.ptr, order.ptr, order_s.ptr,
                                              size_t(n), 0, end_bit, s));
            DevBuf<unsigned char> tmp(std::max<size_t>(tmp_bytes, 1));
            M4G_HIP(rocprim::radix_sort_pairs(tmp.ptr, tmp_bytes, keys.ptr, keys_s.ptr, order.ptr, order_s.ptr,
                                              size_t(n), 0, end_bit, s));
            DevBuf<uint32_t> heads(static_cast<size_t>(n)), pos(static_cast<size_t>(n));
            hipLaunchKernelGGL(k_heads, dim3(grid_for(n)), dim3(256), 0, s, n, keys_s.ptr, heads.ptr);
            size_t scan_bytes = 0;
            M4G_HIP(rocprim::inclusive_scan(nullptr, scan_bytes, heads.ptr, pos.ptr, size_t(n),
                                            rocprim::plus<uint32_t>(), s));
            DevBuf<unsigned char> scan_tmp(std::max<size_t>(scan_bytes, 1));
            M4G_HIP(rocprim::inclusive_scan(scan_tmp.ptr, scan_bytes, heads.ptr, pos.ptr, size_t(n),
                                            rocprim::plus<uint32_t>(), s));
            uint32_t hu = 0;
            M4G_HIP(hipMemcpyAsync(&hu, pos.ptr + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            M4G_HIP(hipStreamSynchronize(s));
            nunique = int64_t(hu);
            DevBuf<int64_t> seg(static_cast<size_t>(nunique));
            hipLaunchKernelGGL(k_segment_starts, dim3(grid_for(n)), dim3(256), 0, s, n, heads.ptr, pos.ptr, seg.ptr);
            ovals.reset(size_t(nunique));
            ocols.reset(size_t(nunique));
            hipLaunchKernelGGL(k_fold, dim3(grid_for(nunique)), dim3(256), 0, s, nunique, n, n_old, ncols, updater,
                               seg.ptr, keys_s.ptr, order_s.ptr, n_old ? existing->values.ptr : nullptr, V, ovals.ptr,
                               ocols.ptr, counts.ptr);
            M4G_HIP(hipGetLastError());
        }
        // row starts: exclusive scan of the per-row counts (nrows + 1 entries, the last count 0)
        A->starts.reset(size_t(nrows) + 1);
        size_t ex_bytes = 0;
        M4G_HIP(rocprim::exclusive_scan(nullptr, ex_bytes, counts.ptr, A->starts.ptr, int32_t(0), size_t(nrows) + 1,
                                        rocprim::plus<int32_t>(), s));
        DevBuf<unsigned char> ex_tmp(std::max<size_t>(ex_bytes, 1));
        M4G_HIP(rocprim::exclusive_scan(ex_tmp.ptr, ex_bytes, counts.ptr, A->starts.ptr, int32_t(0), size_t(nrows) + 1,
                                        rocprim::plus<int32_t>(), s));
        A->nnz = nunique;
        A->indices = std::move(ocols);
        A->values = std::move(ovals);
        if (!A->indices.ptr) {
            A->indices.reset(0);
            A->values.reset(0);
        }
        // the host copy of the pattern (schedules, downloads)
        std::vector<int32_t> s32(static_cast<size_t>(nrows) + 1), i32(static_cast<size_t>(nunique));
        A->starts.download(s32.data(), size_t(nrows) + 1);
        if (nunique) A->indices.download(i32.data(), size_t(nunique));
        A->h_starts.assign(s32.begin(), s32.end());
        A->h_indices.assign(i32.begin(), i32.end());
        sell_build(*A);
        M4G_HIP(hipStreamSynchronize(s));
        *out = A.release();
        return M4G_OK;
    });
}
