// CSR SpMV on MI355X through a sliced-ELL copy of the matrix.
//
// Layout: rows are grouped in slices of 64 (one wavefront, one row per lane).
// Inside a slice the k-th stored entry of the 64 rows is contiguous, so every
// value/column load of a wavefront is one 512 B / 256 B coalesced access, and
// for stencil matrices the x gathers of a wavefront are contiguous as well.
// Each row keeps the reference's stored (column-sorted) order and is summed
// sequentially from zero, so results equal mat_vec_mult.hpp:662-687 bit for
// bit (the library is built with -ffp-contract=off).  Workgroups are mapped so
// that every XCD streams one contiguous range of slices, keeping the x window
// of a range inside that XCD's L2.
#include "m4g_internal.hpp"
#include "m4g_device.hpp"
#include "spmv_dev.hpp"
#include "krylov.hpp"

#include <algorithm>

namespace m4g {

__global__ __launch_bounds__(kSpmvNT) void k_sell_fill(int64_t n, const int64_t* __restrict__ slice_ptr,
                                                       const int64_t* __restrict__ slice_idx,
                                                       const int32_t* __restrict__ starts,
                                                       const int32_t* __restrict__ indices,
                                                       const double* __restrict__ values, uint8_t* __restrict__ idx,
                                                       double* __restrict__ vals, int64_t* __restrict__ pos)
{
    int64_t r = blockIdx.x * int64_t(kSpmvNT) + threadIdx.x;
    int64_t s = r / kWave;
    int lane = int(r % kWave);
    int64_t nslices = (n + kWave - 1) / kWave;
    if (s >= nslices) return;
    int64_t base = slice_ptr[s];
    int width = int((slice_ptr[s + 1] - base) / kWave);
    const int64_t si = slice_idx[s];
    int32_t* wide = reinterpret_cast<int32_t*>(idx + (si >> 1));
    int16_t* narrow = reinterpret_cast<int16_t*>(idx + (si >> 1));
    int64_t b = 0, len = 0;
    if (r < n) {
        b = starts[r];
        len = starts[r + 1] - b;
    }
    for (int k = 0; k < width; k++) {
        int64_t d = base + int64_t(k) * kWave + lane;
        int64_t e = int64_t(k) * kWave + lane;
        if (k < len) {
            if (si & 1) wide[e] = indices[b + k];
            else narrow[e] = int16_t(indices[b + k] - r);
            vals[d] = values[b + k];
            pos[d] = b + k;
        } else {
            if (si & 1) wide[e] = -1;
            else narrow[e] = kSellPad16;
            vals[d] = 0.0;
            pos[d] = -1;
        }
    }
}

__global__ __launch_bounds__(kSpmvNT) void k_sell_refresh(int64_t padded, const int64_t* __restrict__ pos,
                                                          const double* __restrict__ values, double* __restrict__ vals)
{
    for (int64_t d = blockIdx.x * int64_t(kSpmvNT) + threadIdx.x; d < padded; d += int64_t(gridDim.x) * kSpmvNT) {
        int64_t p = pos[d];
        vals[d] = p >= 0 ? values[p] : 0.0;
    }
}

__global__ __launch_bounds__(kSpmvNT) void k_sell_to_f32(int64_t padded, const double* __restrict__ v,
                                                         float* __restrict__ f)
{
    for (int64_t d = blockIdx.x * int64_t(kSpmvNT) + threadIdx.x; d < padded; d += int64_t(gridDim.x) * kSpmvNT)
        f[d] = float(v[d]);
}


void sell_build(m4g_matrix_s& A)
{
    Sell& S = A.sell;
    S.n = A.nrows;
    S.nslices = (A.nrows + kWave - 1) / kWave;
    std::vector<int64_t> sp(S.nslices + 1, 0), si(std::max<int64_t>(S.nslices, 1), 0);
    int maxw = 0;
    int64_t ibytes = 0;
    S.narrow_slices = 0;
    for (int64_t s = 0; s < S.nslices; s++) {
        int64_t r0 = s * kWave, r1 = std::min<int64_t>(A.nrows, r0 + kWave);
        int64_t w = 0;
        bool narrow = true;
        for (int64_t r = r0; r < r1; r++) {
            w = std::max<int64_t>(w, A.h_starts[r + 1] - A.h_starts[r]);
            for (int64_t e = A.h_starts[r]; e < A.h_starts[r + 1]; e++) {
                const int64_t d = A.h_indices[e] - r;
                narrow = narrow && d >= -32767 && d <= 32767;
            }
        }
        maxw = std::max<int>(maxw, int(w));
        sp[s + 1] = sp[s] + w * kWave;
        si[s] = (ibytes << 1) | (narrow ? 0 : 1);
        ibytes += w * kWave * (narrow ? 2 : 4);
        S.narrow_slices += narrow;
    }
    S.padded = sp[S.nslices];
    S.max_width = maxw;
    M4G_REQUIRE(S.padded < (int64_t(1) << 40), M4G_ERR_INVALID, "matrix too large");
    S.slice_ptr.reset(S.nslices + 1);
    S.slice_ptr.upload(sp.data(), S.nslices + 1);
    S.slice_idx.reset(si.size());
    S.slice_idx.upload(si.data(), si.size());
    S.idx.reset(ibytes + 4 * kSellSlack);
    // kSellSlack entries after the last slice keep the pipelined kernel's
    // look-ahead loads of an empty final slice in bounds
    S.vals.reset(S.padded + kSellSlack);
    S.csr_pos.reset(S.padded);
    // slack after the last slice (read by the pipelined kernel for an empty
    // final slice): value 0 and padding offsets
    M4G_HIP(hipMemsetAsync(S.vals.ptr + S.padded, 0, kSellSlack * sizeof(double), current_stream()));
    if (S.narrow_slices == S.nslices)
        M4G_HIP(hipMemsetD16Async(reinterpret_cast<hipDeviceptr_t>(S.idx.ptr + 2 * S.padded),
                                  static_cast<unsigned short>(kSellPad16), kSellSlack, current_stream()));
    S.vals32.release();
    if (S.nslices) {
        int64_t threads = S.nslices * kWave;
        hipLaunchKernelGGL(k_sell_fill, dim3((threads + kSpmvNT - 1) / kSpmvNT), dim3(kSpmvNT), 0, current_stream(),
                           A.nrows, S.slice_ptr.ptr, S.slice_idx.ptr, A.starts.ptr, A.indices.ptr, A.values.ptr,
                           S.idx.ptr,
                           S.vals.ptr, S.csr_pos.ptr);
        M4G_HIP(hipGetLastError());
    }
    // reduction workspace for the fused SpMV + dot
    int64_t groups = (S.nslices + kSliceWaves - 1) / kSliceWaves;
    int64_t chunk = (groups + kXcds - 1) / kXcds;
    A.red_partials.reset(std::max<int64_t>({int64_t(1), chunk * kXcds, int64_t(sell_pipe_grid())}));
    A.red_counter.reset(1);
    M4G_HIP(hipMemsetAsync(A.red_counter.ptr, 0, sizeof(unsigned), current_stream()));
}

void sell_refresh_values(m4g_matrix_s& A)
{
    Sell& S = A.sell;
    if (!S.padded) return;
    int64_t g = std::min<int64_t>((S.padded + kSpmvNT - 1) / kSpmvNT, 16384);
    hipLaunchKernelGGL(k_sell_refresh, dim3(g), dim3(kSpmvNT), 0, current_stream(), S.padded, S.csr_pos.ptr,
                       A.values.ptr, S.vals.ptr);
    M4G_HIP(hipGetLastError());
    if (S.vals32.ptr) sell_build_f32(A);
}

void sell_build_f32(m4g_matrix_s& A)
{
    Sell& S = A.sell;
    if (!S.padded) return;
    if (!S.vals32.ptr) {
        S.vals32.reset(S.padded + kSellSlack);
        M4G_HIP(hipMemsetAsync(S.vals32.ptr + S.padded, 0, kSellSlack * sizeof(float), current_stream()));
    }
    int64_t g = std::min<int64_t>((S.padded + kSpmvNT - 1) / kSpmvNT, 16384);
    hipLaunchKernelGGL(k_sell_to_f32, dim3(g), dim3(kSpmvNT), 0, current_stream(), S.padded, S.vals.ptr,
                       S.vals32.ptr);
    M4G_HIP(hipGetLastError());
}

// Pipelined SpMV for matrices whose slices all use 16-bit offsets and hold
// at most MAXW entries per row: a persistent grid (kPipeBlocksPerCU workgroups
// per CU) in which wavefront g takes slices g, g + G, g + 2G, ... (G = all
// wavefronts, so the whole chip streams one front through the matrix) and
// issues the value/offset loads of its next slice before it waits for the x
// gathers of the current one: HBM loads stay in flight through the gather
// phase.  Row sums as sell_row_sum (stored order, padding skipped).
// MODE 0: w = A v; 1: w += A v; 2: w -= A v; 3: w = b - A v.  DOT: per
// workgroup partial of v.(A v) over its slices.
template <typename T, int MODE, bool DOT, int MAXW>
__global__ __launch_bounds__(kSpmvNT) void k_sell_pipe(int64_t n, int64_t nslices, int64_t padded,
                                                       const int64_t* __restrict__ slice_ptr,
                                                       const uint8_t* __restrict__ idx, const T* __restrict__ vals,
                                                       const T* __restrict__ x, T* __restrict__ w,
                                                       const T* __restrict__ b, double* partials)
{
    __shared__ double scratch[kSliceWaves];
    double dotp = 0;
    struct Pre {
        T prior, xr;
    };
    sell_pipe_run<T, MAXW>(
        n, nslices, padded, slice_ptr, idx, vals, x,
        [&](int64_t rr) {
            Pre p{T(0), T(0)};
            if (MODE == 1 || MODE == 2) p.prior = w[rr];
            if (MODE == 3) p.prior = b[rr];
            if (DOT) p.xr = x[rr];
            return p;
        },
        [&](int64_t row, T acc, const Pre& p) {
            if (MODE == 0) w[row] = acc;
            else if (MODE == 1) w[row] = p.prior + acc;
            else w[row] = p.prior - acc;
            if (DOT) dotp += double(p.xr * acc);
        });
    if (DOT) {
        double blk = block_sum<kSpmvNT>(dotp, scratch);
        if (threadIdx.x == 0) partials[blockIdx.x] = blk;
    }
}

// the pipelined kernels apply: every slice narrow, rows of at most 16 entries
int sell_pipe_width(const m4g_matrix_s& A)
{
    const Sell& S = A.sell;
    if (A.ncols == 0 || S.nslices == 0 || S.narrow_slices != S.nslices || std::getenv("M4G_SPMV_NO_PIPE")) return 0;
    if (S.max_width <= 3) return 3;
    if (S.max_width <= 8) return S.max_width;
    if (S.max_width <= 16) return 16;
    return 0;
}

int sell_pipe_grid()
{
    return std::max(1, device_state().num_cus) * kPipeBlocksPerCU;
}

template <typename T, int MODE, bool DOT>
static void launch_pipe(const m4g_matrix_s& A, int mw, const T* vals, const T* x, T* w, const T* b, double* partials,
                        hipStream_t st)
{
    const Sell& S = A.sell;
#define M4G_PIPE(MW)                                                                                                \
    hipLaunchKernelGGL((k_sell_pipe<T, MODE, DOT, MW>), dim3(sell_pipe_grid()), dim3(kSpmvNT), 0, st, A.nrows, S.nslices, \
                       S.padded, S.slice_ptr.ptr, S.idx.ptr, vals, x, w, b, partials)
    switch (mw) {
    case 3: M4G_PIPE(3); break;
    case 4: M4G_PIPE(4); break;
    case 5: M4G_PIPE(5); break;
    case 6: M4G_PIPE(6); break;
    case 7: M4G_PIPE(7); break;
    case 8: M4G_PIPE(8); break;
    default: M4G_PIPE(16); break;
    }
#undef M4G_PIPE
    M4G_HIP(hipGetLastError());
}

// MODE 0: w = A v; 1: w += A v; 2: w -= A v; 3: w = b - A v.  DOT adds p.q.
template <int MODE, bool DOT>
__global__ __launch_bounds__(kSpmvNT) void k_sell_spmv(int64_t n, int64_t nslices, int64_t chunk,
                                                       const int64_t* __restrict__ slice_ptr,
                                                       const int64_t* __restrict__ slice_idx,
                                                       const uint8_t* __restrict__ idx,
                                                       const double* __restrict__ vals, const double* __restrict__ x,
                                                       double* __restrict__ w, const double* __restrict__ b,
                                                       double* partials, unsigned* counter, double* out)
{
    __shared__ double scratch[kSliceWaves];
    const int64_t g = xcd_logical_block(chunk);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t s = g * kSliceWaves + wave;
    double dotp = 0;
    if (s < nslices) {
        const double acc = sell_row_sum(slice_ptr, slice_idx, idx, vals, x, s, lane);
        const int64_t row = s * kWave + lane;
        if (row < n) {
            if (MODE == 0) w[row] = acc;
            else if (MODE == 1) w[row] = w[row] + acc;
            else if (MODE == 2) w[row] = w[row] - acc;
            else w[row] = b[row] - acc;
            if (DOT) dotp = x[row] * acc;
        }
    }
    if (DOT) {
        double blk = block_sum<kSpmvNT>(dotp, scratch);
        if (threadIdx.x == 0) partials[blockIdx.x] = blk;
    }
}

template <typename T>
__global__ __launch_bounds__(kSpmvNT) void k_sell_residual(int64_t n, int64_t nslices, int64_t chunk,
                                                           const int64_t* __restrict__ slice_ptr,
                                                           const int64_t* __restrict__ slice_idx,
                                                           const uint8_t* __restrict__ idx, const T* __restrict__ vals,
                                                           const T* __restrict__ x, const T* __restrict__ b,
                                                           T* __restrict__ r)
{
    const int64_t g = xcd_logical_block(chunk);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t s = g * kSliceWaves + wave;
    if (s >= nslices) return;
    const T acc = sell_row_sum(slice_ptr, slice_idx, idx, vals, x, s, lane);
    const int64_t row = s * kWave + lane;
    if (row < n) r[row] = b[row] - acc;
}

static void spmv_geometry(const m4g_matrix_s& A, int64_t& chunk, int64_t& grid)
{
    int64_t groups = (A.sell.nslices + kSliceWaves - 1) / kSliceWaves;
    chunk = (groups + kXcds - 1) / kXcds;
    grid = std::max<int64_t>(1, chunk * kXcds);
}

void launch_spmv(const m4g_matrix_s& A, const double* x, double* w, int mode, const double* b, hipStream_t st)
{
    TraceRange tr(kTraceCrsMult);
    if (A.nrows == 0) return;
    if (const int mw = sell_pipe_width(A)) {
        const double* v = A.sell.vals.ptr;
        if (b) launch_pipe<double, 3, false>(A, mw, v, x, w, b, nullptr, st);
        else if (mode == M4G_ASSIGN) launch_pipe<double, 0, false>(A, mw, v, x, w, b, nullptr, st);
        else if (mode == M4G_PLUS) launch_pipe<double, 1, false>(A, mw, v, x, w, b, nullptr, st);
        else launch_pipe<double, 2, false>(A, mw, v, x, w, b, nullptr, st);
        return;
    }
    int64_t chunk, grid;
    spmv_geometry(A, chunk, grid);
    const Sell& S = A.sell;
#define M4G_SPMV_ARGS                                                                                   \
    dim3(grid), dim3(kSpmvNT), 0, st, A.nrows, S.nslices, chunk, S.slice_ptr.ptr, S.slice_idx.ptr, S.idx.ptr, S.vals.ptr, x, w, b, \
        (double*)nullptr, (unsigned*)nullptr, (double*)nullptr
    if (b) hipLaunchKernelGGL((k_sell_spmv<3, false>), M4G_SPMV_ARGS);
    else if (mode == M4G_ASSIGN) hipLaunchKernelGGL((k_sell_spmv<0, false>), M4G_SPMV_ARGS);
    else if (mode == M4G_PLUS) hipLaunchKernelGGL((k_sell_spmv<1, false>), M4G_SPMV_ARGS);
    else hipLaunchKernelGGL((k_sell_spmv<2, false>), M4G_SPMV_ARGS);
#undef M4G_SPMV_ARGS
    M4G_HIP(hipGetLastError());
}

void launch_spmv_dot(const m4g_matrix_s& A, const double* p, double* q, double* out, hipStream_t st)
{
    TraceRange tr(kTraceCrsMult);
    if (A.nrows == 0) {
        M4G_HIP(hipMemsetAsync(out, 0, sizeof(double), st));
        return;
    }
    if (const int mw = sell_pipe_width(A)) {
        M4G_REQUIRE(A.red_partials.count >= size_t(sell_pipe_grid()), M4G_ERR_INVALID, "reduction workspace too small");
        launch_pipe<double, 0, true>(A, mw, A.sell.vals.ptr, p, q, (const double*)nullptr, A.red_partials.ptr, st);
        launch_sum_partials(A.red_partials.ptr, sell_pipe_grid(), out, st);
        M4G_HIP(hipGetLastError());
        return;
    }
    int64_t chunk, grid;
    spmv_geometry(A, chunk, grid);
    const Sell& S = A.sell;
    hipLaunchKernelGGL((k_sell_spmv<0, true>), dim3(grid), dim3(kSpmvNT), 0, st, A.nrows, S.nslices, chunk,
                       S.slice_ptr.ptr, S.slice_idx.ptr, S.idx.ptr, S.vals.ptr, p, q, (const double*)nullptr,
                       A.red_partials.ptr, A.red_counter.ptr, out);
    launch_sum_partials(A.red_partials.ptr, int(grid), out, st);
    M4G_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_pattern_gradient(int64_t n, const int32_t* __restrict__ starts,
                                                         const int32_t* __restrict__ indices,
                                                         const double* __restrict__ lambda,
                                                         const double* __restrict__ x, double* __restrict__ grad)
{
    for (int64_t r = blockIdx.x * int64_t(256) + threadIdx.x; r < n; r += int64_t(gridDim.x) * 256) {
        const double l = lambda[r];
        for (int32_t p = starts[r]; p < starts[r + 1]; p++) grad[p] = -l * x[indices[p]];
    }
}

__global__ __launch_bounds__(256) void k_gather_values(int64_t count, const int64_t* __restrict__ pos,
                                                       const double* __restrict__ src, double* __restrict__ dst)
{
    for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < count; e += int64_t(gridDim.x) * 256)
        dst[e] = src[pos[e]];
}

template <typename T>
void launch_residual_t(const m4g_matrix_s& A, const T* x, const T* b, T* r, hipStream_t st)
{
    if (A.nrows == 0) return;
    int64_t chunk, grid;
    spmv_geometry(A, chunk, grid);
    const Sell& S = A.sell;
    const T* vals = sizeof(T) == 8 ? (const T*)S.vals.ptr : (const T*)S.vals32.ptr;
    M4G_REQUIRE(vals, M4G_ERR_INVALID, "float matrix values not prepared");
    if (const int mw = sell_pipe_width(A)) {
        launch_pipe<T, 3, false>(A, mw, vals, x, r, b, nullptr, st);
        return;
    }
    hipLaunchKernelGGL(k_sell_residual<T>, dim3(grid), dim3(kSpmvNT), 0, st, A.nrows, S.nslices, chunk,
                       S.slice_ptr.ptr, S.slice_idx.ptr, S.idx.ptr, vals, x, b, r);
    M4G_HIP(hipGetLastError());
}

void launch_pattern_gradient(const m4g_matrix_s& A, const double* lambda, const double* x, double* grad,
                             hipStream_t s)
{
    if (!A.nrows) return;
    int g = int(std::min<int64_t>((A.nrows + 255) / 256, 16384));
    hipLaunchKernelGGL(k_pattern_gradient, dim3(g), dim3(256), 0, s, A.nrows, A.starts.ptr, A.indices.ptr, lambda,
                       x, grad);
    M4G_HIP(hipGetLastError());
}

void launch_gather_values(int64_t count, const int64_t* pos, const double* src, double* dst, hipStream_t s)
{
    if (!count) return;
    int g = int(std::min<int64_t>((count + 255) / 256, 16384));
    hipLaunchKernelGGL(k_gather_values, dim3(g), dim3(256), 0, s, count, pos, src, dst);
    M4G_HIP(hipGetLastError());
}

template void launch_residual_t<double>(const m4g_matrix_s&, const double*, const double*, double*, hipStream_t);
template void launch_residual_t<float>(const m4g_matrix_s&, const float*, const float*, float*, hipStream_t);

}  // namespace m4g
