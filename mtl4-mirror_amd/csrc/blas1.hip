// Vector kernels of the CG / BiCGStab loops: dot, squared norm, axpy, xpay,
// fused axpy + squared norm, fill and precision conversion.  All are
// HBM-bound streams; reductions are deterministic two-stage sums: a fixed
// partition into workgroups with a fixed butterfly inside each, then a
// one-workgroup kernel adding the partials in workgroup order (kernel
// ordering makes the partials visible without device-scope fences).
#include "m4g_internal.hpp"
#include "m4g_device.hpp"

namespace m4g {

constexpr int kNT = 256;

static int reduce_grid(int64_t n)
{
    int64_t g = (n + kNT * 4 - 1) / (kNT * 4);
    if (g < 1) g = 1;
    if (g > kReduceBlocks) g = kReduceBlocks;
    return int(g);
}

__global__ __launch_bounds__(kNT) void k_dot(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                             double* partials, unsigned* counter, double* out)
{
    __shared__ double scratch[kNT / kWave];
    double acc = 0;
    for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kNT)
        acc += x[i] * y[i];
    acc = block_sum<kNT>(acc, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kNT) void k_sumsq(int64_t n, const double* __restrict__ x, double* partials,
                                               unsigned* counter, double* out)
{
    __shared__ double scratch[kNT / kWave];
    double acc = 0;
    for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kNT)
        acc += x[i] * x[i];
    acc = block_sum<kNT>(acc, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kNT) void k_axpy_sumsq(int64_t n, double alpha, const double* __restrict__ q,
                                                    double* __restrict__ r, double* partials, unsigned* counter,
                                                    double* out)
{
    __shared__ double scratch[kNT / kWave];
    double acc = 0;
    for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kNT) {
        double ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc += ri * ri;
    }
    acc = block_sum<kNT>(acc, scratch);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kNT) void k_axpy(int64_t n, double alpha, const double* __restrict__ x,
                                              double* __restrict__ y)
{
    for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kNT)
        y[i] = y[i] + alpha * x[i];
}

__global__ __launch_bounds__(kNT) void k_xpay(int64_t n, const double* __restrict__ x, double beta,
                                              double* __restrict__ y)
{
    for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kNT)
        y[i] = x[i] + beta * y[i];
}

template <typename T>
__global__ __launch_bounds__(kNT) void k_fill(int64_t n, T value, T* __restrict__ x)
{
    for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kNT) x[i] = value;
}

template <typename S, typename D>
__global__ __launch_bounds__(kNT) void k_convert(int64_t n, const S* __restrict__ src, D* __restrict__ dst)
{
    for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kNT)
        dst[i] = static_cast<D>(src[i]);
}

// One thread writes count scalars into host-coherent pinned memory and then
// the sequence number, both at system scope (the number released after the
// values), so a host that sees the number sees the values.
__device__ inline void post_scalars(HostPost* box, const double* v, int count, unsigned seq)
{
    for (int i = 0; i < count; i++) __hip_atomic_store(&box->val[i], v[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __hip_atomic_store(&box->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// total of the partials into *out and, with a mailbox, posted to the host
__global__ void __launch_bounds__(256) k_sum_partials(const double* __restrict__ partials, int nblocks, double* out,
                                                      HostPost* box, unsigned seq)
{
    __shared__ double scratch[256 / kWave];
    double acc = 0;
    for (int b = threadIdx.x; b < nblocks; b += 256) acc += partials[b];
    acc = block_sum<256>(acc, scratch);
    if (threadIdx.x == 0) {
        *out = acc;
        if (box) post_scalars(box, &acc, 1, seq);
    }
}

__global__ void __launch_bounds__(64) k_post(const double* __restrict__ src, int count, HostPost* box, unsigned seq)
{
    if (threadIdx.x == 0) post_scalars(box, src, count, seq);
}

void launch_sum_partials(const double* partials, int nblocks, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partials, nblocks, out, nullptr, 0u);
}

void launch_post(const double* src, int count, HostPost* box, unsigned seq, hipStream_t s)
{
    M4G_REQUIRE(count >= 0 && count <= kPostMax, M4G_ERR_INVALID, "launch_post: too many scalars");
    hipLaunchKernelGGL(k_post, dim3(1), dim3(64), 0, s, src, count, box, seq);
    M4G_HIP(hipGetLastError());
}

static int stream_grid(int64_t n)
{
    int64_t g = (n + kNT - 1) / kNT;
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return int(g);
}

int reduce_partials(int64_t n) { return reduce_grid(n); }

void launch_dot_with(int64_t n, const double* x, const double* y, double* partials, double* out, hipStream_t s)
{
    TraceRange tr(kTraceDot);
    hipLaunchKernelGGL(k_dot, dim3(reduce_grid(n)), dim3(kNT), 0, s, n, x, y, partials, nullptr, out);
    launch_sum_partials(partials, reduce_grid(n), out, s);
    M4G_HIP(hipGetLastError());
}

void launch_dot(int64_t n, const double* x, const double* y, double* out, hipStream_t s)
{
    launch_dot_with(n, x, y, reduce_ws().partials, out, s);
}

void launch_dot_post(int64_t n, const double* x, const double* y, double* out, HostPost* box, unsigned seq,
                     hipStream_t s)
{
    TraceRange tr(kTraceDot);
    ReduceWs& ws = reduce_ws();
    hipLaunchKernelGGL(k_dot, dim3(reduce_grid(n)), dim3(kNT), 0, s, n, x, y, ws.partials, nullptr, out);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, ws.partials, reduce_grid(n), out, box, seq);
    M4G_HIP(hipGetLastError());
}

void launch_sumsq(int64_t n, const double* x, double* out, hipStream_t s)
{
    TraceRange tr(kTraceUnaryDot);
    ReduceWs& ws = reduce_ws();
    hipLaunchKernelGGL(k_sumsq, dim3(reduce_grid(n)), dim3(kNT), 0, s, n, x, ws.partials, ws.counters, out);
    launch_sum_partials(ws.partials, reduce_grid(n), out, s);
    M4G_HIP(hipGetLastError());
}

void launch_sumsq_post(int64_t n, const double* x, double* out, HostPost* box, unsigned seq, hipStream_t s)
{
    TraceRange tr(kTraceUnaryDot);
    ReduceWs& ws = reduce_ws();
    hipLaunchKernelGGL(k_sumsq, dim3(reduce_grid(n)), dim3(kNT), 0, s, n, x, ws.partials, ws.counters, out);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, ws.partials, reduce_grid(n), out, box, seq);
    M4G_HIP(hipGetLastError());
}

void launch_axpy_sumsq(int64_t n, double alpha, const double* q, double* r, double* out, hipStream_t s)
{
    TraceRange tr(kTraceVecExpr);
    ReduceWs& ws = reduce_ws();
    hipLaunchKernelGGL(k_axpy_sumsq, dim3(reduce_grid(n)), dim3(kNT), 0, s, n, alpha, q, r, ws.partials,
                       ws.counters, out);
    launch_sum_partials(ws.partials, reduce_grid(n), out, s);
    M4G_HIP(hipGetLastError());
}

void launch_axpy(int64_t n, double alpha, const double* x, double* y, hipStream_t s)
{
    TraceRange tr(kTraceVecExpr);
    if (n == 0) return;
    hipLaunchKernelGGL(k_axpy, dim3(stream_grid(n)), dim3(kNT), 0, s, n, alpha, x, y);
    M4G_HIP(hipGetLastError());
}

void launch_xpay(int64_t n, const double* x, double beta, double* y, hipStream_t s)
{
    TraceRange tr(kTraceVecExpr);
    if (n == 0) return;
    hipLaunchKernelGGL(k_xpay, dim3(stream_grid(n)), dim3(kNT), 0, s, n, x, beta, y);
    M4G_HIP(hipGetLastError());
}

void launch_fill(int64_t n, double value, void* x, int type, hipStream_t s)
{
    if (n == 0) return;
    if (type == M4G_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(stream_grid(n)), dim3(kNT), 0, s, n, value, (double*)x);
    else
        hipLaunchKernelGGL(k_fill<float>, dim3(stream_grid(n)), dim3(kNT), 0, s, n, float(value), (float*)x);
    M4G_HIP(hipGetLastError());
}

void launch_convert(int64_t n, const void* src, int st, void* dst, int dt, hipStream_t s)
{
    if (n == 0) return;
    if (st == M4G_F64 && dt == M4G_F32)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(stream_grid(n)), dim3(kNT), 0, s, n,
                           (const double*)src, (float*)dst);
    else if (st == M4G_F32 && dt == M4G_F64)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(stream_grid(n)), dim3(kNT), 0, s, n,
                           (const float*)src, (double*)dst);
    else
        throw Error(M4G_ERR_INVALID, "launch_convert: unsupported types");
    M4G_HIP(hipGetLastError());
}

}  // namespace m4g
