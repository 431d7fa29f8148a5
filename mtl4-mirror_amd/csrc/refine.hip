// FP32-inner / FP64-refined solves (see m4g_refined_solve in mtl4gpu.h).
//
// Each refinement step computes the FP64 residual r = b - A x with the FP64
// matrix, tests it with the outer iteration control, rounds r to FP32, solves
// A d = r with the FP32 Krylov method (FP32 matrix values, FP32 factors, FP32
// vectors and accumulation) to the inner tolerance, and adds d to x in FP64.
// The FP32 inner solve halves the bytes of every SpMV value and factor entry
// and of every vector stream.
#include "m4g_internal.hpp"
#include "m4g_device.hpp"
#include "pc.hpp"
#include "krylov.hpp"

#include <cmath>

namespace m4g {

__global__ __launch_bounds__(256) void k_add_f32(int64_t n, const float* __restrict__ d, double* __restrict__ x)
{
    for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
        x[i] = x[i] + double(d[i]);
}

static void launch_add_f32(int64_t n, const float* d, double* x, hipStream_t s)
{
    if (!n) return;
    int g = int(std::min<int64_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL(k_add_f32, dim3(g), dim3(256), 0, s, n, d, x);
    M4G_HIP(hipGetLastError());
}

}  // namespace m4g

using namespace m4g;

extern "C" int m4g_refined_solve(int method, m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P,
                                 m4g_iteration* outer, double inner_rtol, int inner_max_iter, m4g_refine_info* info)
{
    return guard([&] {
        M4G_REQUIRE(A && x && b && outer, M4G_ERR_INVALID, "m4g_refined_solve: null argument");
        M4G_REQUIRE(method == M4G_METHOD_CG || method == M4G_METHOD_BICGSTAB, M4G_ERR_INVALID,
                    "m4g_refined_solve: unknown method");
        M4G_REQUIRE(A->nrows == A->ncols, M4G_ERR_NOT_SQUARE, "Matrix is not square");
        M4G_REQUIRE(x->type == M4G_F64 && b->type == M4G_F64, M4G_ERR_INVALID,
                    "m4g_refined_solve: x and b are double");
        M4G_REQUIRE(x->n == A->nrows && b->n == A->nrows, M4G_ERR_SIZE, "m4g_refined_solve: incompatible size");
        M4G_REQUIRE(!P || P->kind <= M4G_PC_DIAGONAL || P->factor_type == M4G_F32, M4G_ERR_INVALID,
                    "m4g_refined_solve: the preconditioner's factors must be float");
        TraceRange tr(kTraceRefinedSolve);
        const int64_t n = A->nrows;
        hipStream_t s = current_stream();
        if (!A->sell.vals32.ptr) sell_build_f32(*A);
        DevBuf<double> r(std::max<int64_t>(n, 1));
        DevBuf<float> r32(std::max<int64_t>(n, 1)), d32(std::max<int64_t>(n, 1));
        ReduceWs& ws = reduce_ws();
        double* dnorm = ws.partials + kReduceBlocks * 16 + 1;
        m4g_refine_info inf{0, 0, 0.0, 0};
        const int i0 = outer->iterations;
        int error = outer->error;
        while (true) {
            launch_spmv(*A, (const double*)x->ptr, r.ptr, M4G_ASSIGN, (const double*)b->ptr, s);   // r = b - A x
            launch_sumsq(n, r.ptr, dnorm, s);
            const double rnorm = std::sqrt(read_scalar(dnorm));
            inf.resid = rnorm;
            const int k = outer->iterations;
            if (outer->history && k < outer->history_capacity) outer->history[k] = rnorm;
            // finished(|r|) of the outer iteration (basic_iteration.hpp:60-65, 85-93)
            const bool conv = outer->norm_r0 == 0 ? rnorm <= outer->atol
                                                   : (rnorm <= outer->rtol * outer->norm_r0 || rnorm <= outer->atol);
            if (conv) break;
            if (k >= outer->max_iter) {
                error = M4G_MAX_ITER;
                break;
            }
            outer->iterations = k + 1;
            launch_convert(n, r.ptr, M4G_F64, r32.ptr, M4G_F32, s);
            launch_fill(n, 0.0, d32.ptr, M4G_F32, s);
            m4g_iteration in{};
            in.max_iter = inner_max_iter;
            in.rtol = inner_rtol;
            in.atol = 0.0;
            in.norm_r0 = rnorm;
            int st = krylov_solve(method, *A, d32.ptr, r32.ptr, M4G_F32, P, &in, default_ws());
            inf.inner_iterations += in.iterations;
            inf.inner_error = in.error;
            if (st != M4G_OK && st != M4G_MAX_ITER) {
                if (info) *info = inf;
                return st;
            }
            launch_add_f32(n, d32.ptr, (double*)x->ptr, s);                  // x += d in FP64
        }
        inf.outer_iterations = outer->iterations - i0;
        outer->resid = inf.resid;
        outer->error = error;
        if (info) *info = inf;
        M4G_HIP(hipStreamSynchronize(s));
        return error;
    });
}
