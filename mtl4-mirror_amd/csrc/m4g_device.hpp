// Device-side helpers shared by the kernels (wave64 reductions, XCD mapping).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace m4g {

constexpr int kWave = 64;
constexpr int kXcds = 8;    // MI355X: workgroup i runs on XCD i % 8

// Barrier for values exchanged through LDS only: waits for this wave's LDS
// writes, not for its outstanding global loads and stores.
__device__ inline void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Butterfly sum over the 64 lanes of a wavefront; every lane gets the total.
// The combination order is fixed, so results are deterministic.
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sum over a workgroup of NT threads; the result is valid in thread 0.
// scratch must hold NT/64 values.
template <int NT, typename T>
__device__ inline T block_sum(T v, T* scratch)
{
    v = wave_sum(v);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (NT == kWave) return v;
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < NT / kWave; w++) t += scratch[w];
    __syncthreads();
    return t;
}

// Logical workgroup index with XCD-contiguous ranges: the hardware places
// workgroup b on XCD b % 8; this mapping gives each XCD one contiguous chunk
// of `chunk` logical groups, so neighbouring rows share an L2.
__device__ inline int64_t xcd_logical_block(int64_t chunk)
{
    const int64_t b = blockIdx.x;
    return (b % kXcds) * chunk + b / kXcds;
}

}  // namespace m4g
