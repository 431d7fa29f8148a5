// Device building blocks of the sliced-ELL SpMV shared by the standalone
// SpMV and the fused kernels of the Krylov loops.
#pragma once

#include "m4g_device.hpp"

namespace m4g {

constexpr int kSliceWaves = 4;                 // slices (wavefronts) per workgroup
constexpr int kSpmvNT = kSliceWaves * kWave;   // 256 threads

// Row sum of lane `lane` in a slice starting at `base` with `width` entries
// per row: sequential in stored order from zero, as mat_vec_mult.hpp:662-687.
// T: value type of the matrix, x and the accumulator.
template <typename T>
__device__ inline T sell_row_sum(const int32_t* __restrict__ cols, const T* __restrict__ vals,
                                 const T* __restrict__ x, int64_t base, int width, int lane)
{
    const int32_t* c = cols + base + lane;
    const T* v = vals + base + lane;
    T acc = 0;
    int k = 0;
    for (; k + 4 <= width; k += 4) {
        int32_t c0 = c[0], c1 = c[kWave], c2 = c[2 * kWave], c3 = c[3 * kWave];
        T v0 = v[0], v1 = v[kWave], v2 = v[2 * kWave], v3 = v[3 * kWave];
        T x0 = c0 >= 0 ? x[c0] : T(0), x1 = c1 >= 0 ? x[c1] : T(0);
        T x2 = c2 >= 0 ? x[c2] : T(0), x3 = c3 >= 0 ? x[c3] : T(0);
        if (c0 >= 0) acc += v0 * x0;
        if (c1 >= 0) acc += v1 * x1;
        if (c2 >= 0) acc += v2 * x2;
        if (c3 >= 0) acc += v3 * x3;
        c += 4 * kWave;
        v += 4 * kWave;
    }
    for (; k < width; k++) {
        int32_t c0 = c[0];
        T v0 = v[0];
        if (c0 >= 0) acc += v0 * x[c0];
        c += kWave;
        v += kWave;
    }
    return acc;
}

}  // namespace m4g
