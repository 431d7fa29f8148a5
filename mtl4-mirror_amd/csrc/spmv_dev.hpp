// Device building blocks of the sliced-ELL SpMV shared by the standalone
// SpMV and the fused kernels of the Krylov loops.
#pragma once

#include "m4g_device.hpp"

namespace m4g {

constexpr int kSliceWaves = 4;                 // slices (wavefronts) per workgroup
constexpr int kSpmvNT = kSliceWaves * kWave;   // 256 threads

// Column storage of a slice: the indices of slice s start at byte
// (slice_idx[s] >> 1) of `idx`.  Bit 0 of slice_idx[s] clear ("narrow"): one
// int16 per entry holding column − row, kSellPad16 for padding — used when
// every entry of the slice lies within ±32767 of its row, which holds for the
// banded/stencil matrices of the solvers and halves the index bytes.  Bit 0
// set ("wide"): int32 columns, −1 for padding.
constexpr int16_t kSellPad16 = -32768;
constexpr int64_t kSellSlack = 64;      // entries allocated after the last slice
constexpr int kPipeBlocksPerCU = 3;     // pipelined SpMV: persistent workgroups per CU

// Row sum of lane `lane` of slice s: sequential in stored order from zero, as
// mat_vec_mult.hpp:662-687; padding entries are skipped, not added as zeros.
// T: value type of the matrix, x and the accumulator.
template <typename T>
__device__ inline T sell_row_sum(const int64_t* __restrict__ slice_ptr, const int64_t* __restrict__ slice_idx,
                                 const uint8_t* __restrict__ idx, const T* __restrict__ vals,
                                 const T* __restrict__ x, int64_t s, int lane)
{
    const int64_t base = slice_ptr[s];
    const int width = int((slice_ptr[s + 1] - base) / kWave);
    const int64_t si = slice_idx[s];
    const T* v = vals + base + lane;
    T acc = 0;
    int k = 0;
    if (si & 1) {
        const int32_t* c = reinterpret_cast<const int32_t*>(idx + (si >> 1)) + lane;
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
    } else {
        // padding lanes load x[0] (always present when the slice has entries)
        const int16_t* d = reinterpret_cast<const int16_t*>(idx + (si >> 1)) + lane;
        const int64_t row = s * kWave + lane;
        for (; k + 4 <= width; k += 4) {
            const int d0 = d[0], d1 = d[kWave], d2 = d[2 * kWave], d3 = d[3 * kWave];
            T v0 = v[0], v1 = v[kWave], v2 = v[2 * kWave], v3 = v[3 * kWave];
            T x0 = x[d0 == kSellPad16 ? 0 : row + d0], x1 = x[d1 == kSellPad16 ? 0 : row + d1];
            T x2 = x[d2 == kSellPad16 ? 0 : row + d2], x3 = x[d3 == kSellPad16 ? 0 : row + d3];
            if (d0 != kSellPad16) acc += v0 * x0;
            if (d1 != kSellPad16) acc += v1 * x1;
            if (d2 != kSellPad16) acc += v2 * x2;
            if (d3 != kSellPad16) acc += v3 * x3;
            d += 4 * kWave;
            v += 4 * kWave;
        }
        for (; k < width; k++) {
            const int d0 = d[0];
            T v0 = v[0];
            T x0 = x[d0 == kSellPad16 ? 0 : row + d0];
            if (d0 != kSellPad16) acc += v0 * x0;
            d += kWave;
            v += kWave;
        }
    }
    return acc;
}

}  // namespace m4g
