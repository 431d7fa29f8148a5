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

// Pipelined slice loop for matrices whose slices are all narrow (offsets of
// slice s start at entry slice_ptr[s] of the int16 index array) with at most
// MAXW entries per row.  A persistent grid: wavefront g takes slices g,
// g + G, g + 2G, ... (G = all wavefronts, one chip-wide front through the
// matrix).  For each slice it issues the x gathers and pre(rr) (per-row
// loads of the caller, rr = the row clamped into range), then the offset and
// value loads of its next slice, then sums in stored order (padding and
// entries past the slice's width skipped, as sell_row_sum) and calls
// epi(row, sum, pre) for rows < n: the next slice's HBM loads are in flight
// while this slice's gathers complete.  Entries k >= a slice's width re-load
// its last entry (an empty slice: the slack after the last slice) and are not
// summed.  (Measured against a two-register-set variant without copies and
// one that forces the sum ahead of the row test: this form streams fastest.)
template <typename T, int MAXW, typename Pre, typename Epi>
__device__ inline void sell_pipe_run(int64_t n, int64_t nslices, int64_t padded, const int64_t* __restrict__ slice_ptr,
                                     const uint8_t* __restrict__ idx, const T* __restrict__ vals,
                                     const T* __restrict__ x, Pre pre, Epi epi)
{
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x % kWave;
    const int64_t stride = int64_t(gridDim.x) * (blockDim.x / kWave);
    int64_t s = int64_t(blockIdx.x) * (blockDim.x / kWave) + wave;
    if (s >= nslices) return;
    const int16_t* i16 = reinterpret_cast<const int16_t*>(idx);
    int64_t base = slice_ptr[s];
    int wd = int((slice_ptr[s + 1] - base) / kWave);
    int16_t d[MAXW];
    T v[MAXW];
#pragma unroll
    for (int k = 0; k < MAXW; k++) {
        const int kk = k < wd ? k : (wd > 0 ? wd - 1 : 0);
        const int16_t t = i16[base + kk * kWave + lane];
        d[k] = wd > 0 ? t : kSellPad16;
        v[k] = vals[base + kk * kWave + lane];
    }
    for (;;) {
        const int64_t sn = s + stride;
        const bool more = sn < nslices;
        const int64_t row = s * kWave + lane;
        const int64_t rr = row < n ? row : n - 1;
        T xv[MAXW];
#pragma unroll
        for (int k = 0; k < MAXW; k++) xv[k] = x[d[k] == kSellPad16 ? 0 : row + d[k]];
        const auto pv = pre(rr);
        // look-ahead: the next slice's offsets and values (this slice again at the end)
        const int64_t sl = more ? sn : s;
        const int64_t bn = slice_ptr[sl];
        const int wdn = int((slice_ptr[sl + 1] - bn) / kWave);
        int16_t dn[MAXW];
        T vn[MAXW];
#pragma unroll
        for (int k = 0; k < MAXW; k++) {
            const int kk = k < wdn ? k : (wdn > 0 ? wdn - 1 : 0);
            const int16_t t = i16[bn + kk * kWave + lane];
            dn[k] = wdn > 0 ? t : kSellPad16;
            vn[k] = vals[bn + kk * kWave + lane];
        }
        T acc = 0;
#pragma unroll
        for (int k = 0; k < MAXW; k++)
            if (k < wd && d[k] != kSellPad16) acc += v[k] * xv[k];
        if (row < n) epi(row, acc, pv);
        if (!more) break;
        s = sn;
        wd = wdn;
#pragma unroll
        for (int k = 0; k < MAXW; k++) {
            d[k] = dn[k];
            v[k] = vn[k];
        }
    }
}

}  // namespace m4g
