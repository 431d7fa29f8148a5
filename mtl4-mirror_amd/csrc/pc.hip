// Preconditioners: identity, diagonal, IC(0) and ILU(0) with their
// triangular sweeps on MI355X.
//
// Factorisation runs on the GPU in natural-order level sets (one launch per
// level; rows of a level are independent), with the reference's per-row
// operation order, so factors equal the CPU result bit for bit:
//  - IC(0) (ic_0.hpp:153-185) is restated left-looking: row i of L = trans(U)
//    lists the U_ki, k < i, in ascending k; U_ki <- U_ki * (1/sqrt(D_k)) and
//    D_i <- D_i - U_ki * U_ki in that order, then U_ii <- 1/sqrt(D_i).  This is
//    the sequence MTL4's right-looking loop applies to each entry.
//  - ILU(0) (ilu_0.hpp:63-78) is already row-by-row (IKJ).
// The sweeps run the tile-pipelined schedules compiled in schedule.cpp.
#include "m4g_internal.hpp"
#include "m4g_device.hpp"
#include "pc.hpp"
#include "krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace m4g {

// ---------------------------------------------------------------------------
// sentinels of the sweep exchange buffers: signalling NaNs, which arithmetic
// never produces (results that are NaN come out quiet)
// ---------------------------------------------------------------------------
template <typename T> struct Sentinel;
template <> struct Sentinel<double> {
    static constexpr uint64_t bits = 0x7FF0DEADBEEF0001ull;
    __device__ static double value() { return __longlong_as_double((long long)bits); }
    __device__ static bool is(double v) { return (uint64_t)__double_as_longlong(v) == bits; }
    __device__ static double canon(double v) { return is(v) ? __longlong_as_double(0x7FF8000000000000ll) : v; }
};
template <> struct Sentinel<float> {
    static constexpr uint32_t bits = 0x7F80BEEFu;
    __device__ static float value() { return __int_as_float((int)bits); }
    __device__ static bool is(float v) { return (uint32_t)__float_as_int(v) == bits; }
    __device__ static float canon(float v) { return is(v) ? __int_as_float(0x7FC00000) : v; }
};

// ---------------------------------------------------------------------------
// factorisation kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ic0_level(int64_t count, const int32_t* __restrict__ level_rows,
                                                   const int64_t* __restrict__ lptr,
                                                   const int32_t* __restrict__ lcol,
                                                   const int64_t* __restrict__ lpos,
                                                   const int64_t* __restrict__ diag_pos, double* __restrict__ U,
                                                   double* __restrict__ invdiag)
{
    int64_t t = blockIdx.x * int64_t(256) + threadIdx.x;
    if (t >= count) return;
    const int32_t i = level_rows[t];
    double D = U[diag_pos[i]];
    for (int64_t e = lptr[i]; e < lptr[i + 1]; e++) {
        const int64_t p = lpos[e];
        const double d = U[p] * invdiag[lcol[e]];   // U_ki *= 1/sqrt(D_k)   (ic_0.hpp:166)
        U[p] = d;
        D = D - d * d;                             // U_ii -= U_ki * U[k][i] (ic_0.hpp:181)
    }
    const double inv = 1.0 / sqrt(D);             // reciprocal(sqrt(U_ii)) (ic_0.hpp:160)
    U[diag_pos[i]] = inv;
    invdiag[i] = inv;
}

__global__ __launch_bounds__(256) void k_ilu0_level(int64_t count, const int32_t* __restrict__ level_rows,
                                                    const int32_t* __restrict__ S, const int32_t* __restrict__ I,
                                                    const int64_t* __restrict__ diag_pos, double* __restrict__ LU,
                                                    double* __restrict__ invdiag)
{
    int64_t t = blockIdx.x * int64_t(256) + threadIdx.x;
    if (t >= count) return;
    const int32_t i = level_rows[t];
    const int64_t dp = diag_pos[i], end = S[i + 1];
    for (int64_t kc = S[i]; kc < dp; kc++) {
        const int32_t k = I[kc];
        const double aik = LU[kc] * invdiag[k];                  // ilu_0.hpp:70
        LU[kc] = aik;
        // LU[k][j] for the later entries j of row i: merge with row k
        int64_t pk = diag_pos[k] + 1;
        const int64_t ke = S[k + 1];
        for (int64_t jc = kc + 1; jc < end; jc++) {
            const int32_t j = I[jc];
            while (pk < ke && I[pk] < j) pk++;
            const double lukj = (pk < ke && I[pk] == j) ? LU[pk] : 0.0;
            LU[jc] = LU[jc] - aik * lukj;                        // ilu_0.hpp:74
        }
    }
    const double inv = 1.0 / LU[dp];                             // ilu_0.hpp:77 and invert_diagonal
    invdiag[i] = inv;
    LU[dp] = inv;
}

// ---------------------------------------------------------------------------
// sweep kernel
// ---------------------------------------------------------------------------
struct SweepArgs {
    int64_t ntiles;
    const int32_t* tile_row_begin;    // [ntiles+1] into rows
    const int32_t* rows;              // [n] global row of each (tile, slot)
    const int32_t* tile_group_begin;  // [ntiles+1] into the group table
    const int4* group_tab;            // [ngroups+1]: first group row, first entry, width, imports needed
    const int2* gr_meta;              // [n]: slot | len << 16, export slot (-1: none)
    const int32_t* ent_src;           // [nent]: LDS slot, or -(tile-local import + 1)
    const int32_t* tile_import_begin; // [ntiles+1] into imports
    const int32_t* imports;           // exchange slot of each import
    int32_t max_tile_rows;            // LDS rows reserved per tile (imports follow them)
    int32_t max_tile_imports;         // LDS import slots reserved per tile (group table follows)
    unsigned* tickets;                // [0] next ticket, [1] finished tiles
    int* fault;                       // set when a bounded wait runs out
    int64_t other_nexport;            // size of the other direction's exchange buffer
    int64_t spin_limit;
    int64_t* trace;                   // optional: [tile][2 + 2 * trace_steps] wall-clock stamps
    int32_t trace_steps;
};

constexpr int kGroupChunk = 256;      // group-table entries staged in LDS at a time
constexpr int kImportDepth = 8;       // imports each importer lane keeps in flight

__device__ inline int wave_min(int v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}

// The data one lane needs for its row of a level group, loaded one level
// ahead so that the loads overlap the previous level's arithmetic.
template <typename FV, int MAXW>
struct RowPre {
    int2 meta;
    FV diag;
    int32_t src[MAXW];
    FV val[MAXW];
};

template <typename FV, bool UNIT, int MAXW>
__device__ inline void row_prefetch(RowPre<FV, MAXW>& p, const SweepArgs& a, const FV* __restrict__ ent_val,
                                    const FV* __restrict__ gr_diag, int4 t0, int4 t1, int q)
{
    const int m = t1.x - t0.x;
    if (q >= m) return;
    p.meta = a.gr_meta[t0.x + q];
    p.diag = UNIT ? FV(1) : gr_diag[t0.x + q];
    const int w = t0.z;
#pragma unroll
    for (int k = 0; k < MAXW; k++)
        if (k < w) {
            const int64_t d = int64_t(t0.y) + int64_t(k) * m + q;
            p.src[k] = a.ent_src[d];
            p.val[k] = ent_val[d];
        }
}

// One row: rr = v_r; rr -= L_rj * y_j in stored order; y_r = rr * d_r
// (lower_trisolve.hpp:191-209 / 172-189, upper_trisolve.hpp:159-176).  Own
// values come from the tile's LDS rows, other tiles' values from the LDS
// import area filled by the importer wavefront.
template <typename FV, typename VT, bool UNIT, int MAXW>
__device__ inline void row_compute(const RowPre<FV, MAXW>& p, const SweepArgs& a, const FV* __restrict__ ent_val,
                                   VT* lds, const VT* imp, VT* xbuf, int4 t0, int m, int q)
{
    const int slot = p.meta.x & 0xFFFF;
    const int len = int(unsigned(p.meta.x) >> 16);
    VT rr = lds[slot];
#pragma unroll
    for (int k = 0; k < MAXW; k++)
        if (k < len) {
            const int32_t src = p.src[k];
            const VT yv = src >= 0 ? lds[src] : imp[-(src + 1)];
            rr -= VT(p.val[k]) * yv;
        }
    for (int k = MAXW; k < len; k++) {   // rows longer than the prefetch window
        const int64_t d = int64_t(t0.y) + int64_t(k) * m + q;
        const int32_t src = a.ent_src[d];
        const VT yv = src >= 0 ? lds[src] : imp[-(src + 1)];
        rr -= VT(ent_val[d]) * yv;
    }
    const VT out = UNIT ? rr : VT(rr * VT(p.diag));
    lds[slot] = out;
    if (p.meta.y >= 0)
        __hip_atomic_store(xbuf + p.meta.y, Sentinel<VT>::canon(out), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Prefetch depth (levels) of the compute wavefront's register ring.
template <int MAXW> struct PrefetchDepth { static constexpr int value = MAXW <= 4 ? 8 : (MAXW <= 8 ? 4 : 2); };

// Workgroup = 2 wavefronts.  Wavefront 0 imports the tile's values from
// earlier tiles (exchange buffer -> LDS) in first-use order, keeping up to
// 64 * kImportDepth polls in flight, and publishes how many are in LDS.
// Wavefront 1 walks the tile's levels: it waits (on LDS) until the imports a
// level needs are present, then computes the level's rows from LDS only; its
// per-row schedule data is prefetched PrefetchDepth levels ahead into a
// register ring.
template <typename FV, typename VT, bool UNIT, int MAXW>
__global__ __launch_bounds__(128) void k_sweep(SweepArgs a, const FV* __restrict__ ent_val,
                                               const FV* __restrict__ gr_diag, const VT* __restrict__ v,
                                               VT* __restrict__ y, VT* xbuf, VT* xbuf_other, const int* skip)
{
    constexpr int D = PrefetchDepth<MAXW>::value;
    if (skip && *skip) return;   // the solve this sweep belongs to has finished
    extern __shared__ unsigned char lds_raw[];
    VT* lds = reinterpret_cast<VT*>(lds_raw);
    VT* imp = lds + a.max_tile_rows;
    int4* tab = reinterpret_cast<int4*>(imp + a.max_tile_imports);   // whole tile's group table (+D sentinel rows)
    __shared__ int tile_sh;
    __shared__ int progress;
    const int tid = threadIdx.x, wave = tid / kWave, lane = tid % kWave;
    if (tid == 0) {
        tile_sh = int(atomicAdd(&a.tickets[0], 1u));
        progress = 0;
    }
    __syncthreads();
    const int t = tile_sh;
    if (a.trace && tid == 0) a.trace[int64_t(t) * (2 + 2 * a.trace_steps)] = int64_t(wall_clock64());
    // the other direction's exchange buffer is free: reset it for its next use
    for (int64_t x = int64_t(t) * 128 + tid; x < a.other_nexport; x += a.ntiles * int64_t(128))
        xbuf_other[x] = Sentinel<VT>::value();
    const int32_t rb = a.tile_row_begin[t], re = a.tile_row_begin[t + 1];
    for (int32_t i0 = rb + tid; i0 < re; i0 += 4 * 128) {   // rows first, then the values: 4 gathers in flight
        int32_t r[4];
#pragma unroll
        for (int u = 0; u < 4; u++) r[u] = i0 + u * 128 < re ? a.rows[i0 + u * 128] : -1;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (r[u] >= 0) lds[i0 + u * 128 - rb] = v[r[u]];
    }
    const int32_t g0 = a.tile_group_begin[t], g1 = a.tile_group_begin[t + 1], ng = g1 - g0;
    for (int i = tid; i <= ng + D; i += 128) {
        // entries past the tile's last group repeat its end, i.e. empty groups
        tab[i] = a.group_tab[g0 + min(i, ng)];
    }
    const int32_t ib = a.tile_import_begin[t], ni = a.tile_import_begin[t + 1] - ib;
    __syncthreads();

    if (wave == 0) {
        // ---- importer ----
        // Probes only the first missing import while it is not there (one
        // load per poll, with back-off), then loads a window of up to
        // 64 * kImportDepth imports at once and publishes the prefix that
        // arrived.  Waiting tiles so put almost no load on memory.
        int head = 0;
        int64_t polls = 0;
        unsigned nap = 1;
        while (head < ni) {
            VT hv = VT(0);
            if (lane == 0)
                hv = __hip_atomic_load(xbuf + a.imports[ib + head], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const bool head_ready = __shfl(int(!Sentinel<VT>::is(hv)), 0, kWave) != 0;
            if (head_ready) {
                nap = 1;
                int32_t xs[kImportDepth];
                VT val[kImportDepth];
#pragma unroll
                for (int u = 0; u < kImportDepth; u++) {
                    const int idx = head + lane + kWave * u;
                    xs[u] = idx < ni ? a.imports[ib + idx] : -1;
                }
#pragma unroll
                for (int u = 0; u < kImportDepth; u++)
                    val[u] = xs[u] >= 0 ? __hip_atomic_load(xbuf + xs[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                        : VT(0);
                int first_missing = ni;   // smallest index of the window that has not arrived
#pragma unroll
                for (int u = 0; u < kImportDepth; u++) {
                    const int idx = head + lane + kWave * u;
                    if (xs[u] >= 0) {
                        if (!Sentinel<VT>::is(val[u])) imp[idx] = val[u];
                        else first_missing = min(first_missing, idx);
                    }
                }
                const int window_end = min(ni, head + kWave * kImportDepth);
                first_missing = wave_min(min(first_missing, window_end));
                head = first_missing;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lane == 0) __hip_atomic_store(&progress, head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                continue;
            }
            ++polls;
            if (polls > a.spin_limit ||
                ((polls & 63) == 0 && __hip_atomic_load(a.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                if (lane == 0) __hip_atomic_store(a.fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            for (unsigned k = 0; k < nap; k++) __builtin_amdgcn_s_sleep(2);
            nap = min(nap * 2, 16u);
        }
    } else {
        // ---- compute ----
        bool failed = false;
        RowPre<FV, MAXW> ring[D];
#pragma unroll
        for (int u = 0; u < D; u++) row_prefetch<FV, UNIT, MAXW>(ring[u], a, ent_val, gr_diag, tab[u], tab[u + 1], lane);
        for (int32_t g = 0; g < ng; g += D) {
#pragma unroll
            for (int u = 0; u < D; u++) {
                const int gi = g + u;
                if (gi < ng) {
                    const int4 t0 = tab[gi];
                    const int m = tab[gi + 1].x - t0.x;
                    const RowPre<FV, MAXW> cur = ring[u];
                    row_prefetch<FV, UNIT, MAXW>(ring[u], a, ent_val, gr_diag, tab[gi + D], tab[gi + D + 1], lane);
                    // imports this level needs
                    int64_t polls = 0;
                    while (!failed &&
                           __hip_atomic_load(&progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < t0.w) {
                        ++polls;
                        if (polls > a.spin_limit ||
                            ((polls & 63) == 0 && __hip_atomic_load(a.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                            failed = true;
                            if (lane == 0) __hip_atomic_store(a.fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    if (a.trace && lane == 0 && gi < a.trace_steps)
                        a.trace[int64_t(t) * (2 + 2 * a.trace_steps) + 2 + 2 * gi] = int64_t(wall_clock64());
                    if (lane < m) row_compute<FV, VT, UNIT, MAXW>(cur, a, ent_val, lds, imp, xbuf, t0, m, lane);
                    for (int qq = lane + kWave; qq < m; qq += kWave) {   // levels wider than the wavefront
                        RowPre<FV, MAXW> extra;
                        row_prefetch<FV, UNIT, MAXW>(extra, a, ent_val, gr_diag, t0, tab[gi + 1], qq);
                        row_compute<FV, VT, UNIT, MAXW>(extra, a, ent_val, lds, imp, xbuf, t0, m, qq);
                    }
                    if (a.trace && lane == 0 && gi < a.trace_steps)
                        a.trace[int64_t(t) * (2 + 2 * a.trace_steps) + 3 + 2 * gi] = int64_t(wall_clock64());
                    // the next level reads these rows from LDS
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
        }
    }
    __syncthreads();
    if (a.trace && tid == 0) a.trace[int64_t(t) * (2 + 2 * a.trace_steps) + 1] = int64_t(wall_clock64());
    for (int32_t i0 = rb + tid; i0 < re; i0 += 4 * 128) {
        int32_t r[4];
#pragma unroll
        for (int u = 0; u < 4; u++) r[u] = i0 + u * 128 < re ? a.rows[i0 + u * 128] : -1;
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (r[u] >= 0) y[r[u]] = lds[i0 + u * 128 - rb];
    }
    if (tid == 0) {
        unsigned done = atomicAdd(&a.tickets[1], 1u);
        if (done == unsigned(a.ntiles) - 1) {   // every tile holds a ticket: reset for the next launch
            a.tickets[0] = 0;
            a.tickets[1] = 0;
        }
    }
}

template <typename T>
__global__ void k_gather(int64_t count, const int64_t* __restrict__ pos, const double* __restrict__ f,
                         T* __restrict__ out, T absent)
{
    for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < count; e += int64_t(gridDim.x) * 256) {
        int64_t p = pos[e];
        out[e] = p >= 0 ? T(f[p]) : absent;
    }
}

template <typename T>
__global__ void k_fill_sentinel(int64_t count, T* __restrict__ x)
{
    for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < count; e += int64_t(gridDim.x) * 256)
        x[e] = Sentinel<T>::value();
}

__global__ void k_diag_solve(int64_t n, const double* __restrict__ inv, const double* __restrict__ x,
                             double* __restrict__ y)
{
    for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < n; e += int64_t(gridDim.x) * 256)
        y[e] = inv[e] * x[e];
}

__global__ void k_diag_setup(int64_t n, const int32_t* __restrict__ S, const int32_t* __restrict__ I,
                             const double* __restrict__ V, double* __restrict__ inv)
{
    for (int64_t r = blockIdx.x * int64_t(256) + threadIdx.x; r < n; r += int64_t(gridDim.x) * 256) {
        double d = 0.0;   // A[i][i] of an absent entry reads 0 (compressed2D.hpp:536-547)
        for (int32_t p = S[r]; p < S[r + 1]; p++)
            if (I[p] == r) d = V[p];
        inv[r] = 1.0 / d;   // reciprocal (diagonal.hpp:44)
    }
}

static int grid_for(int64_t n, int64_t cap = 16384)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, cap)));
}

// ---------------------------------------------------------------------------
// host side of a compiled sweep
// ---------------------------------------------------------------------------
void SweepDev::upload(const SweepPlan& p, const double* fstore_dev, int ftype)
{
    plan_meta = p;
    plan_meta.rows.clear();   // keep only scalars on the host copy
    plan_meta.ent_src.clear();
    plan_meta.imports.clear();
    plan_meta.ent_val_pos.clear();
    plan_meta.gr_slot.clear();
    plan_meta.gr_len.clear();
    plan_meta.gr_export.clear();
    plan_meta.gr_diag_pos.clear();
    auto up32 = [](DevBuf<int32_t>& d, const std::vector<int32_t>& h) {
        d.reset(std::max<size_t>(h.size(), 1));
        d.upload(h.data(), h.size());
    };
    auto up64 = [](DevBuf<int64_t>& d, const std::vector<int64_t>& h) {
        d.reset(std::max<size_t>(h.size(), 1));
        d.upload(h.data(), h.size());
    };
    up32(tile_row_begin, p.tile_row_begin);
    up32(rows, p.rows);
    up32(tile_group_begin, p.tile_group_begin);
    up32(ent_src, p.ent_src);
    up32(tile_import_begin, p.tile_import_begin);
    up32(imports, p.imports);
    up64(ent_val_pos, p.ent_val_pos);
    up64(gr_diag_pos, p.gr_diag_pos);
    M4G_REQUIRE(p.nent < (int64_t(1) << 31) && p.n < (int64_t(1) << 31), M4G_ERR_INVALID, "sweep too large");
    std::vector<int4> tab(p.ngroups + 1);
    for (int64_t g = 0; g <= p.ngroups; g++) {
        tab[g].x = int(p.group_row_begin[g]);
        tab[g].y = int(p.group_ent_begin[g]);
        tab[g].z = g < p.ngroups ? p.group_width[g] : 0;
        tab[g].w = g < p.ngroups ? p.group_need[g] : 0;
    }
    group_tab.reset(tab.size());
    group_tab.upload(tab.data(), tab.size());
    std::vector<int2> meta(std::max<int64_t>(p.n, 1));
    max_len = 0;
    for (int64_t i = 0; i < p.n; i++) {
        M4G_REQUIRE(p.gr_slot[i] < 65536 && p.gr_len[i] < 65536, M4G_ERR_INVALID, "sweep row too long");
        meta[i].x = int(unsigned(p.gr_slot[i]) | (unsigned(p.gr_len[i]) << 16));
        meta[i].y = p.gr_export[i];
        max_len = std::max(max_len, int(p.gr_len[i]));
    }
    gr_meta.reset(meta.size());
    gr_meta.upload(meta.data(), meta.size());
    tickets.reset(2);
    M4G_HIP(hipMemsetAsync(tickets.ptr, 0, 2 * sizeof(unsigned), current_stream()));
    fault.reset(1);
    M4G_HIP(hipMemsetAsync(fault.ptr, 0, sizeof(int), current_stream()));
    refresh_values(fstore_dev, ftype);
}

void SweepDev::refresh_values(const double* f, int ftype)
{
    factor_type = ftype;
    hipStream_t s = current_stream();
    const int64_t nent = plan_meta.nent, n = plan_meta.n;
    if (ftype == M4G_F64) {
        ent_val64.reset(std::max<int64_t>(nent, 1));
        diag64.reset(std::max<int64_t>(n, 1));
        if (nent)
            hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(nent)), dim3(256), 0, s, nent, ent_val_pos.ptr, f,
                               ent_val64.ptr, 0.0);
        if (n)
            hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(256), 0, s, n, gr_diag_pos.ptr, f,
                               diag64.ptr, 1.0);
    } else {
        ent_val32.reset(std::max<int64_t>(nent, 1));
        diag32.reset(std::max<int64_t>(n, 1));
        if (nent)
            hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(nent)), dim3(256), 0, s, nent, ent_val_pos.ptr, f,
                               ent_val32.ptr, 0.0f);
        if (n)
            hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(n)), dim3(256), 0, s, n, gr_diag_pos.ptr, f,
                               diag32.ptr, 1.0f);
    }
    M4G_HIP(hipGetLastError());
}

void SweepDev::ensure_exchange(int vtype)
{
    size_t bytes = size_t(std::max<int64_t>(plan_meta.nexport, 1)) * (vtype == M4G_F64 ? 8 : 4);
    if (xbuf_type == vtype && xbuf.count == bytes) return;
    xbuf.reset(bytes);
    xbuf_type = vtype;
    hipStream_t s = current_stream();
    int64_t cnt = std::max<int64_t>(plan_meta.nexport, 1);
    if (vtype == M4G_F64)
        hipLaunchKernelGGL(k_fill_sentinel<double>, dim3(grid_for(cnt)), dim3(256), 0, s, cnt, (double*)xbuf.ptr);
    else
        hipLaunchKernelGGL(k_fill_sentinel<float>, dim3(grid_for(cnt)), dim3(256), 0, s, cnt, (float*)xbuf.ptr);
    M4G_HIP(hipGetLastError());
}

static int64_t spin_limit()
{
    static int64_t lim = [] {
        const char* e = std::getenv("M4G_SPIN_LIMIT");
        return e ? std::max<int64_t>(1, std::atoll(e)) : int64_t(1) << 20;   // about one second
    }();
    return lim;
}

template <typename FV, typename VT>
static void launch_sweep_t(SweepDev& d, SweepDev& other, const VT* v, VT* y, const int* skip, hipStream_t s)
{
    const SweepPlan& p = d.plan_meta;
    SweepArgs a;
    a.ntiles = p.ntiles;
    a.tile_row_begin = d.tile_row_begin.ptr;
    a.rows = d.rows.ptr;
    a.tile_group_begin = d.tile_group_begin.ptr;
    a.group_tab = d.group_tab.ptr;
    a.gr_meta = d.gr_meta.ptr;
    a.tile_import_begin = d.tile_import_begin.ptr;
    a.imports = d.imports.ptr;
    a.max_tile_rows = int32_t(p.max_tile_rows);
    a.max_tile_imports = int32_t(p.max_tile_imports);
    a.ent_src = d.ent_src.ptr;
    a.tickets = d.tickets.ptr;
    a.fault = d.fault.ptr;
    a.other_nexport = other.plan_meta.nexport;
    a.spin_limit = spin_limit();
    a.trace = nullptr;
    a.trace_steps = 0;
    static const char* trace_path = std::getenv("M4G_SWEEP_TRACE");
    static int traced = 0;
    DevBuf<int64_t> trace_buf;
    if (trace_path && traced < 4) {
        a.trace_steps = int32_t(p.max_steps);
        trace_buf.reset(size_t(p.ntiles) * (2 + 2 * p.max_steps));
        M4G_HIP(hipMemsetAsync(trace_buf.ptr, 0, trace_buf.count * sizeof(int64_t), s));
        a.trace = trace_buf.ptr;
    }
    const FV* ev = std::is_same<FV, double>::value ? (const FV*)d.ent_val64.ptr : (const FV*)d.ent_val32.ptr;
    const FV* dg = std::is_same<FV, double>::value ? (const FV*)d.diag64.ptr : (const FV*)d.diag32.ptr;
    // rows, imports, then the tile's group table (+ prefetch padding), 16-byte aligned
    size_t lds = size_t(p.max_tile_rows + p.max_tile_imports) * sizeof(VT);
    lds = (lds + 15) / 16 * 16 + size_t(p.max_steps + 16 + 1) * sizeof(int4);
    if (p.ntiles == 0) return;
    auto pick = [&](auto unit_tag) {
        constexpr bool U = decltype(unit_tag)::value;
        if (d.max_len <= 4) return k_sweep<FV, VT, U, 4>;
        if (d.max_len <= 8) return k_sweep<FV, VT, U, 8>;
        return k_sweep<FV, VT, U, 16>;
    };
    auto kern = p.unit ? pick(std::true_type()) : pick(std::false_type());
    if (lds + 64 > 65536) {
        // beyond 64 KB of LDS per workgroup (with the static group table) must be requested
        M4G_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    }
    hipLaunchKernelGGL(kern, dim3(p.ntiles), dim3(128), lds, s, a, ev, dg, v, y, (VT*)d.xbuf.ptr,
                       (VT*)other.xbuf.ptr, skip);
    M4G_HIP(hipGetLastError());
    if (a.trace) {
        // trace file: header (ntiles, steps) then the stamps of every tile
        std::vector<int64_t> h(trace_buf.count);
        trace_buf.download(h.data(), h.size());
        std::string path = std::string(trace_path) + "." + std::to_string(traced++);
        if (FILE* f = std::fopen(path.c_str(), "wb")) {
            int64_t hdr[2] = {p.ntiles, p.max_steps};
            std::fwrite(hdr, sizeof(hdr), 1, f);
            std::fwrite(h.data(), sizeof(int64_t), h.size(), f);
            std::fclose(f);
        }
    }
}

// Forward sweep on `first` then backward on `second`, v -> tmp -> y.
void run_sweep_pair(SweepDev& first, SweepDev& second, const void* v, void* tmp, void* y, int vtype,
                    const int* skip, hipStream_t s)
{
    first.ensure_exchange(vtype);
    second.ensure_exchange(vtype);
    M4G_REQUIRE(first.factor_type == second.factor_type, M4G_ERR_INVALID, "sweep factor types differ");
    if (first.factor_type == M4G_F64) {
        M4G_REQUIRE(vtype == M4G_F64, M4G_ERR_INVALID, "double factors need double vectors");
        launch_sweep_t<double, double>(first, second, (const double*)v, (double*)tmp, skip, s);
        launch_sweep_t<double, double>(second, first, (const double*)tmp, (double*)y, skip, s);
    } else if (vtype == M4G_F64) {
        launch_sweep_t<float, double>(first, second, (const double*)v, (double*)tmp, skip, s);
        launch_sweep_t<float, double>(second, first, (const double*)tmp, (double*)y, skip, s);
    } else {
        launch_sweep_t<float, float>(first, second, (const float*)v, (float*)tmp, skip, s);
        launch_sweep_t<float, float>(second, first, (const float*)tmp, (float*)y, skip, s);
    }
}

int check_sweep_fault(SweepDev& d)
{
    int h = 0;
    M4G_HIP(hipMemcpyAsync(&h, d.fault.ptr, sizeof(int), hipMemcpyDeviceToHost, current_stream()));
    M4G_HIP(hipStreamSynchronize(current_stream()));
    return h;
}

// ---------------------------------------------------------------------------
// PC construction
// ---------------------------------------------------------------------------
static double env_double(const char* name, double dflt)
{
    const char* e = std::getenv(name);
    return e ? std::atof(e) : dflt;
}

int64_t sweep_tile_cap(int vtype)
{
    // the tile's vector entries live in LDS: stay within the per-workgroup limit
    // (minus the kernel's static shared words)
    int lds = 65536;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lds = std::min(v, 160 * 1024);
    }
    const int64_t elem = vtype == M4G_F64 ? 8 : 4;
    int64_t cap = std::min<int64_t>(vtype == M4G_F64 ? 8192 : 16384, (lds - 4400) / elem);
    const char* e = std::getenv("M4G_TILE_ROWS");
    if (e) cap = std::min<int64_t>(cap, std::max<int64_t>(64, std::atoll(e)));
    return cap;
}

SweepPlan plan_sweep(const SweepRows& rows, const GridInfo& g, int64_t cap, int vtype)
{
    // a tile's rows and imports share the workgroup's LDS
    int lds = 65536, dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lds = std::min(v, 160 * 1024);
    }
    // sized for double vectors, the widest the sweeps run with
    const int64_t elem = 8;
    (void)vtype;
    SweepPlan p;
    for (;;) {
        TileShape shape = choose_tiles(rows, g, cap, env_double("M4G_SWEEP_HOP_US", 1.5),
                                       env_double("M4G_SWEEP_STEP_US", 0.15));
        p = build_sweep_plan(rows, g, shape, cap);
        if ((p.max_tile_rows + p.max_tile_imports) * elem + (p.max_steps + 17) * 16 + 64 <= lds || cap <= 64) break;
        cap = cap * 3 / 4;
    }
    M4G_REQUIRE((p.max_tile_rows + p.max_tile_imports) * elem + (p.max_steps + 17) * 16 + 64 <= lds, M4G_ERR_INVALID,
                "sweep tile does not fit in LDS");
    if (std::getenv("M4G_VALIDATE_PLANS")) {
        std::string msg = validate_sweep_plan(p, rows);
        M4G_REQUIRE(msg.empty(), M4G_ERR_LOGIC, "sweep plan invalid: " + msg);
    }
    return p;
}

}  // namespace m4g

using namespace m4g;

m4g_pc_s::~m4g_pc_s() = default;

static void factorize_levels(m4g_pc_s& P, const SweepRows& lower_dag, const m4g_matrix_s& A)
{
    // natural-order level sets of the lower DAG
    int64_t nlev = 0;
    std::vector<int32_t> lvl = natural_levels(lower_dag, &nlev);
    std::vector<int64_t> lptr(nlev + 1, 0);
    for (int64_t r = 0; r < P.n; r++) lptr[lvl[r] + 1]++;
    for (int64_t l = 0; l < nlev; l++) lptr[l + 1] += lptr[l];
    std::vector<int32_t> lrows(P.n);
    {
        std::vector<int64_t> f(lptr.begin(), lptr.end() - 1);
        for (int64_t r = 0; r < P.n; r++) lrows[f[lvl[r]]++] = int32_t(r);
    }
    DevBuf<int32_t> d_rows(std::max<int64_t>(P.n, 1));
    d_rows.upload(lrows.data(), P.n);
    DevBuf<double> invdiag(std::max<int64_t>(P.n, 1));
    hipStream_t s = current_stream();
    if (P.kind == M4G_PC_IC0) {
        std::vector<int32_t> lcol(lower_dag.col.begin(), lower_dag.col.end());
        DevBuf<int64_t> d_lptr(P.n + 1), d_lpos(std::max<size_t>(lower_dag.val_pos.size(), 1)),
            d_diag(std::max<int64_t>(P.n, 1));
        DevBuf<int32_t> d_lcol(std::max<size_t>(lcol.size(), 1));
        d_lptr.upload(lower_dag.ptr.data(), P.n + 1);
        d_lpos.upload(lower_dag.val_pos.data(), lower_dag.val_pos.size());
        d_lcol.upload(lcol.data(), lcol.size());
        d_diag.upload(lower_dag.diag_pos.data(), P.n);
        for (int64_t l = 0; l < nlev; l++) {
            int64_t cnt = lptr[l + 1] - lptr[l];
            hipLaunchKernelGGL(k_ic0_level, dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, d_rows.ptr + lptr[l],
                               d_lptr.ptr, d_lcol.ptr, d_lpos.ptr, d_diag.ptr, P.fstore.ptr, invdiag.ptr);
        }
        M4G_HIP(hipGetLastError());
        M4G_HIP(hipStreamSynchronize(s));
    } else {
        DevBuf<int64_t> d_diag(std::max<int64_t>(P.n, 1));
        d_diag.upload(P.diag_pos.data(), P.n);
        for (int64_t l = 0; l < nlev; l++) {
            int64_t cnt = lptr[l + 1] - lptr[l];
            hipLaunchKernelGGL(k_ilu0_level, dim3((cnt + 255) / 256), dim3(256), 0, s, cnt, d_rows.ptr + lptr[l],
                               A.starts.ptr, A.indices.ptr, d_diag.ptr, P.fstore.ptr, invdiag.ptr);
        }
        M4G_HIP(hipGetLastError());
        M4G_HIP(hipStreamSynchronize(s));
    }
    P.factor_levels = nlev;
}

extern "C" {

int m4g_pc_create(m4g_pc* out, m4g_matrix A, int kind, int ftype)
{
    return guard([&] {
        M4G_REQUIRE(out && A, M4G_ERR_INVALID, "m4g_pc_create: null argument");
        M4G_REQUIRE(kind >= M4G_PC_IDENTITY && kind <= M4G_PC_ILU0, M4G_ERR_INVALID, "m4g_pc_create: unknown kind");
        M4G_REQUIRE(ftype == M4G_F64 || ftype == M4G_F32, M4G_ERR_INVALID, "m4g_pc_create: unknown value type");
        M4G_REQUIRE(A->nrows == A->ncols, M4G_ERR_NOT_SQUARE, "Matrix is not square");
        auto P = std::make_unique<m4g_pc_s>();
        P->kind = kind;
        P->factor_type = ftype;
        P->n = A->nrows;
        hipStream_t s = current_stream();
        const int64_t n = A->nrows;
        if (kind == M4G_PC_DIAGONAL) {
            P->inv_diag.reset(std::max<int64_t>(n, 1));
            if (n)
                hipLaunchKernelGGL(k_diag_setup, dim3(grid_for(n)), dim3(256), 0, s, n, A->starts.ptr,
                                   A->indices.ptr, A->values.ptr, P->inv_diag.ptr);
            M4G_HIP(hipGetLastError());
        } else if (kind == M4G_PC_IC0 || kind == M4G_PC_ILU0) {
            GridInfo g = infer_grid(n, A->h_starts.data(), A->h_indices.data());
            P->grid = g;
            SweepRows lo, up;
            if (kind == M4G_PC_IC0) {
                std::vector<int64_t> a_pos;
                M4G_REQUIRE(upper_pattern(n, A->h_starts.data(), A->h_indices.data(), P->fS, P->fI, a_pos),
                            M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                M4G_REQUIRE(ic0_rows(n, P->fS, P->fI, lo, up), M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                // U = upper(A)
                P->fstore.reset(std::max<size_t>(a_pos.size(), 1));
                P->a_pos.reset(std::max<size_t>(a_pos.size(), 1));
                P->a_pos.upload(a_pos.data(), a_pos.size());
                if (!a_pos.empty())
                    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(a_pos.size())), dim3(256), 0, s,
                                       int64_t(a_pos.size()), P->a_pos.ptr, A->values.ptr, P->fstore.ptr, 0.0);
                M4G_HIP(hipGetLastError());
                factorize_levels(*P, lo, *A);
            } else {
                P->fS = A->h_starts;
                P->fI = A->h_indices;
                M4G_REQUIRE(ilu0_rows(n, A->h_starts.data(), A->h_indices.data(), lo, up, P->diag_pos),
                            M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                P->fstore.reset(std::max<int64_t>(A->nnz, 1));
                if (A->nnz)
                    M4G_HIP(hipMemcpyAsync(P->fstore.ptr, A->values.ptr, A->nnz * sizeof(double),
                                           hipMemcpyDeviceToDevice, s));
                factorize_levels(*P, lo, *A);
            }
            int64_t cap = sweep_tile_cap(ftype == M4G_F64 ? M4G_F64 : M4G_F32);
            P->lower.upload(plan_sweep(lo, g, cap, ftype), P->fstore.ptr, ftype);
            P->upper.upload(plan_sweep(up, g, cap, ftype), P->fstore.ptr, ftype);
            P->tmp_bytes = 0;
        }
        M4G_HIP(hipStreamSynchronize(s));
        *out = P.release();
        return M4G_OK;
    });
}

int m4g_pc_refactor(m4g_pc P, m4g_matrix A)
{
    return guard([&] {
        M4G_REQUIRE(P && A, M4G_ERR_INVALID, "m4g_pc_refactor: null argument");
        M4G_REQUIRE(A->nrows == P->n && A->ncols == P->n, M4G_ERR_SIZE, "m4g_pc_refactor: incompatible size");
        hipStream_t s = current_stream();
        const int64_t n = P->n;
        if (P->kind == M4G_PC_DIAGONAL) {
            if (n)
                hipLaunchKernelGGL(k_diag_setup, dim3(grid_for(n)), dim3(256), 0, s, n, A->starts.ptr,
                                   A->indices.ptr, A->values.ptr, P->inv_diag.ptr);
            M4G_HIP(hipGetLastError());
        } else if (P->kind == M4G_PC_IC0 || P->kind == M4G_PC_ILU0) {
            SweepRows lo, up;
            if (P->kind == M4G_PC_IC0) {
                std::vector<int64_t> US, UI, apos;
                M4G_REQUIRE(upper_pattern(n, A->h_starts.data(), A->h_indices.data(), US, UI, apos) && US == P->fS &&
                                UI == P->fI,
                            M4G_ERR_INVALID, "m4g_pc_refactor: pattern differs");
                M4G_REQUIRE(ic0_rows(n, P->fS, P->fI, lo, up), M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                const int64_t nu = P->fS[n];
                if (nu)
                    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(nu)), dim3(256), 0, s, nu, P->a_pos.ptr,
                                       A->values.ptr, P->fstore.ptr, 0.0);
            } else {
                M4G_REQUIRE(A->h_starts == P->fS && A->h_indices == P->fI, M4G_ERR_INVALID,
                            "m4g_pc_refactor: pattern differs");
                std::vector<int64_t> dpos;
                M4G_REQUIRE(ilu0_rows(n, P->fS.data(), P->fI.data(), lo, up, dpos), M4G_ERR_MISSING_DIAGONAL,
                            "missing diagonal");
                if (A->nnz)
                    M4G_HIP(hipMemcpyAsync(P->fstore.ptr, A->values.ptr, A->nnz * sizeof(double),
                                           hipMemcpyDeviceToDevice, s));
            }
            factorize_levels(*P, lo, *A);
            P->lower.refresh_values(P->fstore.ptr, P->factor_type);
            P->upper.refresh_values(P->fstore.ptr, P->factor_type);
            if (P->adj_ready) {
                P->adj_lower.refresh_values(P->fstore.ptr, P->factor_type);
                P->adj_upper.refresh_values(P->fstore.ptr, P->factor_type);
            }
        }
        M4G_HIP(hipStreamSynchronize(s));
        return M4G_OK;
    });
}

int m4g_pc_destroy(m4g_pc P)
{
    return guard([&] {
        if (P) M4G_HIP(hipStreamSynchronize(current_stream()));
        delete P;
        return M4G_OK;
    });
}

}  // extern "C"

namespace m4g {

void* pc_tmp(m4g_pc_s& P, int vtype)
{
    size_t bytes = size_t(std::max<int64_t>(P.n, 1)) * (vtype == M4G_F64 ? 8 : 4);
    if (P.tmp.count != bytes) P.tmp.reset(bytes);
    return P.tmp.ptr;
}

void pc_apply(m4g_pc_s& P, const void* x, void* y, int vtype, bool adjoint, hipStream_t s, const int* skip)
{
    const int64_t n = P.n;
    switch (P.kind) {
    case M4G_PC_IDENTITY:
        if (n && x != y) M4G_HIP(hipMemcpyAsync(y, x, n * (vtype == M4G_F64 ? 8 : 4), hipMemcpyDeviceToDevice, s));
        return;
    case M4G_PC_DIAGONAL:
        M4G_REQUIRE(vtype == M4G_F64, M4G_ERR_INVALID, "diagonal PC: double vectors only");
        if (n)
            hipLaunchKernelGGL(k_diag_solve, dim3(grid_for(n)), dim3(256), 0, s, n, P.inv_diag.ptr,
                               (const double*)x, (double*)y);
        M4G_HIP(hipGetLastError());
        return;
    default: break;
    }
    if (n == 0) return;
    void* tmp = pc_tmp(P, vtype);
    if (adjoint && P.kind == M4G_PC_ILU0) {
        if (!P.adj_ready) {
            SweepRows lo, up;
            M4G_REQUIRE(ilu0_adjoint_rows(n, P.fS.data(), P.fI.data(), lo, up), M4G_ERR_MISSING_DIAGONAL,
                        "missing diagonal");
            int64_t cap = sweep_tile_cap(P.factor_type);
            P.adj_lower.upload(plan_sweep(lo, P.grid, cap, P.factor_type), P.fstore.ptr, P.factor_type);
            P.adj_upper.upload(plan_sweep(up, P.grid, cap, P.factor_type), P.fstore.ptr, P.factor_type);
            P.adj_ready = true;
        }
        run_sweep_pair(P.adj_lower, P.adj_upper, x, tmp, y, vtype, skip, s);
    } else {
        // IC(0): adjoint_solve == solve (ic_0.hpp:97-103)
        run_sweep_pair(P.lower, P.upper, x, tmp, y, vtype, skip, s);
    }
}

void pc_apply_skip(m4g_pc_s& P, const void* x, void* y, int vtype, const int* skip, hipStream_t s)
{
    pc_apply(P, x, y, vtype, false, s, skip);
}

}  // namespace m4g

extern "C" {

static int pc_solve_common(m4g_pc P, m4g_vector x, m4g_vector y, bool adjoint)
{
    return guard([&] {
        M4G_REQUIRE(P && x && y, M4G_ERR_INVALID, "m4g_pc_solve: null argument");
        M4G_REQUIRE(x->n == P->n && y->n == P->n, M4G_ERR_SIZE, "m4g_pc_solve: incompatible size");
        M4G_REQUIRE(x->type == y->type, M4G_ERR_INVALID, "m4g_pc_solve: vector types differ");
        M4G_REQUIRE(x->ptr != y->ptr || P->n == 0 || P->kind <= M4G_PC_DIAGONAL, M4G_ERR_INVALID,
                    "m4g_pc_solve: x and y alias");
        pc_apply(*P, x->ptr, y->ptr, x->type, adjoint, current_stream());
        if (P->kind >= M4G_PC_IC0 && P->n) {
            SweepDev& d = adjoint && P->kind == M4G_PC_ILU0 ? P->adj_lower : P->lower;
            M4G_REQUIRE(check_sweep_fault(d) == 0 && check_sweep_fault(adjoint && P->kind == M4G_PC_ILU0 ? P->adj_upper : P->upper) == 0,
                        M4G_ERR_TIMEOUT, "sweep wait timed out");
        }
        return M4G_OK;
    });
}

int m4g_pc_solve(m4g_pc P, m4g_vector x, m4g_vector y) { return pc_solve_common(P, x, y, false); }

int m4g_pc_apply_async(m4g_pc P, m4g_vector x, m4g_vector y)
{
    return guard([&] {
        M4G_REQUIRE(P && x && y, M4G_ERR_INVALID, "m4g_pc_apply_async: null argument");
        M4G_REQUIRE(x->n == P->n && y->n == P->n && x->type == y->type, M4G_ERR_SIZE,
                    "m4g_pc_apply_async: incompatible vectors");
        M4G_REQUIRE(x->ptr != y->ptr || P->n == 0, M4G_ERR_INVALID, "m4g_pc_apply_async: x and y alias");
        pc_apply(*P, x->ptr, y->ptr, x->type, false, current_stream());
        return M4G_OK;
    });
}
int m4g_pc_adjoint_solve(m4g_pc P, m4g_vector x, m4g_vector y) { return pc_solve_common(P, x, y, true); }

int m4g_pc_factor_nnz(m4g_pc P, int which, int64_t* nnz)
{
    return guard([&] {
        M4G_REQUIRE(P && nnz && (which == 0 || which == 1), M4G_ERR_INVALID, "m4g_pc_factor_nnz: bad argument");
        M4G_REQUIRE(P->kind >= M4G_PC_IC0, M4G_ERR_INVALID, "m4g_pc_factor_nnz: no factors");
        if (P->kind == M4G_PC_IC0) {
            *nnz = P->fS[P->n];
        } else {
            int64_t lower = 0;
            for (int64_t r = 0; r < P->n; r++)
                for (int64_t p = P->fS[r]; p < P->fS[r + 1]; p++) lower += P->fI[p] < r;
            *nnz = which == 0 ? lower : P->fS[P->n] - lower;
        }
        return M4G_OK;
    });
}

int m4g_pc_download_factor(m4g_pc P, int which, int64_t* starts, int64_t* indices, double* values)
{
    return guard([&] {
        M4G_REQUIRE(P && (which == 0 || which == 1), M4G_ERR_INVALID, "m4g_pc_download_factor: bad argument");
        M4G_REQUIRE(P->kind >= M4G_PC_IC0, M4G_ERR_INVALID, "m4g_pc_download_factor: no factors");
        const int64_t n = P->n, fn = P->fS[n];
        std::vector<double> f(fn);
        if (fn) {
            M4G_HIP(hipMemcpyAsync(f.data(), P->fstore.ptr, fn * sizeof(double), hipMemcpyDeviceToHost,
                                   current_stream()));
            M4G_HIP(hipStreamSynchronize(current_stream()));
        }
        if (P->factor_type == M4G_F32)   // values as stored: rounded to float
            for (auto& v : f) v = double(float(v));
        std::vector<int64_t> S(n + 1, 0), I;
        std::vector<double> V;
        if (P->kind == M4G_PC_IC0 && which == 1) {
            S = P->fS;
            I = P->fI;
            V = f;
        } else if (P->kind == M4G_PC_IC0) {
            // L = trans(U), rows sorted
            for (int64_t p = 0; p < fn; p++) S[P->fI[p] + 1]++;
            for (int64_t r = 0; r < n; r++) S[r + 1] += S[r];
            I.assign(fn, 0);
            V.assign(fn, 0);
            std::vector<int64_t> fill(S.begin(), S.end() - 1);
            for (int64_t r = 0; r < n; r++)
                for (int64_t p = P->fS[r]; p < P->fS[r + 1]; p++) {
                    int64_t d = fill[P->fI[p]]++;
                    I[d] = r;
                    V[d] = f[p];
                }
        } else {
            for (int64_t r = 0; r < n; r++) {
                for (int64_t p = P->fS[r]; p < P->fS[r + 1]; p++) {
                    bool low = P->fI[p] < r;
                    if ((which == 0) == low) {
                        I.push_back(P->fI[p]);
                        V.push_back(f[p]);
                    }
                }
                S[r + 1] = int64_t(I.size());
            }
        }
        if (starts) std::copy(S.begin(), S.end(), starts);
        if (indices) std::copy(I.begin(), I.end(), indices);
        if (values) std::copy(V.begin(), V.end(), values);
        return M4G_OK;
    });
}

int m4g_pc_stats_get(m4g_pc P, m4g_pc_stats* st)
{
    return guard([&] {
        M4G_REQUIRE(P && st, M4G_ERR_INVALID, "m4g_pc_stats_get: null argument");
        std::memset(st, 0, sizeof(*st));
        if (P->kind < M4G_PC_IC0) return M4G_OK;
        const SweepPlan& l = P->lower.plan_meta;
        const SweepPlan& u = P->upper.plan_meta;
        st->levels_lower = l.levels;
        st->levels_upper = u.levels;
        st->tiles_lower = l.ntiles;
        st->tiles_upper = u.ntiles;
        st->steps_lower = l.max_steps;
        st->steps_upper = u.max_steps;
        st->grid_nx = P->grid.nx;
        st->grid_ny = P->grid.ny;
        st->grid_nz = P->grid.nz;
        if (l.shape.grid) {
            st->tile_nx = l.shape.e3;
            st->tile_ny = l.shape.e2;
            st->tile_nz = l.shape.e1;
        }
        return M4G_OK;
    });
}

}  // extern "C"
