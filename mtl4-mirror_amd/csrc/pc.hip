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
#include <future>
#include <chrono>
#include <cstdio>
#include <string>
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
                                                   const int32_t* __restrict__ lpos,
                                                   const int32_t* __restrict__ diag_pos, double* __restrict__ U,
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

// The same row sequence with one wavefront per row (rows of at most 64
// entries): lane q holds entry q of row i in a register.  For each lower entry
// k (ascending) the lane of k scales it, and every later lane subtracts a_ik
// times its column's entry in row k's U part (0 where row k has none), found
// among that U part spread over the lanes; the next k's U part is loaded while
// the current one is applied.  Each entry sees the same operations in the
// same order as in k_ilu0_level.
__global__ __launch_bounds__(256) void k_ilu0_level_wave(int64_t count, const int32_t* __restrict__ level_rows,
                                                         const int32_t* __restrict__ S,
                                                         const int32_t* __restrict__ I,
                                                         const int64_t* __restrict__ diag_pos,
                                                         double* __restrict__ LU, double* __restrict__ invdiag)
{
    const int lane = threadIdx.x & 63;
    const int64_t t = blockIdx.x * int64_t(4) + (threadIdx.x >> 6);
    if (t >= count) return;   // the whole wavefront
    const int32_t i = level_rows[t];
    const int32_t s0 = S[i], len = S[i + 1] - s0;
    const int32_t dl = int32_t(diag_pos[i]) - s0;   // lower entries before the diagonal
    const bool mine = lane < len;
    const int32_t col = mine ? I[s0 + lane] : -1;
    double v = mine ? LU[s0 + lane] : 0.0;
    // lane q < dl: bounds of row col_q's U part and its inverted diagonal
    int32_t kb = 0, ke = 0;
    double kinv = 0.0;
    if (lane < dl) {
        kb = int32_t(diag_pos[col]) + 1;
        ke = S[col + 1];
        kinv = invdiag[col];
    }
    int32_t cn = -1;
    double un = 0.0;
    int mn = 0;
    auto load_u = [&](int kk) {
        const int32_t b = __shfl(kb, kk, 64), e = __shfl(ke, kk, 64);
        mn = e - b;
        cn = lane < mn ? I[b + lane] : -1;
        un = lane < mn ? LU[b + lane] : 0.0;
    };
    if (dl > 0) load_u(0);
    for (int kk = 0; kk < dl; kk++) {
        const int32_t ck = cn;
        const double uk = un;
        const int m_n = mn;
        if (kk + 1 < dl) load_u(kk + 1);
        const double kinv_k = __shfl(kinv, kk, 64);
        if (lane == kk) v = v * kinv_k;                         // ilu_0.hpp:70
        const double aik = __shfl(v, kk, 64);
        double lukj = 0.0;
        for (int m = 0; m < m_n; m++) {
            const int32_t cm = __shfl(ck, m, 64);
            const double um = __shfl(uk, m, 64);
            if (cm == col) lukj = um;
        }
        if (lane > kk && mine) v = v - aik * lukj;             // ilu_0.hpp:74
    }
    if (mine) {
        if (lane == dl) {
            const double inv = 1.0 / v;                         // ilu_0.hpp:77 and invert_diagonal
            invdiag[i] = inv;
            v = inv;
        }
        LU[s0 + lane] = v;
    }
}

// ---------------------------------------------------------------------------
// sweep kernel
// ---------------------------------------------------------------------------
// Packed record of one piece (a run of at most W entries of a row).  The
// record is a whole number of 16-byte chunks; the records of a group of m
// pieces are stored chunk-interleaved (chunk c of piece q at c*m + q), so
// each chunk load of a wavefront reads one contiguous span.
template <typename FV, int W>
struct alignas(16) SweepRec {
    uint32_t row;       // LDS byte offset of the row's value
    int32_t need;       // imports of the tile the piece's group needs (same for the whole group)
    uint32_t src[W];    // LDS byte offset of each entry's value (padding: the zero)
    FV diag;            // inverse diagonal; 1 for a unit diagonal or a piece before the last
    FV val[W];          // 0 for padding
};
template <typename FV, int W> constexpr int kRecChunks = int(sizeof(SweepRec<FV, W>) / 16);
// Records reach the compute wavefront through a ring of 2*B group slots in
// LDS filled by two loader wavefronts: loader k writes the batches of B
// groups b = k, k+2, ... into ring half k, with the next batch's loads in
// flight meanwhile (two register buffers of B*C chunk loads, within the 63
// loads a wavefront can count).
template <typename FV, int W> constexpr int kLoadBatch =
    kRecChunks<FV, W> <= 3 ? 8 : (kRecChunks<FV, W> <= 4 ? 6 : (kRecChunks<FV, W> <= 6 ? 4 : 2));

// Host-side mirror of the record size (bytes) for a factor type and width.
inline int rec_chunks(int ftype, int width)
{
    const int fb = ftype == M4G_F64 ? 8 : 4;
    int off = 8 + 4 * width;                 // row, need, src[]
    off = (off + fb - 1) / fb * fb;          // diag alignment
    off += fb * (1 + width);                 // diag, val[]
    return (off + 15) / 16;
}
inline int load_batch(int chunks) { return chunks <= 3 ? 8 : (chunks <= 4 ? 6 : (chunks <= 6 ? 4 : 2)); }

template <typename FV, int W>
union RecBits {
    SweepRec<FV, W> r;
    int4 c[kRecChunks<FV, W>];
};

struct SweepArgs {
    int64_t ntiles;
    const int32_t* tile_row_begin;    // [ntiles+1] into rows
    const int32_t* rows;              // [n] rows of each tile, natural order
    const int32_t* row_slot;          // [n] LDS slot of each entry of rows
    const int32_t* tile_group_begin;  // [ntiles+1] into the group table
    const int4* group_tab;            // [ngroups+1]: first piece, imports needed, exports before, 0
    const int32_t* tile_import_begin; // [ntiles+1] into imports
    const int32_t* imports;           // exchange slot of each import
    const int32_t* tile_export_begin; // [ntiles+1] exchange slots of each tile's exports
    const int32_t* xlist;             // [nexport] LDS slot of the row behind each export
    int32_t max_tile_rows;            // LDS rows reserved per tile (imports follow them)
    int32_t max_tile_imports;         // LDS import slots reserved per tile
    int32_t max_tile_exports;
    int32_t max_steps;                // most groups of a tile
    unsigned* tickets;                // [0] next ticket, [1] finished tiles
    int* fault;                       // set when a bounded wait runs out
    int64_t other_nexport;            // size of the other direction's exchange buffer
    int64_t spin_limit;
    const void* dot_with;             // w (vector type), or null: no dot
    double* dot_partials;             // [ntiles] per tile: sum of w_r * y_r over its rows
};

// Workgroup roles: wavefront 0 imports, 1 computes, 2 and 3 load records,
// 4 exports; all five move rows between HBM and LDS at the start and end.
constexpr int kSweepWaves = 5;
constexpr int kSweepThreads = kSweepWaves * kWave;
constexpr int kImportDepth = 4;       // imports each importer lane polls per window
constexpr int kFillUnroll = 26;       // row values each thread moves at once between HBM and LDS
constexpr int kFillLists = 8;         // import and export list entries per thread in the first fill pass

// Byte offsets of the dynamic LDS regions of a sweep workgroup: header (the
// workgroup's counters), values (rows, imports, one zero), group table, export
// list (16-bit slots), record ring (2 * B group slots).  The workgroup has no
// static LDS, so the byte offsets in the records address LDS directly.
constexpr int kSweepHeader = 64;
struct SweepLds {
    size_t vals, tab, xlist, ring, total;
};
__host__ __device__ inline SweepLds sweep_lds_layout(int rows, int imports, int steps, int exports, int vbytes,
                                                     int C, int B)
{
    SweepLds l;
    l.vals = kSweepHeader;
    l.tab = l.vals + (size_t(rows + imports + 1) * vbytes + 15) / 16 * 16;
    l.xlist = l.tab + size_t(steps + 3 * B + 2) * 16;
    l.ring = l.xlist + (size_t(exports) * 2 + 15) / 16 * 16;
    l.total = l.ring + size_t(2 * B * C * kWave) * 16;
    return l;
}

// Native 4 x 32-bit vector: copies of it stay in registers (HIP's int4 is a
// struct, whose copies from global memory become private-memory memcpys).
typedef int i32x4 __attribute__((ext_vector_type(4)));

// Loads the records of batch b (groups b*B .. b*B+B-1 of the tile) into
// registers: lane q takes piece q of each group.  Groups past the tile's end
// are empty (their table entries repeat its end) and read the record at the
// end; lanes past a group's size reread its last record.  So every load is
// issued and in bounds.
template <int C, int B>
__device__ __attribute__((always_inline)) inline void load_batch_regs(i32x4 (&buf)[B][C],
                                                                      const i32x4* __restrict__ recs, const int4* tab,
                                                                      int b, int lane)
{
#pragma unroll
    for (int j = 0; j < B; j++) {
        const int gi = b * B + j;
        const int first = tab[gi].x, m = tab[gi + 1].x - first;
        const int q = max(0, min(lane, m - 1));
        const i32x4* base = recs + int64_t(first) * C + q;
#pragma unroll
        for (int c = 0; c < C; c++) buf[j][c] = base[c * m];
    }
}

// Writes a batch into its ring half and publishes it.
template <int C, int B>
__device__ __attribute__((always_inline)) inline void store_batch_lds(const i32x4 (&buf)[B][C], i32x4* ring,
                                                                      int half, int b, int* half_batch, int lane)
{
#pragma unroll
    for (int j = 0; j < B; j++) {
        i32x4* slot = ring + (half * B + j) * (C * kWave);
#pragma unroll
        for (int c = 0; c < C; c++) slot[c * kWave + lane] = buf[j][c];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane == 0) __hip_atomic_store(half_batch + half, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// LDS at a byte offset.  The sweep workgroup has no static LDS, so its
// dynamic LDS starts at LDS address 0 and record offsets are addresses.
template <typename T>
__device__ inline __attribute__((address_space(3))) T* lds_at(uint32_t off)
{
    return reinterpret_cast<__attribute__((address_space(3))) T*>(static_cast<uintptr_t>(off));
}

__device__ inline void wait_lds_at_least(int* p, int target)
{
    while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < target) __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// One workgroup per tile (ticket order), five wavefronts with fixed roles,
// communicating through LDS counters:
//  0  importer: polls the tile's imports from earlier tiles (exchange buffer
//     -> LDS) in first-use order; publishes how many have arrived;
//  1  compute: walks the tile's groups from LDS alone; once a group's
//     records and the imports it needs are present, computes one piece per
//     lane:  rr = (slot); rr -= L_rj * y_j in stored order; rr *= d_r
//     (lower_trisolve.hpp:191-209 / 172-189, upper_trisolve.hpp:159-176);
//     publishes how many groups are finished;
//  2,3 loaders: stream the tile's piece records into the LDS ring;
//  4  exporter: stores the finished rows that later tiles read into the
//     exchange buffer.
// The compute wavefront so issues no global memory operation: its pace is
// set by LDS and arithmetic latency only.
template <typename FV, typename VT, bool UNIT, int W>
__global__ __launch_bounds__(kSweepThreads) void k_sweep(SweepArgs a, const int4* __restrict__ recs,
                                                         const VT* __restrict__ v, VT* __restrict__ y, VT* xbuf,
                                                         VT* xbuf_other, const int* skip)
{
    constexpr int C = kRecChunks<FV, W>;
    constexpr int B = kLoadBatch<FV, W>;
    if (skip && *skip) return;   // the solve this sweep belongs to has finished
    extern __shared__ unsigned char lds_raw[];
    // records address LDS from 0: the dynamic segment must start there (this
    // kernel has no static LDS; the compiler folds the test away)
    if (static_cast<uint32_t>(reinterpret_cast<uintptr_t>(
            (__attribute__((address_space(3))) unsigned char*)(lds_raw))) != 0) {
        if (threadIdx.x == 0) __hip_atomic_store(a.fault, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const SweepLds L = sweep_lds_layout(a.max_tile_rows, a.max_tile_imports, a.max_steps, a.max_tile_exports,
                                        int(sizeof(VT)), C, B);
    VT* lds = reinterpret_cast<VT*>(lds_raw + L.vals);   // rows, imports, then a zero that padding entries read
    VT* imp = lds + a.max_tile_rows;
    int4* tab = reinterpret_cast<int4*>(lds_raw + L.tab);
    uint16_t* xl = reinterpret_cast<uint16_t*>(lds_raw + L.xlist);
    int4* ring = reinterpret_cast<int4*>(lds_raw + L.ring);
    int* hdr = reinterpret_cast<int*>(lds_raw);
    int& tile_sh = hdr[0];
    int& progress = hdr[1];       // imports in LDS
    int& consumed = hdr[2];       // groups whose records the compute wavefront no longer needs
    int& finished = hdr[3];       // groups the compute wavefront has finished
    int* half_batch = hdr + 4;    // batch held by each ring half
    const int tid = threadIdx.x, lane = tid % kWave;
    // the role of a wavefront is uniform: keep its branches scalar, so the
    // roles' code is not linearised into one another
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    // Tiles go out in ticket order, so a tile's dependencies are held by
    // workgroups that have started.  Workgroup 0 is dispatched first and
    // takes tile 0 without a ticket: the first tile, whose start-up is on the
    // critical path, then issues its loads without waiting for an atomic
    // that all workgroups contend for.
    if (tid == 0) {
        if (blockIdx.x != 0) tile_sh = 1 + int(atomicAdd(&a.tickets[0], 1u));
        progress = 0;
        consumed = 0;
        finished = 0;
        half_batch[0] = -1;
        half_batch[1] = -1;
        lds[a.max_tile_rows + a.max_tile_imports] = VT(0);
    }
    if (blockIdx.x != 0) __syncthreads();
    const int t = blockIdx.x == 0 ? 0 : tile_sh;
    // the other direction's exchange buffer is free: reset it for its next use
    for (int64_t x = int64_t(t) * kSweepThreads + tid; x < a.other_nexport; x += a.ntiles * int64_t(kSweepThreads))
        xbuf_other[x] = Sentinel<VT>::value();
    const int32_t rb = a.tile_row_begin[t], re = a.tile_row_begin[t + 1];
    const int32_t ib = a.tile_import_begin[t], ni = a.tile_import_begin[t + 1] - ib;
    const int32_t xb = a.tile_export_begin[t], nx = a.tile_export_begin[t + 1] - xb;
    const int32_t g0 = a.tile_group_begin[t], ng = a.tile_group_begin[t + 1] - g0;
    const int nb = (ng + B - 1) / B;   // batches
    const int ntab = (nb + 2) * B + 1;
    // Fill LDS in two round trips: row indices and slots, import slots,
    // export slots and group table entries of the first pass together, then
    // the row values.  Table entries past the tile's last group are empty
    // groups that need nothing and follow every export; each import slot
    // holds its exchange slot until the value arrives.
    {
        int32_t r[kFillUnroll], sl[kFillUnroll], xs[kFillLists], es[kFillLists];
        int4 e = make_int4(0, 0, 0, 0);
#pragma unroll
        for (int u = 0; u < kFillUnroll; u++) {
            const int32_t i = rb + tid + u * kSweepThreads;
            r[u] = i < re ? a.rows[i] : -1;
            sl[u] = i < re ? a.row_slot[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kFillLists; u++) {
            const int i = tid + u * kSweepThreads;
            xs[u] = i < ni ? a.imports[ib + i] : -1;
            es[u] = i < nx ? a.xlist[xb + i] : -1;
        }
        if (tid < ntab) e = a.group_tab[g0 + min(tid, ng)];
        VT x[kFillUnroll];
#pragma unroll
        for (int u = 0; u < kFillUnroll; u++) x[u] = r[u] >= 0 ? v[r[u]] : VT(0);
#pragma unroll
        for (int u = 0; u < kFillLists; u++) {
            if (xs[u] >= 0) reinterpret_cast<int32_t*>(imp + tid + u * kSweepThreads)[0] = xs[u];
            if (es[u] >= 0) xl[tid + u * kSweepThreads] = uint16_t(es[u]);
        }
        if (tid < ntab) {
            if (tid >= ng) e = make_int4(e.x, 0, nx, 0);
            tab[tid] = e;
        }
#pragma unroll
        for (int u = 0; u < kFillUnroll; u++)
            if (r[u] >= 0) lds[sl[u]] = x[u];
    }
    // tiles larger than one pass
    for (int32_t i0 = rb + tid + kFillUnroll * kSweepThreads; i0 < re; i0 += kFillUnroll * kSweepThreads) {
        int32_t r[kFillUnroll], sl[kFillUnroll];
        VT x[kFillUnroll];
#pragma unroll
        for (int u = 0; u < kFillUnroll; u++) {
            const int32_t i = i0 + u * kSweepThreads;
            r[u] = i < re ? a.rows[i] : -1;
            sl[u] = i < re ? a.row_slot[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kFillUnroll; u++) x[u] = r[u] >= 0 ? v[r[u]] : VT(0);
#pragma unroll
        for (int u = 0; u < kFillUnroll; u++)
            if (r[u] >= 0) lds[sl[u]] = x[u];
    }
    for (int i = tid + kFillLists * kSweepThreads; i < ni; i += kSweepThreads)
        reinterpret_cast<int32_t*>(imp + i)[0] = a.imports[ib + i];
    for (int i = tid + kFillLists * kSweepThreads; i < nx; i += kSweepThreads) xl[i] = uint16_t(a.xlist[xb + i]);
    for (int i = tid + kSweepThreads; i < ntab; i += kSweepThreads) {
        int4 e = a.group_tab[g0 + min(i, ng)];
        if (i >= ng) e = make_int4(e.x, 0, nx, 0);
        tab[i] = e;
    }
    __syncthreads();

    if (wave == 0) {
        // ---- importer ----
        // Polls the next imports of the tile, in first-use order, straight
        // from the exchange buffer into LDS and publishes the prefix that has
        // arrived: a window of 64 * kImportDepth after progress, 64 while
        // waiting (fewer loads on memory that is being written).  A poll
        // that finds nothing new touches no LDS (the window's slots stay in
        // registers, the first missing import comes from ballots), so it
        // does not queue behind the compute wavefront's LDS traffic.
        const int32_t* slot_of = reinterpret_cast<const int32_t*>(imp);
        int head = 0;
        int64_t polls = 0;
        bool wide = true;
        int32_t xs[kImportDepth];   // exchange slot of import head + lane + 64u; -1: not polled
        auto read_slots = [&]() {
#pragma unroll
            for (int u = 0; u < kImportDepth; u++) {
                const int idx = head + lane + kWave * u;
                xs[u] = idx < ni && (u == 0 || wide) ? slot_of[idx * (sizeof(VT) / 4)] : -1;
            }
        };
        read_slots();
        while (head < ni) {
            VT val[kImportDepth];
#pragma unroll
            for (int u = 0; u < kImportDepth; u++)
                val[u] = xs[u] >= 0 ? __hip_atomic_load(xbuf + xs[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                    : VT(0);
            // smallest index of the window that has not arrived (uniform)
            int first_missing = min(ni, head + kWave * (wide ? kImportDepth : 1));
#pragma unroll
            for (int u = kImportDepth - 1; u >= 0; u--) {
                const unsigned long long miss = __ballot(xs[u] >= 0 && Sentinel<VT>::is(val[u]));
                if (miss) first_missing = head + kWave * u + int(__builtin_ctzll(miss));
            }
#pragma unroll
            for (int u = 0; u < kImportDepth; u++) {
                const int idx = head + lane + kWave * u;
                if (xs[u] >= 0 && idx < first_missing) imp[idx] = val[u];
            }
            if (first_missing > head) {
                head = first_missing;
                wide = true;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                if (lane == 0) __hip_atomic_store(&progress, head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                read_slots();   // the values written above lie below head
                continue;
            }
            if (wide) {   // nothing new: poll the first 64 only
                wide = false;
#pragma unroll
                for (int u = 1; u < kImportDepth; u++) xs[u] = -1;
            }
            ++polls;
            if (polls > a.spin_limit ||
                ((polls & 63) == 0 && __hip_atomic_load(a.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                // give up: flag the fault and release the compute wavefront
                // (its results are then meaningless; the host reports the fault)
                if (lane == 0) {
                    __hip_atomic_store(a.fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&progress, ni, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    } else if (wave == 1) {
        // ---- compute ----
        // Batch by batch: wait for the batch's records in the ring, then its
        // groups unrolled.  The next group's record is read from LDS while
        // the current group computes.  Lanes past a group's size hold a copy
        // of its last record: they compute the same value and write it to the
        // same slot, so no lane is masked off.
        const i32x4* ringv = reinterpret_cast<const i32x4*>(ring);
        // (No clock reads or other scalar memory operations inside the loop:
        // they would make the compiler wait for every LDS read.)
        int seen = 0;    // imports known to be in LDS
        for (int b = 0; b < nb; b++) {
            __hip_atomic_store(&consumed, b * B, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            while (__hip_atomic_load(half_batch + (b & 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) < b) {
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const i32x4* rs = ringv + (b & 1) * (B * C * kWave) + lane;
            i32x4 rc[C];
#pragma unroll
            for (int c = 0; c < C; c++) rc[c] = rs[c * kWave];
#pragma unroll
            for (int j = 0; j < B; j++) {
                if (b * B + j >= ng) break;
                // the whole record stays allocated until here (its padding
                // bytes too), so no temporary reuses registers a pending
                // read is still writing
#pragma unroll
                for (int c = 0; c < C; c++) asm volatile("" ::"v"(rc[c]));
                SweepRec<FV, W> cur;
                __builtin_memcpy(&cur, rc, sizeof(cur));
                const int need = __builtin_amdgcn_readfirstlane(cur.need);
                if (need > seen) {
                    while ((seen = __hip_atomic_load(&progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) <
                           need) {
                        __builtin_amdgcn_s_sleep(1);
                    }
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                }
                VT yv[W];
#pragma unroll
                for (int k = 0; k < W; k++) yv[k] = *lds_at<VT>(cur.src[k]);
                VT rr = *lds_at<VT>(cur.row);
                // the next group's record, issued after this group's reads
                // so that waiting for those does not wait for it
                __builtin_amdgcn_sched_barrier(0);
                if (j + 1 < B) {
#pragma unroll
                    for (int c = 0; c < C; c++) rc[c] = rs[((j + 1) * C + c) * kWave];
                }
                __builtin_amdgcn_sched_barrier(0);
                // rr -= L_rj * y_j in stored order; padding entries subtract 0 * (+0)
#pragma unroll
                for (int k = 0; k < W; k++) rr = rr - VT(cur.val[k]) * yv[k];
                if (!UNIT) rr = rr * VT(cur.diag);
                *lds_at<VT>(cur.row) = rr;
                // LDS serves a wavefront's operations in order: the exporter
                // that sees this count also sees the row written above
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __hip_atomic_store(&finished, b * B + j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    } else if (wave == 4) {
        // ---- exporter ----
        // Follows the compute wavefront and stores each finished exported row
        // (in finish order, so exchange slots xb, xb+1, ...) for later tiles.
        VT* xout = xbuf + xb;
        int e = 0;
        while (e < nx) {
            const int fin = __hip_atomic_load(&finished, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const int avail = tab[fin].z;
            if (avail == e) {
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            for (int i = e + lane; i < avail; i += kWave)
                __hip_atomic_store(xout + i, Sentinel<VT>::canon(lds[xl[i]]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            e = avail;
        }
    } else {
        // ---- loaders: wavefront 2 fills ring half 0, wavefront 3 half 1 ----
        const int k = wave - 2;
        const i32x4* rv = reinterpret_cast<const i32x4*>(recs);
        i32x4* rr = reinterpret_cast<i32x4*>(ring);
        i32x4 bufA[B][C], bufB[B][C];
        load_batch_regs<C, B>(bufA, rv, tab, k, lane);
        for (int b = k; b < nb; b += 4) {
            load_batch_regs<C, B>(bufB, rv, tab, b + 2, lane);
            wait_lds_at_least(&consumed, (b - 1) * B);   // batch b - 2 (same half) is finished
            store_batch_lds<C, B>(bufA, rr, k, b, half_batch, lane);
            if (b + 2 >= nb) break;
            load_batch_regs<C, B>(bufA, rv, tab, b + 4, lane);
            wait_lds_at_least(&consumed, (b + 1) * B);
            store_batch_lds<C, B>(bufB, rr, k, b + 2, half_batch, lane);
        }
        // drain the batch still in flight here, so the compiler does not
        // carry it into the other roles' code at the join below
        __builtin_amdgcn_s_waitcnt(0);
    }
    __syncthreads();
    // write the tile's rows back; with dot_with, also the tile's partial of
    // dot(w, y): per thread in the vector type, then across the workgroup
    // in double (as the dot kernels sum their grid-stride partials)
    const VT* w = static_cast<const VT*>(a.dot_with);
    VT dacc = 0;
    for (int32_t i0 = rb + tid; i0 < re; i0 += kFillUnroll * kSweepThreads) {
        int32_t r[kFillUnroll], sl[kFillUnroll];
#pragma unroll
        for (int u = 0; u < kFillUnroll; u++) {
            const int32_t i = i0 + u * kSweepThreads;
            r[u] = i < re ? a.rows[i] : -1;
            sl[u] = i < re ? a.row_slot[i] : 0;
        }
        if (w) {
            VT wv[kFillUnroll];
#pragma unroll
            for (int u = 0; u < kFillUnroll; u++) wv[u] = r[u] >= 0 ? w[r[u]] : VT(0);
#pragma unroll
            for (int u = 0; u < kFillUnroll; u++)
                if (r[u] >= 0) {
                    const VT yv = lds[sl[u]];
                    y[r[u]] = yv;
                    dacc += wv[u] * yv;
                }
        } else {
#pragma unroll
            for (int u = 0; u < kFillUnroll; u++)
                if (r[u] >= 0) y[r[u]] = lds[sl[u]];
        }
    }
    if (w) {
        // the record ring is free now: its first words hold the wave sums
        double* dscratch = reinterpret_cast<double*>(lds_raw + L.ring);
        const double tot = block_sum<kSweepThreads>(double(dacc), dscratch);
        if (tid == 0) a.dot_partials[t] = tot;
    }
    if (tid == 0) {
        unsigned done = atomicAdd(&a.tickets[1], 1u);
        if (done == unsigned(a.ntiles) - 1) {   // every tile has started: reset for the next launch
            a.tickets[0] = 0;
            a.tickets[1] = 0;
        }
    }
}

// Packs the records of one group per workgroup from the plan's position
// arrays and the factor storage, with LDS byte offsets for vector entries of
// vbytes bytes.
template <typename FV, int W>
__global__ __launch_bounds__(64) void k_build_recs(const int4* __restrict__ group_tab, const int2* __restrict__ meta,
                                                   const int32_t* __restrict__ src,
                                                   const int32_t* __restrict__ val_pos,
                                                   const int32_t* __restrict__ diag_pos, const double* __restrict__ f,
                                                   int vbytes, int4* __restrict__ recs)
{
    constexpr int C = kRecChunks<FV, W>;
    const int4 g = group_tab[blockIdx.x];
    const int first = g.x, m = group_tab[blockIdx.x + 1].x - first;
    const int q = threadIdx.x;
    if (q >= m) return;
    const int64_t pc = int64_t(first) + q;
    RecBits<FV, W> u;
#pragma unroll
    for (int c = 0; c < C; c++) u.c[c] = make_int4(0, 0, 0, 0);
    u.r.row = uint32_t(kSweepHeader) + uint32_t(meta[pc].x & 0xFFFF) * uint32_t(vbytes);
    u.r.need = g.y;
#pragma unroll
    for (int k = 0; k < W; k++) {
        u.r.src[k] = uint32_t(kSweepHeader) + uint32_t(src[pc * W + k]) * uint32_t(vbytes);
        const int32_t vp = val_pos[pc * W + k];
        u.r.val[k] = vp >= 0 ? FV(f[vp]) : FV(0);
    }
    const int32_t dp = diag_pos[pc];
    u.r.diag = dp >= 0 ? FV(f[dp]) : FV(1);
#pragma unroll
    for (int c = 0; c < C; c++) recs[int64_t(first) * C + int64_t(c) * m + q] = u.c[c];
}

// out_k[r] = f[pos[3r + k]] (0 where the position is -1), k = 0, 1, 2
__global__ void k_gather_strided(int64_t n, const int32_t* __restrict__ pos, const double* __restrict__ f,
                                 double* __restrict__ o0, double* __restrict__ o1, double* __restrict__ o2)
{
    for (int64_t r = blockIdx.x * int64_t(256) + threadIdx.x; r < n; r += int64_t(gridDim.x) * 256) {
        const int32_t a = pos[3 * r], b = pos[3 * r + 1], c = pos[3 * r + 2];
        o0[r] = a >= 0 ? f[a] : 0.0;
        o1[r] = b >= 0 ? f[b] : 0.0;
        o2[r] = c >= 0 ? f[c] : 0.0;
    }
}

template <typename T, typename IT>
__global__ void k_gather(int64_t count, const IT* __restrict__ pos, const double* __restrict__ f,
                         T* __restrict__ out, T absent)
{
    for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < count; e += int64_t(gridDim.x) * 256) {
        const int64_t p = pos[e];
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
    M4G_REQUIRE(p.npieces * int64_t(p.width) < (int64_t(1) << 31) && p.n < (int64_t(1) << 31), M4G_ERR_INVALID,
                "sweep too large");
    auto up32 = [](DevBuf<int32_t>& d, const auto& h) {
        d.reset(std::max<size_t>(h.size(), 1));
        d.upload(h.data(), h.size());
    };
    up32(tile_row_begin, p.tile_row_begin);
    up32(rows, p.rows);
    up32(row_slot, p.row_slot);
    up32(tile_group_begin, p.tile_group_begin);
    up32(tile_import_begin, p.tile_import_begin);
    up32(imports, p.imports);
    up32(tile_export_begin, p.tile_export_begin);
    up32(xlist, p.xlist);
    up32(pc_src, p.pc_src);
    up32(pc_val_pos, p.pc_val_pos);
    up32(pc_diag_pos, p.pc_diag_pos);
    std::vector<int4> tab(p.ngroups + 1);
    for (int64_t g = 0; g <= p.ngroups; g++) {
        tab[g].x = p.group_piece_begin[g];
        tab[g].y = g < p.ngroups ? p.group_need[g] : 0;
        tab[g].z = g < p.ngroups ? p.group_xbefore[g] : 0;
        tab[g].w = 0;
    }
    group_tab.reset(tab.size());
    group_tab.upload(tab.data(), tab.size());
    std::vector<int2> meta(std::max<int64_t>(p.npieces, 1));
    for (int64_t i = 0; i < p.npieces; i++) {
        M4G_REQUIRE(p.pc_slot[i] < 65536 && p.pc_len[i] <= p.width, M4G_ERR_INVALID, "sweep piece out of range");
        meta[i].x = int(unsigned(p.pc_slot[i]) | (unsigned(p.pc_len[i]) << 16));
        meta[i].y = p.pc_export[i];
    }
    pc_meta.reset(meta.size());
    pc_meta.upload(meta.data(), meta.size());
    tickets.reset(2);
    M4G_HIP(hipMemsetAsync(tickets.ptr, 0, 2 * sizeof(unsigned), current_stream()));
    fault.reset(1);
    M4G_HIP(hipMemsetAsync(fault.ptr, 0, sizeof(int), current_stream()));
    plan_meta = p;   // keep only scalars and offsets on the host copy
    plan_meta.rows.clear();
    plan_meta.row_slot.clear();
    plan_meta.pc_slot.clear();
    plan_meta.pc_len.clear();
    plan_meta.pc_export.clear();
    plan_meta.pc_diag_pos.clear();
    plan_meta.pc_src.clear();
    plan_meta.pc_val_pos.clear();
    plan_meta.imports.clear();
    plan_meta.group_piece_begin.clear();
    plan_meta.group_need.clear();
    plan_meta.group_xbefore.clear();
    plan_meta.xlist.clear();
    refresh_values(fstore_dev, ftype);
}

template <typename FV, int W>
static void build_recs_t(SweepDev& d, int vbytes, hipStream_t s)
{
    constexpr int C = kRecChunks<FV, W>;
    const SweepPlan& p = d.plan_meta;
    d.recs.reset(size_t(p.npieces + 1) * C);   // one record of padding: the ring reads past the last group
    M4G_HIP(hipMemsetAsync(d.recs.ptr, 0, d.recs.count * sizeof(int4), s));
    if (p.ngroups)
        hipLaunchKernelGGL((k_build_recs<FV, W>), dim3(p.ngroups), dim3(64), 0, s, d.group_tab.ptr, d.pc_meta.ptr,
                           d.pc_src.ptr, d.pc_val_pos.ptr, d.pc_diag_pos.ptr, d.fsrc, vbytes, d.recs.ptr);
    M4G_HIP(hipGetLastError());
}

template <typename FV>
static void build_recs_w(SweepDev& d, int vbytes, hipStream_t s)
{
    switch (d.plan_meta.width) {
    case 3: build_recs_t<FV, 3>(d, vbytes, s); break;
    case 4: build_recs_t<FV, 4>(d, vbytes, s); break;
    case 8: build_recs_t<FV, 8>(d, vbytes, s); break;
    case 16: build_recs_t<FV, 16>(d, vbytes, s); break;
    default: throw Error(M4G_ERR_LOGIC, "unsupported sweep piece width");
    }
}

void SweepDev::refresh_values(const double* f, int ftype)
{
    // records are packed at the next launch, for that launch's vector type
    factor_type = ftype;
    fsrc = f;
    recs_vbytes = 0;
}

void SweepDev::ensure_recs(int vbytes, hipStream_t s)
{
    if (recs_vbytes == vbytes) return;
    if (factor_type == M4G_F64) build_recs_w<double>(*this, vbytes, s);
    else build_recs_w<float>(*this, vbytes, s);
    recs_vbytes = vbytes;
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


// Polls a sweep's bounded waits make before giving up (M4G_SPIN_LIMIT; read
// at every launch, so a process can change it between solves).
static int64_t spin_limit()
{
    const char* e = std::getenv("M4G_SPIN_LIMIT");
    return e ? std::max<int64_t>(1, std::atoll(e)) : int64_t(1) << 20;   // about one second
}

// Dynamic LDS of a sweep launch (layout in sweep_lds_layout).
static size_t sweep_lds_bytes(const SweepPlan& p, size_t vbytes, int ftype)
{
    const int C = rec_chunks(ftype, p.width);
    return sweep_lds_layout(int(p.max_tile_rows), int(p.max_tile_imports), int(p.max_steps), int(p.max_tile_exports),
                            int(vbytes), C, load_batch(C))
        .total;
}

template <typename FV, typename VT, bool UNIT, int W>
static void launch_sweep_w(SweepDev& d, const SweepArgs& a, const VT* v, VT* y, VT* xo, const int* skip,
                           hipStream_t s)
{
    const SweepPlan& p = d.plan_meta;
    static_assert(sizeof(SweepRec<FV, W>) % 16 == 0, "record size");
    auto kern = k_sweep<FV, VT, UNIT, W>;
    constexpr int ftype = std::is_same<FV, double>::value ? M4G_F64 : M4G_F32;
    const size_t lds = sweep_lds_bytes(p, sizeof(VT), ftype);
    const bool layout_ok = rec_chunks(ftype, W) == kRecChunks<FV, W>;
    M4G_REQUIRE(layout_ok, M4G_ERR_LOGIC, "record layout mismatch");
    // beyond 64 KB of dynamic LDS must be requested, once per kernel and size
    static size_t granted = 0;
    if (lds > 65536 && lds > granted) {
        M4G_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lds)));
        granted = lds;
    }
    hipLaunchKernelGGL(kern, dim3(p.ntiles), dim3(kSweepThreads), lds, s, a, d.recs.ptr, v, y, (VT*)d.xbuf.ptr, xo,
                       skip);
}

template <typename FV, typename VT>
static void launch_sweep_t(SweepDev& d, SweepDev& other, const VT* v, VT* y, const int* skip, hipStream_t s,
                           const VT* dot_with = nullptr, double* dot_partials = nullptr)
{
    const SweepPlan& p = d.plan_meta;
    TraceRange tr(!p.lower ? kTraceUpperTrisolve : p.unit ? kTraceLowerUnitTrisolve : kTraceLowerTrisolve);
    if (p.ntiles == 0) return;
    d.ensure_recs(int(sizeof(VT)), s);
    SweepArgs a;
    a.ntiles = p.ntiles;
    a.tile_row_begin = d.tile_row_begin.ptr;
    a.rows = d.rows.ptr;
    a.row_slot = d.row_slot.ptr;
    a.tile_group_begin = d.tile_group_begin.ptr;
    a.group_tab = d.group_tab.ptr;
    a.tile_import_begin = d.tile_import_begin.ptr;
    a.imports = d.imports.ptr;
    a.tile_export_begin = d.tile_export_begin.ptr;
    a.xlist = d.xlist.ptr;
    a.max_tile_exports = int32_t(p.max_tile_exports);
    a.max_tile_rows = int32_t(p.max_tile_rows);
    a.max_tile_imports = int32_t(p.max_tile_imports);
    a.max_steps = int32_t(p.max_steps);
    a.tickets = d.tickets.ptr;
    a.fault = d.fault.ptr;
    a.other_nexport = other.plan_meta.nexport;
    a.spin_limit = spin_limit();
    a.dot_with = dot_with;
    a.dot_partials = dot_partials;
    VT* xo = (VT*)other.xbuf.ptr;
    auto go = [&](auto unit_tag) {
        constexpr bool U = decltype(unit_tag)::value;
        switch (p.width) {
        case 3: launch_sweep_w<FV, VT, U, 3>(d, a, v, y, xo, skip, s); break;
        case 4: launch_sweep_w<FV, VT, U, 4>(d, a, v, y, xo, skip, s); break;
        case 8: launch_sweep_w<FV, VT, U, 8>(d, a, v, y, xo, skip, s); break;
        case 16: launch_sweep_w<FV, VT, U, 16>(d, a, v, y, xo, skip, s); break;
        default: throw Error(M4G_ERR_LOGIC, "unsupported sweep piece width");
        }
    };
    if (p.unit) go(std::true_type());
    else go(std::false_type());
    M4G_HIP(hipGetLastError());
}


void run_sweep_pair(SweepDev& first, SweepDev& second, const void* v, void* tmp, void* y, int vtype,
                    const int* skip, hipStream_t s, const void* dot_with, double* dot_partials)
{
    M4G_REQUIRE(first.factor_type == second.factor_type, M4G_ERR_INVALID, "sweep factor types differ");
    first.ensure_exchange(vtype);
    second.ensure_exchange(vtype);
    if (first.factor_type == M4G_F64) {
        M4G_REQUIRE(vtype == M4G_F64, M4G_ERR_INVALID, "double factors need double vectors");
        launch_sweep_t<double, double>(first, second, (const double*)v, (double*)tmp, skip, s);
        launch_sweep_t<double, double>(second, first, (const double*)tmp, (double*)y, skip, s,
                                       (const double*)dot_with, dot_partials);
    } else if (vtype == M4G_F64) {
        launch_sweep_t<float, double>(first, second, (const double*)v, (double*)tmp, skip, s);
        launch_sweep_t<float, double>(second, first, (const double*)tmp, (double*)y, skip, s,
                                      (const double*)dot_with, dot_partials);
    } else {
        launch_sweep_t<float, float>(first, second, (const float*)v, (float*)tmp, skip, s);
        launch_sweep_t<float, float>(second, first, (const float*)tmp, (float*)y, skip, s,
                                     (const float*)dot_with, dot_partials);
    }
}

bool pc_grid_factor(m4g_pc_s& P, int64_t max_n, DevBuf<double>& east, DevBuf<double>& north, DevBuf<double>& dinv,
                    hipStream_t s)
{
    const int64_t n = P.n, nx = P.grid.nx, ny = P.grid.ny;
    if (P.grid5 == 0) {
        P.grid5 = -1;
        if (P.kind == M4G_PC_IC0 && P.factor_type == M4G_F64 && n > 0 && n <= max_n && nx > 0 && P.grid.nz == 1 &&
            ny <= 64) {
            // row r = i + nx*j of U holds r, then r+1 (i < nx-1), then r+nx (j < ny-1)
            std::vector<int32_t> pos(size_t(3 * n));
            bool ok = true;
            for (int64_t r = 0; r < n && ok; r++) {
                const int64_t i = r % nx, j = r / nx;
                int64_t q = P.fS[r];
                const int64_t len = 1 + (i < nx - 1) + (j < ny - 1);
                ok = P.fS[r + 1] - q == len && P.fI[q] == r;
                pos[3 * r] = int32_t(q++);
                pos[3 * r + 1] = pos[3 * r + 2] = -1;
                if (ok && i < nx - 1) {
                    ok = P.fI[q] == r + 1;
                    pos[3 * r + 1] = int32_t(q++);
                }
                if (ok && j < ny - 1) {
                    ok = P.fI[q] == r + nx;
                    pos[3 * r + 2] = int32_t(q);
                }
            }
            if (ok) {
                P.grid5_pos.reset(pos.size());
                P.grid5_pos.upload(pos.data(), pos.size());
                P.grid5 = 1;
            }
        }
    }
    if (P.grid5 != 1) return false;
    east.reset(size_t(n));
    north.reset(size_t(n));
    dinv.reset(size_t(n));
    // strided gathers of the three positions of each row (values as the
    // factorisation left them: the diagonal already inverted)
    hipLaunchKernelGGL(k_gather_strided, dim3(grid_for(n)), dim3(256), 0, s, n, P.grid5_pos.ptr, P.fstore.ptr,
                       dinv.ptr, east.ptr, north.ptr);
    M4G_HIP(hipGetLastError());
    return true;
}

// Reads the sweep's fault flag and clears it: a wait that gave up spoils the
// solve that launched it, and only that one.
int check_sweep_fault(SweepDev& d)
{
    int h = 0;
    M4G_HIP(hipMemcpyAsync(&h, d.fault.ptr, sizeof(int), hipMemcpyDeviceToHost, current_stream()));
    M4G_HIP(hipStreamSynchronize(current_stream()));
    if (h) {
        M4G_HIP(hipMemsetAsync(d.fault.ptr, 0, sizeof(int), current_stream()));
        M4G_HIP(hipStreamSynchronize(current_stream()));
    }
    return h;
}

int check_pc_faults(m4g_pc_s& P)
{
    if (P.kind < M4G_PC_IC0 || P.n == 0) return 0;
    int f = check_sweep_fault(P.lower);
    f |= check_sweep_fault(P.upper);
    if (P.adj_ready) {
        f |= check_sweep_fault(P.adj_lower);
        f |= check_sweep_fault(P.adj_upper);
    }
    return f;
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

// LDS bytes a sweep tile may use on the current device: a tile's rows,
// imports, group table and record ring share the workgroup's LDS
static size_t sweep_lds_budget()
{
    int lds = 65536, dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lds = std::min(v, 160 * 1024);
    }
    return size_t(lds) - 64;   // static shared words
}

// host only (no HIP call): safe on any thread
static SweepPlan plan_sweep_budget(const SweepRows& rows, const GridInfo& g, int64_t cap, int ftype, size_t budget,
                                   const std::vector<int32_t>* levels = nullptr)
{
    // sized for double vectors, the widest the sweeps use
    std::vector<int32_t> own;
    int64_t nlev = 0;
    if (!levels) {
        own = natural_levels(rows, &nlev);
        levels = &own;
    } else {
        for (int32_t l : *levels) nlev = std::max<int64_t>(nlev, int64_t(l) + 1);
    }
    SweepPlan p;
    for (;;) {
        TileShape shape = choose_tiles(rows, g, cap, env_double("M4G_SWEEP_HOP_US", kSweepHopUs),
                                       env_double("M4G_SWEEP_STEP_US", kSweepStepUs), levels);
        p = build_sweep_plan(rows, g, shape, cap, 0, nlev);
        if (sweep_lds_bytes(p, 8, ftype) <= budget || cap <= 64) break;
        cap = cap * 3 / 4;
    }
    M4G_REQUIRE(sweep_lds_bytes(p, 8, ftype) <= budget, M4G_ERR_INVALID, "sweep tile does not fit in LDS");
    if (std::getenv("M4G_VALIDATE_PLANS")) {
        std::string msg = validate_sweep_plan(p, rows);
        M4G_REQUIRE(msg.empty(), M4G_ERR_LOGIC, "sweep plan invalid: " + msg);
    }
    return p;
}

SweepPlan plan_sweep(const SweepRows& rows, const GridInfo& g, int64_t cap, int ftype)
{
    return plan_sweep_budget(rows, g, cap, ftype, sweep_lds_budget());
}

void plan_sweep_pair(const SweepRows& lo, const SweepRows& up, const GridInfo& g, int64_t cap, int ftype,
                     SweepPlan& pl, SweepPlan& pu)
{
    // the two directions are independent: plan them on two host threads
    const size_t budget = sweep_lds_budget();
    auto fu = std::async(std::launch::async, [&] { return plan_sweep_budget(up, g, cap, ftype, budget); });
    try {
        pl = plan_sweep_budget(lo, g, cap, ftype, budget);
    } catch (...) {
        fu.wait();
        throw;
    }
    pu = fu.get();
}

// Plans of both sweep directions, built on two host threads from the moment
// the job is made, so the caller's thread can meanwhile run the
// factorisation; finish() uploads them (the lower one while the upper one may
// still be in the making).  lo, up and lo_levels must outlive the job.
class PlanPairJob {
  public:
    PlanPairJob(const SweepRows& lo, const SweepRows& up, const GridInfo& g, int64_t cap, int ftype,
                const std::vector<int32_t>* lo_levels)
        : ftype_(ftype)
    {
        const size_t budget = sweep_lds_budget();   // device query: on the caller's thread
        lo_ = std::async(std::launch::async,
                         [&lo, g, cap, ftype, budget, lo_levels] {
                             return plan_sweep_budget(lo, g, cap, ftype, budget, lo_levels);
                         });
        up_ = std::async(std::launch::async,
                         [&up, g, cap, ftype, budget] { return plan_sweep_budget(up, g, cap, ftype, budget); });
    }
    PlanPairJob(const PlanPairJob&) = delete;
    PlanPairJob& operator=(const PlanPairJob&) = delete;
    ~PlanPairJob()
    {
        if (lo_.valid()) lo_.wait();
        if (up_.valid()) up_.wait();
    }
    void finish(SweepDev& dl, SweepDev& du, const double* fstore)
    {
        dl.upload(lo_.get(), fstore, ftype_);
        du.upload(up_.get(), fstore, ftype_);
    }

  private:
    int ftype_;
    std::future<SweepPlan> lo_, up_;
};

void plan_upload_pair(const SweepRows& lo, const SweepRows& up, const GridInfo& g, int64_t cap, int ftype,
                      SweepDev& dl, SweepDev& du, const double* fstore, const std::vector<int32_t>* lo_levels)
{
    PlanPairJob job(lo, up, g, cap, ftype, lo_levels);
    job.finish(dl, du, fstore);
}

}  // namespace m4g

using namespace m4g;

m4g_pc_s::~m4g_pc_s() = default;

static void factorize_levels(m4g_pc_s& P, const SweepRows& lower_dag, const m4g_matrix_s& A,
                             const std::vector<int32_t>& lvl)
{
    // natural-order level sets of the lower DAG
    int64_t nlev = 0;
    for (int32_t l : lvl) nlev = std::max<int64_t>(nlev, int64_t(l) + 1);
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
        DevBuf<int64_t> d_lptr(P.n + 1);
        DevBuf<int32_t> d_lpos(std::max<size_t>(lower_dag.val_pos.size(), 1)), d_diag(std::max<int64_t>(P.n, 1)),
            d_lcol(std::max<size_t>(lower_dag.col.size(), 1));
        d_lptr.upload(lower_dag.ptr.data(), P.n + 1);
        d_lpos.upload(lower_dag.val_pos.data(), lower_dag.val_pos.size());
        d_lcol.upload(lower_dag.col.data(), lower_dag.col.size());
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
        int64_t max_len = 0;
        for (int64_t r = 0; r < P.n; r++) max_len = std::max(max_len, A.h_starts[r + 1] - A.h_starts[r]);
        for (int64_t l = 0; l < nlev; l++) {
            int64_t cnt = lptr[l + 1] - lptr[l];
            if (max_len <= 64)   // a wavefront per row
                hipLaunchKernelGGL(k_ilu0_level_wave, dim3((cnt + 3) / 4), dim3(256), 0, s, cnt,
                                   d_rows.ptr + lptr[l], A.starts.ptr, A.indices.ptr, d_diag.ptr, P.fstore.ptr,
                                   invdiag.ptr);
            else
                hipLaunchKernelGGL(k_ilu0_level, dim3((cnt + 255) / 256), dim3(256), 0, s, cnt,
                                   d_rows.ptr + lptr[l], A.starts.ptr, A.indices.ptr, d_diag.ptr, P.fstore.ptr,
                                   invdiag.ptr);
        }
        M4G_HIP(hipGetLastError());
        M4G_HIP(hipStreamSynchronize(s));
    }
    P.factor_levels = nlev;
}

// Trace event of a preconditioner application (identity.hpp:37-58,
// diagonal.hpp:59, ic_0.hpp:64-101, ilu.hpp:86-108).
static int pc_event(int kind, bool adjoint)
{
    switch (kind) {
    case M4G_PC_IDENTITY: return adjoint ? kTraceIdAdjoint : kTraceIdSolve;
    case M4G_PC_DIAGONAL: return kTraceDiagSolve;
    case M4G_PC_IC0: return adjoint ? kTraceIc0Adjoint : kTraceIc0Solve;
    default: return adjoint ? kTraceIlu0Adjoint : kTraceIlu0Solve;
    }
}

// Trace event of a preconditioner's construction (diagonal.hpp:39,
// ic_0.hpp:140, ilu_0.hpp:50); 0 for the identity, which has none.
static int pc_setup_event(int kind)
{
    switch (kind) {
    case M4G_PC_DIAGONAL: return kTraceDiagSetup;
    case M4G_PC_IC0: return kTraceIc0Factorize;
    case M4G_PC_ILU0: return kTraceIlu0Factorize;
    default: return 0;
    }
}

extern "C" {

int m4g_pc_create(m4g_pc* out, m4g_matrix A, int kind, int ftype)
{
    return guard([&] {
        M4G_REQUIRE(out && A, M4G_ERR_INVALID, "m4g_pc_create: null argument");
        M4G_REQUIRE(kind >= M4G_PC_IDENTITY && kind <= M4G_PC_ILU0, M4G_ERR_INVALID, "m4g_pc_create: unknown kind");
        M4G_REQUIRE(ftype == M4G_F64 || ftype == M4G_F32, M4G_ERR_INVALID, "m4g_pc_create: unknown value type");
        M4G_REQUIRE(A->nrows == A->ncols, M4G_ERR_NOT_SQUARE, "Matrix is not square");
        TraceRange tr(pc_setup_event(kind));
        auto P = std::make_unique<m4g_pc_s>();
        P->kind = kind;
        P->factor_type = ftype;
        P->n = A->nrows;
        hipStream_t s = current_stream();
        const int64_t n = A->nrows;
        // M4G_PC_TIMING: host time of each construction phase on stderr
        const bool timing = std::getenv("M4G_PC_TIMING") != nullptr;
        auto t_prev = std::chrono::steady_clock::now();
        std::string tlog;
        auto mark = [&](const char* what) {
            if (!timing) return;
            const auto t = std::chrono::steady_clock::now();
            tlog += std::string(what) + "=" +
                    std::to_string(std::chrono::duration<double, std::milli>(t - t_prev).count()) + "ms ";
            t_prev = t;
        };
        if (kind == M4G_PC_DIAGONAL) {
            P->inv_diag.reset(std::max<int64_t>(n, 1));
            if (n)
                hipLaunchKernelGGL(k_diag_setup, dim3(grid_for(n)), dim3(256), 0, s, n, A->starts.ptr,
                                   A->indices.ptr, A->values.ptr, P->inv_diag.ptr);
            M4G_HIP(hipGetLastError());
        } else if (kind == M4G_PC_IC0 || kind == M4G_PC_ILU0) {
            // grid inference runs beside the pattern work on a second host thread
            auto fg = std::async(std::launch::async,
                                 [&] { return infer_grid(n, A->h_starts.data(), A->h_indices.data()); });
            struct Join {
                std::future<GridInfo>& f;
                ~Join() { if (f.valid()) f.wait(); }
            } join_g{fg};
            SweepRows lo, up;
            std::vector<int32_t> lo_levels;   // natural levels of the lower DAG (factorisation and sweep)
            if (kind == M4G_PC_IC0) {
                PlanVec32 a_pos;
                M4G_REQUIRE(upper_pattern(n, A->h_starts.data(), A->h_indices.data(), P->fS, P->fI, a_pos),
                            M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                mark("pattern");
                M4G_REQUIRE(ic0_rows(n, P->fS, P->fI, lo, up), M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                mark("rows");
                // U = upper(A)
                P->fstore.reset(std::max<size_t>(a_pos.size(), 1));
                P->a_pos.reset(std::max<size_t>(a_pos.size(), 1));
                P->a_pos.upload(a_pos.data(), a_pos.size());
                if (!a_pos.empty())
                    hipLaunchKernelGGL((k_gather<double, int32_t>), dim3(grid_for(a_pos.size())), dim3(256), 0, s,
                                       int64_t(a_pos.size()), P->a_pos.ptr, A->values.ptr, P->fstore.ptr, 0.0);
                M4G_HIP(hipGetLastError());
                mark("gather");
            } else {
                P->fS.assign(A->h_starts.begin(), A->h_starts.end());
                P->fI.assign(A->h_indices.begin(), A->h_indices.end());
                M4G_REQUIRE(ilu0_rows(n, A->h_starts.data(), A->h_indices.data(), lo, up, P->diag_pos),
                            M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                mark("rows");
                P->fstore.reset(std::max<int64_t>(A->nnz, 1));
                if (A->nnz)
                    M4G_HIP(hipMemcpyAsync(P->fstore.ptr, A->values.ptr, A->nnz * sizeof(double),
                                           hipMemcpyDeviceToDevice, s));
            }
            lo_levels = natural_levels(lo, nullptr);
            mark("levels");
            const GridInfo g = fg.get();
            mark("grid_wait");
            P->grid = g;
            // the sweep plans need only the pattern: they are built on host
            // threads while this thread runs the factorisation on the device
            const int64_t cap = sweep_tile_cap(ftype == M4G_F64 ? M4G_F64 : M4G_F32);
            PlanPairJob plans(lo, up, g, cap, ftype, &lo_levels);
            factorize_levels(*P, lo, *A, lo_levels);
            mark("factor");
            plans.finish(P->lower, P->upper, P->fstore.ptr);
            mark("plan+upload");
            P->tmp_bytes = 0;
        }
        M4G_HIP(hipStreamSynchronize(s));
        mark("sync");
        if (timing) std::fprintf(stderr, "m4g_pc_create: %s\n", tlog.c_str());
        *out = P.release();
        return M4G_OK;
    });
}

int m4g_pc_refactor(m4g_pc P, m4g_matrix A)
{
    return guard([&] {
        M4G_REQUIRE(P && A, M4G_ERR_INVALID, "m4g_pc_refactor: null argument");
        M4G_REQUIRE(A->nrows == P->n && A->ncols == P->n, M4G_ERR_SIZE, "m4g_pc_refactor: incompatible size");
        TraceRange tr(pc_setup_event(P->kind));
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
                PlanVec64 US, UI;
                PlanVec32 apos;
                M4G_REQUIRE(upper_pattern(n, A->h_starts.data(), A->h_indices.data(), US, UI, apos) && US == P->fS &&
                                UI == P->fI,
                            M4G_ERR_INVALID, "m4g_pc_refactor: pattern differs");
                M4G_REQUIRE(ic0_rows(n, P->fS, P->fI, lo, up), M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                const int64_t nu = P->fS[n];
                if (nu)
                    hipLaunchKernelGGL((k_gather<double, int32_t>), dim3(grid_for(nu)), dim3(256), 0, s, nu, P->a_pos.ptr,
                                       A->values.ptr, P->fstore.ptr, 0.0);
            } else {
                M4G_REQUIRE(std::equal(A->h_starts.begin(), A->h_starts.end(), P->fS.begin(), P->fS.end()) &&
                                std::equal(A->h_indices.begin(), A->h_indices.end(), P->fI.begin(), P->fI.end()),
                            M4G_ERR_INVALID, "m4g_pc_refactor: pattern differs");
                PlanVec64 dpos;
                M4G_REQUIRE(ilu0_rows(n, P->fS.data(), P->fI.data(), lo, up, dpos), M4G_ERR_MISSING_DIAGONAL,
                            "missing diagonal");
                if (A->nnz)
                    M4G_HIP(hipMemcpyAsync(P->fstore.ptr, A->values.ptr, A->nnz * sizeof(double),
                                           hipMemcpyDeviceToDevice, s));
            }
            factorize_levels(*P, lo, *A, natural_levels(lo, nullptr));
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
    TraceRange tr(pc_event(P.kind, adjoint));
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
            plan_upload_pair(lo, up, P.grid, cap, P.factor_type, P.adj_lower, P.adj_upper, P.fstore.ptr);
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

int64_t pc_dot_tiles(const m4g_pc_s& P)
{
    if ((P.kind != M4G_PC_IC0 && P.kind != M4G_PC_ILU0) || P.n == 0) return 0;
    return P.upper.plan_meta.ntiles;
}

void pc_apply_skip_dot(m4g_pc_s& P, const void* x, void* y, int vtype, const int* skip, hipStream_t s,
                       const void* w, double* partials)
{
    M4G_REQUIRE(pc_dot_tiles(P) > 0, M4G_ERR_LOGIC, "pc_apply_skip_dot: no sweeps");
    TraceRange tr(pc_event(P.kind, false));
    run_sweep_pair(P.lower, P.upper, x, pc_tmp(P, vtype), y, vtype, skip, s, w, partials);
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
        M4G_REQUIRE(check_pc_faults(*P) == 0, M4G_ERR_TIMEOUT, "sweep wait timed out");
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
            S.assign(P->fS.begin(), P->fS.end());
            I.assign(P->fI.begin(), P->fI.end());
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
