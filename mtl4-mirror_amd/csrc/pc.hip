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
    const int32_t* tile_row_begin;
    const int32_t* rows;
    const int32_t* tile_group_begin;
    const int64_t* group_row_begin;
    const int64_t* group_ent_begin;
    const int32_t* gr_slot;
    const int32_t* gr_len;
    const int32_t* gr_export;
    const int32_t* ent_src;
    unsigned* tickets;      // [0] next ticket, [1] finished tiles
    int* fault;             // set when a bounded wait runs out
    int64_t other_nexport;  // size of the other direction's exchange buffer
    int64_t spin_limit;
};

// Waits for an exported value.  Bounded: after `limit` polls the wait gives up
// and raises the fault flag, and every later wait of any workgroup returns at
// once, so a schedule fault ends the kernel instead of hanging it.
template <typename VT>
__device__ inline VT exchange_wait(const VT* p, int* fault, int64_t limit)
{
    VT v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!Sentinel<VT>::is(v)) return v;
    for (int64_t it = 0; it < limit; it++) {
        if (__hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return v;
        __builtin_amdgcn_s_sleep(2);
        v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!Sentinel<VT>::is(v)) return v;
    }
    __hip_atomic_store(fault, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v;
}

template <typename FV, typename VT, bool UNIT>
__global__ void k_sweep(SweepArgs a, const FV* __restrict__ ent_val, const FV* __restrict__ gr_diag,
                        const VT* __restrict__ v, VT* __restrict__ y, VT* xbuf, VT* xbuf_other, const int* skip)
{
    if (skip && *skip) return;   // the solve this sweep belongs to has finished
    extern __shared__ unsigned char lds_raw[];
    VT* lds = reinterpret_cast<VT*>(lds_raw);
    __shared__ int tile_sh;
    if (threadIdx.x == 0) tile_sh = int(atomicAdd(&a.tickets[0], 1u));
    __syncthreads();
    const int t = tile_sh;
    const int nt = blockDim.x;
    // the other direction's exchange buffer is free: reset it for its next use
    for (int64_t x = int64_t(t) * nt + threadIdx.x; x < a.other_nexport; x += a.ntiles * int64_t(nt))
        xbuf_other[x] = Sentinel<VT>::value();
    const int32_t rb = a.tile_row_begin[t], re = a.tile_row_begin[t + 1];
    for (int32_t q = rb + threadIdx.x; q < re; q += nt) lds[q - rb] = v[a.rows[q]];
    __syncthreads();
    const int32_t g0 = a.tile_group_begin[t], g1 = a.tile_group_begin[t + 1];
    for (int32_t g = g0; g < g1; g++) {
        const int64_t b = a.group_row_begin[g];
        const int64_t m = a.group_row_begin[g + 1] - b;
        const int64_t eb = a.group_ent_begin[g];
        for (int64_t q = threadIdx.x; q < m; q += nt) {
            const int32_t slot = a.gr_slot[b + q];
            const int32_t len = a.gr_len[b + q];
            VT rr = lds[slot];
            for (int32_t k = 0; k < len; k++) {
                const int64_t d = eb + int64_t(k) * m + q;
                const int32_t src = a.ent_src[d];
                const VT yv = src >= 0 ? lds[src] : exchange_wait(xbuf - (src + 1), a.fault, a.spin_limit);
                rr -= VT(ent_val[d]) * yv;
            }
            VT out = UNIT ? rr : VT(rr * VT(gr_diag[b + q]));
            lds[slot] = out;
            const int32_t ex = a.gr_export[b + q];
            if (ex >= 0)
                __hip_atomic_store(xbuf + ex, Sentinel<VT>::canon(out), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
    for (int32_t q = rb + threadIdx.x; q < re; q += nt) y[a.rows[q]] = lds[q - rb];
    if (threadIdx.x == 0) {
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
    up64(group_row_begin, p.group_row_begin);
    up64(group_ent_begin, p.group_ent_begin);
    up32(gr_slot, p.gr_slot);
    up32(gr_len, p.gr_len);
    up32(gr_export, p.gr_export);
    up32(ent_src, p.ent_src);
    up64(ent_val_pos, p.ent_val_pos);
    up64(gr_diag_pos, p.gr_diag_pos);
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
    a.group_row_begin = d.group_row_begin.ptr;
    a.group_ent_begin = d.group_ent_begin.ptr;
    a.gr_slot = d.gr_slot.ptr;
    a.gr_len = d.gr_len.ptr;
    a.gr_export = d.gr_export.ptr;
    a.ent_src = d.ent_src.ptr;
    a.tickets = d.tickets.ptr;
    a.fault = d.fault.ptr;
    a.other_nexport = other.plan_meta.nexport;
    a.spin_limit = spin_limit();
    const FV* ev = std::is_same<FV, double>::value ? (const FV*)d.ent_val64.ptr : (const FV*)d.ent_val32.ptr;
    const FV* dg = std::is_same<FV, double>::value ? (const FV*)d.diag64.ptr : (const FV*)d.diag32.ptr;
    size_t lds = size_t(p.max_tile_rows) * sizeof(VT);
    if (p.ntiles == 0) return;
    auto kern = p.unit ? k_sweep<FV, VT, true> : k_sweep<FV, VT, false>;
    if (lds > 65536 - 64) {
        // more than 64 KB of dynamic LDS must be requested explicitly
        M4G_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    }
    hipLaunchKernelGGL(kern, dim3(p.ntiles), dim3(p.block), lds, s, a, ev, dg, v, y, (VT*)d.xbuf.ptr,
                       (VT*)other.xbuf.ptr, skip);
    M4G_HIP(hipGetLastError());
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
    int64_t cap = std::min<int64_t>(vtype == M4G_F64 ? 8192 : 16384, (lds - 256) / elem);
    const char* e = std::getenv("M4G_TILE_ROWS");
    if (e) cap = std::min<int64_t>(cap, std::max<int64_t>(64, std::atoll(e)));
    return cap;
}

SweepPlan plan_sweep(const SweepRows& rows, const GridInfo& g, int64_t cap)
{
    TileShape shape = choose_tiles(rows, g, cap, env_double("M4G_SWEEP_HOP_US", 1.5),
                                   env_double("M4G_SWEEP_STEP_US", 0.15));
    SweepPlan p = build_sweep_plan(rows, g, shape, cap);
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
                DevBuf<int64_t> d_apos(std::max<size_t>(a_pos.size(), 1));
                d_apos.upload(a_pos.data(), a_pos.size());
                if (!a_pos.empty())
                    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(a_pos.size())), dim3(256), 0, s,
                                       int64_t(a_pos.size()), d_apos.ptr, A->values.ptr, P->fstore.ptr, 0.0);
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
            P->lower.upload(plan_sweep(lo, g, cap), P->fstore.ptr, ftype);
            P->upper.upload(plan_sweep(up, g, cap), P->fstore.ptr, ftype);
            P->tmp_bytes = 0;
        }
        M4G_HIP(hipStreamSynchronize(s));
        *out = P.release();
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
            P.adj_lower.upload(plan_sweep(lo, P.grid, cap), P.fstore.ptr, P.factor_type);
            P.adj_upper.upload(plan_sweep(up, P.grid, cap), P.fstore.ptr, P.factor_type);
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
