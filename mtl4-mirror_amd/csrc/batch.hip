// Batched IC(0)-PCG: many independent SPD systems, several right-hand sides
// each, factorised by one launch and solved by one persistent launch.
//
// Every (system, right-hand side) pair runs the reference's CG with an IC(0)
// preconditioner (cg.hpp:75-112, ic_0.hpp:135-186 / 68-103) and
// basic_iteration's convergence test (basic_iteration.hpp:43-92), with the
// reference's per-row operation order in the factorisation and the sweeps.
//
// Work item = (system, group of R right-hand sides).  A persistent grid of
// workgroups takes items from an atomic queue in order of decreasing system
// size (longest first balances the tail).  A workgroup runs the whole solve of
// its item: SpMV and vector updates over rows, the two triangular sweeps level
// by level (natural-order levels of the factor's DAG, a workgroup barrier per
// level), and per-RHS reductions as fixed-order trees in LDS.  Vectors are
// interleaved by right-hand side (entry (row i, rhs j) at i*nrhs + j), so the
// R lanes that share a row touch consecutive addresses.
#include "m4g_internal.hpp"
#include "m4g_device.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

namespace m4g {

// Offsets of one system in the concatenated arrays.
struct BatchSys {
    int64_t n, row0;          // rows; first row in the concatenated row space
    int64_t a_st, a_nz;       // A: starts (n+1 local pointers), first entry
    int64_t u_st, u_nz;       // U = upper(A) with the inverted diagonal first in each row
    int64_t l_st, l_nz;       // L = trans(U) strictly lower: column k and U index of each entry
    int64_t lvl_l, nlev_l;    // levels of the lower sweep: first pointer, count
    int64_t lvl_u, nlev_u;    // levels of the upper sweep
    int64_t rows_l, rows_u;   // first entry of each sweep's rows in level order
};

// Each sweep's rows in level order, packed for the solve: the row, its
// dependencies (fixed width, -1 = none) and their factor values, and the
// inverted diagonal.  Position q of a system is at row0 + q (times the width).
struct BatchSweep {
    const int32_t* row;       // [n_total]
    const int32_t* ci;        // [n_total * w]
    const double* val;        // [n_total * w]
    const double* diag;       // [n_total]
    int w;
};

constexpr int kBatchMaxR = 8;

}  // namespace m4g

struct m4g_batch_s {
    int64_t nsys = 0, n_total = 0, nnz_total = 0, u_total = 0;
    std::vector<m4g::BatchSys> h_sys;
    m4g::DevBuf<m4g::BatchSys> sys;
    m4g::DevBuf<int32_t> a_st, a_ci, u_st, u_ci, l_st, l_ci, l_up, lv_ptr, lv_rows;
    m4g::DevBuf<int64_t> u_apos;
    m4g::DevBuf<double> a_v, u_v;
    m4g::DevBuf<double> work;           // r, z, p, q and the sweep intermediate of every (system, rhs)
    int wl = 0, wu = 0;                 // widths of the packed lower / upper sweeps
    m4g::DevBuf<int32_t> lq_row, lq_ci, lq_up, lq_dp, uq_row, uq_ci, uq_up, uq_dp;   // positions into U
    m4g::DevBuf<int64_t> q_ubase;       // [n_total] first U entry of each position's system
    m4g::DevBuf<double> lq_val, lq_diag, uq_val, uq_diag;
    bool factored = false;
    int last_block = 0;                 // block size of the last solve
};

namespace m4g {

// ---------------------------------------------------------------------------
// factorisation: one workgroup per system, one barrier per level
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_batch_ic0(const BatchSys* __restrict__ sys, const int32_t* __restrict__ l_st,
                                                   const int32_t* __restrict__ l_ci, const int32_t* __restrict__ l_up,
                                                   const int32_t* __restrict__ u_st, const int32_t* __restrict__ lv_ptr,
                                                   const int32_t* __restrict__ lv_rows, double* __restrict__ u_v)
{
    const BatchSys S = sys[blockIdx.x];
    double* U = u_v + S.u_nz;
    const int32_t* ust = u_st + S.u_st;
    const int32_t* lst = l_st + S.l_st;
    const int32_t* lci = l_ci + S.l_nz;
    const int32_t* lup = l_up + S.l_nz;
    const int32_t* lp = lv_ptr + S.lvl_l;
    const int32_t* rows = lv_rows + S.rows_l;
    for (int64_t l = 0; l < S.nlev_l; l++) {
        for (int32_t q = lp[l] + threadIdx.x; q < lp[l + 1]; q += 256) {
            const int32_t i = rows[q];
            double D = U[ust[i]];
            for (int32_t e = lst[i]; e < lst[i + 1]; e++) {
                const int32_t p = lup[e];
                const double d = U[p] * U[ust[lci[e]]];   // U_ki *= 1/sqrt(D_k)   (ic_0.hpp:166)
                U[p] = d;
                D = D - d * d;                            // U_ii -= U_ki * U_ki   (ic_0.hpp:181)
            }
            U[ust[i]] = 1.0 / sqrt(D);                    // reciprocal(sqrt(U_ii)) (ic_0.hpp:160)
        }
        __syncthreads();
    }
}

__global__ void k_batch_gather(int64_t count, const int64_t* __restrict__ pos, const double* __restrict__ a,
                               double* __restrict__ u)
{
    for (int64_t e = blockIdx.x * int64_t(256) + threadIdx.x; e < count; e += int64_t(gridDim.x) * 256)
        u[e] = a[pos[e]];
}

// Factor values and inverted diagonals into the packed sweeps (after the
// factorisation): val[e] = U[up[e]] (0 where there is no entry), diag[q] = U[dp[q]].
__global__ void k_batch_pack(int64_t count, int w, const int32_t* __restrict__ up, const int32_t* __restrict__ dp,
                             const int64_t* __restrict__ ubase, const double* __restrict__ u, double* __restrict__ val,
                             double* __restrict__ diag)
{
    for (int64_t q = blockIdx.x * int64_t(256) + threadIdx.x; q < count; q += int64_t(gridDim.x) * 256) {
        const double* U = u + ubase[q];
        diag[q] = U[dp[q]];
        for (int k = 0; k < w; k++) {
            const int32_t p = up[q * w + k];
            val[q * w + k] = p >= 0 ? U[p] : 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// persistent CG
// ---------------------------------------------------------------------------
struct BatchCgArgs {
    BatchSweep lo, up;
    const BatchSys* sys;
    const int32_t *a_st, *a_ci, *u_st, *u_ci, *l_st, *l_ci, *l_up, *lv_ptr, *lv_rows;
    const double *a_v, *u_v;
    const int4* items;         // in processing order: system, first rhs, rhs count
    int32_t nitems, nrhs;
    unsigned* queue;
    const double* b;
    double* x;
    double *r, *z, *p, *q, *t;   // work vectors, same layout as x
    int max_iter;
    double rtol, atol;
    int* out_iters;            // [nsys * nrhs]
    double* out_resid;
    int* out_error;
};

// Per-RHS sums over the workgroup: thread t holds the partial of RHS t % R
// (NT is a multiple of R); a fixed-order tree in LDS leaves sum j in red[j].
template <int NT>
__device__ inline void rhs_sums(double v, double* red, int R)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int stride = NT / 2; stride >= R; stride >>= 1) {
        if (int(threadIdx.x) < stride) red[threadIdx.x] += red[threadIdx.x + stride];
        __syncthreads();
    }
}

// One triangular sweep, level by level, over a system's packed rows:
// out_i = (in_i - sum_k val_k * out_{ci_k}) * diag_i in stored order.  The
// metadata loads of a row are independent of each other, so a level costs two
// dependent memory round trips (row data, then the neighbours' values) and a
// workgroup barrier.
template <typename At>
__device__ inline void sweep_levels(const BatchSweep& w, const BatchSys& S, const int32_t* lp, int64_t nlev, int r0,
                                    int rstep, bool run, const double* in, double* out, At at)
{
    const int32_t* row = w.row + S.row0;
    const int32_t* ci = w.ci + S.row0 * w.w;
    const double* val = w.val + S.row0 * w.w;
    const double* diag = w.diag + S.row0;
    for (int64_t l = 0; l < nlev; l++) {
        if (run)
            for (int32_t q = lp[l] + r0; q < lp[l + 1]; q += rstep) {
                const int32_t i = row[q];
                double acc = in[at(i)];
                for (int k = 0; k < w.w; k++) {
                    const int32_t c = ci[int64_t(q) * w.w + k];
                    if (c >= 0) acc -= val[int64_t(q) * w.w + k] * out[at(c)];
                }
                out[at(i)] = acc * diag[q];
            }
        __syncthreads();
    }
}

template <int NT>
__global__ __launch_bounds__(NT) void k_batch_cg(BatchCgArgs a)
{
    __shared__ double red[NT];
    __shared__ double norm_r0[kBatchMaxR], rho[kBatchMaxR], rho_1[kBatchMaxR], alpha[kBatchMaxR], beta[kBatchMaxR];
    __shared__ double resid[kBatchMaxR];
    __shared__ int iters[kBatchMaxR], state[kBatchMaxR];   // state: 0 running, 1 finished
    __shared__ int item_sh, active_sh;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) item_sh = int(atomicAdd(a.queue, 1u));
        __syncthreads();
        const int item = item_sh;
        __syncthreads();
        if (item >= a.nitems) break;
        const int4 it = a.items[item];
        const int s = it.x, R = it.z;       // R divides NT
        const int jj = tid % R;             // this thread's right-hand side within the item
        const int rstep = NT / R;           // rows advanced per pass
        const int r0 = tid / R;
        const BatchSys S = a.sys[s];
        const int n = int(S.n);
        const int rhs = it.y + jj;          // global right-hand side of this thread
        const int64_t base = S.row0 * a.nrhs;
        auto at = [&](int i) { return base + int64_t(i) * a.nrhs + rhs; };
        const int32_t* ast = a.a_st + S.a_st;
        const int32_t* aci = a.a_ci + S.a_nz;
        const double* av = a.a_v + S.a_nz;
        const int32_t* lpl = a.lv_ptr + S.lvl_l;                // level pointers of the two sweeps
        const int32_t* lpu = a.lv_ptr + S.lvl_u;

        // r = b - A x (compute_summand.hpp:42-60) and ||b|| (basic_iteration ctor)
        double bb = 0.0;
        for (int i = r0; i < n; i += rstep) {
            double sum = 0.0;
            for (int32_t e = ast[i]; e < ast[i + 1]; e++) sum += av[e] * a.x[base + int64_t(aci[e]) * a.nrhs + rhs];
            const double bi = a.b[at(i)];
            a.r[at(i)] = bi - sum;
            bb += bi * bi;
        }
        rhs_sums<NT>(bb, red, R);
        if (tid < R) {
            norm_r0[tid] = sqrt(red[tid]);
            iters[tid] = 0;
            state[tid] = 0;
            rho_1[tid] = 0.0;
        }
        __syncthreads();
        for (;;) {
            // ---- while (!iter.finished(r)) { ++iter; ...  (two_norm(r)) ----
            double rr = 0.0;
            for (int i = r0; i < n; i += rstep) {
                const double ri = a.r[at(i)];
                rr += ri * ri;
            }
            rhs_sums<NT>(rr, red, R);
            if (tid == 0) {
                int active = 0;
                for (int j = 0; j < R; j++) {
                    if (state[j]) continue;
                    const double res = sqrt(red[j]);
                    resid[j] = res;
                    const bool conv = norm_r0[j] == 0 ? res <= a.atol : (res <= a.rtol * norm_r0[j] || res <= a.atol);
                    if (conv || iters[j] >= a.max_iter) {
                        state[j] = 1;
                        const int idx = s * a.nrhs + it.y + j;
                        a.out_iters[idx] = iters[j];
                        a.out_resid[idx] = res;
                        a.out_error[idx] = conv ? 0 : M4G_MAX_ITER;
                    } else {
                        iters[j] += 1;
                        active = 1;
                    }
                }
                active_sh = active;
            }
            __syncthreads();
            if (!active_sh) break;
            const bool run = state[jj] == 0;
            // ---- z = solve(P, r): lower sweep on L = trans(U) (lower_trisolve.hpp:191-209):
            //      t_i = (r_i - sum_k L_ik t_k) * L_ii, k ascending; then the upper
            //      sweep on U (upper_trisolve.hpp:159-176): z_i = (t_i - sum_j U_ij z_j) * U_ii
            sweep_levels(a.lo, S, lpl, S.nlev_l, r0, rstep, run, a.r, a.t, at);
            sweep_levels(a.up, S, lpu, S.nlev_u, r0, rstep, run, a.t, a.z, at);
            // ---- rho = dot(r, z) ----
            double rz = 0.0;
            if (run)
                for (int i = r0; i < n; i += rstep) rz += a.r[at(i)] * a.z[at(i)];
            rhs_sums<NT>(rz, red, R);
            if (tid < R) {
                rho[tid] = red[tid];
                beta[tid] = iters[tid] <= 1 ? 0.0 : red[tid] / rho_1[tid];
            }
            __syncthreads();
            // ---- p = z (first) or p = z + beta p ----
            if (run) {
                const bool first = iters[jj] <= 1;
                const double bj = beta[jj];
                for (int i = r0; i < n; i += rstep)
                    a.p[at(i)] = first ? a.z[at(i)] : a.z[at(i)] + bj * a.p[at(i)];
            }
            __syncthreads();
            // ---- q = A p, alpha = rho / dot(p, q) ----
            double pq = 0.0;
            if (run)
                for (int i = r0; i < n; i += rstep) {
                    double sum = 0.0;
                    for (int32_t e = ast[i]; e < ast[i + 1]; e++) sum += av[e] * a.p[at(aci[e])];
                    a.q[at(i)] = sum;
                    pq += a.p[at(i)] * sum;
                }
            rhs_sums<NT>(pq, red, R);
            if (tid < R) alpha[tid] = rho[tid] / red[tid];
            __syncthreads();
            // ---- x += alpha p; r -= alpha q; rho_1 = rho ----
            if (run) {
                const double al = alpha[jj];
                for (int i = r0; i < n; i += rstep) {
                    a.x[at(i)] += al * a.p[at(i)];
                    a.r[at(i)] -= al * a.q[at(i)];
                }
            }
            if (tid < R) rho_1[tid] = rho[tid];
            __syncthreads();
        }
    }
}

}  // namespace m4g

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace m4g {

// Natural-order levels: level(r) = 1 + max level of r's dependencies.  Rows
// are listed level by level (ascending row within a level for the lower
// sweep, descending for the upper one).
static void levels_of(int64_t n, const std::vector<int32_t>& st, const std::vector<int32_t>& ci, bool lower,
                      std::vector<int32_t>& ptr, std::vector<int32_t>& rows)
{
    std::vector<int32_t> lvl(n, 0);
    int32_t maxl = -1;
    for (int64_t q = 0; q < n; q++) {
        const int64_t r = lower ? q : n - 1 - q;
        int32_t l = 0;
        for (int32_t e = st[r]; e < st[r + 1]; e++) l = std::max(l, lvl[ci[e]] + 1);
        lvl[r] = l;
        maxl = std::max(maxl, l);
    }
    ptr.assign(maxl + 2, 0);
    for (int64_t r = 0; r < n; r++) ptr[lvl[r] + 1]++;
    for (int32_t l = 0; l <= maxl; l++) ptr[l + 1] += ptr[l];
    rows.assign(n, 0);
    std::vector<int32_t> f(ptr.begin(), ptr.end() - 1);
    for (int64_t q = 0; q < n; q++) {
        const int64_t r = lower ? q : n - 1 - q;
        rows[f[lvl[r]]++] = int32_t(r);
    }
}

// Launches the persistent CG with block size NT on the given items.
static void launch_batch_cg(m4g_batch_s& B, int NT, const BatchCgArgs& a, int grid, hipStream_t s)
{
    switch (NT) {
    case 256: hipLaunchKernelGGL(k_batch_cg<256>, dim3(grid), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL(k_batch_cg<512>, dim3(grid), dim3(512), 0, s, a); break;
    case 1024: hipLaunchKernelGGL(k_batch_cg<1024>, dim3(grid), dim3(1024), 0, s, a); break;
    default: throw Error(M4G_ERR_INVALID, "batch block size must be 256, 512 or 1024");
    }
    M4G_HIP(hipGetLastError());
}

// Runs the persistent CG on the given systems: each split into items of R
// right-hand sides (R: a power of two dividing nrhs, at most
// rhs_per_item[s]), largest systems first.
static void batch_run(m4g_batch_s& B, const std::vector<int>& systems, const int* rhs_per_item, int nrhs, int block,
                      double* x, const double* b, int max_iter, double rtol, double atol, int* out_it,
                      double* out_res, int* out_err, hipStream_t s)
{
    std::vector<int> order(systems);
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return B.h_sys[p].n > B.h_sys[q].n; });
    std::vector<int4> items;
    for (int sidx : order) {
        const int want = std::max(1, std::min(rhs_per_item ? rhs_per_item[sidx] : kBatchMaxR, kBatchMaxR));
        int R = 1;
        while (R * 2 <= want && nrhs % (R * 2) == 0) R *= 2;
        for (int j = 0; j < nrhs; j += R) items.push_back(make_int4(sidx, j, R, 0));
    }
    if (items.empty()) return;
    const size_t nv = size_t(B.n_total) * nrhs;
    if (B.work.count < 5 * nv) B.work.reset(5 * nv);
    DevBuf<int4> d_items(items.size());
    d_items.upload(items.data(), items.size());
    DevBuf<unsigned> queue(1);
    M4G_HIP(hipMemsetAsync(queue.ptr, 0, sizeof(unsigned), s));
    BatchCgArgs a;
    a.lo = BatchSweep{B.lq_row.ptr, B.lq_ci.ptr, B.lq_val.ptr, B.lq_diag.ptr, B.wl};
    a.up = BatchSweep{B.uq_row.ptr, B.uq_ci.ptr, B.uq_val.ptr, B.uq_diag.ptr, B.wu};
    a.sys = B.sys.ptr;
    a.a_st = B.a_st.ptr;
    a.a_ci = B.a_ci.ptr;
    a.u_st = B.u_st.ptr;
    a.u_ci = B.u_ci.ptr;
    a.l_st = B.l_st.ptr;
    a.l_ci = B.l_ci.ptr;
    a.l_up = B.l_up.ptr;
    a.lv_ptr = B.lv_ptr.ptr;
    a.lv_rows = B.lv_rows.ptr;
    a.a_v = B.a_v.ptr;
    a.u_v = B.u_v.ptr;
    a.items = d_items.ptr;
    a.nitems = int32_t(items.size());
    a.nrhs = nrhs;
    a.queue = queue.ptr;
    a.b = b;
    a.x = x;
    a.r = B.work.ptr;
    a.z = B.work.ptr + nv;
    a.p = B.work.ptr + 2 * nv;
    a.q = B.work.ptr + 3 * nv;
    a.t = B.work.ptr + 4 * nv;
    a.max_iter = max_iter;
    a.rtol = rtol;
    a.atol = atol;
    a.out_iters = out_it;
    a.out_resid = out_res;
    a.out_error = out_err;
    // persistent grid: as many workgroups as fit at once (at most one per item)
    const int per_cu = std::max(1, 2048 / block);
    const int grid = int(std::min<int64_t>(int64_t(items.size()), int64_t(device_state().num_cus) * per_cu));
    launch_batch_cg(B, block, a, std::max(1, grid), s);
    M4G_HIP(hipStreamSynchronize(s));   // items and queue are released on return
}

}  // namespace m4g

using namespace m4g;

extern "C" {

int m4g_batch_create(m4g_batch* out, int64_t nsys, const int64_t* sizes, const int64_t* starts,
                     const int64_t* indices, const double* values)
{
    return guard([&] {
        M4G_REQUIRE(out && nsys > 0 && sizes && starts && indices && values, M4G_ERR_INVALID,
                    "m4g_batch_create: null argument");
        auto B = std::make_unique<m4g_batch_s>();
        B->nsys = nsys;
        std::vector<int32_t> a_st, a_ci, u_st, u_ci, l_st, l_ci, l_up, lv_ptr, lv_rows;
        std::vector<int64_t> u_apos;
        std::vector<double> a_v;
        int64_t row0 = 0, a_nz = 0, u_nz = 0, l_nz = 0, sp = 0;
        std::vector<int32_t> rows_upper_all;
        std::vector<BatchSys> sys(nsys);
        for (int64_t s = 0; s < nsys; s++) {
            const int64_t n = sizes[s];
            M4G_REQUIRE(n > 0 && n < (int64_t(1) << 30), M4G_ERR_INVALID, "m4g_batch_create: bad system size");
            const int64_t* S = starts + sp;
            const int64_t nnz = S[n] - S[0];
            BatchSys& d = sys[s];
            d.n = n;
            d.row0 = row0;
            d.a_st = int64_t(a_st.size());
            d.a_nz = a_nz;
            // A (local starts from 0)
            for (int64_t r = 0; r <= n; r++) a_st.push_back(int32_t(S[r] - S[0]));
            for (int64_t e = 0; e < nnz; e++) {
                const int64_t c = indices[a_nz + e];
                M4G_REQUIRE(c >= 0 && c < n, M4G_ERR_INVALID, "m4g_batch_create: column out of range");
                a_ci.push_back(int32_t(c));
                a_v.push_back(values[a_nz + e]);
            }
            for (int64_t r = 0; r < n; r++)
                for (int64_t e = S[r] - S[0] + 1; e < S[r + 1] - S[0]; e++)
                    M4G_REQUIRE(indices[a_nz + e] > indices[a_nz + e - 1], M4G_ERR_LOGIC,
                                "Matrix entries must be sorted for this.");
            // U = upper(A): diagonal first (ic_0.hpp:148)
            d.u_st = int64_t(u_st.size());
            d.u_nz = u_nz;
            std::vector<int32_t> ust(n + 1, 0), uci;
            for (int64_t r = 0; r < n; r++) {
                bool first = true;
                for (int64_t e = S[r] - S[0]; e < S[r + 1] - S[0]; e++) {
                    const int64_t c = indices[a_nz + e];
                    if (c < r) continue;
                    M4G_REQUIRE(!first || c == r, M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                    first = false;
                    uci.push_back(int32_t(c));
                    u_apos.push_back(a_nz + e);
                }
                M4G_REQUIRE(!first, M4G_ERR_MISSING_DIAGONAL, "missing diagonal");
                ust[r + 1] = int32_t(uci.size());
            }
            // L = trans(U) strictly lower: row i lists U_ki, k < i ascending
            d.l_st = int64_t(l_st.size());
            d.l_nz = l_nz;
            std::vector<int32_t> lst(n + 1, 0);
            for (int64_t r = 0; r < n; r++)
                for (int32_t e = ust[r] + 1; e < ust[r + 1]; e++) lst[uci[e] + 1]++;
            for (int64_t r = 0; r < n; r++) lst[r + 1] += lst[r];
            std::vector<int32_t> lci(lst[n]), lup(lst[n]);
            {
                std::vector<int32_t> f(lst.begin(), lst.end() - 1);
                for (int64_t r = 0; r < n; r++)
                    for (int32_t e = ust[r] + 1; e < ust[r + 1]; e++) {
                        const int32_t q = f[uci[e]]++;
                        lci[q] = int32_t(r);
                        lup[q] = e;
                    }
            }
            // levels: lower DAG on L, upper DAG on strict U
            std::vector<int32_t> lp, lr, up, ur;
            levels_of(n, lst, lci, true, lp, lr);
            std::vector<int32_t> sust(n + 1, 0), suci;   // strictly upper pattern
            for (int64_t r = 0; r < n; r++) {
                for (int32_t e = ust[r] + 1; e < ust[r + 1]; e++) suci.push_back(uci[e]);
                sust[r + 1] = int32_t(suci.size());
            }
            levels_of(n, sust, suci, false, up, ur);
            d.lvl_l = int64_t(lv_ptr.size());
            d.nlev_l = int64_t(lp.size()) - 1;
            lv_ptr.insert(lv_ptr.end(), lp.begin(), lp.end());
            d.lvl_u = int64_t(lv_ptr.size());
            d.nlev_u = int64_t(up.size()) - 1;
            lv_ptr.insert(lv_ptr.end(), up.begin(), up.end());
            d.rows_l = int64_t(lv_rows.size());
            lv_rows.insert(lv_rows.end(), lr.begin(), lr.end());
            d.rows_u = int64_t(lv_rows.size());
            lv_rows.insert(lv_rows.end(), ur.begin(), ur.end());
            u_st.insert(u_st.end(), ust.begin(), ust.end());
            u_ci.insert(u_ci.end(), uci.begin(), uci.end());
            l_st.insert(l_st.end(), lst.begin(), lst.end());
            l_ci.insert(l_ci.end(), lci.begin(), lci.end());
            l_up.insert(l_up.end(), lup.begin(), lup.end());
            row0 += n;
            a_nz += nnz;
            u_nz += int64_t(uci.size());
            l_nz += int64_t(lci.size());
            sp += n + 1;
        }
        M4G_REQUIRE(a_nz < (int64_t(1) << 31) && u_nz < (int64_t(1) << 31), M4G_ERR_INVALID,
                    "m4g_batch_create: batch too large");
        // packed sweeps: rows in level order with fixed-width dependency lists
        int wl = 1, wu = 1;
        for (const BatchSys& d : sys)
            for (int64_t i = 0; i < d.n; i++) {
                wl = std::max(wl, l_st[d.l_st + i + 1] - l_st[d.l_st + i]);
                wu = std::max(wu, u_st[d.u_st + i + 1] - u_st[d.u_st + i] - 1);
            }
        std::vector<int32_t> lq_row(row0), lq_ci(row0 * wl, -1), lq_up(row0 * wl, -1), lq_dp(row0);
        std::vector<int32_t> uq_row(row0), uq_ci(row0 * wu, -1), uq_up(row0 * wu, -1), uq_dp(row0);
        std::vector<int64_t> q_ubase(row0);
        for (const BatchSys& d : sys)
            for (int64_t q = 0; q < d.n; q++) {
                const int64_t g = d.row0 + q;
                q_ubase[g] = d.u_nz;
                // lower sweep: L row i (U_ki, k ascending), diagonal U_ii
                const int32_t i = lv_rows[d.rows_l + q];
                lq_row[g] = i;
                lq_dp[g] = u_st[d.u_st + i];
                for (int32_t e = l_st[d.l_st + i], k = 0; e < l_st[d.l_st + i + 1]; e++, k++) {
                    lq_ci[g * wl + k] = l_ci[d.l_nz + e];
                    lq_up[g * wl + k] = l_up[d.l_nz + e];
                }
                // upper sweep: U row j after its diagonal, ascending columns
                const int32_t j = lv_rows[d.rows_u + q];
                uq_row[g] = j;
                uq_dp[g] = u_st[d.u_st + j];
                for (int32_t e = u_st[d.u_st + j] + 1, k = 0; e < u_st[d.u_st + j + 1]; e++, k++) {
                    uq_ci[g * wu + k] = u_ci[d.u_nz + e];
                    uq_up[g * wu + k] = e;
                }
            }
        B->n_total = row0;
        B->nnz_total = a_nz;
        B->u_total = u_nz;
        B->h_sys = sys;
        auto up32 = [](DevBuf<int32_t>& d, const std::vector<int32_t>& h) {
            d.reset(std::max<size_t>(h.size(), 1));
            d.upload(h.data(), h.size());
        };
        B->sys.reset(nsys);
        B->sys.upload(sys.data(), nsys);
        up32(B->a_st, a_st);
        up32(B->a_ci, a_ci);
        up32(B->u_st, u_st);
        up32(B->u_ci, u_ci);
        up32(B->l_st, l_st);
        up32(B->l_ci, l_ci);
        up32(B->l_up, l_up);
        up32(B->lv_ptr, lv_ptr);
        up32(B->lv_rows, lv_rows);
        B->u_apos.reset(std::max<size_t>(u_apos.size(), 1));
        B->u_apos.upload(u_apos.data(), u_apos.size());
        B->a_v.reset(std::max<size_t>(a_v.size(), 1));
        B->a_v.upload(a_v.data(), a_v.size());
        B->u_v.reset(std::max<int64_t>(u_nz, 1));
        B->wl = wl;
        B->wu = wu;
        up32(B->lq_row, lq_row);
        up32(B->lq_ci, lq_ci);
        up32(B->lq_up, lq_up);
        up32(B->lq_dp, lq_dp);
        up32(B->uq_row, uq_row);
        up32(B->uq_ci, uq_ci);
        up32(B->uq_up, uq_up);
        up32(B->uq_dp, uq_dp);
        B->q_ubase.reset(std::max<int64_t>(row0, 1));
        B->q_ubase.upload(q_ubase.data(), q_ubase.size());
        B->lq_val.reset(std::max<int64_t>(row0 * wl, 1));
        B->lq_diag.reset(std::max<int64_t>(row0, 1));
        B->uq_val.reset(std::max<int64_t>(row0 * wu, 1));
        B->uq_diag.reset(std::max<int64_t>(row0, 1));
        M4G_HIP(hipStreamSynchronize(current_stream()));
        *out = B.release();
        return M4G_OK;
    });
}

int m4g_batch_destroy(m4g_batch B)
{
    return guard([&] {
        delete B;
        return M4G_OK;
    });
}

int m4g_batch_dims(m4g_batch B, int64_t* nsys, int64_t* n_total, int64_t* nnz_total)
{
    return guard([&] {
        M4G_REQUIRE(B, M4G_ERR_INVALID, "null batch");
        if (nsys) *nsys = B->nsys;
        if (n_total) *n_total = B->n_total;
        if (nnz_total) *nnz_total = B->nnz_total;
        return M4G_OK;
    });
}

int m4g_batch_factor_ic0(m4g_batch B)
{
    return guard([&] {
        M4G_REQUIRE(B, M4G_ERR_INVALID, "null batch");
        hipStream_t s = current_stream();
        const int64_t cnt = B->u_total;
        if (cnt)
            hipLaunchKernelGGL(k_batch_gather, dim3(int(std::min<int64_t>((cnt + 255) / 256, 16384))), dim3(256), 0,
                               s, cnt, B->u_apos.ptr, B->a_v.ptr, B->u_v.ptr);
        hipLaunchKernelGGL(k_batch_ic0, dim3(int(B->nsys)), dim3(256), 0, s, B->sys.ptr, B->l_st.ptr, B->l_ci.ptr,
                           B->l_up.ptr, B->u_st.ptr, B->lv_ptr.ptr, B->lv_rows.ptr, B->u_v.ptr);
        // factor values into the packed sweeps
        const int pg = int(std::min<int64_t>((B->n_total + 255) / 256, 16384));
        hipLaunchKernelGGL(k_batch_pack, dim3(pg), dim3(256), 0, s, B->n_total, B->wl, B->lq_up.ptr, B->lq_dp.ptr,
                           B->q_ubase.ptr, B->u_v.ptr, B->lq_val.ptr, B->lq_diag.ptr);
        hipLaunchKernelGGL(k_batch_pack, dim3(pg), dim3(256), 0, s, B->n_total, B->wu, B->uq_up.ptr, B->uq_dp.ptr,
                           B->q_ubase.ptr, B->u_v.ptr, B->uq_val.ptr, B->uq_diag.ptr);
        M4G_HIP(hipGetLastError());
        M4G_HIP(hipStreamSynchronize(s));
        B->factored = true;
        return M4G_OK;
    });
}

int m4g_batch_download_factor(m4g_batch B, double* values)
{
    return guard([&] {
        M4G_REQUIRE(B && values && B->factored, M4G_ERR_INVALID, "m4g_batch_download_factor: not factored");
        B->u_v.download(values, size_t(B->u_total));
        return M4G_OK;
    });
}

int m4g_batch_cg_ic0(m4g_batch B, int nrhs, m4g_vector x, m4g_vector b, int max_iter, double rtol, double atol,
                     const int* rhs_per_item, int block, int* iterations, double* resid, int* error)
{
    return guard([&] {
        M4G_REQUIRE(B && x && b && iterations && resid && error, M4G_ERR_INVALID, "m4g_batch_cg_ic0: null argument");
        M4G_REQUIRE(B->factored, M4G_ERR_INVALID, "m4g_batch_cg_ic0: call m4g_batch_factor_ic0 first");
        M4G_REQUIRE(nrhs >= 1, M4G_ERR_INVALID, "m4g_batch_cg_ic0: nrhs");
        M4G_REQUIRE(x->type == M4G_F64 && b->type == M4G_F64 && x->n == B->n_total * nrhs && b->n == x->n,
                    M4G_ERR_SIZE, "m4g_batch_cg_ic0: x and b need n_total * nrhs doubles");
        if (block == 0) block = 512;
        std::vector<int> all(B->nsys);
        std::iota(all.begin(), all.end(), 0);
        DevBuf<int> d_it(size_t(B->nsys) * nrhs), d_err(size_t(B->nsys) * nrhs);
        DevBuf<double> d_res(size_t(B->nsys) * nrhs);
        batch_run(*B, all, rhs_per_item, nrhs, block, static_cast<double*>(x->ptr),
                  static_cast<const double*>(b->ptr), max_iter, rtol, atol, d_it.ptr, d_res.ptr, d_err.ptr,
                  current_stream());
        d_it.download(iterations, d_it.count);
        d_res.download(resid, d_res.count);
        d_err.download(error, d_err.count);
        B->last_block = block;
        return M4G_OK;
    });
}

int m4g_batch_autotune(m4g_batch B, int nrhs, m4g_vector b, int trial_iters, int* rhs_per_item, int* block,
                       double* trial_ms)
{
    return guard([&] {
        M4G_REQUIRE(B && b && rhs_per_item && block, M4G_ERR_INVALID, "m4g_batch_autotune: null argument");
        M4G_REQUIRE(B->factored, M4G_ERR_INVALID, "m4g_batch_autotune: call m4g_batch_factor_ic0 first");
        M4G_REQUIRE(b->type == M4G_F64 && b->n == B->n_total * nrhs, M4G_ERR_SIZE, "m4g_batch_autotune: b size");
        if (trial_iters <= 0) trial_iters = 20;
        hipStream_t s = current_stream();
        // size classes: thirds of the systems by size
        std::vector<int> order(B->nsys);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return B->h_sys[p].n < B->h_sys[q].n; });
        const int nclass = int(std::min<int64_t>(3, B->nsys));
        std::vector<std::vector<int>> cls(nclass);
        for (int64_t k = 0; k < B->nsys; k++) cls[size_t(k * nclass / B->nsys)].push_back(order[k]);
        const int blocks[3] = {256, 512, 1024};
        std::vector<int> Rs;
        for (int R = kBatchMaxR; R >= 1; R /= 2)
            if (nrhs % R == 0) Rs.push_back(R);
        DevBuf<double> xs(size_t(B->n_total) * nrhs);
        DevBuf<int> d_it(size_t(B->nsys) * nrhs), d_err(size_t(B->nsys) * nrhs);
        DevBuf<double> d_res(size_t(B->nsys) * nrhs);
        std::vector<int> all(B->nsys);
        std::iota(all.begin(), all.end(), 0);
        hipEvent_t e0, e1;
        M4G_HIP(hipEventCreate(&e0));
        M4G_HIP(hipEventCreate(&e1));
        // one trial: trial_iters iterations of the whole batch
        auto trial = [&](const std::vector<int>& rpi, int blk) {
            M4G_HIP(hipMemsetAsync(xs.ptr, 0, xs.count * sizeof(double), s));
            M4G_HIP(hipEventRecord(e0, s));
            batch_run(*B, all, rpi.data(), nrhs, blk, xs.ptr, static_cast<const double*>(b->ptr), trial_iters, 0.0,
                      0.0, d_it.ptr, d_res.ptr, d_err.ptr, s);
            M4G_HIP(hipEventRecord(e1, s));
            M4G_HIP(hipEventSynchronize(e1));
            float ms = 0;
            M4G_HIP(hipEventElapsedTime(&ms, e0, e1));
            return double(ms);
        };
        std::vector<double> t;
        // stage 1: every block size with one grouping for all systems
        int best_b = 0;
        size_t best_r = 0;
        double best = 1e300;
        for (int bi = 0; bi < 3; bi++)
            for (size_t ri = 0; ri < Rs.size(); ri++) {
                const double ms = trial(std::vector<int>(B->nsys, Rs[ri]), blocks[bi]);
                t.push_back(ms);
                if (ms < best) {
                    best = ms;
                    best_b = bi;
                    best_r = ri;
                }
            }
        // stage 2: with that block size, each size class tries the other groupings
        std::vector<int> cur(B->nsys, Rs[best_r]);
        for (int c = 0; c < nclass; c++) {
            int keep = Rs[best_r];
            for (size_t ri = 0; ri < Rs.size(); ri++) {
                if (Rs[ri] == Rs[best_r]) {
                    t.push_back(best);
                    continue;
                }
                std::vector<int> rpi = cur;
                for (int sidx : cls[c]) rpi[sidx] = Rs[ri];
                const double ms = trial(rpi, blocks[best_b]);
                t.push_back(ms);
                if (ms < best) {
                    best = ms;
                    keep = Rs[ri];
                }
            }
            for (int sidx : cls[c]) cur[sidx] = keep;
        }
        M4G_HIP(hipEventDestroy(e0));
        M4G_HIP(hipEventDestroy(e1));
        *block = blocks[best_b];
        for (int64_t k = 0; k < B->nsys; k++) rhs_per_item[k] = cur[k];
        if (trial_ms)
            for (size_t k = 0; k < t.size(); k++) trial_ms[k] = t[k];
        B->last_block = *block;
        return M4G_OK;
    });
}

}  // extern "C"
