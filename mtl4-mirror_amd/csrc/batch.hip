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
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <sys/stat.h>
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
    int64_t reach_l, reach_u; // largest (end of a level) - (position of a dependency of that level)
    int64_t level_width;      // rows of the widest level of either sweep
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
    int W = 0;                          // width of the packed sweeps; rows of A hold 2W+1 entries
    m4g::DevBuf<int32_t> a_eci;         // A as fixed-width rows (-1 = none)
    m4g::DevBuf<double> a_ev;
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
    const int32_t* a_ci;       // A as fixed-width rows: [n_total * (2W+1)], -1 = none
    const double* a_v;
    const int32_t* lv_ptr;
    const int4* items;         // in processing order: system, first rhs, rhs count
    int32_t nitems, nrhs;
    unsigned* queue;
    const double* b;
    double* x;
    double *r, *z, *p, *q, *t;   // work vectors, same layout as x
    int max_iter;
    const int* sys_max_iter;   // per-system iteration limit (autotuner trials) or null
    double rtol, atol;
    int ring_doubles;          // dynamic LDS of the sweep ring (a power of two)
    double* dump;              // [grid * NT] stores of unused sweep pairs (never read)
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

typedef __attribute__((address_space(3))) double lds_double;   // LDS-qualified, so accesses are ds_* ops


// The inputs of one (position, rhs) pair of a sweep that do not depend on
// earlier levels: the row, its dependencies (positions in the sweep's level
// order, -1 = none), factor values and inverted diagonal.
template <int W>
struct SweepMeta {
    int32_t q, i;
    int32_t c[W];
    double v[W];
    double d;
};

// One triangular sweep of one system, level by level:
// out_i = (in_i - sum_k val_k * out_{c_k}) * diag_i in stored order.  Results
// also go to an LDS ring indexed by level-order position (ring_pos slots per
// rhs, plus one scratch slot).  DOT: also accumulates dv_i * out_i into `dot`.
//
// Pipelined form, for systems whose levels fit in one pass of the block and
// whose dependencies all lie within ring_pos positions of the end of their
// level: a level costs a barrier and an LDS round trip.  Each thread's pair of
// level l + D is loaded while level l runs and its right-hand side (which
// needs the row) two levels ahead; all loads are unconditional (clamped to
// position 0, results of unused pairs go to the scratch slot) and the level
// loop is unrolled D times so slot indices are constants: the compiler's wait
// counts then never hold a level for a load issued after the level began.
template <int W, bool DOT, typename At>
__device__ inline void sweep_ring_pipe(const BatchSweep& sw, const BatchSys& S, const int32_t* lp, int64_t nlev,
                                       int r0, int jj, int R, bool run, const double* in, double* out, At at,
                                       lds_double* ring, int ring_pos, const double* dv, double& dot, double* dump)
{
    constexpr int D = W <= 4 ? 4 : 3;
    const int32_t* row = sw.row + S.row0;
    const int32_t* ci = sw.ci + S.row0 * W;
    const double* val = sw.val + S.row0 * W;
    const double* diag = sw.diag + S.row0;
    const int mask = ring_pos - 1;
    auto meta = [&](int32_t q) {
        SweepMeta<W> m;
        m.q = q;
        m.i = row[q];
#pragma unroll
        for (int k = 0; k < W; k++) {
            m.c[k] = ci[int64_t(q) * W + k];
            m.v[k] = val[int64_t(q) * W + k];
        }
        m.d = diag[q];
        return m;
    };
    // this thread's pair of level l, or position 0 (not used); the level
    // bounds are uniform (clamped, so read unconditionally)
    auto first = [&](int64_t l, bool& ok) -> int32_t {
        const int32_t b0 = lp[l < nlev ? l : nlev], b1 = lp[l < nlev ? l + 1 : nlev];
        ok = run & (b0 + r0 < b1);
        return ok ? b0 + r0 : 0;
    };
    const int dump_slot = ring_pos, zero_slot = ring_pos + 1;
    SweepMeta<W> m[D];
    bool ok[D];
    double x[D], y[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        m[k] = meta(first(k, ok[k]));
        x[k] = 0.0;
        y[k] = 0.0;
    }
    x[0] = in[at(m[0].i)];
    x[1] = in[at(m[1].i)];
    if (DOT) {
        y[0] = dv[at(m[0].i)];
        y[1] = dv[at(m[1].i)];
    }
    for (int64_t l0 = 0; l0 < nlev; l0 += D) {
#pragma unroll
        for (int u = 0; u < D; u++) {
            const int64_t l = l0 + u;
            if (l < nlev) {
                const int u2 = (u + 2) % D;   // a constant once unrolled
                x[u2] = in[at(m[u2].i)];     // right-hand side of level l + 2
                if (DOT) y[u2] = dv[at(m[u2].i)];
                // padding entries (c = -1, value 0) read the zero slot:
                // acc - 0 * 0 == acc exactly, so every entry is applied
                // without a branch; an unused pair writes the dump slots
                double acc = x[u];
#pragma unroll
                for (int k = 0; k < W; k++) {
                    const int32_t c = m[u].c[k];
                    const double yv = ring[(c >= 0 ? (c & mask) : zero_slot) * R + jj];
                    acc -= m[u].v[k] * yv;
                }
                const double res = acc * m[u].d;
                ring[(ok[u] ? (m[u].q & mask) : dump_slot) * R + jj] = res;
                *(ok[u] ? &out[at(m[u].i)] : &dump[threadIdx.x]) = res;
                if (DOT) dot = ok[u] ? dot + y[u] * res : dot;
                m[u] = meta(first(l + D, ok[u]));   // level l + D
                lds_barrier();
            }
        }
    }
}

// General form: any level width, dependencies beyond the ring read from
// `out` behind full barriers; loads are issued in the level that uses them.
template <int W, bool DOT, typename At>
__device__ inline void sweep_ring_general(const BatchSweep& sw, const BatchSys& S, const int32_t* lp, int64_t nlev,
                                          int r0, int rstep, int jj, int R, bool run, const double* in, double* out,
                                          At at, lds_double* ring, int ring_pos, const double* dv, double& dot)
{
    const int32_t* row = sw.row + S.row0;
    const int32_t* ci = sw.ci + S.row0 * W;
    const double* val = sw.val + S.row0 * W;
    const double* diag = sw.diag + S.row0;
    const int mask = ring_pos - 1;
    for (int64_t l = 0; l < nlev; l++) {
        const int32_t lend = lp[l + 1];
        if (run)
            for (int32_t q = lp[l] + r0; q < lend; q += rstep) {
                const int32_t i = row[q];
                double acc = in[at(i)];
                for (int k = 0; k < W; k++) {
                    const int32_t c = ci[int64_t(q) * W + k];
                    if (c >= 0) {
                        const double yv = c >= lend - ring_pos ? ring[(c & mask) * R + jj] : out[at(row[c])];
                        acc -= val[int64_t(q) * W + k] * yv;
                    }
                }
                const double res = acc * diag[q];
                ring[(q & mask) * R + jj] = res;
                out[at(i)] = res;
                if (DOT) dot += dv[at(i)] * res;
            }
        __syncthreads();
    }
}

template <int NT, int W, bool DOT, typename At>
__device__ inline void sweep_ring(const BatchSweep& sw, const BatchSys& S, const int32_t* lp, int64_t nlev,
                                  int64_t reach, int r0, int rstep, int jj, int R, bool run, const double* in,
                                  double* out, At at, lds_double* ring, int ring_pos, const double* dv, double& dot,
                                  double* dump)
{
    if (reach <= ring_pos && S.level_width * R <= NT)
        sweep_ring_pipe<W, DOT>(sw, S, lp, nlev, r0, jj, R, run, in, out, at, ring, ring_pos, dv, dot, dump);
    else
        sweep_ring_general<W, DOT>(sw, S, lp, nlev, r0, rstep, jj, R, run, in, out, at, ring, ring_pos, dv, dot);
}

// Row loops of the vector kernels: each thread takes rows r0, r0 + rstep, ...
// in groups of kRowsInFlight so that the loads of a group are issued together.
constexpr int kRowsInFlight = 2;

template <int NT, int W>
// lv_ptr is read-only for the whole launch (a separate __restrict__ argument
// so its uniform reads become scalar loads that never wait on vector memory)
__global__ __launch_bounds__(NT) void k_batch_cg(BatchCgArgs a, const int32_t* __restrict__ lv_ptr)
{
    constexpr int E = 2 * W + 1;   // entries per row of A
    extern __shared__ double ring[];
    __shared__ double red[NT];
    __shared__ double norm_r0[kBatchMaxR], rho[kBatchMaxR], rho_1[kBatchMaxR], alpha[kBatchMaxR], beta[kBatchMaxR];
    __shared__ double resid[kBatchMaxR];
    __shared__ int iters[kBatchMaxR], state[kBatchMaxR];   // state: 0 running, 1 finished
    __shared__ int item_sh, active_sh;
    const int tid = threadIdx.x;
    for (;;) {
        if (tid == 0) item_sh = int(atomicAdd(a.queue, 1u));
        __syncthreads();
        // uniform from here on: item, system and level pointers live in scalar registers
        const int item = __builtin_amdgcn_readfirstlane(item_sh);
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
        const int ring_pos = a.ring_doubles / R;
        lds_double* ring3 = (lds_double*)ring;
        // two slots after the ring: dump (writes of unused pairs), zero (padding reads)
        if (tid < 2 * kBatchMaxR) ring3[a.ring_doubles + tid] = 0.0;
        double* dump = a.dump + int64_t(blockIdx.x) * NT;
        auto at = [&](int i) { return base + int64_t(i) * a.nrhs + rhs; };
        const int32_t* aci = a.a_ci + S.row0 * E;
        const double* av = a.a_v + S.row0 * E;
        const int32_t* lpl = lv_ptr + S.lvl_l;                  // level pointers of the two sweeps
        const int32_t* lpu = lv_ptr + S.lvl_u;
        // (A v)_i for the rows of a group, entries in stored order (mat_vec_mult.hpp:662-687)
        // (rows loaded together: as many as keep the registers within bounds)
        auto spmv_rows = [&](const double* v, const int* ii, double* out) {
            constexpr int RS = E <= 5 ? 2 : 1;
#pragma unroll
            for (int u0 = 0; u0 < kRowsInFlight; u0 += RS) {
                int32_t c[RS][E];
                double w[RS][E], xv[RS][E];
#pragma unroll
                for (int u = 0; u < RS; u++)
#pragma unroll
                    for (int k = 0; k < E; k++) {
                        const int i = ii[u0 + u] < n ? ii[u0 + u] : n - 1;
                        c[u][k] = aci[int64_t(i) * E + k];
                        w[u][k] = av[int64_t(i) * E + k];
                    }
#pragma unroll
                for (int u = 0; u < RS; u++)
#pragma unroll
                    for (int k = 0; k < E; k++) xv[u][k] = v[at(c[u][k] >= 0 ? c[u][k] : 0)];
#pragma unroll
                for (int u = 0; u < RS; u++) {
                    double sum = 0.0;
#pragma unroll
                    for (int k = 0; k < E; k++)
                        if (c[u][k] >= 0) sum += w[u][k] * xv[u][k];
                    out[u0 + u] = sum;
                }
            }
        };

        // r = b - A x (compute_summand.hpp:42-60) and ||b|| (basic_iteration ctor);
        // rr: this thread's part of |r|^2 for the next convergence test
        double bb = 0.0, rr = 0.0;
        for (int i0 = r0; i0 < n; i0 += kRowsInFlight * rstep) {
            int ii[kRowsInFlight];
            double ax[kRowsInFlight], bv[kRowsInFlight];
#pragma unroll
            for (int u = 0; u < kRowsInFlight; u++) {
                ii[u] = i0 + u * rstep;
                bv[u] = a.b[at(ii[u] < n ? ii[u] : n - 1)];
            }
            spmv_rows(a.x, ii, ax);
#pragma unroll
            for (int u = 0; u < kRowsInFlight; u++)
                if (ii[u] < n) {
                    const double ri = bv[u] - ax[u];
                    a.r[at(ii[u])] = ri;
                    bb += bv[u] * bv[u];
                    rr += ri * ri;
                }
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
            // ---- while (!iter.finished(r)) { ++iter; ...  (two_norm(r), summed
            //      when r was written) ----
            rhs_sums<NT>(rr, red, R);
            if (tid == 0) {
                int active = 0;
                for (int j = 0; j < R; j++) {
                    if (state[j]) continue;
                    const double res = sqrt(red[j]);
                    resid[j] = res;
                    const bool conv = norm_r0[j] == 0 ? res <= a.atol : (res <= a.rtol * norm_r0[j] || res <= a.atol);
                    if (conv || iters[j] >= (a.sys_max_iter ? a.sys_max_iter[s] : a.max_iter)) {
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
            //      rho = dot(r, z) is summed as z is produced
            double rz = 0.0;
            sweep_ring<NT, W, false>(a.lo, S, lpl, S.nlev_l, S.reach_l, r0, rstep, jj, R, run, a.r, a.t, at, ring3,
                                     ring_pos, nullptr, rz, dump);
            __syncthreads();   // t complete in global memory; the ring is reused
            sweep_ring<NT, W, true>(a.up, S, lpu, S.nlev_u, S.reach_u, r0, rstep, jj, R, run, a.t, a.z, at, ring3,
                                    ring_pos, a.r, rz, dump);
            // ---- p = z (first) or p = z + beta p ----
            rhs_sums<NT>(rz, red, R);
            if (tid < R) {
                rho[tid] = red[tid];
                beta[tid] = iters[tid] <= 1 ? 0.0 : red[tid] / rho_1[tid];
            }
            __syncthreads();
            if (run) {
                const bool first = iters[jj] <= 1;
                const double bj = beta[jj];
                for (int i0 = r0; i0 < n; i0 += kRowsInFlight * rstep) {
                    double zv[kRowsInFlight], pv[kRowsInFlight];
#pragma unroll
                    for (int u = 0; u < kRowsInFlight; u++) {
                        const int i = i0 + u * rstep < n ? i0 + u * rstep : n - 1;
                        zv[u] = a.z[at(i)];
                        pv[u] = a.p[at(i)];
                    }
#pragma unroll
                    for (int u = 0; u < kRowsInFlight; u++)
                        if (i0 + u * rstep < n) a.p[at(i0 + u * rstep)] = first ? zv[u] : zv[u] + bj * pv[u];
                }
            }
            __syncthreads();
            // ---- q = A p, alpha = rho / dot(p, q) ----
            double pq = 0.0;
            if (run)
                for (int i0 = r0; i0 < n; i0 += kRowsInFlight * rstep) {
                    int ii[kRowsInFlight];
                    double qv[kRowsInFlight], pv[kRowsInFlight];
#pragma unroll
                    for (int u = 0; u < kRowsInFlight; u++) {
                        ii[u] = i0 + u * rstep;
                        pv[u] = a.p[at(ii[u] < n ? ii[u] : n - 1)];
                    }
                    spmv_rows(a.p, ii, qv);
#pragma unroll
                    for (int u = 0; u < kRowsInFlight; u++)
                        if (ii[u] < n) {
                            a.q[at(ii[u])] = qv[u];
                            pq += pv[u] * qv[u];
                        }
                }
            rhs_sums<NT>(pq, red, R);
            if (tid < R) alpha[tid] = rho[tid] / red[tid];
            __syncthreads();
            // ---- x += alpha p; r -= alpha q (and |r|^2); rho_1 = rho ----
            rr = 0.0;
            if (run) {
                const double al = alpha[jj];
                for (int i0 = r0; i0 < n; i0 += kRowsInFlight * rstep) {
                    double xv[kRowsInFlight], pv[kRowsInFlight], rv[kRowsInFlight], qv[kRowsInFlight];
#pragma unroll
                    for (int u = 0; u < kRowsInFlight; u++) {
                        const int i = i0 + u * rstep < n ? i0 + u * rstep : n - 1;
                        xv[u] = a.x[at(i)];
                        pv[u] = a.p[at(i)];
                        rv[u] = a.r[at(i)];
                        qv[u] = a.q[at(i)];
                    }
#pragma unroll
                    for (int u = 0; u < kRowsInFlight; u++)
                        if (i0 + u * rstep < n) {
                            const double ri = rv[u] - al * qv[u];
                            a.x[at(i0 + u * rstep)] = xv[u] + al * pv[u];
                            a.r[at(i0 + u * rstep)] = ri;
                            rr += ri * ri;
                        }
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

// LDS ring of the sweeps per workgroup size: with the kernel's static LDS,
// 2048 threads' worth of workgroups fit in a CU's 160 KB
static int ring_doubles_for(int NT)
{
    return NT == 1024 ? 8192 : NT == 512 ? 4096 : 2048;
}

template <int NT>
static void launch_batch_cg_w(int W, const BatchCgArgs& a, int grid, hipStream_t s)
{
    const size_t lds = size_t(a.ring_doubles + 2 * kBatchMaxR) * sizeof(double);   // ring + dump and zero slots
    switch (W) {
    case 2: hipLaunchKernelGGL((k_batch_cg<NT, 2>), dim3(grid), dim3(NT), lds, s, a, a.lv_ptr); break;
    case 4: hipLaunchKernelGGL((k_batch_cg<NT, 4>), dim3(grid), dim3(NT), lds, s, a, a.lv_ptr); break;
    case 8: hipLaunchKernelGGL((k_batch_cg<NT, 8>), dim3(grid), dim3(NT), lds, s, a, a.lv_ptr); break;
    default: hipLaunchKernelGGL((k_batch_cg<NT, 16>), dim3(grid), dim3(NT), lds, s, a, a.lv_ptr); break;
    }
}

// Launches the persistent CG with block size NT on the given items.
static void launch_batch_cg(m4g_batch_s& B, int NT, const BatchCgArgs& a, int grid, hipStream_t s)
{
    switch (NT) {
    case 256: launch_batch_cg_w<256>(B.W, a, grid, s); break;
    case 512: launch_batch_cg_w<512>(B.W, a, grid, s); break;
    case 1024: launch_batch_cg_w<1024>(B.W, a, grid, s); break;
    default: throw Error(M4G_ERR_INVALID, "batch block size must be 256, 512 or 1024");
    }
    M4G_HIP(hipGetLastError());
}

// Runs the persistent CG on the given systems: each split into items of R
// right-hand sides (R: a power of two dividing nrhs, at most
// rhs_per_item[s]), largest systems first.
static void batch_run(m4g_batch_s& B, const std::vector<int>& systems, const int* rhs_per_item, int nrhs, int block,
                      double* x, const double* b, int max_iter, double rtol, double atol, int* out_it,
                      double* out_res, int* out_err, hipStream_t s, const int* sys_max_iter = nullptr)
{
    std::vector<int> order(systems);
    std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return B.h_sys[p].n > B.h_sys[q].n; });
    std::vector<int4> items;
    for (int sidx : order) {
        const int want = std::max(1, std::min(rhs_per_item ? rhs_per_item[sidx] : kBatchMaxR, kBatchMaxR));
        // at most as many rhs as keep a sweep level within one pass of the block
        int R = 1;
        while (R * 2 <= want && nrhs % (R * 2) == 0 && R * 2 * B.h_sys[sidx].level_width <= block) R *= 2;
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
    a.lo = BatchSweep{B.lq_row.ptr, B.lq_ci.ptr, B.lq_val.ptr, B.lq_diag.ptr, B.W};
    a.up = BatchSweep{B.uq_row.ptr, B.uq_ci.ptr, B.uq_val.ptr, B.uq_diag.ptr, B.W};
    a.sys = B.sys.ptr;
    a.a_ci = B.a_eci.ptr;
    a.a_v = B.a_ev.ptr;
    a.lv_ptr = B.lv_ptr.ptr;
    a.ring_doubles = ring_doubles_for(block);
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
    a.sys_max_iter = sys_max_iter;
    a.rtol = rtol;
    a.atol = atol;
    a.out_iters = out_it;
    a.out_resid = out_res;
    a.out_error = out_err;
    // persistent grid: as many workgroups as fit at once (at most one per item)
    const int per_cu = std::max(1, 2048 / block);
    const int grid = int(std::min<int64_t>(int64_t(items.size()), int64_t(device_state().num_cus) * per_cu));
    DevBuf<double> dump(size_t(std::max(1, grid)) * block);
    a.dump = dump.ptr;
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
        // packed sweeps: rows in level order with fixed-width dependency
        // lists (positions in the sweep's level order); A as fixed-width rows
        int wmax = 1;
        for (const BatchSys& d : sys)
            for (int64_t i = 0; i < d.n; i++) {
                wmax = std::max(wmax, l_st[d.l_st + i + 1] - l_st[d.l_st + i]);
                wmax = std::max(wmax, u_st[d.u_st + i + 1] - u_st[d.u_st + i] - 1);
                wmax = std::max(wmax, (a_st[d.a_st + i + 1] - a_st[d.a_st + i]) / 2);
            }
        int W = 2;
        while (W < wmax) W *= 2;
        M4G_REQUIRE(W <= 16, M4G_ERR_INVALID, "m4g_batch_create: rows of more than 33 entries are not supported");
        const int E = 2 * W + 1;
        std::vector<int32_t> lq_row(row0), lq_ci(row0 * W, -1), lq_up(row0 * W, -1), lq_dp(row0);
        std::vector<int32_t> uq_row(row0), uq_ci(row0 * W, -1), uq_up(row0 * W, -1), uq_dp(row0);
        std::vector<int32_t> a_eci(row0 * E, -1);
        std::vector<double> a_ev(row0 * E, 0.0);
        std::vector<int64_t> q_ubase(row0);
        for (BatchSys& d : sys) {
            std::vector<int32_t> posl(d.n), posu(d.n);
            for (int64_t q = 0; q < d.n; q++) {
                posl[lv_rows[d.rows_l + q]] = int32_t(q);
                posu[lv_rows[d.rows_u + q]] = int32_t(q);
            }
            for (int64_t i = 0; i < d.n; i++)
                for (int32_t e = a_st[d.a_st + i], k = 0; e < a_st[d.a_st + i + 1]; e++, k++) {
                    a_eci[(d.row0 + i) * E + k] = a_ci[d.a_nz + e];
                    a_ev[(d.row0 + i) * E + k] = a_v[d.a_nz + e];
                }
            for (int64_t q = 0; q < d.n; q++) {
                const int64_t g = d.row0 + q;
                q_ubase[g] = d.u_nz;
                // lower sweep: L row i (U_ki, k ascending), diagonal U_ii
                const int32_t i = lv_rows[d.rows_l + q];
                lq_row[g] = i;
                lq_dp[g] = u_st[d.u_st + i];
                for (int32_t e = l_st[d.l_st + i], k = 0; e < l_st[d.l_st + i + 1]; e++, k++) {
                    lq_ci[g * W + k] = posl[l_ci[d.l_nz + e]];
                    lq_up[g * W + k] = l_up[d.l_nz + e];
                }
                // upper sweep: U row j after its diagonal, ascending columns
                const int32_t j = lv_rows[d.rows_u + q];
                uq_row[g] = j;
                uq_dp[g] = u_st[d.u_st + j];
                for (int32_t e = u_st[d.u_st + j] + 1, k = 0; e < u_st[d.u_st + j + 1]; e++, k++) {
                    uq_ci[g * W + k] = posu[u_ci[d.u_nz + e]];
                    uq_up[g * W + k] = e;
                }
            }
            // reach of each sweep: how far before the end of its level a
            // dependency may lie (decides whether the LDS ring holds it)
            auto reach = [&](int64_t lvl, int64_t nlev, const std::vector<int32_t>& qci) {
                int64_t rmax = 0;
                for (int64_t l = 0; l < nlev; l++) {
                    const int32_t lb = lv_ptr[lvl + l], le = lv_ptr[lvl + l + 1];
                    for (int32_t q = lb; q < le; q++)
                        for (int k = 0; k < W; k++) {
                            const int32_t c = qci[(d.row0 + q) * W + k];
                            if (c >= 0) rmax = std::max<int64_t>(rmax, le - c);
                        }
                }
                return rmax;
            };
            d.reach_l = reach(d.lvl_l, d.nlev_l, lq_ci);
            d.reach_u = reach(d.lvl_u, d.nlev_u, uq_ci);
            d.level_width = 1;
            for (int64_t l = 0; l < d.nlev_l; l++)
                d.level_width = std::max<int64_t>(d.level_width, lv_ptr[d.lvl_l + l + 1] - lv_ptr[d.lvl_l + l]);
            for (int64_t l = 0; l < d.nlev_u; l++)
                d.level_width = std::max<int64_t>(d.level_width, lv_ptr[d.lvl_u + l + 1] - lv_ptr[d.lvl_u + l]);
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
        B->W = W;
        up32(B->a_eci, a_eci);
        B->a_ev.reset(std::max<size_t>(a_ev.size(), 1));
        B->a_ev.upload(a_ev.data(), a_ev.size());
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
        B->lq_val.reset(std::max<int64_t>(row0 * W, 1));
        B->lq_diag.reset(std::max<int64_t>(row0, 1));
        B->uq_val.reset(std::max<int64_t>(row0 * W, 1));
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
        TraceRange tr(kTraceBatchIc0Factorize);
        hipStream_t s = current_stream();
        const int64_t cnt = B->u_total;
        if (cnt)
            hipLaunchKernelGGL(k_batch_gather, dim3(int(std::min<int64_t>((cnt + 255) / 256, 16384))), dim3(256), 0,
                               s, cnt, B->u_apos.ptr, B->a_v.ptr, B->u_v.ptr);
        hipLaunchKernelGGL(k_batch_ic0, dim3(int(B->nsys)), dim3(256), 0, s, B->sys.ptr, B->l_st.ptr, B->l_ci.ptr,
                           B->l_up.ptr, B->u_st.ptr, B->lv_ptr.ptr, B->lv_rows.ptr, B->u_v.ptr);
        // factor values into the packed sweeps
        const int pg = int(std::min<int64_t>((B->n_total + 255) / 256, 16384));
        hipLaunchKernelGGL(k_batch_pack, dim3(pg), dim3(256), 0, s, B->n_total, B->W, B->lq_up.ptr, B->lq_dp.ptr,
                           B->q_ubase.ptr, B->u_v.ptr, B->lq_val.ptr, B->lq_diag.ptr);
        hipLaunchKernelGGL(k_batch_pack, dim3(pg), dim3(256), 0, s, B->n_total, B->W, B->uq_up.ptr, B->uq_dp.ptr,
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
        TraceRange tr(kTraceBatchCgIc0);
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

// ---------------------------------------------------------------------------
// Autotune cache.  The tuner's answer depends on the batch only through its
// size classes (thirds of the systems by size: count, smallest and largest n
// of each), nrhs, the trial length, the packed sweeps' width and the device,
// so that tuple is the key.  Answers are kept for the process and in a JSON
// Lines file, one object per line:
//   {"key": "...", "block": 1024, "rhs_per_class": [8, 8, 4]}
// at $M4G_AUTOTUNE_CACHE, else $XDG_CACHE_HOME/mtl4gpu/batch_autotune.jsonl,
// else $HOME/.cache/mtl4gpu/batch_autotune.jsonl; M4G_AUTOTUNE_CACHE=off keeps
// the answers in the process only.  A file that cannot be read or written is
// ignored (the tuner then runs as without a cache).
// ---------------------------------------------------------------------------
struct TuneAnswer {
    int block;
    std::vector<int> rpc;   // rhs per item of each size class
};

static std::mutex g_tune_mutex;
static std::map<std::string, TuneAnswer> g_tune_mem;

static std::string tune_cache_path()
{
    const char* e = std::getenv("M4G_AUTOTUNE_CACHE");
    if (e && *e) return std::string(e) == "off" ? std::string() : std::string(e);
    std::string dir;
    if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x) dir = x;
    else if (const char* h = std::getenv("HOME"); h && *h) dir = std::string(h) + "/.cache";
    else return std::string();
    return dir + "/mtl4gpu/batch_autotune.jsonl";
}

static void tune_cache_load(const std::string& path)
{
    FILE* f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!f) return;
    char line[1024];
    while (std::fgets(line, sizeof line, f)) {
        char key[512];
        int block = 0, r[3] = {0, 0, 0};
        const int got = std::sscanf(line, " {\"key\": \"%511[^\"]\", \"block\": %d, \"rhs_per_class\": [%d, %d, %d]}",
                                    key, &block, &r[0], &r[1], &r[2]);
        if (got < 3 || (block != 256 && block != 512 && block != 1024)) continue;   // not an entry of ours
        TuneAnswer a{block, {}};
        for (int c = 0; c < got - 2; c++) {
            if (r[c] < 1 || r[c] > kBatchMaxR) { a.rpc.clear(); break; }
            a.rpc.push_back(r[c]);
        }
        if (!a.rpc.empty()) g_tune_mem[key] = a;   // later lines win
    }
    std::fclose(f);
}

static void tune_cache_append(const std::string& path, const std::string& key, const TuneAnswer& a)
{
    if (path.empty()) return;
    // create the two directory levels under the cache root if needed
    for (size_t pos = path.find('/', 1); pos != std::string::npos; pos = path.find('/', pos + 1))
        (void)::mkdir(path.substr(0, pos).c_str(), 0755);
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return;
    std::string r;
    for (size_t c = 0; c < a.rpc.size(); c++) r += (c ? ", " : "") + std::to_string(a.rpc[c]);
    std::fprintf(f, "{\"key\": \"%s\", \"block\": %d, \"rhs_per_class\": [%s]}\n", key.c_str(), a.block, r.c_str());
    std::fclose(f);
}

static std::string tune_key(const m4g_batch_s& B, const std::vector<std::vector<int>>& cls, int nrhs,
                            int trial_iters)
{
    hipDeviceProp_t prop;
    M4G_HIP(hipGetDeviceProperties(&prop, device_state().device));
    std::string k = std::string(prop.gcnArchName) + "/cu" + std::to_string(prop.multiProcessorCount) + "/nrhs" +
                    std::to_string(nrhs) + "/trial" + std::to_string(trial_iters) + "/w" + std::to_string(B.W);
    for (const auto& c : cls) {
        int64_t lo = INT64_MAX, hi = 0;
        for (int s : c) {
            lo = std::min(lo, B.h_sys[s].n);
            hi = std::max(hi, B.h_sys[s].n);
        }
        k += "/" + std::to_string(c.size()) + ":" + std::to_string(lo) + "-" + std::to_string(hi);
    }
    for (char& ch : k)
        if (ch == '"' || ch == '\\' || ch == ' ') ch = '_';
    return k;
}

int m4g_batch_autotune(m4g_batch B, int nrhs, m4g_vector b, int trial_iters, int* rhs_per_item, int* block,
                       double* trial_ms)
{
    return guard([&] {
        M4G_REQUIRE(B && b && rhs_per_item && block, M4G_ERR_INVALID, "m4g_batch_autotune: null argument");
        M4G_REQUIRE(B->factored, M4G_ERR_INVALID, "m4g_batch_autotune: call m4g_batch_factor_ic0 first");
        M4G_REQUIRE(b->type == M4G_F64 && b->n == B->n_total * nrhs, M4G_ERR_SIZE, "m4g_batch_autotune: b size");
        if (trial_iters <= 0) trial_iters = 40;
        TraceRange tr(kTraceBatchAutotune);
        hipStream_t s = current_stream();
        // size classes: thirds of the systems by size
        std::vector<int> order(B->nsys);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int p, int q) { return B->h_sys[p].n < B->h_sys[q].n; });
        const int nclass = int(std::min<int64_t>(3, B->nsys));
        std::vector<std::vector<int>> cls(nclass);
        for (int64_t k = 0; k < B->nsys; k++) cls[size_t(k * nclass / B->nsys)].push_back(order[k]);
        const std::string key = tune_key(*B, cls, nrhs, trial_iters);
        const std::string path = tune_cache_path();
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            auto hit = g_tune_mem.find(key);
            if (hit == g_tune_mem.end()) {
                tune_cache_load(path);
                hit = g_tune_mem.find(key);
            }
            if (hit != g_tune_mem.end() && int(hit->second.rpc.size()) == nclass) {
                // a cached answer runs no trials and leaves trial_ms untouched
                *block = hit->second.block;
                for (int c = 0; c < nclass; c++)
                    for (int sidx : cls[c]) rhs_per_item[sidx] = hit->second.rpc[c];
                B->last_block = *block;
                return M4G_OK;
            }
        }
        const int blocks[3] = {256, 512, 1024};
        std::vector<int> Rs;
        for (int R = kBatchMaxR; R >= 1; R /= 2)
            if (nrhs % R == 0) Rs.push_back(R);
        DevBuf<double> xs(size_t(B->n_total) * nrhs);
        DevBuf<int> d_it(size_t(B->nsys) * nrhs), d_err(size_t(B->nsys) * nrhs);
        DevBuf<double> d_res(size_t(B->nsys) * nrhs);
        std::vector<int> all(B->nsys);
        std::iota(all.begin(), all.end(), 0);
        // trials run each system for trial_iters * sqrt(n / n_max) iterations:
        // IC(0)-CG on a 2D problem needs iterations in proportion to its side,
        // so a trial keeps the real run's shape (long largest systems, short
        // small ones) at a fraction of its cost
        int64_t nmax = 1;
        for (const BatchSys& d : B->h_sys) nmax = std::max(nmax, d.n);
        std::vector<int> caps(B->nsys);
        for (int64_t k = 0; k < B->nsys; k++)
            caps[k] = std::max(2, int(std::ceil(trial_iters * std::sqrt(double(B->h_sys[k].n) / double(nmax)))));
        DevBuf<int> d_caps(caps.size());
        d_caps.upload(caps.data(), caps.size());
        hipEvent_t e0, e1;
        M4G_HIP(hipEventCreate(&e0));
        M4G_HIP(hipEventCreate(&e1));
        // one trial: trial_iters iterations of the whole batch
        auto trial = [&](const std::vector<int>& rpi, int blk) {
            M4G_HIP(hipMemsetAsync(xs.ptr, 0, xs.count * sizeof(double), s));
            M4G_HIP(hipEventRecord(e0, s));
            batch_run(*B, all, rpi.data(), nrhs, blk, xs.ptr, static_cast<const double*>(b->ptr), trial_iters, 0.0,
                      0.0, d_it.ptr, d_res.ptr, d_err.ptr, s, d_caps.ptr);
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
        {
            TuneAnswer ans{blocks[best_b], {}};
            for (int c = 0; c < nclass; c++) ans.rpc.push_back(cur[cls[c].front()]);
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            g_tune_mem[key] = ans;
            tune_cache_append(path, key, ans);
        }
        if (trial_ms)
            for (size_t k = 0; k < t.size(); k++) trial_ms[k] = t[k];
        B->last_block = *block;
        return M4G_OK;
    });
}

}  // extern "C"
