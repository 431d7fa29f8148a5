// Device-resident CG and BiCGStab (itl/krylov/cg.hpp:39-112,
// itl/krylov/bicgstab.hpp:27-78) with the iteration control of
// itl::basic_iteration (basic_iteration.hpp:24-165) evaluated on the GPU.
//
// Every scalar of the method (rho, alpha, omega, ...) lives in a device state
// block.  Each reduction kernel finishes in its last workgroup, which applies
// the method's scalar update and, where the reference tests convergence, the
// test itself.  Kernels of an iteration that starts after the state says
// "finished" return at once, so the host can queue iterations in batches and
// only polls the finished flag between batches; results and iteration counts
// are those of the sequential method.
#include "m4g_internal.hpp"
#include "m4g_device.hpp"
#include "spmv_dev.hpp"
#include "pc.hpp"
#include "krylov.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace m4g {

constexpr int kNT = 256;
constexpr int kRedBlocks = 1024;
constexpr int kFoldMaxBlocks = 256;   // the CG folds its finalizes at this many reduction workgroups or fewer

// ---------------------------------------------------------------------------
// iteration control on the device
// ---------------------------------------------------------------------------
__device__ inline void record(KState* st, double resid)
{
    st->resid = resid;                                  // basic_iteration.hpp:83
    if (st->history && st->i < st->history_cap) st->history[st->i] = resid;
}

__device__ inline bool converged(const KState* st, double r)
{
    if (st->norm_r0 == 0) return r <= st->atol;        // :90-91
    return r <= st->rtol * st->norm_r0 || r <= st->atol;   // :92
}

// finished(r): converged, or check_max (:43-48).  Returns true when finished.
__device__ inline bool finished(KState* st, double resid)
{
    record(st, resid);
    if (converged(st, resid)) {
        st->done = 1;
        return true;
    }
    if (st->i >= st->max_iter) {
        st->error = M4G_MAX_ITER;
        st->done = 1;
        return true;
    }
    return false;
}

// while (!finished(r)) { ++iter; ... }
__device__ inline bool top_check(KState* st, double resid)
{
    if (finished(st, resid)) return true;
    st->i += 1;
    return false;
}

// ---------------------------------------------------------------------------
// reductions: K sums per thread -> per-block partials -> last block
// ---------------------------------------------------------------------------
template <int K>
struct Sums {
    double v[K];
};

// Each reduction kernel stores one partial per workgroup with plain stores;
// a one-workgroup finalize kernel (ordered after it on the stream, so the
// partials are visible without device-scope fences) adds them in workgroup
// order and applies the method's scalar update.
template <int NT, int K>
__device__ inline void store_partials(Sums<K>& s, double* partials)
{
    __shared__ double scratch[NT / kWave];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double v = block_sum<NT>(s.v[k], scratch);
        if (threadIdx.x == 0) partials[k * gridDim.x + blockIdx.x] = v;
    }
}

enum FinOp {
    FIN_CG_START,      // rr = r.r; finished(sqrt|rr|)             (cg.hpp:54-55, 94-95)
    FIN_CG_RHO,        // rho = r.z                                 (cg.hpp:97)
    FIN_CG_ALPHA,      // alpha_1 = p.q; alpha = rho / alpha_1      (cg.hpp:104-105)
    FIN_CG_UPDATE,     // rho_1 = rho; rr = r.r; finished(sqrt|rr|) (cg.hpp:108-109, 95)
    FIN_BICG_START,    // finished(|r|); rho_1 = rtilde.r; tests    (bicgstab.hpp:43-53)
    FIN_BICG_GAMMA,    // gamma = rtilde.v; alpha = rho_1 / gamma   (bicgstab.hpp:58-61)
    FIN_BICG_S,        // finished(|s|) -> early exit               (bicgstab.hpp:64)
    FIN_BICG_OMEGA,    // omega = t.s / t.t                         (bicgstab.hpp:70)
    FIN_BICG_UPDATE    // rho_2 = rho_1; then the loop top          (bicgstab.hpp:75, 43-53)
};

__device__ inline void bicg_top(KState* st, double rr, double rtr);

template <int K, int OP>
__global__ __launch_bounds__(256) void k_fin(const double* __restrict__ partials, int nblocks, KState* st,
                                             int identity_pc)
{
    if (OP != FIN_CG_START && OP != FIN_BICG_START && st->done) return;   // the producer skipped too
    __shared__ double scratch[256 / kWave];
    double tot[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        double acc = 0;
        for (int b = threadIdx.x; b < nblocks; b += 256) acc += partials[k * nblocks + b];
        tot[k] = block_sum<256>(acc, scratch);
    }
    if (threadIdx.x != 0) return;
    switch (OP) {
    case FIN_CG_START:
        st->rr = tot[0];
        st->rho = tot[0];
        top_check(st, sqrt(fabs(tot[0])));
        break;
    case FIN_CG_RHO: st->rho = tot[0]; break;
    case FIN_CG_ALPHA:
        st->alpha_1 = tot[0];
        st->alpha = st->rho / st->alpha_1;
        break;
    case FIN_CG_UPDATE:
        st->rho_1 = st->rho;
        st->rr = tot[0];
        if (identity_pc) st->rho = tot[0];
        top_check(st, sqrt(fabs(tot[0])));
        break;
    case FIN_BICG_START: bicg_top(st, tot[0], tot[K - 1]); break;
    case FIN_BICG_GAMMA:
        st->gamma = tot[0];
        if (st->gamma == 0.0) {
            st->error = M4G_ERR_ORTHOGONALITY;
            st->done = 1;
            break;
        }
        st->alpha = st->rho_1 / st->gamma;
        break;
    case FIN_BICG_S:
        if (finished(st, sqrt(tot[0]))) st->mid_exit = 1;
        break;
    case FIN_BICG_OMEGA: st->omega = tot[0] / tot[K - 1]; break;
    case FIN_BICG_UPDATE:
        st->rho_2 = st->rho_1;
        bicg_top(st, tot[0], tot[K - 1]);
        break;
    }
}

#define GRID_STRIDE(i, n) for (int64_t i = blockIdx.x * int64_t(kNT) + threadIdx.x; i < (n); i += int64_t(gridDim.x) * kNT)

// ---------------------------------------------------------------------------
// CG kernels
// ---------------------------------------------------------------------------
// rr = r.r; initial finished() test.  For the identity PC rho = rr (cg.hpp:54).
template <typename T>
__global__ __launch_bounds__(kNT) void k_cg_start(int64_t n, const T* __restrict__ r, KState* st, double* partials,
                                                   unsigned* counter)
{
    Sums<1> s{{0}};
    GRID_STRIDE(i, n) s.v[0] += double(r[i]) * double(r[i]);
    store_partials<kNT, 1>(s, partials);
}

template <typename T>
__device__ inline double prec_mul(T a, T b)
{
    return double(a) * double(b);
}

// rho = dot(r, z) (cg.hpp:97).  T=float accumulates products in float per
// thread (FP32 accumulation of the inner solve), partials in double.
template <typename T>
__global__ __launch_bounds__(kNT) void k_cg_rho(int64_t n, const T* __restrict__ r, const T* __restrict__ z,
                                                 KState* st, double* partials, unsigned* counter)
{
    if (st->done) return;
    Sums<1> s{{0}};
    T acc = 0;
    GRID_STRIDE(i, n) acc += r[i] * z[i];
    s.v[0] = double(acc);
    store_partials<kNT, 1>(s, partials);
}

// diagonal PC fused with the dot: z = inv_diag .* r; rho = r.z
template <typename T>
__global__ __launch_bounds__(kNT) void k_cg_diag_rho(int64_t n, const double* __restrict__ inv, const T* __restrict__ r,
                                                      T* __restrict__ z, KState* st, double* partials, unsigned* counter)
{
    if (st->done) return;
    Sums<1> s{{0}};
    T acc = 0;
    GRID_STRIDE(i, n)
    {
        T zi = T(inv[i]) * r[i];
        z[i] = zi;
        acc += r[i] * zi;
    }
    s.v[0] = double(acc);
    store_partials<kNT, 1>(s, partials);
}

// p = z (first) or p = z + (rho/rho_1) * p (cg.hpp:99-102; no-PC: :57-60)
template <typename T>
__global__ __launch_bounds__(kNT) void k_cg_p(int64_t n, const T* __restrict__ z, T* __restrict__ p, const KState* st)
{
    if (st->done) return;
    if (st->i <= 1) {                                    // iter.first()
        GRID_STRIDE(i, n) p[i] = z[i];
    } else {
        const T s = T(st->rho / st->rho_1);
        GRID_STRIDE(i, n) p[i] = z[i] + s * p[i];
    }
}

// (lazy(q) = A*p) || (lazy(alpha_1) = lazy_dot(p, q)); alpha = rho / alpha_1 (cg.hpp:104-105)
template <typename T>
__global__ __launch_bounds__(kSpmvNT) void k_cg_spmv_dot(int64_t n, int64_t nslices, int64_t chunk,
                                                          const int64_t* __restrict__ slice_ptr,
                                                          const int64_t* __restrict__ slice_idx,
                                                          const uint8_t* __restrict__ idx, const T* __restrict__ vals,
                                                          const T* __restrict__ p, T* __restrict__ q, KState* st,
                                                          double* partials, unsigned* counter)
{
    if (st->done) return;
    const int64_t g = xcd_logical_block(chunk);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t s = g * kSliceWaves + wave;
    Sums<1> acc{{0}};
    if (s < nslices) {
        const T qv = sell_row_sum(slice_ptr, slice_idx, idx, vals, p, s, lane);
        const int64_t row = s * kWave + lane;
        if (row < n) {
            q[row] = qv;
            acc.v[0] = double(p[row] * qv);
        }
    }
    store_partials<kSpmvNT, 1>(acc, partials);
}

// The same with the pipelined slice loop (all-narrow matrices): per-thread
// partials summed over the wavefront's slices.
template <typename T, int MAXW>
__global__ __launch_bounds__(kSpmvNT) void k_cg_spmv_pipe(int64_t n, int64_t nslices, int64_t padded,
                                                           const int64_t* __restrict__ slice_ptr,
                                                           const uint8_t* __restrict__ idx, const T* __restrict__ vals,
                                                           const T* __restrict__ p, T* __restrict__ q, KState* st,
                                                           double* partials)
{
    if (st->done) return;
    Sums<1> acc{{0}};
    sell_pipe_run<T, MAXW>(
        n, nslices, padded, slice_ptr, idx, vals, p, [&](int64_t rr) { return p[rr]; },
        [&](int64_t row, T qv, T pr) {
            q[row] = qv;
            acc.v[0] += double(pr * qv);
        });
    store_partials<kSpmvNT, 1>(acc, partials);
}

// x += alpha*p; rho_1 = rho; (lazy(r) -= alpha*q) || (lazy(rr) = lazy_unary_dot(r)); then the
// loop test finished(sqrt|rr|) (cg.hpp:107-109, 95-96)
template <typename T>
__global__ __launch_bounds__(kNT) void k_cg_update(int64_t n, const T* __restrict__ p, const T* __restrict__ q,
                                                    T* __restrict__ x, T* __restrict__ r, KState* st, int identity_pc,
                                                    double* partials, unsigned* counter)
{
    if (st->done) return;
    const T alpha = T(st->alpha);
    T acc = 0;
    GRID_STRIDE(i, n)
    {
        x[i] = x[i] + alpha * p[i];
        T ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc += ri * ri;
    }
    Sums<1> s{{double(acc)}};
    store_partials<kNT, 1>(s, partials);
}

// ---------------------------------------------------------------------------
// CG with the rho and alpha finalizes folded into their consumers (small n)
// ---------------------------------------------------------------------------
// The total that k_fin<1, OP> forms from `nblocks` partials -- the same loop
// and the same tree over 256 threads, hence the same value -- in every thread
// of a 256-thread workgroup.  Each consumer workgroup forms it itself, which
// saves a launch when the partials are few.
__device__ inline double fin_total(const double* __restrict__ partials, int nblocks, double* scratch, double* bcast)
{
    double acc = 0;
    for (int b = threadIdx.x; b < nblocks; b += 256) acc += partials[b];
    const double t = block_sum<256>(acc, scratch);
    if (threadIdx.x == 0) *bcast = t;
    __syncthreads();
    return *bcast;
}

// rho = dot(r, z) from k_cg_rho's partials (FIN_CG_RHO), then
// p = z (first) or z + (rho / rho_1) p   (cg.hpp:97-102)
template <typename T>
__global__ __launch_bounds__(kNT) void k_cg_p_fold(int64_t n, const T* __restrict__ z, T* __restrict__ p, KState* st,
                                                    const double* __restrict__ rho_partials, int nblocks)
{
    if (st->done) return;
    __shared__ double scratch[kNT / kWave];
    __shared__ double bcast;
    const double rho = fin_total(rho_partials, nblocks, scratch, &bcast);
    if (blockIdx.x == 0 && threadIdx.x == 0) st->rho = rho;
    if (st->i <= 1) {                                    // iter.first()
        GRID_STRIDE(i, n) p[i] = z[i];
    } else {
        const T sc = T(rho / st->rho_1);
        GRID_STRIDE(i, n) p[i] = z[i] + sc * p[i];
    }
}

// alpha_1 = dot(p, q) from the SpMV's partials and alpha = rho / alpha_1
// (FIN_CG_ALPHA), then k_cg_update's x, r and r.r partials (into out_partials)
template <typename T>
__global__ __launch_bounds__(kNT) void k_cg_update_fold(int64_t n, const T* __restrict__ p, const T* __restrict__ q,
                                                         T* __restrict__ x, T* __restrict__ r, KState* st,
                                                         const double* __restrict__ pq_partials, int nblocks,
                                                         double* out_partials)
{
    if (st->done) return;
    __shared__ double scratch[kNT / kWave];
    __shared__ double bcast;
    const double alpha_1 = fin_total(pq_partials, nblocks, scratch, &bcast);
    const double alpha_d = st->rho / alpha_1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->alpha_1 = alpha_1;
        st->alpha = alpha_d;
    }
    const T alpha = T(alpha_d);
    T acc = 0;
    GRID_STRIDE(i, n)
    {
        x[i] = x[i] + alpha * p[i];
        T ri = r[i] - alpha * q[i];
        r[i] = ri;
        acc += ri * ri;
    }
    Sums<1> sm{{double(acc)}};
    store_partials<kNT, 1>(sm, out_partials);
}

// ---------------------------------------------------------------------------
// IC(0)-PCG of a small 2D grid system in one workgroup
// ---------------------------------------------------------------------------
// For a system of at most kGridMaxN rows on an nx x ny grid (ny <= 64) whose
// IC(0) factor has the 5-point pattern, a whole batch of CG iterations runs
// in one workgroup with no launch between the steps.  The factor (U's east
// and north entries and the inverted diagonal) and the sweep vector stay in
// LDS; x, r and p stay in registers (rows tid, tid + 512, ...).  One
// wavefront runs both sweeps with lane j on grid line j: at step s it
// finishes cell i = s - j, whose two dependencies are its own previous result
// (the cell before it on the line) and lane j-1's previous one, passed by a
// DPP lane shift -- a step is the dependent arithmetic of the row and no
// memory round trip.  Each row applies its entries in the order of the
// multi-kernel sweeps (L = trans(U): north then west, times the inverted
// diagonal; U: east then north), absent boundary entries as 0 * (+0) with
// both factors +0, which leaves every value unchanged.  Scalars, convergence
// test and history follow the multi-kernel CG (cg.hpp:75-112); only the
// order in which the dots are summed differs.
constexpr int kGridNT = 512;
constexpr int kGridRows = 8;   // rows per thread
constexpr int64_t kGridMaxN = int64_t(kGridNT) * kGridRows;

struct GridCgArgs {
    int32_t n, nx, ny, iters;
    const int32_t* S;      // CSR of A, whose rows have the 5-point pattern
    const double* A;
    const double* east;    // [n] U[r][r+1], 0 where absent
    const double* north;   // [n] U[r][r+nx], 0 where absent
    const double* dinv;    // [n] inverted diagonal of U
    double *x, *r, *p;
    KState* st;
};

// LDS doubles of the kernel: 8 of slack, sweep vector, east (one leading 0),
// north (nx leading 0s), inverted diagonal, one dump slot per lane, 8 zeros.
constexpr int kGridSlack = 8;
inline int64_t grid_cg_lds_doubles(int64_t n, int64_t nx) { return kGridSlack + 4 * n + nx + 1 + kWave + 8; }

// Lane l receives lane l-1's value (lane 0: +0); lane l receives lane l+1's
// (lane 63: +0).  DPP wave_shr:1 / wave_shl:1 on the two halves, lanes
// without a source reading 0 (bound_ctrl).
__device__ inline double lane_from_prev(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp(int(b), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), 0x138, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ inline double lane_from_next(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp(int(b), 0x130, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp(int(b >> 32), 0x130, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Both sweeps in place over V, lane j on grid line j (j < ny), step s on cell
// i = s - j.  Lower sweep on L = trans(U) (lower_trisolve.hpp:191-209):
//   t_r = ((v_r - U[r-nx][r] t_{r-nx}) - U[r-1][r] t_{r-1}) * dinv_r,
// the first operand from lane j-1, the second the lane's own last result.
// Upper sweep on U, steps backwards (upper_trisolve.hpp:159-176):
//   z_r = ((t_r - U[r][r+1] z_{r+1}) - U[r][r+nx] z_{r+nx}) * dinv_r.
// Eight steps a turn.  Lane j reads step s's operands at element
// nx*j + s - j of each array, so a turn's are eight consecutive elements
// from a per-lane pointer that moves by eight a turn; the steps run on
// registers (the dependent chain is the row's arithmetic and one lane
// shift), with the next turn's loads issued before them.  Steps outside a
// lane's line read neighbouring lines or the slack around the arrays (at
// most 7 elements before V and 6 past D): a step before the line starts
// gives +0 (the value of a missing neighbour), a step after it ends gives a
// value no lane reads, and neither is stored (the dump slot takes it).
// Lanes past the last line read eight zeros Z and give +0.  The step count
// is rounded up to a multiple of eight.
constexpr int kSweepTurn = 8;

template <bool LOWER>
__device__ inline void grid_sweep(double* V, const double* E, const double* N, const double* D, double* dump,
                                  const double* Z, int nx, int ny, int j)
{
    const int steps = nx + ny - 1;
    const int turns = (steps + kSweepTurn - 1) / kSweepTurn;
    const bool line = j < ny;
    const int S = steps - 1;
    // element of step s: nx*j + s - j; lower s = 8k + u, upper s = S - 8k - u
    const int q0 = LOWER ? nx * j - j : nx * j - j + S - (kSweepTurn - 1);
    const int stride = line ? (LOWER ? kSweepTurn : -kSweepTurn) : 0;
    const double* pv = line ? V + q0 : Z;
    const double* pf1 = line ? (LOWER ? N + q0 : E + 1 + q0) : Z;   // U[r-nx][r] / U[r][r+1]
    const double* pf2 = line ? (LOWER ? E + q0 : N + nx + q0) : Z;  // U[r-1][r]  / U[r][r+nx]
    const double* pd = line ? D + q0 : Z;
    double y = 0.0;   // the lane's last result
    auto load = [&](int k, double (&v)[kSweepTurn], double (&f1)[kSweepTurn], double (&f2)[kSweepTurn],
                    double (&d)[kSweepTurn]) {
        const int o = k * stride;
#pragma unroll
        for (int u = 0; u < kSweepTurn; u++) {
            const int e = o + (LOWER ? u : kSweepTurn - 1 - u);
            v[u] = pv[e];
            f1[u] = pf1[e];
            f2[u] = pf2[e];
            d[u] = pd[e];
        }
    };
    auto run = [&](int k, const double (&v)[kSweepTurn], const double (&f1)[kSweepTurn],
                   const double (&f2)[kSweepTurn], const double (&d)[kSweepTurn]) {
        const int o = k * stride;
        const int i0 = LOWER ? kSweepTurn * k - j : S - kSweepTurn * k - j;   // cell of step 0
#pragma unroll
        for (int u = 0; u < kSweepTurn; u++) {
            const int i = LOWER ? i0 + u : i0 - u;
            const double nb = LOWER ? lane_from_prev(y) : lane_from_next(y);
            double acc = LOWER ? v[u] - f1[u] * nb : v[u] - f1[u] * y;
            acc = LOWER ? acc - f2[u] * y : acc - f2[u] * nb;
            const double yn = acc * d[u];
            y = (LOWER ? i < 0 : i >= nx) ? 0.0 : yn;
            double* at = line && i >= 0 && i < nx ? V + (q0 + o + (LOWER ? u : kSweepTurn - 1 - u)) : dump + j;
            *at = y;
        }
    };
    // the next turn's operands are loaded before a turn's steps (two register
    // sets, no copies)
    double cv[kSweepTurn], c1[kSweepTurn], c2[kSweepTurn], cd[kSweepTurn];
    double nv[kSweepTurn], n1[kSweepTurn], n2[kSweepTurn], nd[kSweepTurn];
    load(0, cv, c1, c2, cd);
    for (int k = 0; k < turns; k += 2) {
        if (k + 1 < turns) load(k + 1, nv, n1, n2, nd);
        run(k, cv, c1, c2, cd);
        if (k + 1 >= turns) break;
        if (k + 2 < turns) load(k + 2, cv, c1, c2, cd);
        run(k + 1, nv, n1, n2, nd);
    }
}

// Sum over the workgroup in a fixed order; every thread gets the total.
__device__ inline double grid_cg_sum(double v, double* red)
{
    v = wave_sum(v);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) red[wave] = v;
    lds_barrier();
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kGridNT / kWave; w++) t += red[w];
    lds_barrier();   // red is reused by the next sum
    return t;
}

__global__ __launch_bounds__(kGridNT) void k_cg_ic0_grid(GridCgArgs a)
{
    extern __shared__ double lds[];
    __shared__ double red[kGridNT / kWave];
    __shared__ KState ks;
    const int n = a.n, nx = a.nx, tid = threadIdx.x;
    double* V = lds + kGridSlack;
    double* E = V + n;        // E[1 + r] = U[r][r+1]
    double* N = E + n + 1;    // N[nx + r] = U[r][r+nx]
    double* D = N + n + nx;
    double* dump = D + n;
    double* Z = dump + kWave;   // eight zeros
    for (int k = tid; k < n; k += kGridNT) {
        E[1 + k] = a.east[k];
        N[nx + k] = a.north[k];
        D[k] = a.dinv[k];
    }
    if (tid == 0) E[0] = 0.0;
    for (int k = tid; k < nx; k += kGridNT) N[k] = 0.0;
    if (tid < kGridSlack) lds[tid] = 0.0;
    if (tid < 8) Z[tid] = 0.0;
    if (tid == 0) ks = *a.st;
    double x[kGridRows], r[kGridRows], p[kGridRows], q[kGridRows];
    int gi[kGridRows], gj[kGridRows], ge[kGridRows];   // grid position and first entry of each row
#pragma unroll
    for (int u = 0; u < kGridRows; u++) {
        const int row = tid + u * kGridNT;
        const bool in = row < n;
        x[u] = in ? a.x[row] : 0.0;
        r[u] = in ? a.r[row] : 0.0;
        p[u] = in ? a.p[row] : 0.0;
        gi[u] = in ? row % nx : 0;
        gj[u] = in ? row / nx : 0;
        ge[u] = in ? a.S[row] : 0;
    }
    lds_barrier();
    for (int it = 0; it < a.iters && !ks.done; it++) {
        // z = solve(P, r)   (cg.hpp:97)
#pragma unroll
        for (int u = 0; u < kGridRows; u++)
            if (tid + u * kGridNT < n) V[tid + u * kGridNT] = r[u];
        lds_barrier();
        if (tid < kWave) {
            grid_sweep<true>(V, E, N, D, dump, Z, nx, a.ny, tid);
            grid_sweep<false>(V, E, N, D, dump, Z, nx, a.ny, tid);
        }
        lds_barrier();
        // A's entries for q = A p, loaded now so they arrive while rho and p
        // are formed: south (r-nx), west, centre, east, north (r+nx), 0 where
        // the grid has no such neighbour
        double cs[kGridRows], cw[kGridRows], cc[kGridRows], ce[kGridRows], cn[kGridRows];
#pragma unroll
        for (int u = 0; u < kGridRows; u++) {
            const int row = tid + u * kGridNT;
            const bool in = row < n, hs = in && gj[u] > 0, hw = in && gi[u] > 0, he = in && gi[u] < nx - 1,
                       hn = in && gj[u] < a.ny - 1;
            const int es = ge[u], ew = es + hs, ec = ew + hw, ee = ec + 1, en = ee + he;
            cs[u] = hs ? a.A[es] : 0.0;
            cw[u] = hw ? a.A[ew] : 0.0;
            cc[u] = in ? a.A[ec] : 0.0;
            ce[u] = he ? a.A[ee] : 0.0;
            cn[u] = hn ? a.A[en] : 0.0;
        }
        // rho = dot(r, z)
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < kGridRows; u++)
            if (tid + u * kGridNT < n) acc += r[u] * V[tid + u * kGridNT];
        const double rho = grid_cg_sum(acc, red);
        // p = z (first) or z + (rho / rho_1) p   (cg.hpp:99-102)
        const bool first = ks.i <= 1;
        const double beta = first ? 0.0 : rho / ks.rho_1;
#pragma unroll
        for (int u = 0; u < kGridRows; u++) {
            const int row = tid + u * kGridNT;
            if (row < n) {
                const double z = V[row];
                p[u] = first ? z : z + beta * p[u];
                V[row] = p[u];
            }
        }
        lds_barrier();
        // q = A p; alpha = rho / dot(p, q)   (cg.hpp:104-105)
        acc = 0.0;
#pragma unroll
        for (int u = 0; u < kGridRows; u++) {
            const int row = tid + u * kGridNT;
            q[u] = 0.0;
            if (row < n) {   // the row's entries in stored order (mat_vec_mult.hpp:662-687)
                double sum = 0.0;
                if (gj[u] > 0) sum += cs[u] * V[row - nx];
                if (gi[u] > 0) sum += cw[u] * V[row - 1];
                sum += cc[u] * V[row];
                if (gi[u] < nx - 1) sum += ce[u] * V[row + 1];
                if (gj[u] < a.ny - 1) sum += cn[u] * V[row + nx];
                q[u] = sum;
                acc += p[u] * sum;
            }
        }
        const double alpha_1 = grid_cg_sum(acc, red);
        const double alpha = rho / alpha_1;
        // x += alpha p; r -= alpha q; rr = dot(r, r)   (cg.hpp:107-109)
        acc = 0.0;
#pragma unroll
        for (int u = 0; u < kGridRows; u++)
            if (tid + u * kGridNT < n) {
                x[u] = x[u] + alpha * p[u];
                const double ri = r[u] - alpha * q[u];
                r[u] = ri;
                acc += ri * ri;
            }
        const double rr = grid_cg_sum(acc, red);
        if (tid == 0) {   // on a private copy: the shared one is written with LDS stores only
            KState t = ks;
            t.rho = rho;
            t.alpha_1 = alpha_1;
            t.alpha = alpha;
            t.rho_1 = rho;
            t.rr = rr;
            top_check(&t, sqrt(fabs(rr)));   // finished(r) or ++iter   (cg.hpp:95)
            ks = t;
        }
        lds_barrier();
    }
#pragma unroll
    for (int u = 0; u < kGridRows; u++) {
        const int row = tid + u * kGridNT;
        if (row < n) {
            a.x[row] = x[u];
            a.r[row] = r[u];
            a.p[row] = p[u];
        }
    }
    if (tid == 0) *a.st = ks;
}

// ---------------------------------------------------------------------------
// BiCGStab kernels (bicgstab.hpp:40-76)
// ---------------------------------------------------------------------------
// Loop-top: finished(two_norm(r)); ++iter; rho_1 = dot(rtilde, r) with its
// zero test; the omega test and beta for the p update (:43-53).
__device__ inline void bicg_top(KState* st, double rr, double rtr)
{
    if (top_check(st, sqrt(rr))) return;
    st->rho_1 = rtr;
    if (st->rho_1 == 0.0) {
        st->error = M4G_ERR_ORTHOGONALITY;
        st->done = 1;
        return;
    }
    if (st->i > 1) {
        if (st->omega == 0.0) {
            st->error = M4G_ERR_ORTHOGONALITY;
            st->done = 1;
            return;
        }
        st->beta = (st->rho_1 / st->rho_2) * (st->alpha / st->omega);
    }
}

template <typename T>
__global__ __launch_bounds__(kNT) void k_bicg_start(int64_t n, const T* __restrict__ r, T* __restrict__ rtilde,
                                                     KState* st, double* partials, unsigned* counter)
{
    Sums<2> s{{0, 0}};
    T a0 = 0, a1 = 0;
    GRID_STRIDE(i, n)
    {
        T ri = r[i];
        rtilde[i] = ri;                                   // rtilde = r (:41)
        a0 += ri * ri;
        a1 += ri * ri;
    }
    s.v[0] = double(a0);
    s.v[1] = double(a1);
    store_partials<kNT, 2>(s, partials);
}

// p = r (first) or p = r + beta*(p - omega*v) (:48-54)
template <typename T>
__global__ __launch_bounds__(kNT) void k_bicg_p(int64_t n, const T* __restrict__ r, const T* __restrict__ v,
                                                 T* __restrict__ p, const KState* st)
{
    if (st->done) return;
    if (st->i <= 1) {
        GRID_STRIDE(i, n) p[i] = r[i];
    } else {
        const T beta = T(st->beta), omega = T(st->omega);
        GRID_STRIDE(i, n) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
}

// v = A*phat; gamma = dot(rtilde, v) with its zero test; alpha = rho_1/gamma (:56-61)
template <typename T>
__global__ __launch_bounds__(kSpmvNT) void k_bicg_spmv_gamma(int64_t n, int64_t nslices, int64_t chunk,
                                                              const int64_t* __restrict__ slice_ptr,
                                                              const int64_t* __restrict__ slice_idx,
                                                              const uint8_t* __restrict__ idx,
                                                              const T* __restrict__ vals, const T* __restrict__ phat,
                                                              const T* __restrict__ rtilde, T* __restrict__ v,
                                                              KState* st, double* partials, unsigned* counter)
{
    if (st->done) return;
    const int64_t g = xcd_logical_block(chunk);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t s = g * kSliceWaves + wave;
    Sums<1> acc{{0}};
    if (s < nslices) {
        const T vv = sell_row_sum(slice_ptr, slice_idx, idx, vals, phat, s, lane);
        const int64_t row = s * kWave + lane;
        if (row < n) {
            v[row] = vv;
            acc.v[0] = double(rtilde[row] * vv);
        }
    }
    store_partials<kSpmvNT, 1>(acc, partials);
}

template <typename T, int MAXW>
__global__ __launch_bounds__(kSpmvNT) void k_bicg_spmv_gamma_pipe(int64_t n, int64_t nslices, int64_t padded,
                                                                   const int64_t* __restrict__ slice_ptr,
                                                                   const uint8_t* __restrict__ idx,
                                                                   const T* __restrict__ vals,
                                                                   const T* __restrict__ phat,
                                                                   const T* __restrict__ rtilde, T* __restrict__ v,
                                                                   KState* st, double* partials)
{
    if (st->done) return;
    Sums<1> acc{{0}};
    sell_pipe_run<T, MAXW>(
        n, nslices, padded, slice_ptr, idx, vals, phat, [&](int64_t rr) { return rtilde[rr]; },
        [&](int64_t row, T vv, T rt) {
            v[row] = vv;
            acc.v[0] += double(rt * vv);
        });
    store_partials<kSpmvNT, 1>(acc, partials);
}

// s = r - alpha*v; finished(two_norm(s)) -> early exit flag (:62-67)
template <typename T>
__global__ __launch_bounds__(kNT) void k_bicg_s(int64_t n, const T* __restrict__ r, const T* __restrict__ v,
                                                 T* __restrict__ s_, KState* st, double* partials, unsigned* counter)
{
    if (st->done) return;
    const T alpha = T(st->alpha);
    T acc = 0;
    GRID_STRIDE(i, n)
    {
        T si = r[i] - alpha * v[i];
        s_[i] = si;
        acc += si * si;
    }
    Sums<1> s{{double(acc)}};
    store_partials<kNT, 1>(s, partials);
}

// early exit: x += alpha*phat, once (:65)
template <typename T>
__global__ __launch_bounds__(kNT) void k_bicg_exit(int64_t n, const T* __restrict__ phat, T* __restrict__ x, KState* st,
                                                    unsigned* counter)
{
    if (st->mid_exit != 1) return;
    const T alpha = T(st->alpha);
    GRID_STRIDE(i, n) x[i] = x[i] + alpha * phat[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(counter, 1u) == gridDim.x - 1) {
            st->mid_exit = 2;
            *counter = 0;
        }
    }
}

// t = A*shat; omega = dot(t, s) / dot(t, t) (:69-70)
template <typename T>
__global__ __launch_bounds__(kSpmvNT) void k_bicg_spmv_omega(int64_t n, int64_t nslices, int64_t chunk,
                                                              const int64_t* __restrict__ slice_ptr,
                                                              const int64_t* __restrict__ slice_idx,
                                                              const uint8_t* __restrict__ idx,
                                                              const T* __restrict__ vals, const T* __restrict__ shat,
                                                              const T* __restrict__ s_, T* __restrict__ t_, KState* st,
                                                              double* partials, unsigned* counter)
{
    if (st->done) return;
    const int64_t g = xcd_logical_block(chunk);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t s = g * kSliceWaves + wave;
    Sums<2> acc{{0, 0}};
    if (s < nslices) {
        const T tv = sell_row_sum(slice_ptr, slice_idx, idx, vals, shat, s, lane);
        const int64_t row = s * kWave + lane;
        if (row < n) {
            t_[row] = tv;
            acc.v[0] = double(tv * s_[row]);
            acc.v[1] = double(tv * tv);
        }
    }
    store_partials<kSpmvNT, 2>(acc, partials);
}

template <typename T, int MAXW>
__global__ __launch_bounds__(kSpmvNT) void k_bicg_spmv_omega_pipe(int64_t n, int64_t nslices, int64_t padded,
                                                                   const int64_t* __restrict__ slice_ptr,
                                                                   const uint8_t* __restrict__ idx,
                                                                   const T* __restrict__ vals,
                                                                   const T* __restrict__ shat,
                                                                   const T* __restrict__ s_, T* __restrict__ t_,
                                                                   KState* st, double* partials)
{
    if (st->done) return;
    Sums<2> acc{{0, 0}};
    sell_pipe_run<T, MAXW>(
        n, nslices, padded, slice_ptr, idx, vals, shat, [&](int64_t rr) { return s_[rr]; },
        [&](int64_t row, T tv, T sv) {
            t_[row] = tv;
            acc.v[0] += double(tv * sv);
            acc.v[1] += double(tv * tv);
        });
    store_partials<kSpmvNT, 2>(acc, partials);
}

// x += omega*shat + alpha*phat; r = s - omega*t; rho_2 = rho_1; loop top (:72-75, 43-53)
template <typename T>
__global__ __launch_bounds__(kNT) void k_bicg_update(int64_t n, const T* __restrict__ shat, const T* __restrict__ phat,
                                                      const T* __restrict__ s_, const T* __restrict__ t_,
                                                      const T* __restrict__ rtilde, T* __restrict__ x,
                                                      T* __restrict__ r, KState* st, double* partials, unsigned* counter)
{
    if (st->done) return;
    const T alpha = T(st->alpha), omega = T(st->omega);
    T a0 = 0, a1 = 0;
    GRID_STRIDE(i, n)
    {
        x[i] = x[i] + (omega * shat[i] + alpha * phat[i]);
        T ri = s_[i] - omega * t_[i];
        r[i] = ri;
        a0 += ri * ri;
        a1 += rtilde[i] * ri;
    }
    Sums<2> s{{double(a0), double(a1)}};
    store_partials<kNT, 2>(s, partials);
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
static int red_grid(int64_t n)
{
    return int(std::max<int64_t>(1, std::min<int64_t>((n + kNT * 4 - 1) / (kNT * 4), kRedBlocks)));
}

void KrylovWs::prepare(int64_t spmv_grid, int history_cap)
{
    const int64_t need = 2 * std::max<int64_t>(spmv_grid, kRedBlocks);
    if (int64_t(partials.count) < need) partials.reset(need);
    if (!counter.ptr) {
        counter.reset(4);
        M4G_HIP(hipMemsetAsync(counter.ptr, 0, 4 * sizeof(unsigned), current_stream()));
    }
    if (!state.ptr) state.reset(1);
    if (history_cap > 0 && int64_t(history.count) < history_cap) history.reset(history_cap);
    if (!host_state) M4G_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_state), 2 * sizeof(KState), hipHostMallocDefault));
}

KrylovWs::~KrylovWs()
{
    if (host_state) (void)hipHostFree(host_state);
}

// launches KERNEL<T, MAXW> for the pipelined SpMV's width mw
#define M4G_PIPE_LAUNCH(KERNEL, T, mw, ...)                                                                      \
    do {                                                                                                       \
        switch (mw) {                                                                                          \
        case 3: hipLaunchKernelGGL((KERNEL<T, 3>), __VA_ARGS__); break;                                       \
        case 4: hipLaunchKernelGGL((KERNEL<T, 4>), __VA_ARGS__); break;                                       \
        case 5: hipLaunchKernelGGL((KERNEL<T, 5>), __VA_ARGS__); break;                                       \
        case 6: hipLaunchKernelGGL((KERNEL<T, 6>), __VA_ARGS__); break;                                       \
        case 7: hipLaunchKernelGGL((KERNEL<T, 7>), __VA_ARGS__); break;                                       \
        case 8: hipLaunchKernelGGL((KERNEL<T, 8>), __VA_ARGS__); break;                                       \
        default: hipLaunchKernelGGL((KERNEL<T, 16>), __VA_ARGS__); break;                                     \
        }                                                                                                      \
    } while (0)

static void spmv_geom(const m4g_matrix_s& A, int64_t& chunk, int64_t& grid)
{
    int64_t groups = (A.sell.nslices + kSliceWaves - 1) / kSliceWaves;
    chunk = (groups + kXcds - 1) / kXcds;
    grid = std::max<int64_t>(1, chunk * kXcds);
}

// True when every row r = i + nx*j of A holds exactly its 5-point neighbours on
// the (2D) grid g, columns ascending: r-nx, r-1, r, r+1, r+nx where inside.
static bool grid5_pattern(const m4g_matrix_s& A, const GridInfo& g)
{
    const int64_t n = A.nrows, nx = g.nx, ny = g.ny;
    if (nx <= 0 || g.nz != 1 || nx * ny != n || n > kGridMaxN) return false;
    for (int64_t r = 0; r < n; r++) {
        const int64_t i = r % nx, j = r / nx;
        int64_t e = A.h_starts[r];
        const int64_t want[5] = {j > 0 ? r - nx : -1, i > 0 ? r - 1 : -1, r, i < nx - 1 ? r + 1 : -1,
                                 j < ny - 1 ? r + nx : -1};
        for (int64_t c : want)
            if (c >= 0 && (e >= A.h_starts[r + 1] || A.h_indices[e++] != c)) return false;
        if (e != A.h_starts[r + 1]) return false;
    }
    return true;
}

// Dynamic LDS a workgroup may request on this device.
static size_t grid_cg_lds_limit()
{
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
        return 65536;
    return size_t(std::min(v, 160 * 1024)) - 1024;   // the kernel's static shared data
}

static int batch_size()
{
    static int b = [] {
        const char* e = std::getenv("M4G_ITER_BATCH");
        return e ? std::max(1, std::atoi(e)) : 8;
    }();
    return b;
}

static void init_state(KState& h, const m4g_iteration* it, double* hist_dev)
{
    std::memset(&h, 0, sizeof(h));
    h.norm_r0 = it->norm_r0;
    h.rtol = it->rtol;
    h.atol = it->atol;
    h.i = it->iterations;
    h.error = it->error;
    h.max_iter = it->max_iter;
    h.history = it->history ? hist_dev : nullptr;
    h.history_cap = it->history ? it->history_capacity : 0;
}

// M4G_KRYLOV_GRAPH=1: replay the iteration batches as a hipGraph.  Off by
// default: the GPU time of a solve does not improve (the eager launches
// already run back to back; 128^3 IC(0)-PCG 43.36 ms eager, 43.63 ms
// replayed), only the host's work per batch shrinks to one launch.
static bool graph_batches()
{
    const char* e = std::getenv("M4G_KRYLOV_GRAPH");
    return e && std::atoi(e) != 0;
}

// A batch of body() calls captured once as a hipGraph and replayed.  The
// first batch runs eagerly (it makes the lazy allocations, which a capture
// cannot), the second is captured; a stream that cannot be captured (the
// null stream, a capture already running) keeps the eager launches.  The
// iteration kernels take only device pointers and read their scalars from
// the device state, so every replay is the same work as an eager batch, and
// no memset is recorded (the sweep kernels reset their own tickets).
struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchGraph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
    template <typename Body>
    bool capture(hipStream_t s, int calls, Body& body)
    {
        if (!s) return false;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        try {
            for (int k = 0; k < calls; k++) body();
        } catch (...) {
            // nothing recorded has run: the caller launches the batch eagerly,
            // where an error that was not the capture's shows again
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            return false;
        }
        if (hipStreamEndCapture(s, &graph) != hipSuccess || !graph) {
            (void)hipGetLastError();
            return false;
        }
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            exec = nullptr;
            return false;
        }
        return true;
    }
};

// Runs `body()` in batches (calls_per_poll calls, default one per iteration
// of a batch) until the device state reports finished; polls the flag of
// batch b while batch b+1 is queued.  With graph_batches(), a batch from the
// second on is one hipGraph launch.
template <typename Body>
static void run_batches(KrylovWs& ws, const m4g_iteration* it, Body&& body, int calls_per_poll = 0)
{
    hipStream_t s = current_stream();
    hipEvent_t ev[2];
    M4G_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    M4G_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    const int B = calls_per_poll > 0 ? calls_per_poll : batch_size();
    int slot = 0, pending = -1;
    int64_t launched = 0;
    bool finished = false;
    // guard: never queue more batches than max_iter allows (+1 for the check)
    const int64_t cap = int64_t(std::max(it->max_iter - it->iterations, 0)) + B + 1;
    BatchGraph bg;
    bool use_graph = false, tried = !graph_batches();
    try {
        while (!finished) {
            if (use_graph) {
                M4G_HIP(hipGraphLaunch(bg.exec, s));
            } else if (launched > 0 && !tried) {
                tried = true;
                use_graph = bg.capture(s, B, body);
                if (use_graph) M4G_HIP(hipGraphLaunch(bg.exec, s));
                else
                    for (int k = 0; k < B; k++) body();
            } else {
                for (int k = 0; k < B; k++) body();
            }
            launched += B;
            M4G_HIP(hipMemcpyAsync(ws.host_state + slot, ws.state.ptr, sizeof(KState), hipMemcpyDeviceToHost, s));
            M4G_HIP(hipEventRecord(ev[slot], s));
            if (pending >= 0) {
                M4G_HIP(hipEventSynchronize(ev[pending]));
                if (ws.host_state[pending].done) finished = true;
                if (it->interrupt && *it->interrupt) finished = true;
            }
            pending = slot;
            slot ^= 1;
            if (launched > cap) {
                M4G_HIP(hipEventSynchronize(ev[pending]));
                finished = true;
            }
        }
        M4G_HIP(hipStreamSynchronize(s));
    } catch (...) {
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
        throw;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
}

static int finish_state(KrylovWs& ws, m4g_iteration* it, int extra_fault)
{
    KState h;
    M4G_HIP(hipMemcpy(&h, ws.state.ptr, sizeof(KState), hipMemcpyDeviceToHost));
    const int i0 = it->iterations;
    if (it->history && h.history) {
        int hi = std::min(h.i, it->history_capacity - 1);
        if (hi >= i0)
            M4G_HIP(hipMemcpy(it->history + i0, ws.history.ptr + i0, size_t(hi - i0 + 1) * sizeof(double),
                              hipMemcpyDeviceToHost));
    }
    it->iterations = h.i;
    it->resid = h.resid;
    bool interrupted = !h.done && it->interrupt && *it->interrupt;
    M4G_REQUIRE(extra_fault == 0, M4G_ERR_TIMEOUT, "sweep wait timed out");
    if (interrupted) {
        it->error = M4G_ERR_INTERRUPTED;
        return M4G_ERR_INTERRUPTED;
    }
    M4G_REQUIRE(h.done, M4G_ERR_INVALID, "solver stopped before the iteration finished");
    it->error = h.error;
    return h.error;
}

template <typename T>
static int cg_t(m4g_matrix_s& A, void* xp, const void* bp, m4g_pc_s* P, m4g_iteration* it, KrylovWs& ws)
{
    const int64_t n = A.nrows;
    hipStream_t st = current_stream();
    const int vt = sizeof(T) == 8 ? M4G_F64 : M4G_F32;
    const T* vals = sizeof(T) == 8 ? (const T*)A.sell.vals.ptr : (const T*)A.sell.vals32.ptr;
    M4G_REQUIRE(vals, M4G_ERR_INVALID, "float matrix values not prepared");
    int64_t chunk, sgrid;
    spmv_geom(A, chunk, sgrid);
    const int pw = sell_pipe_width(A);
    // the persistent SpMV grid, but no more workgroups than one slice per
    // wavefront needs: on small systems its p.q partials then stay few
    // enough for k_cg_update_fold to form alpha itself
    const int pgrid = int(std::min<int64_t>(sell_pipe_grid(), std::max<int64_t>(1, (A.sell.nslices + kSliceWaves - 1) /
                                                                                          kSliceWaves)));
    ws.prepare(std::max<int64_t>(sgrid, pgrid), it->history ? it->history_capacity : 0);
    KState h;
    init_state(h, it, ws.history.ptr);
    M4G_HIP(hipMemcpyAsync(ws.state.ptr, &h, sizeof(KState), hipMemcpyHostToDevice, st));
    DevBuf<T> p(n), q(n), r(n), z(n);
    T* x = (T*)xp;
    const T* b = (const T*)bp;
    KState* S = ws.state.ptr;
    double* part = ws.partials.ptr;
    unsigned* cnt = ws.counter.ptr;
    const int rg = red_grid(n);
    // r = b - A*x (cg.hpp:93)
    launch_residual_t<T>(A, x, b, r.ptr, st);
    const bool identity = !P || P->kind == M4G_PC_IDENTITY;
    const int idpc = int(!P || P->kind == M4G_PC_IDENTITY);
#define FIN(K, OP, G) hipLaunchKernelGGL((k_fin<K, OP>), dim3(1), dim3(256), 0, st, (const double*)part, int(G), S, idpc)
    hipLaunchKernelGGL(k_cg_start<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)r.ptr, S, part, cnt);
    FIN(1, FIN_CG_START, rg);
    M4G_HIP(hipGetLastError());
    // with few reduction workgroups the rho and alpha finalizes are formed by
    // every workgroup of their consumer (same sums, one launch less each);
    // the r.r partials of the folded update go to the second half of `part`
    const char* fold_env = std::getenv("M4G_CG_FOLD");   // M4G_CG_FOLD=0: separate finalizes always
    const bool fold_ok = !(fold_env && std::atoi(fold_env) == 0);
    const bool fold = rg <= kFoldMaxBlocks && fold_ok;
    const int64_t part_half = std::max<int64_t>(std::max<int64_t>(sgrid, pgrid), kRedBlocks);
    double* part2 = part + part_half;
    // with a sweep preconditioner, rho = r.z comes from the last sweep: one
    // partial per tile, written as the tile stores its rows of z
    // (M4G_CG_SWEEP_DOT=0: a separate dot kernel)
    const char* sdot_env = std::getenv("M4G_CG_SWEEP_DOT");
    const int64_t dot_tiles = P && !(sdot_env && std::atoi(sdot_env) == 0) ? pc_dot_tiles(*P) : 0;
    const bool sweep_dot = dot_tiles > 0 && dot_tiles <= part_half;
    const int rho_blocks = sweep_dot ? int(dot_tiles) : rg;
    const bool fold_rho = rho_blocks <= kFoldMaxBlocks && fold_ok;
    auto body = [&] {
        const T* zz = r.ptr;
        bool rho_partials = false;
        if (!identity) {
            if (P->kind == M4G_PC_DIAGONAL) {
                hipLaunchKernelGGL(k_cg_diag_rho<T>, dim3(rg), dim3(kNT), 0, st, n, (const double*)P->inv_diag.ptr,
                                   (const T*)r.ptr, z.ptr, S, part, cnt);
            } else if (sweep_dot) {
                pc_apply_skip_dot(*P, r.ptr, z.ptr, vt, &S->done, st, r.ptr, part);
            } else {
                pc_apply_skip(*P, r.ptr, z.ptr, vt, &S->done, st);
                hipLaunchKernelGGL(k_cg_rho<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)r.ptr, (const T*)z.ptr, S,
                                   part, cnt);
            }
            if (!fold_rho) FIN(1, FIN_CG_RHO, rho_blocks);
            rho_partials = true;
            zz = z.ptr;
        }
        if (fold_rho && rho_partials)
            hipLaunchKernelGGL(k_cg_p_fold<T>, dim3(rg), dim3(kNT), 0, st, n, zz, p.ptr, S, (const double*)part,
                               rho_blocks);
        else
            hipLaunchKernelGGL(k_cg_p<T>, dim3(rg), dim3(kNT), 0, st, n, zz, p.ptr, (const KState*)S);
        int pq_blocks;
        if (pw) {
            M4G_PIPE_LAUNCH(k_cg_spmv_pipe, T, pw, dim3(pgrid), dim3(kSpmvNT), 0, st, n, A.sell.nslices,
                            A.sell.padded, A.sell.slice_ptr.ptr, A.sell.idx.ptr, vals, (const T*)p.ptr, q.ptr, S, part);
            pq_blocks = pgrid;
        } else {
            hipLaunchKernelGGL(k_cg_spmv_dot<T>, dim3(sgrid), dim3(kSpmvNT), 0, st, n, A.sell.nslices, chunk,
                               A.sell.slice_ptr.ptr, A.sell.slice_idx.ptr, A.sell.idx.ptr, vals, (const T*)p.ptr,
                               q.ptr, S, part, cnt);
            pq_blocks = int(sgrid);
        }
        if (fold) {
            hipLaunchKernelGGL(k_cg_update_fold<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)p.ptr, (const T*)q.ptr,
                               x, r.ptr, S, (const double*)part, pq_blocks, part2);
            hipLaunchKernelGGL((k_fin<1, FIN_CG_UPDATE>), dim3(1), dim3(256), 0, st, (const double*)part2, int(rg), S,
                               idpc);
        } else {
            FIN(1, FIN_CG_ALPHA, pq_blocks);
            hipLaunchKernelGGL(k_cg_update<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)p.ptr, (const T*)q.ptr, x,
                               r.ptr, S, int(identity), part, cnt);
            FIN(1, FIN_CG_UPDATE, rg);
        }
        M4G_HIP(hipGetLastError());
    };
    // a small 2D grid with an FP64 IC(0) factor: whole batches of iterations
    // in one workgroup (k_cg_ic0_grid)
    bool grid = false;
    if constexpr (sizeof(T) == 8) {
        DevBuf<double> east, north, dinv;
        const char* env = std::getenv("M4G_GRID_CG");   // M4G_GRID_CG=0: the multi-kernel path
        const bool enabled = !(env && std::atoi(env) == 0);
        const size_t lds = P && enabled ? size_t(grid_cg_lds_doubles(n, P->grid.nx)) * sizeof(double) : 0;
        if (P && enabled && lds <= grid_cg_lds_limit() && grid5_pattern(A, P->grid) &&
            pc_grid_factor(*P, kGridMaxN, east, north, dinv, st)) {
            grid = true;
            static size_t granted = 0;
            if (lds > 65536 && lds > granted) {
                M4G_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cg_ic0_grid),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
                granted = lds;
            }
            // four batches' worth of iterations per launch (the launch and the
            // LDS fill cost more than an iteration here)
            GridCgArgs ga{int32_t(n), int32_t(P->grid.nx), int32_t(P->grid.ny), 4 * batch_size(), A.starts.ptr,
                          A.values.ptr, east.ptr, north.ptr, dinv.ptr, (double*)x, (double*)r.ptr, (double*)p.ptr, S};
            run_batches(ws, it, [&] {
                hipLaunchKernelGGL(k_cg_ic0_grid, dim3(1), dim3(kGridNT), lds, st, ga);
                M4G_HIP(hipGetLastError());
            }, 1);
        }
    }
    if (!grid) run_batches(ws, it, body);
#undef FIN
    int fault = 0;
    if (P) fault = check_pc_faults(*P);
    return finish_state(ws, it, fault);
}

template <typename T>
static int bicgstab_t(m4g_matrix_s& A, void* xp, const void* bp, m4g_pc_s* P, m4g_iteration* it, KrylovWs& ws)
{
    const int64_t n = A.nrows;
    hipStream_t st = current_stream();
    const int vt = sizeof(T) == 8 ? M4G_F64 : M4G_F32;
    const T* vals = sizeof(T) == 8 ? (const T*)A.sell.vals.ptr : (const T*)A.sell.vals32.ptr;
    M4G_REQUIRE(vals, M4G_ERR_INVALID, "float matrix values not prepared");
    int64_t chunk, sgrid;
    spmv_geom(A, chunk, sgrid);
    const int pw = sell_pipe_width(A);
    const int pgrid = sell_pipe_grid();
    ws.prepare(std::max<int64_t>(sgrid, pgrid), it->history ? it->history_capacity : 0);
    KState h;
    init_state(h, it, ws.history.ptr);
    M4G_HIP(hipMemcpyAsync(ws.state.ptr, &h, sizeof(KState), hipMemcpyHostToDevice, st));
    DevBuf<T> p(n), phat(n), s_(n), shat(n), t_(n), v(n), r(n), rtilde(n);
    T* x = (T*)xp;
    const T* b = (const T*)bp;
    KState* S = ws.state.ptr;
    double* part = ws.partials.ptr;
    unsigned* cnt = ws.counter.ptr;
    const int rg = red_grid(n);
    launch_residual_t<T>(A, x, b, r.ptr, st);                       // r = b - A*x (:40)
    const int idpc = 0;
#define FIN(K, OP, G) hipLaunchKernelGGL((k_fin<K, OP>), dim3(1), dim3(256), 0, st, (const double*)part, int(G), S, idpc)
    hipLaunchKernelGGL(k_bicg_start<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)r.ptr, rtilde.ptr, S, part, cnt);
    FIN(2, FIN_BICG_START, rg);
    M4G_HIP(hipGetLastError());
    const bool identity = !P || P->kind == M4G_PC_IDENTITY;
    auto apply = [&](const T* in, T* out) {
        if (identity) M4G_HIP(hipMemcpyAsync(out, in, n * sizeof(T), hipMemcpyDeviceToDevice, st));
        else pc_apply_skip(*P, in, out, vt, &S->done, st);
    };
    auto body = [&] {
        hipLaunchKernelGGL(k_bicg_p<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)r.ptr, (const T*)v.ptr, p.ptr,
                           (const KState*)S);
        apply(p.ptr, phat.ptr);                                      // phat = solve(M, p) (:55)
        if (pw) {
            M4G_PIPE_LAUNCH(k_bicg_spmv_gamma_pipe, T, pw, dim3(pgrid), dim3(kSpmvNT), 0, st, n, A.sell.nslices,
                            A.sell.padded, A.sell.slice_ptr.ptr, A.sell.idx.ptr, vals, (const T*)phat.ptr,
                            (const T*)rtilde.ptr, v.ptr, S, part);
            FIN(1, FIN_BICG_GAMMA, pgrid);
        } else {
            hipLaunchKernelGGL(k_bicg_spmv_gamma<T>, dim3(sgrid), dim3(kSpmvNT), 0, st, n, A.sell.nslices, chunk,
                               A.sell.slice_ptr.ptr, A.sell.slice_idx.ptr, A.sell.idx.ptr, vals, (const T*)phat.ptr,
                               (const T*)rtilde.ptr, v.ptr, S, part, cnt);
            FIN(1, FIN_BICG_GAMMA, sgrid);
        }
        hipLaunchKernelGGL(k_bicg_s<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)r.ptr, (const T*)v.ptr, s_.ptr, S,
                           part, cnt);
        FIN(1, FIN_BICG_S, rg);
        hipLaunchKernelGGL(k_bicg_exit<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)phat.ptr, x, S, cnt + 1);
        apply(s_.ptr, shat.ptr);                                     // shat = solve(M, s) (:68)
        if (pw) {
            M4G_PIPE_LAUNCH(k_bicg_spmv_omega_pipe, T, pw, dim3(pgrid), dim3(kSpmvNT), 0, st, n, A.sell.nslices,
                            A.sell.padded, A.sell.slice_ptr.ptr, A.sell.idx.ptr, vals, (const T*)shat.ptr,
                            (const T*)s_.ptr, t_.ptr, S, part);
            FIN(2, FIN_BICG_OMEGA, pgrid);
        } else {
            hipLaunchKernelGGL(k_bicg_spmv_omega<T>, dim3(sgrid), dim3(kSpmvNT), 0, st, n, A.sell.nslices, chunk,
                               A.sell.slice_ptr.ptr, A.sell.slice_idx.ptr, A.sell.idx.ptr, vals, (const T*)shat.ptr,
                               (const T*)s_.ptr, t_.ptr, S, part, cnt);
            FIN(2, FIN_BICG_OMEGA, sgrid);
        }
        hipLaunchKernelGGL(k_bicg_update<T>, dim3(rg), dim3(kNT), 0, st, n, (const T*)shat.ptr, (const T*)phat.ptr,
                           (const T*)s_.ptr, (const T*)t_.ptr, (const T*)rtilde.ptr, x, r.ptr, S, part, cnt);
        FIN(2, FIN_BICG_UPDATE, rg);
        M4G_HIP(hipGetLastError());
    };
    run_batches(ws, it, body);
#undef FIN
    int fault = 0;
    if (P) fault = check_pc_faults(*P);
    return finish_state(ws, it, fault);
}

int krylov_solve(int method, m4g_matrix_s& A, void* x, const void* b, int vtype, m4g_pc_s* P, m4g_iteration* it,
                 KrylovWs& ws)
{
    if (vtype == M4G_F32 && !A.sell.vals32.ptr) sell_build_f32(A);
    if (method == M4G_METHOD_CG)
        return vtype == M4G_F64 ? cg_t<double>(A, x, b, P, it, ws) : cg_t<float>(A, x, b, P, it, ws);
    return vtype == M4G_F64 ? bicgstab_t<double>(A, x, b, P, it, ws) : bicgstab_t<float>(A, x, b, P, it, ws);
}

KrylovWs& default_ws()
{
    static std::map<int, std::unique_ptr<KrylovWs>> m;
    static std::mutex mu;
    DeviceState& d = device_state();
    std::lock_guard<std::mutex> lock(mu);
    auto& p = m[d.device];
    if (!p) p = std::make_unique<KrylovWs>();
    return *p;
}

}  // namespace m4g

using namespace m4g;

extern "C" {

static int solve_common(int method, m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* it)
{
    return guard([&] {
        M4G_REQUIRE(A && x && b && it, M4G_ERR_INVALID, "solver: null argument");
        M4G_REQUIRE(A->nrows == A->ncols, M4G_ERR_NOT_SQUARE, "Matrix is not square");
        M4G_REQUIRE(x->n == A->ncols && b->n == A->nrows, M4G_ERR_SIZE, "solver: incompatible size");
        M4G_REQUIRE(x->type == b->type, M4G_ERR_INVALID, "solver: x and b types differ");
        M4G_REQUIRE(!P || P->n == A->nrows, M4G_ERR_SIZE, "solver: preconditioner size differs");
        M4G_REQUIRE(it->max_iter >= 0, M4G_ERR_INVALID, "solver: negative max_iter");
        if (it->history) M4G_REQUIRE(it->history_capacity > 0, M4G_ERR_INVALID, "solver: history capacity");
        // cg.hpp:44 (no preconditioner) and :84, bicgstab.hpp:34
        TraceRange tr(method == M4G_METHOD_CG ? (P ? kTraceCg : kTraceCgNoPc) : kTraceBicgstab);
        return krylov_solve(method, *A, x->ptr, b->ptr, x->type, P, it, default_ws());
    });
}

int m4g_cg(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* it)
{
    return solve_common(M4G_METHOD_CG, A, x, b, P, it);
}

int m4g_bicgstab(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* it)
{
    return solve_common(M4G_METHOD_BICGSTAB, A, x, b, P, it);
}

}  // extern "C"
