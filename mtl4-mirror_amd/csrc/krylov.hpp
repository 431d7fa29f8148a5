// Device state and workspace of the Krylov drivers.
#pragma once

#include "m4g_internal.hpp"

struct m4g_pc_s;

namespace m4g {

// Scalars of the method and the iteration object, resident on the device.
struct KState {
    double rho, rho_1, rho_2, rr, alpha, alpha_1, beta, omega, gamma, resid;
    double norm_r0, rtol, atol;
    double* history;
    int history_cap;
    int i, done, error, max_iter, mid_exit;
};

struct KrylovWs {
    DevBuf<KState> state;
    DevBuf<double> partials;
    DevBuf<unsigned> counter;
    DevBuf<double> history;
    KState* host_state = nullptr;   // pinned, two slots
    void prepare(int64_t spmv_grid, int history_cap);
    KrylovWs() = default;
    KrylovWs(const KrylovWs&) = delete;
    ~KrylovWs();
};

KrylovWs& default_ws();

// x (vtype) solved in place; returns the iteration error code.
int krylov_solve(int method, m4g_matrix_s& A, void* x, const void* b, int vtype, m4g_pc_s* P, m4g_iteration* it,
                 KrylovWs& ws);

// r = b - A*x in value type T (spmv.hip)
template <typename T>
void launch_residual_t(const m4g_matrix_s& A, const T* x, const T* b, T* r, hipStream_t s);

// PC application that is skipped on the device when *skip != 0 (pc.hip)
void pc_apply_skip(m4g_pc_s& P, const void* x, void* y, int vtype, const int* skip, hipStream_t s);

// Tiles of the PC's last sweep: the number of partials pc_apply_skip_dot
// leaves; 0 when the PC has no sweeps.
int64_t pc_dot_tiles(const m4g_pc_s& P);

// pc_apply_skip of an IC(0)/ILU(0) PC that also stores, per tile of its last
// sweep, the partial sum of w_r * y_r over the tile's rows into partials.
void pc_apply_skip_dot(m4g_pc_s& P, const void* x, void* y, int vtype, const int* skip, hipStream_t s,
                       const void* w, double* partials);

}  // namespace m4g
