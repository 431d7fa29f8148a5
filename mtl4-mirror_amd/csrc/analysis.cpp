// Host-only analysis entry points: plan construction, validation and CPU
// emulation of the tile-pipelined sweeps (see mtl4gpu_analysis.h).
#include "../../include/mtl4gpu_analysis.h"
#include "schedule.hpp"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace m4g {
void set_last_error(const std::string& msg);
int64_t analysis_default_cap(int vtype) { return vtype == M4G_F64 ? 8192 : 16384; }
}  // namespace m4g

using namespace m4g;


extern "C" {

int m4g_analysis_grid(int64_t n, const int64_t* S, const int64_t* I, int64_t* nx, int64_t* ny, int64_t* nz)
{
    GridInfo g = infer_grid(n, S, I);
    if (nx) *nx = g.nx;
    if (ny) *ny = g.ny;
    if (nz) *nz = g.nz;
    return M4G_OK;
}

int m4g_analysis_pc_solve(int kind, int adjoint, int64_t n, const int64_t* S, const int64_t* I,
                          const double* factor, int ftype, int vtype, const void* rhs, void* out, int64_t cap,
                          m4g_pc_stats* stats)
{
    try {
        if (kind != M4G_PC_IC0 && kind != M4G_PC_ILU0) {
            set_last_error("m4g_analysis_pc_solve: kind must be IC0 or ILU0");
            return M4G_ERR_INVALID;
        }
        if (cap <= 0) cap = analysis_default_cap(vtype);
        SweepRows lo, up;
        if (kind == M4G_PC_IC0) {
            PlanVec64 US, UI;
            PlanVec32 apos;
            if (!upper_pattern(n, S, I, US, UI, apos) || !ic0_rows(n, US, UI, lo, up)) {
                set_last_error("missing diagonal");
                return M4G_ERR_MISSING_DIAGONAL;
            }
        } else {
            PlanVec64 dpos;
            bool ok = adjoint ? ilu0_adjoint_rows(n, S, I, lo, up) : ilu0_rows(n, S, I, lo, up, dpos);
            if (!ok) {
                set_last_error("missing diagonal");
                return M4G_ERR_MISSING_DIAGONAL;
            }
        }
        GridInfo g = infer_grid(n, S, I);
        TileShape sl = choose_tiles(lo, g, cap, kSweepHopUs, kSweepStepUs),
                  su = choose_tiles(up, g, cap, kSweepHopUs, kSweepStepUs);
        SweepPlan pl = build_sweep_plan(lo, g, sl, cap), pu = build_sweep_plan(up, g, su, cap);
        std::string m1 = validate_sweep_plan(pl, lo), m2 = validate_sweep_plan(pu, up);
        if (!m1.empty() || !m2.empty()) {
            set_last_error("invalid plan: " + m1 + " / " + m2);
            return M4G_ERR_LOGIC;
        }
        size_t fn = 0;
        for (auto p : lo.val_pos) fn = std::max<size_t>(fn, size_t(p) + 1);
        for (auto p : up.val_pos) fn = std::max<size_t>(fn, size_t(p) + 1);
        for (auto p : lo.diag_pos) fn = std::max<size_t>(fn, size_t(p + 1));
        for (auto p : up.diag_pos) fn = std::max<size_t>(fn, size_t(p + 1));
        if (ftype == M4G_F64 && vtype == M4G_F64) {
            std::vector<double> tmp(n);
            emulate_sweep<double, double>(pl, factor, (const double*)rhs, tmp.data());
            emulate_sweep<double, double>(pu, factor, tmp.data(), (double*)out);
        } else {
            std::vector<float> f32(fn);
            for (size_t i = 0; i < fn; i++) f32[i] = float(factor[i]);
            if (vtype == M4G_F64) {
                std::vector<double> tmp(n);
                emulate_sweep<float, double>(pl, f32.data(), (const double*)rhs, tmp.data());
                emulate_sweep<float, double>(pu, f32.data(), tmp.data(), (double*)out);
            } else {
                std::vector<float> tmp(n);
                emulate_sweep<float, float>(pl, f32.data(), (const float*)rhs, tmp.data());
                emulate_sweep<float, float>(pu, f32.data(), tmp.data(), (float*)out);
            }
        }
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->levels_lower = pl.levels;
            stats->levels_upper = pu.levels;
            stats->tiles_lower = pl.ntiles;
            stats->tiles_upper = pu.ntiles;
            stats->steps_lower = pl.max_steps;
            stats->steps_upper = pu.max_steps;
            stats->grid_nx = g.nx;
            stats->grid_ny = g.ny;
            stats->grid_nz = g.nz;
            if (sl.grid) {
                stats->tile_nx = sl.e3;
                stats->tile_ny = sl.e2;
                stats->tile_nz = sl.e1;
            }
        }
        return M4G_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return M4G_ERR_LOGIC;
    }
}

}  // extern "C"
