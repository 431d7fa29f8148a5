// Host-only analysis entry points: plan construction, validation and CPU
// emulation of the tile-pipelined sweeps and (M4G_LINE_SWEEP=1) of the
// line-pipelined ones (see mtl4gpu_analysis.h).
#include "../../include/mtl4gpu_analysis.h"
#include "schedule.hpp"

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

namespace {

// CPU emulation of k_line_sweep's schedule: bands in order, then steps, then
// lanes, with the kernel's operations (selects for absent entries, the
// previous lane's result of the previous step as the across-line value, band
// k-1's line 64k-1 for lane 0).  Returns false on an index outside a buffer.
template <typename FV, typename VT>
bool emulate_line(const LinePlan& p, bool unit, const FV* f, const VT* v, VT* y)
{
    const int64_t plane = p.nbands * p.steps * 64;
    std::vector<VT> last(p.nx), prev_last(p.nx);
    for (int64_t k = 0; k < p.nbands; k++) {
        std::vector<VT> yw(64, VT(0)), ynew(64);
        for (int64_t st = 0; st < p.steps; st++) {
            for (int64_t lane = 0; lane < 64; lane++) {
                const int64_t j = 64 * k + lane, i = st - lane;
                const bool act = j < p.ny && i >= 0 && i < p.nx;
                const int64_t c = (k * p.steps + st) * 64 + lane;
                auto coef = [&](int64_t plane_no, FV absent) {
                    const int64_t q = p.pos[plane_no * plane + c];
                    return q >= 0 ? f[q] : absent;
                };
                const FV w = coef(0, FV(0)), sv = coef(1, FV(0)), d = coef(2, FV(1));
                const int64_t q = j * p.nx + i, r = p.rev ? p.n - 1 - q : q;
                if (act && (r < 0 || r >= p.n)) return false;
                const VT x0 = k > 0 && st < p.nx ? prev_last[st] : VT(0);
                const VT ys = lane == 0 ? x0 : yw[lane - 1];
                VT rr = act ? v[r] : VT(0);
                const bool has_w = i > 0, has_s = j > 0;
                if (p.s_first) {
                    rr = has_s ? rr - VT(sv) * ys : rr;
                    rr = has_w ? rr - VT(w) * yw[lane] : rr;
                } else {
                    rr = has_w ? rr - VT(w) * yw[lane] : rr;
                    rr = has_s ? rr - VT(sv) * ys : rr;
                }
                if (!unit) rr = rr * VT(d);
                if (act) {
                    y[r] = rr;
                    if (lane == 63) last[i] = rr;
                }
                ynew[lane] = rr;
            }
            yw.swap(ynew);
        }
        prev_last.swap(last);
    }
    return true;
}

template <typename FV, typename VT>
bool emulate_line_pair(const LinePlan& pl, const LinePlan& pu, bool unit_l, bool unit_u, const FV* f, const VT* v,
                       VT* y, int64_t n)
{
    std::vector<VT> tmp(n);
    return emulate_line<FV, VT>(pl, unit_l, f, v, tmp.data()) && emulate_line<FV, VT>(pu, unit_u, f, tmp.data(), y);
}

}  // namespace

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
            std::vector<int64_t> US, UI, apos;
            if (!upper_pattern(n, S, I, US, UI, apos) || !ic0_rows(n, US, UI, lo, up)) {
                set_last_error("missing diagonal");
                return M4G_ERR_MISSING_DIAGONAL;
            }
        } else {
            std::vector<int64_t> dpos;
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
        // M4G_LINE_SWEEP=1: the line-pipelined schedule where both directions have it
        LinePlan ll, lu;
        const char* le = std::getenv("M4G_LINE_SWEEP");
        const bool line = le && le[0] == '1' && g.nx >= 32 && plan_line_sweep(lo, g, 8, ll) &&
                          plan_line_sweep(up, g, 8, lu);
        size_t fn = 0;
        for (auto p : lo.val_pos) fn = std::max<size_t>(fn, size_t(p) + 1);
        for (auto p : up.val_pos) fn = std::max<size_t>(fn, size_t(p) + 1);
        for (auto p : lo.diag_pos) fn = std::max<size_t>(fn, size_t(p + 1));
        for (auto p : up.diag_pos) fn = std::max<size_t>(fn, size_t(p + 1));
        if (line) {
            bool ok;
            if (ftype == M4G_F64 && vtype == M4G_F64) {
                ok = emulate_line_pair<double, double>(ll, lu, pl.unit, pu.unit, factor, (const double*)rhs,
                                                       (double*)out, n);
            } else {
                std::vector<float> f32(fn);
                for (size_t i = 0; i < fn; i++) f32[i] = float(factor[i]);
                ok = vtype == M4G_F64 ? emulate_line_pair<float, double>(ll, lu, pl.unit, pu.unit, f32.data(),
                                                                         (const double*)rhs, (double*)out, n)
                                      : emulate_line_pair<float, float>(ll, lu, pl.unit, pu.unit, f32.data(),
                                                                        (const float*)rhs, (float*)out, n);
            }
            if (!ok) {
                set_last_error("line plan indexes outside the vectors");
                return M4G_ERR_LOGIC;
            }
        } else if (ftype == M4G_F64 && vtype == M4G_F64) {
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
            if (line) stats->line_bands = ll.nbands;
        }
        return M4G_OK;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return M4G_ERR_LOGIC;
    }
}

}  // extern "C"
