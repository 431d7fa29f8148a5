// Host-side schedules of the triangular sweeps and factorisations.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace m4g {

// Natural (lexicographic) grid inferred from a matrix pattern: row
// r = i + nx*(j + ny*k) and every stored entry couples cells at Chebyshev
// distance <= 1.  nx == 0 means no grid was found.
struct GridInfo {
    int64_t nx = 0, ny = 0, nz = 0;
};
GridInfo infer_grid(int64_t n, const int64_t* starts, const int64_t* indices);

// One row of a sweep: entries in the order they are subtracted, with the
// factor-storage index of each value and of the diagonal (-1: unit diagonal).
struct SweepRows {
    int64_t n = 0;
    bool lower = true;                    // dependencies point to smaller rows
    std::vector<int64_t> ptr;             // [n+1]
    std::vector<int64_t> col;             // dependency row of each entry
    std::vector<int64_t> val_pos;         // factor storage index of each entry
    std::vector<int64_t> diag_pos;        // [n] factor storage index of the diagonal, -1 = unit
};

// Tile shape in the skewed grid coordinates t1 = k, t2 = j + a*k,
// t3 = i + b*j + c*k (all dependencies are monotone in each coordinate).
struct TileShape {
    bool grid = false;
    int64_t a = 0, b = 0, c = 0;
    int64_t e1 = 1, e2 = 1, e3 = 1;       // extents in t1, t2, t3
    int64_t rows_per_block = 0;           // row-block tiles when !grid
};

// Compiled sweep: tiles in ticket (topological) order; inside a tile, level
// groups in order; inside a group, rows processed concurrently.
struct SweepPlan {
    int64_t n = 0;
    bool lower = true, unit = false;
    int block = 64;                       // workgroup size
    int64_t ntiles = 0, ngroups = 0, nent = 0, nexport = 0;
    int64_t max_tile_rows = 0, max_steps = 0, levels = 0;
    TileShape shape;
    std::vector<int32_t> tile_row_begin;  // [ntiles+1] into rows
    std::vector<int32_t> rows;            // [n] global row of each (tile, slot)
    std::vector<int32_t> tile_group_begin;// [ntiles+1] into groups
    std::vector<int32_t> tile_export_begin;// [ntiles+1] export slots owned by each tile
    std::vector<int64_t> group_row_begin; // [ngroups+1] into the per-group-row arrays
    std::vector<int64_t> group_ent_begin; // [ngroups+1] into the entry arrays
    std::vector<int32_t> group_width;     // [ngroups] max entries of a row in the group
    std::vector<int32_t> gr_slot;         // [n] LDS slot of the row in its tile
    std::vector<int32_t> gr_len;          // [n] number of entries
    std::vector<int32_t> gr_export;       // [n] export slot or -1
    std::vector<int64_t> gr_diag_pos;     // [n] factor storage index of the diagonal or -1
    std::vector<int32_t> ent_src;         // [nent] LDS slot (>= 0) or -(tile-local import + 1)
    std::vector<int32_t> tile_import_begin;  // [ntiles+1] into imports
    std::vector<int32_t> imports;         // exchange slot of each import, per tile in first-use order
    std::vector<int32_t> group_need;      // [ngroups] imports of the tile needed before the group runs
    int64_t max_tile_imports = 0, nimports = 0;
    std::vector<int64_t> ent_val_pos;     // [nent] factor storage index (-1: padding)
};

// Natural-order level of every row (0-based) of a lower (or upper) DAG.
std::vector<int32_t> natural_levels(const SweepRows& s, int64_t* nlevels);

// Chooses a tile shape for the sweep (cost model: pipeline hops * hop_us +
// in-tile levels * step_us) subject to max_tile_rows.
TileShape choose_tiles(const SweepRows& s, const GridInfo& g, int64_t max_tile_rows, double hop_us, double step_us);

SweepPlan build_sweep_plan(const SweepRows& s, const GridInfo& g, const TileShape& shape, int64_t max_tile_rows);

// Checks every invariant the sweep kernel relies on (bounds, ordering of
// dependencies before their use, tile order acyclic).  Returns "" when valid.
std::string validate_sweep_plan(const SweepPlan& p, const SweepRows& s);

// Executes the plan on the CPU with the kernel's operation order.
// FV: factor value type, VT: vector / accumulation type.
template <typename FV, typename VT>
void emulate_sweep(const SweepPlan& p, const FV* factor, const VT* v, VT* y);

}  // namespace m4g

namespace m4g {

// Row descriptions of the sweeps of each factorisation.  Patterns are CSR with
// int64 indices; *_pos index the factor storage:
//  - IC(0): storage = upper(A) in CSR (inverted diagonal first per row);
//  - ILU(0): storage = A's CSR (L strictly lower, U upper, diagonal inverted).
// All return false on a missing diagonal.
bool upper_pattern(int64_t n, const int64_t* S, const int64_t* I, std::vector<int64_t>& US,
                   std::vector<int64_t>& UI, std::vector<int64_t>& a_pos);
bool ic0_rows(int64_t n, const std::vector<int64_t>& US, const std::vector<int64_t>& UI, SweepRows& lower,
              SweepRows& upper);
bool ilu0_rows(int64_t n, const int64_t* S, const int64_t* I, SweepRows& lower, SweepRows& upper,
               std::vector<int64_t>& diag_pos);
// Sweeps of ilu::adjoint_solve (ilu.hpp:105-115): lower on U^T (inverse
// diagonal, ascending rows) then unit upper on L^T (descending rows).
bool ilu0_adjoint_rows(int64_t n, const int64_t* S, const int64_t* I, SweepRows& lower, SweepRows& upper);

}  // namespace m4g
