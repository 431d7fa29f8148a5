// Host-side schedules of the triangular sweeps and factorisations.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace m4g {

// Large host blocks (at least kLargeBlock bytes) on 2 MiB boundaries with
// transparent huge pages requested: the planning arrays are touched once,
// and first-touch faults of 4 KiB pages cost more than filling them.
constexpr size_t kLargeBlock = size_t(4) << 20;
void* large_block_alloc(size_t bytes);
void large_block_free(void* p) noexcept;

// Allocator whose value-initialisation leaves trivial elements uninitialised:
// the plan's large arrays are written in parallel, and each page is first
// touched by the thread that fills it.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <typename U>
    NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
    T* allocate(size_t n)
    {
        if (n * sizeof(T) < kLargeBlock) return std::allocator<T>::allocate(n);
        return static_cast<T*>(large_block_alloc(n * sizeof(T)));
    }
    void deallocate(T* p, size_t n) noexcept
    {
        if (n * sizeof(T) < kLargeBlock) std::allocator<T>::deallocate(p, n);
        else large_block_free(p);
    }
    template <typename U>
    void construct(U* p) noexcept
    {
        ::new (static_cast<void*>(p)) U;
    }
    template <typename U, typename... A>
    void construct(U* p, A&&... a)
    {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
using PlanVec32 = std::vector<int32_t, NoInitAlloc<int32_t>>;
using PlanVec64 = std::vector<int64_t, NoInitAlloc<int64_t>>;

// Natural (lexicographic) grid inferred from a matrix pattern: row
// r = i + nx*(j + ny*k) and every stored entry couples cells at Chebyshev
// distance <= 1.  nx == 0 means no grid was found.
struct GridInfo {
    int64_t nx = 0, ny = 0, nz = 0;
};
GridInfo infer_grid(int64_t n, const int64_t* starts, const int64_t* indices);

// One row of a sweep: entries in the order they are subtracted, with the
// factor-storage index of each value and of the diagonal (-1: unit diagonal).
// Rows and factor-storage indices are 32-bit: the builders below reject a
// matrix of 2^31 or more rows or stored entries.
struct SweepRows {
    int64_t n = 0;
    bool lower = true;                    // dependencies point to smaller rows
    PlanVec64 ptr;                        // [n+1]
    PlanVec32 col;                        // dependency row of each entry
    PlanVec32 val_pos;                    // factor storage index of each entry
    PlanVec32 diag_pos;                   // [n] factor storage index of the diagonal, -1 = unit
};

// Tile shape in the skewed grid coordinates t1 = k, t2 = j + a*k,
// t3 = i + b*j + c*k (all dependencies are monotone in each coordinate).
struct TileShape {
    bool grid = false;
    int64_t a = 0, b = 0, c = 0;
    int64_t e1 = 1, e2 = 1, e3 = 1;       // extents in t1, t2, t3
    int64_t rows_per_block = 0;           // row-block tiles when !grid
};

// Compiled sweep: tiles in ticket (topological) order; inside a tile, groups
// of at most 64 pieces in order; the pieces of a group are independent and run
// one per lane of a wavefront.
//
// A piece is a run of at most `width` consecutive entries of one row.  A row
// with more entries has several pieces in successive groups, each continuing
// the running value the previous one left in the row's LDS slot; the last
// piece applies the diagonal and exports the result.  This keeps the per-row
// operation order while every piece has the same fixed-size record.
struct SweepPlan {
    int64_t n = 0;
    bool lower = true, unit = false;
    int width = 4;                        // entries per piece
    int64_t ntiles = 0, ngroups = 0, npieces = 0, nexport = 0;
    int64_t max_tile_rows = 0, max_steps = 0, levels = 0;
    TileShape shape;
    std::vector<int32_t> tile_row_begin;  // [ntiles+1] into rows
    PlanVec32 rows;                       // [n] rows of each tile, in natural order
    PlanVec32 row_slot;                   // [n] LDS slot of each entry of rows
    std::vector<int32_t> tile_group_begin;// [ntiles+1] into groups
    std::vector<int32_t> tile_export_begin;// [ntiles+1] export slots owned by each tile
    std::vector<int32_t> xlist;           // [nexport] LDS slot of the row behind each export slot
    std::vector<int32_t> group_xbefore;   // [ngroups] exports of the tile finished before the group
    int64_t max_tile_exports = 0;
    std::vector<int32_t> group_piece_begin;// [ngroups+1] into the piece arrays
    std::vector<int32_t> group_need;      // [ngroups] imports of the tile needed before the group runs
    PlanVec32 pc_slot;                    // [npieces] LDS slot of the piece's row in its tile
    PlanVec32 pc_len;                     // [npieces] entries of the piece (<= width)
    PlanVec32 pc_export;                  // [npieces] export slot (last piece of an exported row) or -1
    PlanVec32 pc_diag_pos;                // [npieces] factor index of the diagonal; -1: none
    PlanVec32 pc_src;                     // [npieces*width] LDS index of each entry's value:
                                          //   slot (< max_tile_rows) or max_tile_rows + tile-local import
    PlanVec32 pc_val_pos;                 // [npieces*width] factor index; -1: padding
                                          // (factor storage is limited to 2^31 - 1 entries)
    std::vector<int32_t> tile_import_begin;  // [ntiles+1] into imports
    std::vector<int32_t> imports;         // exchange slot of each import, per tile in first-use order
    int64_t max_tile_imports = 0, nimports = 0;
};

// Entries per piece for rows of at most max_len entries.
int sweep_piece_width(int64_t max_len);

// Natural-order level of every row (0-based) of a lower (or upper) DAG.
std::vector<int32_t> natural_levels(const SweepRows& s, int64_t* nlevels);

// Default cost-model constants (microseconds): latency of a value crossing
// from one tile to the next, and of one wavefront pass over a group.
constexpr double kSweepHopUs = 3.0;
constexpr double kSweepStepUs = 0.2;

// Chooses a tile shape for the sweep (cost model: pipeline hops * hop_us +
// in-tile levels * step_us) subject to max_tile_rows.  levels: natural_levels
// of s when the caller has them (else computed here).
TileShape choose_tiles(const SweepRows& s, const GridInfo& g, int64_t max_tile_rows, double hop_us, double step_us,
                       const std::vector<int32_t>* levels = nullptr);

// levels: the natural level count of s when known (else computed here).
// Construction runs on plan_threads() host threads; the plan does not depend
// on their number.
SweepPlan build_sweep_plan(const SweepRows& s, const GridInfo& g, const TileShape& shape, int64_t max_tile_rows,
                           int width = 0, int64_t levels = -1);
int plan_threads();

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
// All return false on a missing diagonal, and throw std::length_error for a
// matrix of 2^31 or more rows or stored entries.
bool upper_pattern(int64_t n, const int64_t* S, const int64_t* I, PlanVec64& US, PlanVec64& UI, PlanVec32& a_pos);
bool ic0_rows(int64_t n, const PlanVec64& US, const PlanVec64& UI, SweepRows& lower, SweepRows& upper);
bool ilu0_rows(int64_t n, const int64_t* S, const int64_t* I, SweepRows& lower, SweepRows& upper,
               PlanVec64& diag_pos);
// Sweeps of ilu::adjoint_solve (ilu.hpp:105-115): lower on U^T (inverse
// diagonal, ascending rows) then unit upper on L^T (descending rows).
bool ilu0_adjoint_rows(int64_t n, const int64_t* S, const int64_t* I, SweepRows& lower, SweepRows& upper);

}  // namespace m4g
