// Preconditioner handle and compiled sweeps (device side).
#pragma once

#include "m4g_internal.hpp"
#include "schedule.hpp"

namespace m4g {

// A compiled sweep on the device.  Each piece of the plan becomes one packed
// record (layout in pc.hip); records are rebuilt from the position arrays when
// the factor values change.
struct SweepDev {
    SweepPlan plan_meta;                  // scalars and tile offsets (piece arrays dropped)
    DevBuf<int32_t> tile_row_begin, rows, row_slot, tile_group_begin, tile_import_begin, imports;
    DevBuf<int32_t> tile_export_begin, xlist;
    DevBuf<int4> group_tab;               // per group (+1): first piece, imports needed, exports before
    DevBuf<int2> pc_meta;                 // per piece: slot | len << 16, export slot
    DevBuf<int32_t> pc_src;               // per piece and entry: LDS index
    DevBuf<int32_t> pc_val_pos, pc_diag_pos; // per piece and entry / per piece: factor storage index, -1 none
    DevBuf<int4> recs;                    // packed records for factor_type and recs_vbytes
    const double* fsrc = nullptr;         // factor storage the records are packed from
    int recs_vbytes = 0;                  // vector entry size the records address (0: stale)
    DevBuf<unsigned char> xbuf;           // exchange buffer (double or float entries)
    int xbuf_type = -1;
    DevBuf<unsigned> tickets;
    DevBuf<int> fault;
    int factor_type = M4G_F64;

    void upload(const SweepPlan& p, const double* fstore_dev, int ftype);
    void refresh_values(const double* fstore_dev, int ftype);
    void ensure_recs(int vbytes, hipStream_t s);
    void ensure_exchange(int vtype);
};

// Both sweeps, v -> tmp -> y.  With dot_with (a vector of the same type), the
// second sweep also stores each tile's partial of dot(dot_with, y) into
// dot_partials[tile] (second.plan_meta.ntiles entries).
void run_sweep_pair(SweepDev& first, SweepDev& second, const void* v, void* tmp, void* y, int vtype,
                    const int* skip, hipStream_t s, const void* dot_with = nullptr, double* dot_partials = nullptr);
int check_sweep_fault(SweepDev& d);
int64_t sweep_tile_cap(int vtype);
SweepPlan plan_sweep(const SweepRows& rows, const GridInfo& g, int64_t cap, int ftype);
// both directions of a preconditioner, planned on two host threads
void plan_sweep_pair(const SweepRows& lo, const SweepRows& up, const GridInfo& g, int64_t cap, int ftype,
                     SweepPlan& pl, SweepPlan& pu);
// both directions planned, the lower uploaded while the upper is planned
void plan_upload_pair(const SweepRows& lo, const SweepRows& up, const GridInfo& g, int64_t cap, int ftype,
                      SweepDev& dl, SweepDev& du, const double* fstore,
                      const std::vector<int32_t>* lo_levels = nullptr);

}  // namespace m4g

struct m4g_pc_s {
    int kind = M4G_PC_IDENTITY;
    int factor_type = M4G_F64;
    int64_t n = 0;
    m4g::GridInfo grid;
    // factor storage in double: IC(0) -> upper(A) CSR; ILU(0) -> A's CSR
    m4g::PlanVec64 fS, fI;                // pattern of the storage
    m4g::PlanVec64 diag_pos;              // ILU(0): storage index of each diagonal
    m4g::DevBuf<double> fstore;
    m4g::DevBuf<int32_t> a_pos;           // IC(0): CSR index in A of each upper(A) entry
    m4g::DevBuf<double> inv_diag;         // diagonal PC
    m4g::SweepDev lower, upper;           // solve: lower then upper
    m4g::SweepDev adj_lower, adj_upper;   // ILU(0) adjoint_solve
    bool adj_ready = false;
    m4g::DevBuf<unsigned char> tmp;       // intermediate vector of a PC application
    size_t tmp_bytes = 0;
    int64_t factor_levels = 0;
    // IC(0) of a 2D 5-point grid: storage positions of each row's diagonal,
    // east and north entries (-1: absent), for the one-workgroup CG
    int grid5 = 0;                        // 0 not checked, 1 applies, -1 does not
    m4g::DevBuf<int32_t> grid5_pos;       // [3n]
    ~m4g_pc_s();
};

namespace m4g {
// y = P^{-1} x (or the adjoint) on stream s; vtype is the vectors' value type.
void pc_apply(m4g_pc_s& P, const void* x, void* y, int vtype, bool adjoint, hipStream_t s,
              const int* skip = nullptr);
// Reads and clears the fault flags of every sweep of P (nonzero: a bounded
// wait of a sweep launched since the last check gave up).
int check_pc_faults(m4g_pc_s& P);
// The IC(0) factor of a 2D 5-point grid of at most max_n rows and at most 64
// lines as the one-workgroup CG reads it: U's east and north entries (0 where
// absent) and the inverted diagonal per row.  False (nothing written) when P
// is not such a factor.
bool pc_grid_factor(m4g_pc_s& P, int64_t max_n, DevBuf<double>& east, DevBuf<double>& north, DevBuf<double>& dinv,
                    hipStream_t s);
}  // namespace m4g
