// Host-side trace ranges around the library's operations, for rocprofv3's
// marker trace (rocprofv3 --marker-trace -- <program>).
//
// The events are the reference's Vampir trace events: the same numbers and
// names (boost/numeric/mtl/interface/vpt.cpp:95-368), opened where the
// reference opens them (a `vampir_trace<N> tracer;` at the top of dot,
// two_norm, mat_cvec_mult, the preconditioners' factorize / solve /
// adjoint_solve, the triangular solves and every Krylov solver).  The level
// rule is the reference's too (vpt.hpp:36-38): event N is recorded when
// level * 1000 < N, with the level from the environment variable
// M4G_TRACE_LEVEL in place of the compile-time MTL_VPT_LEVEL.  Unset (the
// default), nothing is recorded, the tracing library is never loaded, and a
// range costs one compare.
//
// The ranges are host ranges, as Vampir's are: they span the host work and the
// launches of an operation; kernels that run on after the range closes show in
// the kernel trace.
#pragma once

#include <atomic>
#include <climits>

namespace m4g {

// Events with a number above this are recorded (INT_MAX: tracing off).
extern std::atomic<int> g_trace_threshold;

// Events of the reference's table used here (vpt.cpp), and the library's own
// for operations the table has no entry for (numbers in x900 ranges, which the
// table leaves free).
enum TraceEvent : int {
    kTraceVectorCopy = 2001,        // gen_vector_copy
    kTraceDot = 2003,               // dot
    kTraceVecExpr = 2017,           // dyn_vec_expr
    kTraceSetZero = 2029,           // Vector_set_zero
    kTraceTwoNorm = 2039,           // two_norm
    kTraceUnaryDot = 2041,          // unary_dot
    kTraceCrsMult = 3049,           // crs_cvec_mult
    kTraceIdSolve = 5032,           // pc::id.solve
    kTraceIdAdjoint = 5034,         // pc::id.adjoint_solve
    kTraceIc0Factorize = 5035,      // ic_0::factorize
    kTraceIc0Solve = 5036,          // ic_0::solve
    kTraceIlu0Factorize = 5038,     // ilu_0::factorize
    kTraceIlu0Solve = 5039,         // ilu_0::solve
    kTraceIlu0Adjoint = 5040,       // ilu_0::adjoint_solve
    kTraceIc0Adjoint = 5044,        // ic_0::adjoint_solve
    kTraceUpperTrisolve = 5046,     // upper_trisolve_crs_compact
    kTraceLowerTrisolve = 5047,     // lower_trisolve_crs_compact
    kTraceLowerUnitTrisolve = 5048, // lower_unit_trisolve_crs_compact
    kTraceDiagSetup = 5050,         // diagonal::setup
    kTraceDiagSolve = 5051,         // diagonal::solve
    kTraceBatchIc0Factorize = 5901, // batch::ic_0::factorize (library's own)
    kTraceCgNoPc = 7001,            // cg_without_pc
    kTraceCg = 7002,                // cg
    kTraceBicg = 7003,              // bicg
    kTraceBicgstab = 7004,          // bicgstab
    kTraceBicgstab2 = 7005,         // bicgstab_2
    kTraceBicgstabEll = 7006,       // bicgstab_ell
    kTraceCgs = 7007,               // cgs
    kTraceQmr = 7008,               // qmr
    kTraceTfqmr = 7009,             // tfqmr
    kTraceIdrS = 7010,              // idr_s
    kTraceGmres = 7901,             // gmres (library's own)
    kTraceGmresFull = 7902,         // gmres_full (library's own)
    kTraceRefinedSolve = 7903,      // refined_solve (library's own)
    kTraceBatchCgIc0 = 7904,        // batch::cg_ic0 (library's own)
    kTraceBatchAutotune = 7905,     // batch::autotune (library's own)
};

// Name of an event (nullptr for a number not in the table).
const char* trace_name(int event);
void trace_push(int event);
void trace_pop();

// Opens the event's range for the lifetime of the object when the level
// records it.
class TraceRange {
  public:
    explicit TraceRange(int event) : on_(event > g_trace_threshold.load(std::memory_order_relaxed))
    {
        if (on_) trace_push(event);
    }
    ~TraceRange()
    {
        if (on_) trace_pop();
    }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;

  private:
    bool on_;
};

}  // namespace m4g
