// Trace ranges through roctx (the ROCm tools' marker API), loaded on demand.
#include "trace.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>

namespace m4g {

namespace {

int threshold_from_env()
{
    const char* e = std::getenv("M4G_TRACE_LEVEL");
    if (!e || !*e) return INT_MAX;
    char* end = nullptr;
    long level = std::strtol(e, &end, 10);
    if (*end != '\0' || level < 0 || level > 9) {
        std::fprintf(stderr, "mtl4gpu: M4G_TRACE_LEVEL=%s is not a level 0..9; tracing stays off\n", e);
        return INT_MAX;
    }
    return int(level) * 1000;
}

using PushFn = int (*)(const char*);
using PopFn = int (*)();
PushFn g_push = nullptr;
PopFn g_pop = nullptr;
std::once_flag g_load_once;

// Loads roctx the first time a range opens.  Without it tracing turns off
// (with one message) and the library carries on.
void load_roctx()
{
    const char* libs[] = {"librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so", "libroctx64.so.4",
                          "libroctx64.so"};
    for (const char* name : libs) {
        void* h = dlopen(name, RTLD_NOW | RTLD_GLOBAL);
        if (!h) continue;
        auto push = reinterpret_cast<PushFn>(dlsym(h, "roctxRangePushA"));
        auto pop = reinterpret_cast<PopFn>(dlsym(h, "roctxRangePop"));
        if (push && pop) {
            g_push = push;
            g_pop = pop;
            return;
        }
    }
    std::fprintf(stderr, "mtl4gpu: M4G_TRACE_LEVEL is set but no roctx library loads; tracing is off\n");
    g_trace_threshold.store(INT_MAX, std::memory_order_relaxed);
}

}  // namespace

std::atomic<int> g_trace_threshold{threshold_from_env()};

const char* trace_name(int event)
{
    switch (event) {
    case kTraceVectorCopy: return "gen_vector_copy";
    case kTraceDot: return "dot";
    case kTraceVecExpr: return "dyn_vec_expr";
    case kTraceSetZero: return "Vector_set_zero";
    case kTraceTwoNorm: return "two_norm";
    case kTraceUnaryDot: return "unary_dot";
    case kTraceCrsMult: return "crs_cvec_mult";
    case kTraceIdSolve: return "pc::id.solve";
    case kTraceIdAdjoint: return "pc::id.adjoint_solve";
    case kTraceIc0Factorize: return "ic_0::factorize";
    case kTraceIc0Solve: return "ic_0::solve";
    case kTraceIlu0Factorize: return "ilu_0::factorize";
    case kTraceIlu0Solve: return "ilu_0::solve";
    case kTraceIlu0Adjoint: return "ilu_0::adjoint_solve";
    case kTraceIc0Adjoint: return "ic_0::adjoint_solve";
    case kTraceUpperTrisolve: return "upper_trisolve_crs_compact";
    case kTraceLowerTrisolve: return "lower_trisolve_crs_compact";
    case kTraceLowerUnitTrisolve: return "lower_unit_trisolve_crs_compact";
    case kTraceDiagSetup: return "diagonal::setup";
    case kTraceDiagSolve: return "diagonal::solve";
    case kTraceBatchIc0Factorize: return "batch::ic_0::factorize";
    case kTraceCgNoPc: return "cg_without_pc";
    case kTraceCg: return "cg";
    case kTraceBicg: return "bicg";
    case kTraceBicgstab: return "bicgstab";
    case kTraceBicgstab2: return "bicgstab_2";
    case kTraceBicgstabEll: return "bicgstab_ell";
    case kTraceCgs: return "cgs";
    case kTraceQmr: return "qmr";
    case kTraceTfqmr: return "tfqmr";
    case kTraceIdrS: return "idr_s";
    case kTraceGmres: return "gmres";
    case kTraceGmresFull: return "gmres_full";
    case kTraceRefinedSolve: return "refined_solve";
    case kTraceBatchCgIc0: return "batch::cg_ic0";
    case kTraceBatchAutotune: return "batch::autotune";
    default: return nullptr;
    }
}

void trace_push(int event)
{
    std::call_once(g_load_once, load_roctx);
    const char* name = trace_name(event);
    if (g_push && name) g_push(name);
    else if (g_push) g_push("unnamed");
}

void trace_pop()
{
    if (g_pop) g_pop();
}

}  // namespace m4g
