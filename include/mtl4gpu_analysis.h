/* mtl4gpu_analysis.h -- host-only analysis entry points of libmtl4gpu.
 *
 * These functions need no GPU.  They expose the schedule compiler of the
 * triangular sweeps so that its output can be checked on the CPU: they build
 * the tile-pipelined plans exactly as m4g_pc_create does, validate them, and
 * execute them on the host with the kernel's operation order.  No reference
 * counterpart: the reference runs its sweeps sequentially
 * (lower_trisolve.hpp:172-209, upper_trisolve.hpp:159-176).
 */
#ifndef MTL4GPU_ANALYSIS_H
#define MTL4GPU_ANALYSIS_H

#include <stdint.h>
#include "mtl4gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Infers the natural-order grid (nx, ny, nz) of a CSR pattern; zeros if none. */
int m4g_analysis_grid(int64_t n, const int64_t* starts, const int64_t* indices, int64_t* nx, int64_t* ny,
                      int64_t* nz);

/* Emulates P.solve(rhs) (adjoint = 0) or P.adjoint_solve(rhs) (adjoint = 1) of
 * an IC(0) (kind = M4G_PC_IC0) or ILU(0) (kind = M4G_PC_ILU0) preconditioner.
 *   starts/indices: pattern of A (rows sorted);
 *   factor: the factor storage in double -- IC(0): upper(A) in CSR with the
 *           inverted diagonal first in each row; ILU(0): A's CSR holding L
 *           (strict lower) and U (upper, inverted diagonal);
 *   factor_value_type: M4G_F64, or M4G_F32 to round factor values to float;
 *   vector_value_type: M4G_F64 (rhs/out are double) or M4G_F32 (float);
 *   max_tile_rows: 0 for the library default.
 * Returns M4G_ERR_LOGIC (message in m4g_last_error) when a plan is invalid. */
int m4g_analysis_pc_solve(int kind, int adjoint, int64_t n, const int64_t* starts, const int64_t* indices,
                          const double* factor, int factor_value_type, int vector_value_type,
                          const void* rhs, void* out, int64_t max_tile_rows, m4g_pc_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* MTL4GPU_ANALYSIS_H */
