/* mtl4gpu_gen.h -- host-only generators of the synthetic inputs.
 *
 * m4g_gen_laplacian2d restates mtl::mat::laplacian_setup
 * (mtl/matrix/laplacian_setup.hpp:25-44).  The other stencils have no
 * reference counterpart; they define the benchmark matrices:
 *   - m4g_gen_poisson3d: 7-point Poisson, diagonal 6, neighbours -1,
 *     row = i + nx*(j + ny*k);
 *   - m4g_gen_convdiff27: 27-point convection-diffusion, per-cell kappa ~
 *     lognormal(0,1), harmonic face means, Dirichlet boundary, central
 *     convection with cell Peclet numbers pe*(1, 0.5, 0.25) on face neighbours;
 *   - m4g_gen_varcoef5: 2D 5-point variable-coefficient diffusion, per-cell
 *     kappa ~ lognormal(0,1), harmonic face means, Dirichlet boundary;
 *   - m4g_gen_edge_laplacian: 2D 5-point operator A(theta) = sum over edges
 *     theta_e (e_i - e_j)(e_i - e_j)^T plus boundary edges theta_b e_i e_i^T.
 * Every function fills CSR arrays with sorted rows (starts[n+1], indices[nnz],
 * values[nnz]); pass NULL arrays to query nnz only.
 * The RNG is splitmix64 on (seed, stream, index), so every value is
 * reproducible independently of evaluation order. */
#ifndef MTL4GPU_GEN_H
#define MTL4GPU_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int64_t m4g_gen_laplacian2d(int64_t m, int64_t n, int64_t* starts, int64_t* indices, double* values);
int64_t m4g_gen_poisson3d(int64_t nx, int64_t ny, int64_t nz, int64_t* starts, int64_t* indices, double* values);
int64_t m4g_gen_convdiff27(int64_t nx, int64_t ny, int64_t nz, uint64_t seed, double pe, int64_t* starts,
                           int64_t* indices, double* values);
int64_t m4g_gen_varcoef5(int64_t nx, int64_t ny, uint64_t seed, int64_t* starts, int64_t* indices, double* values);

/* Edge operator: n_edges = interior + boundary edges.  With theta == NULL only
 * the pattern is produced.  edge_i/edge_j (optional, [n_edges]) receive the
 * endpoints, edge_j = -1 for a boundary edge.  assembly_{ptr,idx,coef}
 * (optional) receive the CSR map entry -> (edge, +-1) with which
 * values[p] = sum coef * theta[edge]: assembly_ptr[nnz+1], assembly_idx and
 * assembly_coef of length 4 * n_edges at most.  Returns nnz; *n_edges set. */
int64_t m4g_gen_edge_laplacian(int64_t nx, int64_t ny, const double* theta, int64_t* n_edges, int64_t* starts,
                               int64_t* indices, double* values, int64_t* edge_i, int64_t* edge_j,
                               int64_t* assembly_ptr, int64_t* assembly_idx, double* assembly_coef);

/* Random numbers: out[k] for k in [0, count) from (seed, stream, offset + k). */
void m4g_rng_uniform(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, double lo, double hi, double* out);
void m4g_rng_normal(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, double* out);
void m4g_rng_lognormal(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, double* out);
/* Integers uniform in [lo, hi]. */
void m4g_rng_int(uint64_t seed, uint64_t stream, int64_t offset, int64_t count, int64_t lo, int64_t hi, int64_t* out);

#ifdef __cplusplus
}
#endif

#endif /* MTL4GPU_GEN_H */
