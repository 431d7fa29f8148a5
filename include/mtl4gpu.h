/* mtl4gpu.h -- C ABI of the MI355X (gfx950) sparse iterative-solve library.
 *
 * The library replaces MTL4/ITL's CPU path for CSR SpMV on compressed2D,
 * IC(0)/ILU(0) factorisation and their triangular solves, and the vector
 * kernels inside itl::cg and itl::bicgstab.  Every entry point below names the
 * reference interface it stands in for (paths relative to the reference
 * checkout; mtl/ = boost/numeric/mtl, itl/ = boost/numeric/itl).
 *
 * Conventions
 *   - Every function returns an int status (M4G_OK = 0).  Status values that
 *     mirror the reference: 1 = "Too many iterations." (basic_iteration.hpp:43-48);
 *     M4G_ERR_NOT_SQUARE, M4G_ERR_MISSING_DIAGONAL, M4G_ERR_ORTHOGONALITY and
 *     M4G_ERR_LOGIC mirror mtl::matrix_not_square, mtl::missing_diagonal,
 *     itl::unexpected_orthogonality and mtl::logic_error.  Negative values are
 *     runtime failures.  m4g_last_error() returns a message for the calling
 *     thread's last failure.
 *   - Handles own their device memory.  Host pointers are borrowed for the
 *     duration of the call only.  m4g_vector_wrap borrows device memory.
 *   - All work is queued on the library stream of the current device (or on
 *     the stream given to m4g_set_stream).  Calls returning host scalars, and
 *     downloads, synchronise that stream.
 *   - One preconditioner handle must not be applied from two streams at once
 *     (its sweep exchange buffers are per handle, as the reference's PC
 *     solves share a static temporary, ic_0.hpp:72).
 */
#ifndef MTL4GPU_H
#define MTL4GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
    M4G_OK = 0,
    M4G_MAX_ITER = 1,               /* basic_iteration::check_max, basic_iteration.hpp:43-48 */
    M4G_ERR_NOT_SQUARE = 2,         /* mtl::matrix_not_square (ic_0.hpp:147, ilu_0.hpp:49) */
    M4G_ERR_MISSING_DIAGONAL = 3,   /* mtl::missing_diagonal (lower_trisolve.hpp:198-200) */
    M4G_ERR_ORTHOGONALITY = 4,      /* itl::unexpected_orthogonality (bicgstab.hpp:46,51,59) */
    M4G_ERR_LOGIC = 5,              /* mtl::logic_error (unsorted rows, dense input) */
    M4G_ERR_SIZE = 6,               /* mtl::incompatible_size */
    M4G_ERR_INVALID = 7,            /* null handle, bad enum, bad argument */
    M4G_ERR_INTERRUPTED = 8,        /* interruptible_iteration flag set (interruptible_iteration.hpp:37-40) */
    M4G_ERR_FILE = 9,               /* mtl::file_not_found (matrix_market.hpp:50-60, exception.hpp:232-236) */
    M4G_ERR_FORMAT = 10,            /* mtl::runtime_error from the Matrix Market reader (matrix_market.hpp:172-215) */
    M4G_BREAKDOWN = 11,             /* iter.fail(k >= 2, "... breakdown"): the reference returns k (cgs.hpp:40, bicg.hpp:44, qmr.hpp:51-95, gmres.hpp:114, bicgstab_ell.hpp:112); it->error holds k */
    M4G_ERR_SINGULAR = 12,          /* mtl::matrix_singular from lu_solve (lu.hpp:83; idr_s.hpp:78) */
    M4G_ERR_HIP = -1,               /* HIP runtime error */
    M4G_ERR_NO_DEVICE = -2,         /* no GPU visible */
    M4G_ERR_TIMEOUT = -3            /* a bounded device-side wait ran out (schedule fault) */
};

/* value types (Collection<T>::value_type) */
enum { M4G_F64 = 0, M4G_F32 = 1 };
/* index types of host CSR arrays: int32, int64 (= std::size_t, parameter.hpp:33), uint32 (TIM/ timing programs) */
enum { M4G_IDX_I32 = 0, M4G_IDX_I64 = 1, M4G_IDX_U32 = 2 };
/* assign modes of w = A*v, w += A*v, w -= A*v (mtl/operation/assign_mode.hpp:22-110) */
enum { M4G_ASSIGN = 0, M4G_PLUS = 1, M4G_MINUS = 2 };
/* preconditioner kinds: itl::pc::identity, diagonal, ic_0, ilu_0 */
enum { M4G_PC_IDENTITY = 0, M4G_PC_DIAGONAL = 1, M4G_PC_IC0 = 2, M4G_PC_ILU0 = 3 };
/* Krylov methods for the refinement / batched entry points */
enum { M4G_METHOD_CG = 0, M4G_METHOD_BICGSTAB = 1 };

typedef struct m4g_matrix_s* m4g_matrix;   /* device-resident mtl::compressed2D */
typedef struct m4g_vector_s* m4g_vector;   /* device-resident mtl::dense_vector */
typedef struct m4g_pc_s*     m4g_pc;       /* device-resident itl::pc::* */

/* ---- library state -------------------------------------------------------- */
const char* m4g_last_error(void);
/* Selects the HIP device for this thread's subsequent calls (default 0). */
int m4g_set_device(int device);
/* Queues subsequent work of this thread on the given hipStream_t (NULL = library stream). */
int m4g_set_stream(void* hip_stream);
int m4g_synchronize(void);
/* The hipStream_t the calling thread's work is queued on (for timing events). */
void* m4g_get_stream(void);
/* Library build string (arch, flags). */
const char* m4g_version(void);

/* ---- matrices: mtl::compressed2D (mtl/matrix/compressed2D.hpp:258-653) ------ */
/* compressed2D(m, n, nnz, starts, indices, data): copies the three host arrays
 * (compressed2D.hpp:355-364).  Rows must be sorted by column without
 * duplicates, as the inserter leaves them (compressed2D.hpp:920-986); an
 * unsorted row returns M4G_ERR_LOGIC.  values: double or float (value_type). */
int m4g_matrix_create(m4g_matrix* A, int64_t nrows, int64_t ncols, int64_t nnz,
                      const void* starts, const void* indices, const void* values,
                      int index_type, int value_type);
int m4g_matrix_destroy(m4g_matrix A);
/* num_rows, num_cols, nnz */
int m4g_matrix_dims(m4g_matrix A, int64_t* nrows, int64_t* ncols, int64_t* nnz);
/* ref_major / ref_minor / data copies (compressed2D.hpp:610-625, 648), int64 indices. */
int m4g_matrix_download(m4g_matrix A, int64_t* starts, int64_t* indices, double* values);
/* Replaces the values (same pattern, CSR order); values_device != 0 means a device pointer. */
int m4g_matrix_set_values(m4g_matrix A, const double* values, int values_device);

/* ---- vectors: mtl::dense_vector (mtl/vector/dense_vector.hpp:51-270) -------- */
int m4g_vector_create(m4g_vector* v, int64_t n, int value_type);
/* Borrows n elements of device memory at device_ptr (not freed by destroy). */
int m4g_vector_wrap(m4g_vector* v, void* device_ptr, int64_t n, int value_type);
int m4g_vector_destroy(m4g_vector v);
int m4g_vector_size(m4g_vector v, int64_t* n);
void* m4g_vector_device_ptr(m4g_vector v);
int m4g_vector_upload(m4g_vector v, const void* host);     /* host array of the vector's type */
int m4g_vector_download(m4g_vector v, void* host);
int m4g_vector_fill(m4g_vector v, double value);          /* v = value */
int m4g_vector_copy(m4g_vector dst, m4g_vector src);      /* dst = src, converting precision */

/* ---- level-1 kernels ------------------------------------------------------ */
/* dot(x, y) = sum x_i * y_i (mtl/operation/dot.hpp:62-256) */
int m4g_dot(m4g_vector x, m4g_vector y, double* result);
/* dot(x, y) as m4g_dot, for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  It calls no hipMalloc,
 * hipFree or other allocating routine: the reduction partials and the result
 * live in a file-scope device array of 1056 doubles (8448 bytes) per GPU,
 * held until the process ends; its address is looked up when the device is
 * set up (m4g_set_device or the first handle made on it), and a device that
 * is not set up is an error.  Any n fits (the reduction uses at most 1024
 * partials); a size beyond the scratch is rejected with M4G_ERR_SIZE.  The
 * result is bit-identical to m4g_dot's.  Every call on a GPU shares that
 * array, so calls that overlap on one GPU (several host threads or streams)
 * race. */
int m4g_dot_noalloc(m4g_vector x, m4g_vector y, double* result);
/* two_norm(x) (mtl/operation/two_norm.hpp:33-61) */
int m4g_two_norm(m4g_vector x, double* result);
/* unary_dot(x) = sum x_i^2 (mtl/operation/unary_dot.hpp:29-65) */
int m4g_unary_dot(m4g_vector x, double* result);
/* y += alpha * x (vec_vec_aop_expr.hpp, e.g. cg.hpp:107) */
int m4g_axpy(double alpha, m4g_vector x, m4g_vector y);
/* y = x + beta * y (cg.hpp:102) */
int m4g_xpay(m4g_vector x, double beta, m4g_vector y);
/* (lazy(r) -= alpha*q) || (lazy(rr) = lazy_unary_dot(r)) (cg.hpp:109) */
int m4g_axpy_unary_dot(double alpha, m4g_vector q, m4g_vector r, double* rr);

/* ---- SpMV: smat_cvec_mult (mtl/operation/mat_vec_mult.hpp:619-689) ---------- */
/* w = A*v, w += A*v, w -= A*v (crtp_base_vector.hpp:93-104, 192-198, 269-275) */
int m4g_spmv(m4g_matrix A, m4g_vector v, m4g_vector w, int mode);
/* r = b - A*x (compute_summand.hpp:42-60 then element-wise minus; cg.hpp:93) */
int m4g_residual(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_vector r);
/* (lazy(q) = A*p) || (lazy(pq) = lazy_dot(p, q)) (cg.hpp:104; fused_expr.hpp:42-91) */
int m4g_spmv_dot(m4g_matrix A, m4g_vector p, m4g_vector q, double* pq);

/* ---- preconditioners (itl/pc/) ---------------------------------------------- */
/* ic_0<Matrix, Value>(A) (ic_0.hpp:39-194), ilu_0<Matrix, Value>(A) (ilu_0.hpp:85-92,
 * ilu.hpp:33-133), diagonal<Matrix> (diagonal.hpp:28-103), identity (identity.hpp:24-71).
 * factor_value_type is Value: M4G_F64, or M4G_F32 to store the factors in float
 * after factorising in double (ic_0.hpp:122-130; ilu_0.hpp:80-81). */
int m4g_pc_create(m4g_pc* P, m4g_matrix A, int kind, int factor_value_type);
int m4g_pc_destroy(m4g_pc P);
/* P.solve(x, y) / solve(P, x) (ic_0.hpp:78-87, 257-262; ilu.hpp:83-91, 189-195) */
int m4g_pc_solve(m4g_pc P, m4g_vector x, m4g_vector y);
/* P.adjoint_solve(x, y) (ic_0.hpp:97-103; ilu.hpp:105-115) */
int m4g_pc_adjoint_solve(m4g_pc P, m4g_vector x, m4g_vector y);
/* Queues P.solve(x, y) without waiting or checking the sweeps' wait
 * counters (for timing loops); the next m4g_pc_solve reports any fault. */
int m4g_pc_apply_async(m4g_pc P, m4g_vector x, m4g_vector y);
/* get_L() / get_U() (ic_0.hpp:106-107; ilu.hpp:118-119): which = 0 for L, 1 for U.
 * IC(0): L = trans(U), inverted diagonal last in each row; U: inverted diagonal first.
 * ILU(0): L strict lower (unit diagonal implied); U with inverted diagonal first. */
int m4g_pc_factor_nnz(m4g_pc P, int which, int64_t* nnz);
int m4g_pc_download_factor(m4g_pc P, int which, int64_t* starts, int64_t* indices, double* values);

/* Schedule statistics of a factor's sweeps (no reference counterpart; diagnostics). */
typedef struct {
    int64_t levels_lower, levels_upper;   /* natural-order DAG depth of each sweep */
    int64_t tiles_lower, tiles_upper;     /* workgroup tiles of the pipelined sweeps */
    int64_t steps_lower, steps_upper;     /* max in-tile levels over tiles */
    int64_t grid_nx, grid_ny, grid_nz;    /* inferred grid (0 if none) */
    int64_t tile_nx, tile_ny, tile_nz;    /* tile box (0 for row-block tiles) */
} m4g_pc_stats;
int m4g_pc_stats_get(m4g_pc P, m4g_pc_stats* stats);

/* ---- iteration control: itl::basic_iteration (basic_iteration.hpp:24-165) ---- */
typedef struct {
    int max_iter;          /* in */
    double rtol;           /* in: relative tolerance */
    double atol;           /* in: absolute tolerance */
    double norm_r0;        /* in: basic_iteration(r0, ...) computes |two_norm(r0)|; the tests pass b */
    int iterations;        /* in/out: counter i (kept across calls, like the iteration object) */
    int error;             /* in/out: error code; like basic_iteration's member it is only ever set */
    double resid;          /* out: last residual given to finished() */
    double* history;       /* optional host array: history[i] = resid at the check of iteration i */
    int history_capacity;
    const volatile int* interrupt; /* optional host flag polled between iteration batches */
} m4g_iteration;

/* itl::cg(A, x, b, P, iter) (cg.hpp:75-112; identity PC -> cg.hpp:39-72).
 * x is in/out.  Returns the iteration's error code (0 / 1) or an error status. */
int m4g_cg(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* iter);
/* itl::bicgstab(A, x, b, M, iter) (bicgstab.hpp:27-78). */
int m4g_bicgstab(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* iter);

/* itl::cgs(A, x, b, M, iter) (krylov/cgs.hpp:24-62).  Returns the iteration's
 * error: 0, 1 (too many iterations) or M4G_BREAKDOWN when rho_1 was 0 (the
 * reference records error 2 and continues; it->error holds 2).  Scalars are
 * computed on the host from device reductions (one read-back per dot). */
int m4g_cgs(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc M, m4g_iteration* it);
/* itl::bicg(A, x, b, M, iter) (krylov/bicg.hpp:27-68): solve and
 * adjoint_solve of M, A * p and adjoint(A) * p_tilde (an explicit transpose).
 * M4G_BREAKDOWN (it->error = 2) when rho_1 = dot(z_tilde, z) is 0. */
int m4g_bicg(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc M, m4g_iteration* it);
/* itl::tfqmr(A, x, b, L, R, iter) (krylov/tfqmr.hpp:32-99) with left
 * preconditioner L and right preconditioner R (NULL: identity); breakdowns
 * (sigma = 0, rho = 0) return error 1 as the reference does. */
int m4g_tfqmr(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it);
/* itl::qmr(A, x, b, L, R, iter) (krylov/qmr.hpp:29-109); R may be NULL
 * (identity).  Like the reference, the loop tests rho = two_norm(solve(L,
 * v_tilde)), the Lanczos vector's norm; breakdowns #1 (rho = 0) returns 1,
 * #2-#6 return M4G_BREAKDOWN with it->error = 2..6. */
int m4g_qmr(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it);
/* itl::gmres_full(A, x, b, L, R, iter) (krylov/gmres.hpp:37-119): min(n,
 * max_iter - iterations) Arnoldi steps, then one correction of x; the
 * Hessenberg matrix's rotations are applied as the reference applies them
 * (host side, with the reference's lu_solve).  M4G_BREAKDOWN (error 2) when no
 * leading block of H is regular. */
int m4g_gmres_full(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it);
/* itl::gmres(A, x, b, L, R, iter, restart) (krylov/gmres.hpp:122-137):
 * gmres_full on a copy of the iteration with `restart` more steps, then
 * update_progress, until finished; restart >= 1. */
int m4g_gmres(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it, int64_t restart);
/* itl::bicgstab_2(A, x, b, L, iter) (krylov/bicgstab_2.hpp:35-130). */
int m4g_bicgstab_2(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_iteration* it);
/* itl::bicgstab_ell(A, x, b, L, R, iter, l) (krylov/bicgstab_ell.hpp:37-135),
 * 1 <= l <= 64; omega = 0 returns M4G_BREAKDOWN with it->error = 3.  As in
 * the reference, the inner test checks r_hat[j] and leaves before that step's
 * x += alpha * u_hat[0]. */
int m4g_bicgstab_ell(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it, int64_t l);
/* itl::idr_s(A, x, b, L, R, iter, s) (krylov/idr_s.hpp:35-103), s <= 64; L and
 * R are ignored, as in the reference.  The shadow space P is random(P): column
 * j is filled by srand(17 * ++counter) and rand() (mtl::seed<double>,
 * random.hpp:29-38), so results depend on the C library generator and on the
 * process-wide counter below.  A singular s x s system returns
 * M4G_ERR_SINGULAR. */
int m4g_idr_s(m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc L, m4g_pc R, m4g_iteration* it, int64_t s);
/* mtl::seed<double>::counter (random.hpp:29-38): read / set the counter that
 * idr_s's random(P) increments once per column. */
int m4g_random_seed_counter(void);
void m4g_set_random_seed_counter(int counter);

/* ---- assembly on the device (compressed2D_inserter semantics) -------------- */
/* Updaters of mtl/operation/update.hpp:20-115: the first value of a position
 * stores init(v) (v; -v for minus), each later one applies the operation. */
enum { M4G_UPDATE_STORE = 0, M4G_UPDATE_PLUS = 1, M4G_UPDATE_MINUS = 2, M4G_UPDATE_TIMES = 3 };
/* mat::inserter<compressed2D, Updater> (compressed2D.hpp:662-1151) fed with
 * `count` triplets (rows[i], cols[i], values[i]) in insertion order, built on
 * the device: entries of one position are folded in insertion order; with
 * `existing` != NULL its entries are present before the first insertion (the
 * inserter keeps a matrix's entries), and *out is a new matrix (existing is
 * left unchanged).  on_device = 1: the three arrays are device pointers.
 * Rows come out sorted; an index out of range returns M4G_ERR_INVALID. */
int m4g_matrix_assemble(m4g_matrix* out, int64_t nrows, int64_t ncols, int64_t count, const int64_t* rows,
                        const int64_t* cols, const double* values, int on_device, int updater, m4g_matrix existing);

/* ---- differentiation of x = A(theta)^{-1} b --------------------------------- */
/* No reference counterpart.  The adjoint of the solve is A^T lambda = dL/dx,
 * which runs through the same Krylov kernels on the transposed matrix (or on A
 * itself when it is symmetric); the gradient on the stored pattern is then
 * dL/dA_ij = -lambda_i * x_j. */
/* Re-runs the numeric factorisation of P for A's current values (same pattern
 * as at creation); the sweep schedules are reused. */
int m4g_pc_refactor(m4g_pc P, m4g_matrix A);
/* AT = trans(A) as a new device matrix (rows sorted). */
int m4g_matrix_transpose(m4g_matrix A, m4g_matrix* AT);
/* grad[p] = -lambda[row(p)] * x[col(p)] for every stored entry p of A (CSR order);
 * grad is a vector of length nnz(A). */
int m4g_pattern_gradient(m4g_matrix A, m4g_vector lambda, m4g_vector x, m4g_vector grad);

/* ---- mixed precision: FP32 inner solves inside FP64 iterative refinement ---- */
/* No reference counterpart (the reference's "mixed precision" is FP32 factor
 * storage with FP64 arithmetic, ic_0.hpp:122-130, which m4g_pc_create with
 * M4G_F32 factors and double vectors reproduces).  Here the matrix values, the
 * factors of P (created with M4G_F32) and the inner Krylov vectors are FP32
 * with FP32 accumulation; FP64 is used only for r = b - A*x, the update
 * x += d and the convergence norms:
 *     repeat: r = b - A x (FP64); finished(|r|)?; solve A d = r in FP32 with
 *             `method` + P to inner_rtol relative to |r|; x += d (FP64)
 * `outer` is the iteration control of the refinement loop (its history holds
 * |b - A x| per refinement step). */
typedef struct {
    int outer_iterations;   /* refinement steps taken */
    int inner_iterations;   /* FP32 Krylov iterations, summed */
    double resid;           /* final |b - A x| in FP64 */
    int inner_error;        /* last inner solve's error code */
} m4g_refine_info;
int m4g_refined_solve(int method, m4g_matrix A, m4g_vector x, m4g_vector b, m4g_pc P, m4g_iteration* outer,
                      double inner_rtol, int inner_max_iter, m4g_refine_info* info);

/* ---- batched IC(0)-PCG ----------------------------------------------------
 * Many independent SPD systems with nrhs right-hand sides each: every
 * (system, rhs) pair runs itl::cg with itl::pc::ic_0 (cg.hpp:75-112,
 * ic_0.hpp:135-186) and basic_iteration(b, max_iter, rtol, atol)
 * (basic_iteration.hpp:24-92).  No single reference call does this; the
 * reference's equivalent is a loop over cg() calls, one per system and rhs.
 * Systems are given as concatenated CSR: sizes[s] rows each, starts = each
 * system's n_s + 1 row pointers one after another, indices (local columns,
 * sorted per row) and values concatenated. */
typedef struct m4g_batch_s* m4g_batch;
int m4g_batch_create(m4g_batch* B, int64_t nsys, const int64_t* sizes, const int64_t* starts,
                     const int64_t* indices, const double* values);
int m4g_batch_destroy(m4g_batch B);
int m4g_batch_dims(m4g_batch B, int64_t* nsys, int64_t* n_total, int64_t* nnz_total);
/* IC(0) factors of every system in one launch (the reference's per-row order) */
int m4g_batch_factor_ic0(m4g_batch B);
/* factor storage: each system's upper(A) rows, inverted diagonal first (ic_0.hpp:148-185) */
int m4g_batch_download_factor(m4g_batch B, double* values);
/* One persistent launch solving every (system, rhs).  x (in/out) and b hold
 * n_total * nrhs doubles; entry (system s, row i, rhs j) is at
 * (row0(s) + i) * nrhs + j with row0 the prefix sum of sizes.  rhs_per_item[s]
 * (or NULL: 8) right-hand sides of system s share one workgroup; block is the
 * workgroup size (256, 512, 1024; 0 = 512).  Per (s, j) at s * nrhs + j:
 * iterations, final residual norm, error (0 converged, 1 too many iterations). */
int m4g_batch_cg_ic0(m4g_batch B, int nrhs, m4g_vector x, m4g_vector b, int max_iter, double rtol,
                     double atol, const int* rhs_per_item, int block, int* iterations, double* resid, int* error);
/* Autotuner: times capped runs of the whole batch (system s runs
 * trial_iters * sqrt(n_s / n_max) iterations, trial_iters 0: 40), first
 * for every block size (256, 512, 1024) with one rhs grouping (8, 4, 2, 1
 * dividing nrhs) for all systems, then, with the fastest block size, lets each
 * third of the systems by size try the other groupings.  Returns the block
 * size and each system's grouping.  trial_ms (optional, up to 24 entries)
 * receives the trial times: 3 * groupings of the first stage, then 3 *
 * groupings of the second.  Answers are cached, keyed by the size classes
 * (count, smallest and largest n of each third), nrhs, trial_iters, the packed
 * sweeps' width and the device: in the process and in a JSON Lines file
 * ($M4G_AUTOTUNE_CACHE, else $XDG_CACHE_HOME or $HOME/.cache, under
 * mtl4gpu/batch_autotune.jsonl; M4G_AUTOTUNE_CACHE=off: process only).  A
 * cached answer runs no trials and leaves trial_ms untouched. */
int m4g_batch_autotune(m4g_batch B, int nrhs, m4g_vector b, int trial_iters, int* rhs_per_item, int* block,
                       double* trial_ms);

#ifdef __cplusplus
}
#endif

#endif /* MTL4GPU_H */
