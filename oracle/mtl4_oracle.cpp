// CPU oracle for the sparse iterative-solve path of MTL4/ITL.
//
// TEST INFRASTRUCTURE ONLY.  This file restates, loop by loop, the reference
// algorithms on the path (CSR SpMV, dot/two_norm/unary_dot, the fused CG
// loops, IC(0) and ILU(0) factorisation, the compact-CSR triangular sweeps,
// CG, BiCGStab and the iteration control).  It is linked only by tests/,
// by __graft_entry__.smoke() as a checker and by bench.py's cpu_baseline
// leg.  The product (the HIP library) never calls it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference checkout; mtl/ = boost/numeric/mtl, itl/ = boost/numeric/itl).
// Summation orders, operand orders and reciprocal-then-multiply sequences are
// kept exactly, so results are bit-identical to an MTL4 build without FMA
// contraction (build this file with -ffp-contract=off).
//
// Index type: int64_t, standing in for MTL4's default std::size_t
// (mtl/matrix/parameter.hpp:33).

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

typedef int64_t idx_t;

enum {
    ORC_OK = 0,
    ORC_MAX_ITER = 1,                  // basic_iteration.hpp:43-48
    ORC_ERR_NOT_SQUARE = 2,            // mtl::matrix_not_square
    ORC_ERR_MISSING_DIAGONAL = 3,      // mtl::missing_diagonal
    ORC_ERR_ORTHOGONALITY = 4,         // itl::unexpected_orthogonality
};

// ---------------------------------------------------------------------------
// Iteration control: itl/iteration/basic_iteration.hpp:24-165
// ---------------------------------------------------------------------------
struct orc_iteration {
    int max_iter;
    double rtol, atol, norm_r0;
    int i, error, is_finished;
    double resid;
    double* history;      // optional: resid at every finished() call, indexed by i
    int history_cap;
};

static bool orc_converged(orc_iteration* it, double r)
{
    it->resid = r;                                   // basic_iteration.hpp:83
    if (it->history && it->i < it->history_cap) it->history[it->i] = r;
    if (it->norm_r0 == 0)                            // :90-91
        return r <= it->atol;
    return r <= it->rtol * it->norm_r0 || r <= it->atol;   // :92
}

static bool orc_finished(orc_iteration* it, double r)
{
    if (orc_converged(it, r)) return (it->is_finished = 1);  // :60-64
    if (it->i >= it->max_iter) {                           // check_max :43-48
        it->error = ORC_MAX_ITER;
        it->is_finished = 1;
    }
    return it->is_finished;
}

// ---------------------------------------------------------------------------
// Reductions
// ---------------------------------------------------------------------------

// dot(v1, v2) -> sfunctor::dot<4>: mtl/operation/dot.hpp:136-147 with the
// index distribution of dot_aux (mm/loop1.hpp:21-31): element i+k goes to
// partial k; the tail goes to partial 0; combine ((t0+t1)+(t2+t3)) + 0.
static double dot4(idx_t n, const double* a, const double* b)
{
    double t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0, t5 = 0, t6 = 0, t7 = 0;
    idx_t nb = 4 * (n / 4);
    for (idx_t i = 0; i < nb; i += 4) {
        t0 += a[i] * b[i];
        t1 += a[i + 1] * b[i + 1];
        t2 += a[i + 2] * b[i + 2];
        t3 += a[i + 3] * b[i + 3];
    }
    for (idx_t i = nb; i < n; i++) t0 += a[i] * b[i];
    return ((t0 + t1) + (t2 + t3)) + ((t4 + t5) + (t6 + t7));
}

// reduction<U, two_norm_functor>: mtl/vector/reduction.hpp:154-173; the
// finish tree of impl::reduction (:51-57) gives t0 + (t1 + (t2 + ...)).
template <int U>
static double sq_reduction(idx_t n, const double* v)
{
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    idx_t nb = U * (n / U);
    for (idx_t i = 0; i < nb; i += U)
        for (int k = 0; k < U; k++) t[k] += v[i + k] * v[i + k];   // squared_abs
    for (idx_t i = nb; i < n; i++) t[0] += v[i] * v[i];
    double acc = t[U - 1];
    for (int k = U - 2; k >= 0; k--) acc = t[k] + acc;   // t[k] += t[k+1] chain
    return acc;
}

// two_norm: mtl/operation/two_norm.hpp:52-57 (reduction<4>, then sqrt).
static double two_norm(idx_t n, const double* v) { return std::sqrt(sq_reduction<4>(n, v)); }
// unary_dot: mtl/operation/unary_dot.hpp:56-59 (reduction<8>).
static double unary_dot(idx_t n, const double* v) { return sq_reduction<8>(n, v); }

// ---------------------------------------------------------------------------
// SpMV: mtl/operation/mat_vec_mult.hpp:619-689 (row-major compressed2D).
// Each row is summed sequentially in stored order starting from zero; the
// 4-row blocking only interleaves independent rows, so it does not change any
// result.  Assign::first_update: '=' stores, '+=' adds, '-=' subtracts
// (mtl/operation/assign_mode.hpp).  The very-sparse / adaptive variants
// (:421-449, :551-617) compute the same per-row sums.
// ---------------------------------------------------------------------------
template <typename MV>
static inline double row_sum(const idx_t* S, const idx_t* I, const MV* D, const double* v, idx_t r)
{
    double tmp = 0;
    for (idx_t j = S[r]; j != S[r + 1]; ++j) tmp += D[j] * v[I[j]];
    return tmp;
}

// ---------------------------------------------------------------------------
// IC(0): itl/pc/ic_0.hpp:135-186.  U = upper(A); for each row k:
// U_kk <- 1/sqrt(U_kk); for each i > k in row k: U_ki <- U_ki * U_kk and
// U_ii <- U_ii - U_ki * U[k][i] (the only j in (k, i] of row i is j = i).
// ---------------------------------------------------------------------------
static idx_t lookup(const idx_t* S, const idx_t* I, idx_t r, idx_t c)
{
    // compressed2D::operator() finds (r, c) by search; rows are sorted.
    // Returns the storage index or -1 (value 0), compressed2D.hpp:181-204, 536-547.
    const idx_t* b = I + S[r];
    const idx_t* e = I + S[r + 1];
    const idx_t* p = std::lower_bound(b, e, c);
    if (p != e && *p == c) return p - I;
    return -1;
}

extern "C" {

// ------------------------------- generators --------------------------------

// mtl/matrix/laplacian_setup.hpp:25-44 on an m x n grid, row = i*n + j.
// The inserter keeps each row sorted by column (compressed2D.hpp:945-986).
idx_t orc_laplacian_nnz(int m, int n)
{
    idx_t nnz = 0;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++)
            nnz += 1 + (j < n - 1) + (i < m - 1) + (j > 0) + (i > 0);
    return nnz;
}

void orc_laplacian_setup(int m, int n, idx_t* S, idx_t* I, double* D)
{
    idx_t pos = 0;
    S[0] = 0;
    for (int i = 0; i < m; i++)
        for (int j = 0; j < n; j++) {
            idx_t row = (idx_t)i * n + j;
            // sorted column order of the five inserted entries
            if (i > 0)     { I[pos] = row - n; D[pos++] = -1.0; }
            if (j > 0)     { I[pos] = row - 1; D[pos++] = -1.0; }
            I[pos] = row; D[pos++] = 4.0;
            if (j < n - 1) { I[pos] = row + 1; D[pos++] = -1.0; }
            if (i < m - 1) { I[pos] = row + n; D[pos++] = -1.0; }
            S[row + 1] = pos;
        }
}

// ------------------------------- kernels -----------------------------------

// w = A*v (mode 0), w += A*v (mode 1), w -= A*v (mode 2)
void orc_spmv(idx_t n, const idx_t* S, const idx_t* I, const double* D,
              const double* v, double* w, int mode)
{
    for (idx_t r = 0; r < n; r++) {
        double tmp = row_sum(S, I, D, v, r);
        if (mode == 0) w[r] = tmp;
        else if (mode == 1) w[r] += tmp;
        else w[r] -= tmp;
    }
}

// r = b - A*x: compute_summand materialises A*x into a temporary
// (mtl/operation/compute_summand.hpp:42-60), then element-wise minus.
void orc_residual(idx_t n, const idx_t* S, const idx_t* I, const double* D,
                  const double* x, const double* b, double* r)
{
    for (idx_t i = 0; i < n; i++) r[i] = b[i] - row_sum(S, I, D, x, i);
}

double orc_dot(idx_t n, const double* a, const double* b) { return dot4(n, a, b); }
double orc_two_norm(idx_t n, const double* v) { return two_norm(n, v); }
double orc_unary_dot(idx_t n, const double* v) { return unary_dot(n, v); }

// (lazy(q) = A*p) || (lazy(a) = lazy_dot(p, q)): fused_expr.hpp:56-77 unrolled
// (row_mat_cvec_index_evaluator :35-44; dot_index_evaluator :28-39): partial
// i%4, tail into partial 0, result tmp[0]+tmp[1]+tmp[2]+tmp[3] left to right.
double orc_fused_spmv_dot(idx_t n, const idx_t* S, const idx_t* I, const double* D,
                          const double* p, double* q)
{
    double t[4] = {0, 0, 0, 0};
    idx_t sb = n >> 2 << 2;
    for (idx_t i = 0; i < sb; i += 4)
        for (int k = 0; k < 4; k++) {
            q[i + k] = row_sum(S, I, D, p, i + k);
            t[k] += p[i + k] * q[i + k];
        }
    for (idx_t i = sb; i < n; i++) {
        q[i] = row_sum(S, I, D, p, i);
        t[0] += p[i] * q[i];
    }
    return t[0] + t[1] + t[2] + t[3];
}

// (lazy(r) -= alpha*q) || (lazy(rr) = lazy_unary_dot(r)): vec_vec_aop_expr
// :167-171 and reduction_index_evaluator :29-41: partials i%4, tail into
// partial 0, result (t0+t1)+(t2+t3).
double orc_fused_axpy_unary_dot(idx_t n, double alpha, const double* q, double* r)
{
    double t[4] = {0, 0, 0, 0};
    idx_t sb = n >> 2 << 2;
    for (idx_t i = 0; i < sb; i += 4)
        for (int k = 0; k < 4; k++) {
            r[i + k] -= alpha * q[i + k];
            t[k] += r[i + k] * r[i + k];
        }
    for (idx_t i = sb; i < n; i++) {
        r[i] -= alpha * q[i];
        t[0] += r[i] * r[i];
    }
    return (t[0] + t[1]) + (t[2] + t[3]);
}

// ------------------------------- IC(0) -------------------------------------

// Number of entries of upper(A) (U = upper(A), ic_0.hpp:148).
idx_t orc_upper_nnz(idx_t n, const idx_t* S, const idx_t* I)
{
    idx_t c = 0;
    for (idx_t r = 0; r < n; r++)
        for (idx_t j = S[r]; j < S[r + 1]; j++) c += (I[j] >= r);
    return c;
}

// ic_0::factorizer::factorize (ic_0.hpp:135-186) in double; U gets the
// inverted diagonal first in each row.  Returns ORC_ERR_MISSING_DIAGONAL when
// a row of upper(A) does not start with its diagonal (:157).
int orc_ic0_factor(idx_t n, const idx_t* S, const idx_t* I, const double* D,
                   idx_t* US, idx_t* UI, double* UD)
{
    // U = upper(A)
    idx_t pos = 0;
    US[0] = 0;
    for (idx_t r = 0; r < n; r++) {
        for (idx_t j = S[r]; j < S[r + 1]; j++)
            if (I[j] >= r) { UI[pos] = I[j]; UD[pos] = D[j]; pos++; }
        US[r + 1] = pos;
    }
    for (idx_t k = 0; k < n; k++) {
        idx_t ic = US[k], iend = US[k + 1];
        if (ic == iend || UI[ic] != k) return ORC_ERR_MISSING_DIAGONAL;
        double inv_dia = 1.0 / std::sqrt(UD[ic]);      // reciprocal(sqrt()) :160
        UD[ic] = inv_dia;
        for (++ic; ic != iend; ++ic) {
            double d = UD[ic] * inv_dia;                // :166
            UD[ic] = d;
            idx_t i = UI[ic];
            // row i of U: columns in (k, i] -> only the diagonal (i, i)
            for (idx_t jc = US[i]; jc < US[i + 1]; jc++) {
                idx_t j = UI[jc];
                if (j <= k) continue;
                if (j > i) break;
                idx_t kj = lookup(US, UI, k, j);
                double ukj = kj >= 0 ? UD[kj] : 0.0;    // U[k][j]
                UD[jc] = UD[jc] - d * ukj;              // :181
            }
        }
    }
    return ORC_OK;
}

// CSR transpose as done by the inserter copy of trans(U) (ic_0.hpp:57):
// rows of the result are sorted by column.
void orc_csr_transpose(idx_t n, const idx_t* S, const idx_t* I, const double* D,
                       idx_t* TS, idx_t* TI, double* TD)
{
    std::vector<idx_t> cnt(n + 1, 0);
    for (idx_t j = 0; j < S[n]; j++) cnt[I[j] + 1]++;
    for (idx_t r = 0; r < n; r++) cnt[r + 1] += cnt[r];
    for (idx_t r = 0; r <= n; r++) TS[r] = cnt[r];
    for (idx_t r = 0; r < n; r++)
        for (idx_t j = S[r]; j < S[r + 1]; j++) {
            idx_t c = I[j], p = cnt[c]++;
            TI[p] = r;
            TD[p] = D[j];
        }
}

}  // extern "C"

// Lower sweep, compact CSR, inverted diagonal stored last
// (lower_trisolve.hpp:191-209, version 6).  FV = factor value type.
template <typename FV>
static int lower_v6(idx_t n, const idx_t* S, const idx_t* I, const FV* D, const double* v, double* w)
{
    for (idx_t r = 0; r < n; r++) {
        idx_t j0 = S[r], j1 = S[r + 1];
        if (j0 == j1) return ORC_ERR_MISSING_DIAGONAL;
        --j1;
        if (I[j1] != r) return ORC_ERR_MISSING_DIAGONAL;
        FV dia = D[j1];
        double rr = v[r];
        for (; j0 != j1; ++j0) rr -= D[j0] * w[I[j0]];
        w[r] = rr * dia;
    }
    return ORC_OK;
}

// Lower sweep, compact CSR, unit diagonal (lower_trisolve.hpp:172-189,
// version 5): starts at ref_major[1], so row 0 is taken to be empty.
template <typename FV>
static void lower_v5(idx_t n, const idx_t* S, const idx_t* I, const FV* D, const double* v, double* w)
{
    if (n == 0) return;
    idx_t j1 = S[1];
    for (idx_t r = 0; r != n; ++r) {
        idx_t j0 = j1;
        j1 = S[r + 1];
        double rr = v[r];
        for (; j0 != j1; ++j0) rr -= D[j0] * w[I[j0]];
        w[r] = rr;
    }
}

// Upper sweep, compact CSR, inverted diagonal first
// (upper_trisolve.hpp:159-176, version 3, with crs_row_init :151-156).
template <typename FV>
static void upper_v3(idx_t n, const idx_t* S, const idx_t* I, const FV* D, const double* v, double* w)
{
    for (idx_t r = n; r-- > 0;) {
        idx_t j0 = S[r];
        const idx_t cj1 = S[r + 1];
        double rr = v[r], dia;
        dia = D[j0++];
        for (; j0 != cj1; ++j0) rr -= D[j0] * w[I[j0]];
        w[r] = rr * dia;                      // row_update, inverse_diagonal :102
    }
}

extern "C" {

// ic_0::solve (ic_0.hpp:78-87): lower sweep on L = trans(U), then upper on U.
int orc_ic0_solve(idx_t n, const idx_t* LS, const idx_t* LI, const double* LD,
                  const idx_t* US, const idx_t* UI, const double* UD,
                  const double* x, double* y)
{
    std::vector<double> y0(n);
    int st = lower_v6(n, LS, LI, LD, x, y0.data());
    if (st) return st;
    upper_v3(n, US, UI, UD, y0.data(), y);
    return ORC_OK;
}

// Same with factors rounded to float (ic_0<Matrix, float>: factorise in A's
// type, then U = U_tmp converts, ic_0.hpp:122-130); sweeps accumulate in
// the double output vector (lower_trisolve.hpp:202-207).
int orc_ic0_solve_f32(idx_t n, const idx_t* LS, const idx_t* LI, const float* LD,
                      const idx_t* US, const idx_t* UI, const float* UD,
                      const double* x, double* y)
{
    std::vector<double> y0(n);
    int st = lower_v6(n, LS, LI, LD, x, y0.data());
    if (st) return st;
    upper_v3(n, US, UI, UD, y0.data(), y);
    return ORC_OK;
}

// ------------------------------- ILU(0) ------------------------------------

idx_t orc_strict_lower_nnz(idx_t n, const idx_t* S, const idx_t* I)
{
    idx_t c = 0;
    for (idx_t r = 0; r < n; r++)
        for (idx_t j = S[r]; j < S[r + 1]; j++) c += (I[j] < r);
    return c;
}

// ilu_0_factorizer::factorize (ilu_0.hpp:45-82) + invert_diagonal
// (mtl/operation/invert_diagonal.hpp:47-54); L = strict_lower(LU) and
// U = upper(LU).  LU[k][j] of an absent entry reads 0 (compressed2D.hpp:536-547).
int orc_ilu0_factor(idx_t n, const idx_t* S, const idx_t* I, const double* D,
                    idx_t* LS, idx_t* LI, double* LD, idx_t* US, idx_t* UI, double* UD)
{
    idx_t nnz = S[n];
    std::vector<double> LU(D, D + nnz);
    std::vector<double> inv_dia(n, 0.0);
    for (idx_t i = 0; i < n; i++) {
        // a row without its diagonal would run the k-loop past k == i
        if (lookup(S, I, i, i) < 0) return ORC_ERR_MISSING_DIAGONAL;
        for (idx_t kc = S[i], kend = S[i + 1]; kc != kend; ++kc) {
            idx_t k = I[kc];
            if (k == i) break;
            double aik = LU[kc] * inv_dia[k];                       // :70
            LU[kc] = aik;
            for (idx_t jc = kc + 1; jc != kend; ++jc) {
                idx_t kj = lookup(S, I, k, I[jc]);
                double lukj = kj >= 0 ? LU[kj] : 0.0;
                LU[jc] = LU[jc] - aik * lukj;                        // :74
            }
        }
        idx_t ii = lookup(S, I, i, i);
        if (ii < 0) return ORC_ERR_MISSING_DIAGONAL;
        inv_dia[i] = 1.0 / LU[ii];                                   // :77
    }
    for (idx_t i = 0; i < n; i++) {                                  // invert_diagonal
        idx_t ii = lookup(S, I, i, i);
        LU[ii] = 1.0 / LU[ii];
    }
    idx_t lp = 0, up = 0;
    LS[0] = US[0] = 0;
    for (idx_t r = 0; r < n; r++) {
        for (idx_t j = S[r]; j < S[r + 1]; j++) {
            if (I[j] < r) { LI[lp] = I[j]; LD[lp++] = LU[j]; }
            else          { UI[up] = I[j]; UD[up++] = LU[j]; }
        }
        LS[r + 1] = lp;
        US[r + 1] = up;
    }
    return ORC_OK;
}

// ilu::solve (ilu.hpp:83-91): unit lower sweep (v5) then upper (v3).
void orc_ilu0_solve(idx_t n, const idx_t* LS, const idx_t* LI, const double* LD,
                    const idx_t* US, const idx_t* UI, const double* UD,
                    const double* x, double* y)
{
    std::vector<double> y0(n);
    lower_v5(n, LS, LI, LD, x, y0.data());
    upper_v3(n, US, UI, UD, y0.data(), y);
}

void orc_ilu0_solve_f32(idx_t n, const idx_t* LS, const idx_t* LI, const float* LD,
                        const idx_t* US, const idx_t* UI, const float* UD,
                        const double* x, double* y)
{
    std::vector<double> y0(n);
    lower_v5(n, LS, LI, LD, x, y0.data());
    upper_v3(n, US, UI, UD, y0.data(), y);
}

// ilu::adjoint_solve (ilu.hpp:105-115): lower sweep on adjoint(U) with the
// inverted diagonal (column-major generic version 4, lower_trisolve.hpp:155-170)
// followed by the unit upper sweep on adjoint(L) (column-major version 2,
// upper_trisolve.hpp:131-149).  Both scatter along rows of U and L.
void orc_ilu0_adjoint_solve(idx_t n, const idx_t* LS, const idx_t* LI, const double* LD,
                            const idx_t* US, const idx_t* UI, const double* UD,
                            const double* x, double* y)
{
    std::vector<double> w(x, x + n);
    // adjoint(U) column r == row r of U: diagonal first, then entries c > r
    for (idx_t r = 0; r < n; r++) {
        idx_t j = US[r];
        double rr = w[r] *= UD[j];
        for (++j; j < US[r + 1]; ++j) w[UI[j]] -= UD[j] * rr;
    }
    // adjoint(L) column r == row r of L (strict lower, unit diagonal), r descending
    for (idx_t i = 0; i < n; i++) y[i] = w[i];
    for (idx_t r = n; r-- > 0;) {
        double rr = y[r];
        for (idx_t j = LS[r]; j < LS[r + 1]; ++j) y[LI[j]] -= LD[j] * rr;
    }
}

// ------------------------------- solvers -----------------------------------

// Preconditioner kinds for the drivers.
enum { ORC_PC_IDENTITY = 0, ORC_PC_DIAGONAL = 1, ORC_PC_IC0 = 2, ORC_PC_ILU0 = 3 };

struct orc_pc {
    int kind;
    idx_t n;
    const double* inv_diag;                               // diagonal
    const idx_t *LS, *LI; const double* LD;               // ic0: L = trans(U); ilu0: strict lower
    const idx_t *US, *UI; const double* UD;
};

static void pc_solve(const orc_pc* P, const double* x, double* y)
{
    idx_t n = P->n;
    switch (P->kind) {
    case ORC_PC_IDENTITY: for (idx_t i = 0; i < n; i++) y[i] = x[i]; break;
    case ORC_PC_DIAGONAL:                                // itl/pc/diagonal.hpp:60-67
        for (idx_t i = 0; i < n; i++) y[i] = P->inv_diag[i] * x[i];
        break;
    case ORC_PC_IC0: orc_ic0_solve(n, P->LS, P->LI, P->LD, P->US, P->UI, P->UD, x, y); break;
    case ORC_PC_ILU0: orc_ilu0_solve(n, P->LS, P->LI, P->LD, P->US, P->UI, P->UD, x, y); break;
    }
}

// diagonal PC setup: inv_diag[i] = reciprocal(A[i][i]) (itl/pc/diagonal.hpp:36-45).
void orc_diagonal_setup(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* inv_diag)
{
    for (idx_t i = 0; i < n; i++) {
        idx_t p = lookup(S, I, i, i);
        inv_diag[i] = 1.0 / (p >= 0 ? D[p] : 0.0);
    }
}

// itl::cg: without PC (cg.hpp:39-72) when the PC is the identity (:80-82),
// otherwise the left-preconditioned variant (:75-112).
int orc_cg(idx_t n, const idx_t* S, const idx_t* I, const double* D,
           double* x, const double* b, const orc_pc* P, orc_iteration* it)
{
    std::vector<double> p(n), q(n), r(n), z(n);
    orc_residual(n, S, I, D, x, b, r.data());              // r = b - A*x
    if (P->kind == ORC_PC_IDENTITY) {
        double rho = 0, rho_1 = 0, alpha = 0, alpha_1 = 0;
        rho = dot4(n, r.data(), r.data());
        while (!orc_finished(it, std::sqrt(std::fabs(rho)))) {
            ++it->i;
            if (it->i <= 1) for (idx_t k = 0; k < n; k++) p[k] = r[k];
            else {
                double s = rho / rho_1;
                for (idx_t k = 0; k < n; k++) p[k] = r[k] + s * p[k];
            }
            alpha_1 = orc_fused_spmv_dot(n, S, I, D, p.data(), q.data());
            alpha = rho / alpha_1;
            for (idx_t k = 0; k < n; k++) x[k] += alpha * p[k];
            rho_1 = rho;
            rho = orc_fused_axpy_unary_dot(n, alpha, q.data(), r.data());
        }
        return it->error;
    }
    double rho = 0, rho_1 = 0, rr, alpha = 0, alpha_1;
    rr = dot4(n, r.data(), r.data());                        // :94
    while (!orc_finished(it, std::sqrt(std::fabs(rr)))) {    // :95
        ++it->i;
        pc_solve(P, r.data(), z.data());                     // :97 (not fused, Q8)
        rho = dot4(n, r.data(), z.data());
        if (it->i <= 1) for (idx_t k = 0; k < n; k++) p[k] = z[k];
        else {
            double s = rho / rho_1;                          // :102
            for (idx_t k = 0; k < n; k++) p[k] = z[k] + s * p[k];
        }
        alpha_1 = orc_fused_spmv_dot(n, S, I, D, p.data(), q.data());   // :104
        alpha = rho / alpha_1;
        for (idx_t k = 0; k < n; k++) x[k] += alpha * p[k];  // :107
        rho_1 = rho;
        rr = orc_fused_axpy_unary_dot(n, alpha, q.data(), r.data());   // :109
    }
    return it->error;
}

// itl::bicgstab (bicgstab.hpp:27-78).
int orc_bicgstab(idx_t n, const idx_t* S, const idx_t* I, const double* D,
                 double* x, const double* b, const orc_pc* P, orc_iteration* it)
{
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0, gamma, omega = 0;
    std::vector<double> p(n), phat(n), s(n), shat(n), t(n), v(n), r(n), rtilde(n);
    orc_residual(n, S, I, D, x, b, r.data());
    rtilde = r;
    while (!orc_finished(it, two_norm(n, r.data()))) {
        ++it->i;
        rho_1 = dot4(n, rtilde.data(), r.data());
        if (rho_1 == 0.0) return ORC_ERR_ORTHOGONALITY;
        if (it->i <= 1) p = r;
        else {
            if (omega == 0.0) return ORC_ERR_ORTHOGONALITY;
            beta = (rho_1 / rho_2) * (alpha / omega);
            for (idx_t k = 0; k < n; k++) p[k] = r[k] + beta * (p[k] - omega * v[k]);
        }
        pc_solve(P, p.data(), phat.data());
        orc_spmv(n, S, I, D, phat.data(), v.data(), 0);
        gamma = dot4(n, rtilde.data(), v.data());
        if (gamma == 0.0) return ORC_ERR_ORTHOGONALITY;
        alpha = rho_1 / gamma;
        for (idx_t k = 0; k < n; k++) s[k] = r[k] - alpha * v[k];
        if (orc_finished(it, two_norm(n, s.data()))) {
            for (idx_t k = 0; k < n; k++) x[k] += alpha * phat[k];
            break;
        }
        pc_solve(P, s.data(), shat.data());
        orc_spmv(n, S, I, D, shat.data(), t.data(), 0);
        omega = dot4(n, t.data(), s.data()) / dot4(n, t.data(), t.data());
        for (idx_t k = 0; k < n; k++) x[k] += omega * shat[k] + alpha * phat[k];
        for (idx_t k = 0; k < n; k++) r[k] = s[k] - omega * t[k];
        rho_2 = rho_1;
    }
    return it->error;
}

// adjoint_solve(P, x) (pc/solver.hpp:22-52): identity, diagonal (real) and
// IC(0) (ic_0.hpp:97-103) solve as usual; ILU(0) runs ilu::adjoint_solve.
static void pc_adjoint_solve(const orc_pc* P, const double* x, double* y)
{
    if (P->kind == ORC_PC_ILU0) orc_ilu0_adjoint_solve(P->n, P->LS, P->LI, P->LD, P->US, P->UI, P->UD, x, y);
    else pc_solve(P, x, y);
}

// w = adjoint(A) * v for a real CSR A: the transposed view is traversed by
// A's rows, each row scattering into w (entries of w summed in row order of A).
static void spmv_adjoint(idx_t n, const idx_t* S, const idx_t* I, const double* D, const double* v, double* w)
{
    for (idx_t k = 0; k < n; k++) w[k] = 0.0;
    for (idx_t r = 0; r < n; r++)
        for (idx_t e = S[r]; e < S[r + 1]; e++) w[I[e]] += D[e] * v[r];
}

// itl::cgs (cgs.hpp:24-62).  A zero rho_1 records error 2 ("cgs breakdown")
// and, as in the reference, the loop goes on.
int orc_cgs(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b, const orc_pc* P,
            orc_iteration* it)
{
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
    std::vector<double> p(n), phat(n), q(n), qhat(n), vhat(n), u(n), uhat(n), r(n), rtilde(n);
    orc_residual(n, S, I, D, x, b, r.data());
    rtilde = r;
    while (!orc_finished(it, two_norm(n, r.data()))) {
        ++it->i;
        rho_1 = dot4(n, rtilde.data(), r.data());
        if (rho_1 == 0.0) it->error = 2;
        if (it->i <= 1) {
            u = r;
            p = r;
        } else {
            beta = rho_1 / rho_2;
            for (idx_t k = 0; k < n; k++) u[k] = r[k] + beta * q[k];
            for (idx_t k = 0; k < n; k++) p[k] = u[k] + beta * (q[k] + beta * p[k]);
        }
        pc_solve(P, p.data(), phat.data());
        orc_spmv(n, S, I, D, phat.data(), vhat.data(), 0);
        alpha = rho_1 / dot4(n, rtilde.data(), vhat.data());
        for (idx_t k = 0; k < n; k++) q[k] = u[k] - alpha * vhat[k];
        for (idx_t k = 0; k < n; k++) u[k] += q[k];
        pc_solve(P, u.data(), uhat.data());
        for (idx_t k = 0; k < n; k++) x[k] += alpha * uhat[k];
        orc_spmv(n, S, I, D, uhat.data(), qhat.data(), 0);
        for (idx_t k = 0; k < n; k++) r[k] -= alpha * qhat[k];
        rho_2 = rho_1;
    }
    return it->error;
}

// itl::bicg (bicg.hpp:27-68).  A zero rho_1 returns error 2 ("bicg breakdown").
int orc_bicg(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b, const orc_pc* P,
             orc_iteration* it)
{
    double rho_1 = 0, rho_2 = 0, alpha = 0, beta = 0;
    std::vector<double> r(n), z(n), p(n), q(n), r_tilde(n), z_tilde(n), p_tilde(n), q_tilde(n);
    orc_residual(n, S, I, D, x, b, r.data());
    r_tilde = r;
    while (!orc_finished(it, two_norm(n, r.data()))) {
        ++it->i;
        pc_solve(P, r.data(), z.data());
        pc_adjoint_solve(P, r_tilde.data(), z_tilde.data());
        rho_1 = dot4(n, z_tilde.data(), z.data());
        if (rho_1 == 0.0) return (it->error = 2);
        if (it->i <= 1) {
            p = z;
            p_tilde = z_tilde;
        } else {
            beta = rho_1 / rho_2;
            for (idx_t k = 0; k < n; k++) p[k] = z[k] + beta * p[k];
            for (idx_t k = 0; k < n; k++) p_tilde[k] = z_tilde[k] + beta * p_tilde[k];
        }
        orc_spmv(n, S, I, D, p.data(), q.data(), 0);
        spmv_adjoint(n, S, I, D, p_tilde.data(), q_tilde.data());
        alpha = rho_1 / dot4(n, p_tilde.data(), q.data());
        for (idx_t k = 0; k < n; k++) x[k] += alpha * p[k];
        for (idx_t k = 0; k < n; k++) r[k] -= alpha * q[k];
        for (idx_t k = 0; k < n; k++) r_tilde[k] -= alpha * q_tilde[k];
        rho_2 = rho_1;
    }
    return it->error;
}

// itl::tfqmr (tfqmr.hpp:32-99) with left preconditioner L and right
// preconditioner R (R = NULL: identity).  Breakdowns return error 1.
int orc_tfqmr(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b,
              const orc_pc* L, const orc_pc* R, orc_iteration* it)
{
    const orc_pc ident = {ORC_PC_IDENTITY, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (!R) R = &ident;
    double theta = 0, eta = 0, tau, rho, rhon, sigma, alpha, beta, c;
    std::vector<double> rt(n), r(n), u1(n), u2(n), y1(n), y2(n), w(n), d(n, 0.0), v(n), tmp(n);
    pc_solve(R, x, tmp.data());                               // rt = b - A * solve(R, x)
    orc_residual(n, S, I, D, tmp.data(), b, rt.data());
    pc_solve(L, rt.data(), r.data());
    if (orc_finished(it, two_norm(n, rt.data()))) return it->error;
    y1 = r;
    w = r;
    pc_solve(R, y1.data(), tmp.data());
    orc_spmv(n, S, I, D, tmp.data(), rt.data(), 0);
    pc_solve(L, rt.data(), u1.data());
    v = u1;
    tau = two_norm(n, r.data());
    rho = tau * tau;
    while (!orc_finished(it, tau)) {
        ++it->i;
        sigma = dot4(n, r.data(), v.data());
        if (sigma == 0.0) return (it->error = 1);
        alpha = rho / sigma;
        for (int j = 1; j < 3; j++) {
            if (j == 1) {
                for (idx_t k = 0; k < n; k++) w[k] -= alpha * u1[k];
                const double s = theta * theta * eta / alpha;
                for (idx_t k = 0; k < n; k++) d[k] = y1[k] + s * d[k];
            } else {
                for (idx_t k = 0; k < n; k++) y2[k] = y1[k] - alpha * v[k];
                pc_solve(R, y2.data(), tmp.data());
                orc_spmv(n, S, I, D, tmp.data(), rt.data(), 0);
                pc_solve(L, rt.data(), u2.data());
                for (idx_t k = 0; k < n; k++) w[k] -= alpha * u2[k];
                const double s = theta * theta * eta / alpha;
                for (idx_t k = 0; k < n; k++) d[k] = y2[k] + s * d[k];
            }
            theta = two_norm(n, w.data()) / tau;
            c = 1.0 / std::sqrt(1.0 + theta * theta);
            tau *= theta * c;
            eta = c * c * alpha;
            for (idx_t k = 0; k < n; k++) x[k] += eta * d[k];
        }
        if (rho == 0.0) return (it->error = 1);
        rhon = dot4(n, r.data(), w.data());
        beta = rhon / rho;
        rho = rhon;
        for (idx_t k = 0; k < n; k++) y1[k] = w[k] + beta * y2[k];
        pc_solve(R, y1.data(), tmp.data());
        orc_spmv(n, S, I, D, tmp.data(), rt.data(), 0);
        pc_solve(L, rt.data(), u1.data());
        for (idx_t k = 0; k < n; k++) v[k] = u1[k] + beta * (u2[k] + beta * v[k]);
        // rt = A * x - b (tfqmr.hpp:94) is not used afterwards
    }
    pc_solve(R, x, tmp.data());                               // shift back: x = solve(R, x)
    for (idx_t k = 0; k < n; k++) x[k] = tmp[k];
    return it->error;
}

// itl::qmr (qmr.hpp:29-109) with left preconditioner L and right R (NULL:
// identity).  trans(A) * q is the row-order scatter of spmv_adjoint.
// Breakdowns return their codes 1..6 as the reference's iter.fail(k, ...).
int orc_qmr(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b, const orc_pc* L,
            const orc_pc* R, orc_iteration* it)
{
    const orc_pc ident = {ORC_PC_IDENTITY, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (!R) R = &ident;
    double beta, gamma = 1.0, gamma_1, delta, eta = -1.0, ep = 1.0, rho_1, theta = 0.0, theta_1;
    std::vector<double> r(n), v_tld(n), y(n), w_tld(n), z(n), v(n), w(n), y_tld(n), z_tld(n), p(n), q(n), p_tld(n),
        d(n), s(n);
    orc_residual(n, S, I, D, x, b, r.data());
    v_tld = r;
    pc_solve(L, v_tld.data(), y.data());
    w_tld = r;
    pc_adjoint_solve(R, w_tld.data(), z.data());
    if (orc_finished(it, two_norm(n, r.data()))) return it->error;
    double rho = two_norm(n, y.data()), xi = two_norm(n, z.data());
    while (!orc_finished(it, rho)) {
        ++it->i;
        if (rho == 0.0) return (it->error = 1);
        if (xi == 0.0) return (it->error = 2);
        for (idx_t k = 0; k < n; k++) v[k] = v_tld[k] / rho;
        for (idx_t k = 0; k < n; k++) y[k] /= rho;
        for (idx_t k = 0; k < n; k++) w[k] = w_tld[k] / xi;
        for (idx_t k = 0; k < n; k++) z[k] /= xi;
        delta = dot4(n, z.data(), y.data());
        if (delta == 0.0) return (it->error = 3);
        pc_solve(R, y.data(), y_tld.data());
        pc_adjoint_solve(L, z.data(), z_tld.data());
        if (it->i <= 1) {
            p = y_tld;
            q = z_tld;
        } else {
            const double sp = (xi * delta) / ep, sq = (rho * delta) / ep;
            for (idx_t k = 0; k < n; k++) p[k] = y_tld[k] - sp * p[k];
            for (idx_t k = 0; k < n; k++) q[k] = z_tld[k] - sq * q[k];
        }
        orc_spmv(n, S, I, D, p.data(), p_tld.data(), 0);
        ep = dot4(n, q.data(), p_tld.data());
        if (ep == 0.0) return (it->error = 4);
        beta = ep / delta;
        if (beta == 0.0) return (it->error = 5);
        for (idx_t k = 0; k < n; k++) v_tld[k] = p_tld[k] - beta * v[k];
        pc_solve(L, v_tld.data(), y.data());
        rho_1 = rho;
        rho = two_norm(n, y.data());
        spmv_adjoint(n, S, I, D, q.data(), w_tld.data());
        for (idx_t k = 0; k < n; k++) w_tld[k] = w_tld[k] - beta * w[k];
        pc_adjoint_solve(R, w_tld.data(), z.data());
        xi = two_norm(n, z.data());
        gamma_1 = gamma;
        theta_1 = theta;
        theta = rho / (gamma_1 * beta);
        gamma = 1.0 / std::sqrt(1.0 + theta * theta);
        if (gamma == 0.0) return (it->error = 6);
        eta = -eta * rho_1 * gamma * gamma / (beta * gamma_1 * gamma_1);
        if (it->i <= 1) {
            for (idx_t k = 0; k < n; k++) d[k] = eta * p[k];
            for (idx_t k = 0; k < n; k++) s[k] = eta * p_tld[k];
        } else {
            const double f = theta_1 * theta_1 * gamma * gamma;
            for (idx_t k = 0; k < n; k++) d[k] = eta * p[k] + f * d[k];
            for (idx_t k = 0; k < n; k++) s[k] = eta * p_tld[k] + f * s[k];
        }
        for (idx_t k = 0; k < n; k++) x[k] += d[k];
        for (idx_t k = 0; k < n; k++) r[k] -= s[k];
    }
    return it->error;
}

// ---- small dense algebra of gmres / idr_s ----------------------------------

// lu_solve(A, b) (mtl/operation/lu.hpp:132-140): LU with partial pivoting
// (lu.hpp:63-88: first row of largest |a| in the column, rows swapped, a zero
// pivot among the first m-1 throws matrix_singular), then
// upper_trisolve(upper(LU), unit_lower_trisolve(strict_lower(LU), P b)).
// A is m x m row-major; returns false for matrix_singular.
static bool dense_lu_solve(int m, std::vector<double> A, const double* b, double* y)
{
    std::vector<int> P(m);
    for (int i = 0; i < m; i++) P[i] = i;
    for (int i = 0; i + 1 < m; i++) {
        int rmax = i;
        for (int r = i + 1; r < m; r++)
            if (std::fabs(A[size_t(r) * m + i]) > std::fabs(A[size_t(rmax) * m + i])) rmax = r;
        if (rmax != i) {
            for (int c = 0; c < m; c++) std::swap(A[size_t(i) * m + c], A[size_t(rmax) * m + c]);
            std::swap(P[i], P[rmax]);
        }
        if (std::fabs(A[size_t(i) * m + i]) <= 0.0) return false;
        for (int r = i + 1; r < m; r++) A[size_t(r) * m + i] /= A[size_t(i) * m + i];
        for (int r = i + 1; r < m; r++)
            for (int c = i + 1; c < m; c++) A[size_t(r) * m + c] -= A[size_t(r) * m + i] * A[size_t(i) * m + c];
    }
    std::vector<double> w(m);
    for (int r = 0; r < m; r++) {
        double t = b[P[r]];
        for (int c = 0; c < r; c++) t -= A[size_t(r) * m + c] * w[c];
        w[r] = t;
    }
    for (int r = m - 1; r >= 0; r--) {
        double t = w[r];
        for (int c = r + 1; c < m; c++) t -= A[size_t(r) * m + c] * y[c];
        y[r] = t / A[size_t(r) * m + r];
    }
    return true;
}

// mat::givens::set_rotation (mtl/operation/givens.hpp:40-56)
static void givens_rotation(double a, double b, double& c, double& s)
{
    if (b == 0.0) {
        c = 1.0;
        s = 0.0;
    } else if (std::fabs(b) > std::fabs(a)) {
        const double t = -a / b;
        s = 1.0 / std::sqrt(1.0 + t * t);
        c = s * t;
    } else {
        const double t = -b / a;
        c = 1.0 / std::sqrt(1.0 + t * t);
        s = c * t;
    }
}

// mat::givens::trafo(k) (givens.hpp:66-84) on the (rows x cols) row-major H
// with G = [c s; -s c]: rows k, k+1 become trans(G) * H[k:k+2][:], then
// columns k, k+1 become H[:][k:k+2] * G.
static void givens_trafo(std::vector<double>& H, int rows, int cols, int k, double c, double s)
{
    for (int j = 0; j < cols; j++) {
        const double h0 = H[size_t(k) * cols + j], h1 = H[size_t(k + 1) * cols + j];
        H[size_t(k) * cols + j] = c * h0 + (-s) * h1;
        H[size_t(k + 1) * cols + j] = s * h0 + c * h1;
    }
    if (k + 1 >= cols) return;   // the column block lies outside H
    for (int i = 0; i < rows; i++) {
        const double h0 = H[size_t(i) * cols + k], h1 = H[size_t(i) * cols + k + 1];
        H[size_t(i) * cols + k] = h0 * c + h1 * (-s);
        H[size_t(i) * cols + k + 1] = h0 * s + h1 * c;
    }
}

// itl::gmres_full (gmres.hpp:37-119): kmax = min(n, max_iter - iterations)
// Arnoldi steps with modified Gram-Schmidt and one re-orthogonalisation, the
// reference's Givens handling (for i < k a rotation built from column k-1,
// applied from both sides, then the new rotation of column k), H y = g by
// lu_solve on the largest regular leading block, x += solve(R, V y) and the
// final test on the true residual b - A x.
int orc_gmres_full(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b,
                   const orc_pc* L, const orc_pc* R, orc_iteration* it)
{
    const orc_pc ident = {ORC_PC_IDENTITY, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (!R) R = &ident;
    const long room = long(it->max_iter) - long(it->i);
    const int kmax = int(room < 0 ? n : std::min<long>(long(n), room));
    const int rows = kmax + 1, cols = kmax;
    std::vector<double> r0(n), r(n), va(n), va0(n), tmp(n);
    std::vector<std::vector<double>> V(size_t(kmax) + 1, std::vector<double>(n, 0.0));
    std::vector<double> s(size_t(kmax) + 1, 0.0), c(size_t(kmax) + 1, 0.0), g(size_t(kmax) + 1, 0.0),
        y(size_t(kmax), 0.0), H(size_t(rows) * std::max(cols, 1), 0.0);
    auto h = [&](int i, int j) -> double& { return H[size_t(i) * cols + j]; };
    orc_residual(n, S, I, D, x, b, r0.data());
    pc_solve(L, r0.data(), r.data());
    double rho = g[0] = two_norm(n, r.data());
    if (orc_finished(it, rho)) return it->error;
    for (idx_t e = 0; e < n; e++) V[0][e] = r[e] / rho;
    int k;
    for (k = 0; k < kmax; ++k, ++it->i) {
        pc_solve(R, V[k].data(), tmp.data());
        orc_spmv(n, S, I, D, tmp.data(), va0.data(), 0);
        pc_solve(L, va0.data(), va.data());
        V[k + 1] = va;
        for (int j = 0; j < k + 1; j++) {
            h(j, k) = dot4(n, V[j].data(), V[k + 1].data());
            for (idx_t e = 0; e < n; e++) V[k + 1][e] -= h(j, k) * V[j][e];
        }
        h(k + 1, k) = two_norm(n, V[k + 1].data());
        for (int j = 0; j < k + 1; j++) {
            const double hr = dot4(n, V[k + 1].data(), V[j].data());
            h(j, k) += hr;
            for (idx_t e = 0; e < n; e++) V[k + 1][e] -= hr * V[j][e];
        }
        h(k + 1, k) = two_norm(n, V[k + 1].data());
        if (h(k + 1, k) != 0.0) {
            const double inv = 1. / h(k + 1, k);
            for (idx_t e = 0; e < n; e++) V[k + 1][e] *= inv;
        }
        for (int i = 0; i < k; i++) {
            double gc, gs;
            givens_rotation(h(i, k - 1), h(i + 1, k - 1), gc, gs);
            givens_trafo(H, rows, cols, i, gc, gs);
        }
        const double nu = std::sqrt(h(k, k) * h(k, k) + h(k + 1, k) * h(k + 1, k));
        if (nu != 0.0) {
            c[k] = h(k, k) / nu;
            s[k] = -h(k + 1, k) / nu;
            h(k, k) = c[k] * h(k, k) - s[k] * h(k + 1, k);
            h(k + 1, k) = 0;
            const double w1 = c[k] * g[k] - s[k] * g[k + 1], w2 = s[k] * g[k] + c[k] * g[k + 1];
            g[k] = w1;
            g[k + 1] = w2;
        }
        rho = std::fabs(g[k + 1]);
    }
    while (k > 0 && std::fabs(g[k - 1]) <= it->atol) k--;
    int m = k;
    for (; m > 0; --m) {
        std::vector<double> Hm(size_t(m) * m);
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) Hm[size_t(i) * m + j] = h(i, j);
        if (dense_lu_solve(m, Hm, g.data(), y.data())) break;
    }
    if (m == 0) return (it->error = 2);   // "GMRES did not find any direction to correct x"
    std::vector<double> Vy(n, 0.0);
    for (int j = 0; j < m; j++)
        for (idx_t e = 0; e < n; e++) Vy[e] += V[j][e] * y[j];
    pc_solve(R, Vy.data(), tmp.data());
    for (idx_t e = 0; e < n; e++) x[e] += tmp[e];
    orc_residual(n, S, I, D, x, b, r.data());
    orc_finished(it, two_norm(n, r.data()));   // iter.terminate(r)
    return it->error;
}

// itl::gmres with restart (gmres.hpp:122-137): gmres_full on a copy of the
// iteration limited to `restart` more steps, then update_progress
// (basic_iteration.hpp:145-155) until the iteration is finished.
int orc_gmres(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b,
              const orc_pc* L, const orc_pc* R, orc_iteration* it, int restart)
{
    if (restart < 1) return -1;
    do {
        orc_iteration inner = *it;
        inner.max_iter = std::min(it->i + restart, it->max_iter);
        orc_gmres_full(n, S, I, D, x, b, L, R, &inner);
        it->i = inner.i;
        it->resid = inner.resid;
        if (inner.error > 1) {
            it->error = inner.error;
            it->is_finished = 1;
        } else {
            orc_finished(it, it->resid);
        }
    } while (!it->is_finished);
    return it->error;
}

// orthogonalize_factors (mtl/operation/orth.hpp:61-85, unfused branch) on
// v[0..m): tau[0][0] = v0.v0; for j: tau[i][j] = (v_i.v_j) / tau[i][i],
// v_j -= tau[i][j] v_i (i < j), tau[j][j] = v_j.v_j.  tau is m x m row-major.
static void orthogonalize_factors(idx_t n, std::vector<std::vector<double>*>& v, std::vector<double>& tau)
{
    const int m = int(v.size());
    tau.assign(size_t(m) * m, 0.0);
    if (m == 0) return;
    tau[0] = dot4(n, v[0]->data(), v[0]->data());
    for (int j = 1; j < m; j++) {
        for (int i = 0; i < j; i++) {
            const double t = dot4(n, v[i]->data(), v[j]->data()) / tau[size_t(i) * m + i];
            tau[size_t(i) * m + j] = t;
            for (idx_t e = 0; e < n; e++) (*v[j])[e] -= t * (*v[i])[e];
        }
        tau[size_t(j) * m + j] = dot4(n, v[j]->data(), v[j]->data());
    }
}

// itl::bicgstab_ell (bicgstab_ell.hpp:37-135) with left L and right R (NULL:
// identity); itl::bicgstab_2 (bicgstab_2.hpp:35-130) is this with l = 2 and
// R = identity.  The MR part orthogonalises r_hat[1..l] in place (the
// sub-vector r_hat[1:] is a view).  omega == 0 returns error 3.
int orc_bicgstab_ell(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b,
                     const orc_pc* L, const orc_pc* R, orc_iteration* it, int l)
{
    const orc_pc ident = {ORC_PC_IDENTITY, n, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (!R) R = &ident;
    if (l < 1) return -1;
    std::vector<double> x0(n, 0.0), y(n), tmp(n), r0_tilde(n);
    std::vector<std::vector<double>> r_hat(size_t(l) + 1, std::vector<double>(n)),
        u_hat(size_t(l) + 1, std::vector<double>(n));
    r_hat[0].assign(b, b + n);
    if (two_norm(n, x) != 0.0) {
        orc_spmv(n, S, I, D, x, tmp.data(), 0);
        for (idx_t e = 0; e < n; e++) r_hat[0][e] -= tmp[e];
        x0.assign(x, x + n);
        for (idx_t e = 0; e < n; e++) x[e] = 0.0;
    }
    const double nr = two_norm(n, r_hat[0].data());
    for (idx_t e = 0; e < n; e++) r0_tilde[e] = r_hat[0][e] / nr;
    pc_solve(L, r_hat[0].data(), y.data());
    r_hat[0] = y;
    std::fill(u_hat[0].begin(), u_hat[0].end(), 0.0);
    double rho_0 = 1.0, rho_1 = 0.0, alpha = 0.0, Gamma = 0.0, beta = 0.0, omega = 1.0;
    std::vector<double> tau, gamma(size_t(l) + 1), gamma_a(size_t(l) + 1), gamma_aa(size_t(l) + 1);
    auto finish_x = [&]() {
        pc_solve(R, x, tmp.data());
        for (idx_t e = 0; e < n; e++) x[e] = tmp[e] + x0[e];
    };
    while (!orc_finished(it, two_norm(n, r_hat[0].data()))) {
        ++it->i;
        rho_0 = -omega * rho_0;
        for (int j = 0; j < l; ++j) {
            rho_1 = dot4(n, r0_tilde.data(), r_hat[j].data());
            beta = alpha * rho_1 / rho_0;
            rho_0 = rho_1;
            for (int i = 0; i <= j; ++i)
                for (idx_t e = 0; e < n; e++) u_hat[i][e] = r_hat[i][e] - beta * u_hat[i][e];
            pc_solve(R, u_hat[j].data(), tmp.data());
            orc_spmv(n, S, I, D, tmp.data(), y.data(), 0);
            pc_solve(L, y.data(), u_hat[j + 1].data());
            Gamma = dot4(n, r0_tilde.data(), u_hat[j + 1].data());
            alpha = rho_0 / Gamma;
            for (int i = 0; i <= j; ++i)
                for (idx_t e = 0; e < n; e++) r_hat[i][e] -= alpha * u_hat[i + 1][e];
            if (orc_finished(it, two_norm(n, r_hat[j].data()))) {
                finish_x();
                return it->error;
            }
            pc_solve(R, r_hat[j].data(), r_hat[j + 1].data());
            orc_spmv(n, S, I, D, r_hat[j + 1].data(), y.data(), 0);
            pc_solve(L, y.data(), r_hat[j + 1].data());
            for (idx_t e = 0; e < n; e++) x[e] += alpha * u_hat[0][e];
        }
        std::vector<std::vector<double>*> tail;
        for (int j = 1; j <= l; j++) tail.push_back(&r_hat[j]);
        orthogonalize_factors(n, tail, tau);   // tau[1..l][1..l] of the reference, 0-based here
        auto T = [&](int i, int j) { return tau[size_t(i - 1) * l + (j - 1)]; };
        for (int j = 1; j <= l; ++j) gamma_a[j] = dot4(n, r_hat[j].data(), r_hat[0].data()) / T(j, j);
        gamma[l] = gamma_a[l];
        omega = gamma[l];
        if (omega == 0.0) return (it->error = 3);   // "bicg breakdown #2"
        for (int j = l - 1; j > 0; --j) {
            double sum = 0.0;
            for (int i = j + 1; i <= l; ++i) sum += T(j, i) * gamma[i];
            gamma[j] = gamma_a[j] - sum;
        }
        // gamma_aa[1..l-1] = strict_upper(tau[1..l-1][1..l-1]) * gamma[2..l] + gamma[2..l]
        for (int j = 1; j < l; ++j) {
            double sum = 0.0;
            for (int i = j + 1; i < l; ++i) sum += T(j, i) * gamma[i + 1];
            gamma_aa[j] = sum + gamma[j + 1];
        }
        for (idx_t e = 0; e < n; e++) x[e] += gamma[1] * r_hat[0][e];
        for (idx_t e = 0; e < n; e++) r_hat[0][e] -= gamma_a[l] * r_hat[l][e];
        for (idx_t e = 0; e < n; e++) u_hat[0][e] -= gamma[l] * u_hat[l][e];
        for (int j = 1; j < l; ++j) {
            for (idx_t e = 0; e < n; e++) u_hat[0][e] -= gamma[j] * u_hat[j][e];
            for (idx_t e = 0; e < n; e++) x[e] += gamma_aa[j] * r_hat[j][e];
            for (idx_t e = 0; e < n; e++) r_hat[0][e] -= gamma_a[j] * r_hat[j][e];
        }
    }
    finish_x();
    return it->error;
}

// ---- assembly: compressed2D_inserter --------------------------------------

// mat::inserter<compressed2D, Updater> (mtl/matrix/compressed2D.hpp:662-1151,
// modify :955-985): the matrix's entries are present first; each inserted
// value either updates a present entry with the updater (update.hpp:20-115:
// store x = v, plus x += v, minus x -= v, times x *= v) or creates it with
// updater.init(v) (v; -v for minus).  Rows are sorted by column at the end.
// updater: 0 store, 1 plus, 2 minus, 3 times.  With outS == NULL only the
// entry count is returned.
idx_t orc_assemble(idx_t nrows, idx_t count, const idx_t* rows, const idx_t* cols, const double* vals, int updater,
                   const idx_t* exS, const idx_t* exI, const double* exD, idx_t* outS, idx_t* outI, double* outD)
{
    std::vector<std::vector<std::pair<idx_t, double>>> R(static_cast<size_t>(nrows));
    auto find = [&](idx_t r, idx_t c) -> double* {
        for (auto& e : R[size_t(r)])
            if (e.first == c) return &e.second;
        return nullptr;
    };
    if (exS)
        for (idx_t r = 0; r < nrows; r++)
            for (idx_t j = exS[r]; j < exS[r + 1]; j++) R[size_t(r)].push_back({exI[j], exD[j]});
    for (idx_t i = 0; i < count; i++) {
        double* x = find(rows[i], cols[i]);
        const double v = vals[i];
        if (x) {
            if (updater == 0) *x = v;
            else if (updater == 1) *x += v;
            else if (updater == 2) *x -= v;
            else *x *= v;
        } else {
            R[size_t(rows[i])].push_back({cols[i], updater == 2 ? -v : v});
        }
    }
    idx_t nnz = 0;
    for (auto& row : R) nnz += idx_t(row.size());
    if (!outS) return nnz;
    idx_t k = 0;
    outS[0] = 0;
    for (idx_t r = 0; r < nrows; r++) {
        auto row = R[size_t(r)];
        std::sort(row.begin(), row.end(), [](const std::pair<idx_t, double>& a, const std::pair<idx_t, double>& b) {
            return a.first < b.first;
        });
        for (auto& e : row) {
            outI[k] = e.first;
            outD[k] = e.second;
            k++;
        }
        outS[r + 1] = k;
    }
    return nnz;
}

// ---- matrix-free operator ---------------------------------------------------

// poisson2D_dirichlet::mult (mtl/matrix/poisson2D_dirichlet.hpp:27-80), loop by
// loop: the inner domain in blocks of four columns then the remainder, the
// four borders, the four corners; mode 0 assign, 1 +=, 2 -= (Assign::apply).
void orc_poisson2d_dirichlet_mult(int m, int n, const double* v, double* w, int mode)
{
    auto apply = [&](idx_t k, double x) {
        if (mode == 0) w[k] = x;
        else if (mode == 1) w[k] += x;
        else w[k] -= x;
    };
    const int nb = n < 3 ? 1 : (n - 2) / 4 * 4 + 1;
    for (int i = 1; i < m - 1; i++) {
        const int kmax = i * n + nb;
        for (int k = i * n + 1; k < kmax; k += 4) {
            const double v0 = v[k], v1 = v[k + 1], v2 = v[k + 2], v3 = v[k + 3];
            apply(k, 4 * v0 - v[k - n] - v[k + n] - v[k - 1] - v1);
            apply(k + 1, 4 * v1 - v[k - n + 1] - v[k + n + 1] - v0 - v2);
            apply(k + 2, 4 * v2 - v[k - n + 2] - v[k + n + 2] - v1 - v3);
            apply(k + 3, 4 * v3 - v[k - n + 3] - v[k + n + 3] - v2 - v[k + 4]);
        }
        for (int j = nb, k = i * n + j; j < n - 1; j++, k++) apply(k, 4 * v[k] - v[k - n] - v[k + n] - v[k - 1] - v[k + 1]);
    }
    for (int j = 1; j < n - 1; j++) apply(j, 4 * v[j] - v[j + n] - v[j - 1] - v[j + 1]);
    for (int j = 1, k = (m - 1) * n + j; j < n - 1; j++, k++) apply(k, 4 * v[k] - v[k - n] - v[k - 1] - v[k + 1]);
    for (int i = 1, k = n; i < m - 1; i++, k += n) apply(k, 4 * v[k] - v[k - n] - v[k + n] - v[k + 1]);
    for (int i = 1, k = n + n - 1; i < m - 1; i++, k += n) apply(k, 4 * v[k] - v[k - n] - v[k + n] - v[k - 1]);
    apply(0, 4 * v[0] - v[1] - v[n]);
    apply(n - 1, 4 * v[n - 1] - v[n - 2] - v[2 * n - 1]);
    apply(idx_t(m - 1) * n, 4 * v[(m - 1) * n] - v[(m - 2) * n] - v[(m - 1) * n + 1]);
    apply(idx_t(m) * n - 1, 4 * v[m * n - 1] - v[m * n - 2] - v[m * n - n - 1]);
}

// ---- ell_matrix and sparse_banded -----------------------------------------

// ell_matrix_inserter's conversion (mtl/matrix/ell_matrix.hpp:146-176) of a
// sorted CSR: stride = num_rows rounded up to 32 (set_stride, :43-44), slots =
// longest row; slot i of row r at r + i * stride; padding repeats the row's
// last column index (0 for an empty row) with value 0.  With indices == NULL
// only slots and stride are returned.
void orc_ell_layout(idx_t nrows, const idx_t* S, const idx_t* I, const double* D, idx_t* slots_out,
                    idx_t* stride_out, idx_t* indices, double* data)
{
    const idx_t stride = (nrows + 31) / 32 * 32;
    idx_t slots = 0;
    for (idx_t r = 0; r < nrows; r++) slots = std::max(slots, S[r + 1] - S[r]);
    *slots_out = slots;
    *stride_out = stride;
    if (!indices) return;
    for (idx_t r = 0; r < nrows; r++) {
        idx_t k = r, patch_index = 0;
        for (idx_t j = S[r]; j < S[r + 1]; ++j, k += stride) {
            patch_index = indices[k] = I[j];
            data[k] = D[j];
        }
        for (idx_t j = 0; j < slots - (S[r + 1] - S[r]); ++j, k += stride) {
            indices[k] = patch_index;
            data[k] = 0.0;
        }
    }
}

// smat_cvec_mult(ell_matrix) (mtl/operation/mat_vec_mult.hpp:692-704);
// mode 0 assign, 1 +=, 2 -= (Assign::first_update).
void orc_ell_spmv(idx_t nrows, idx_t stride, idx_t slots, const idx_t* indices, const double* data,
                  const double* v, double* w, int mode)
{
    for (idx_t r = 0; r < nrows; ++r) {
        double s = 0.0;
        for (idx_t k = r, i = 0; i < slots; ++i, k += stride) s += data[k] * v[indices[k]];
        if (mode == 0) w[r] = s;
        else if (mode == 1) w[r] += s;
        else w[r] -= s;
    }
}

// sparse_banded_inserter (mtl/matrix/sparse_banded.hpp:168-258) fed with a
// CSR's entries in row order: a new offset is inserted at its sorted place
// (find_dia :150-156, insert_dia :188-198), update_store keeps the last value,
// the destructor lays the diagonals out row-major (:210-219).  With bands ==
// NULL only the band count is returned.
void orc_banded_layout(idx_t nrows, const idx_t* S, const idx_t* I, const double* D, idx_t* nb_out, idx_t* bands,
                       double* data)
{
    std::vector<idx_t> bs;
    std::vector<std::vector<double>> diagonals;
    for (idx_t r = 0; r < nrows; r++)
        for (idx_t j = S[r]; j < S[r + 1]; j++) {
            const idx_t dia = I[j] - r;
            size_t b = 0;
            while (b < bs.size() && dia > bs[b]) b++;
            if (b == bs.size() || dia != bs[b]) {
                diagonals.push_back(std::vector<double>(size_t(nrows), 0.0));
                bs.push_back(dia);
                for (size_t i = diagonals.size() - 1; i > b; i--) {
                    std::swap(diagonals[i - 1], diagonals[i]);
                    std::swap(bs[i - 1], bs[i]);
                }
            }
            diagonals[b][size_t(r)] = D[j];
        }
    *nb_out = idx_t(bs.size());
    if (!bands) return;
    for (size_t b = 0; b < bs.size(); b++) bands[b] = bs[b];
    double* p = data;
    for (idx_t r = 0; r < nrows; r++)
        for (size_t b = 0; b < bs.size(); b++) *p++ = diagonals[b][size_t(r)];
}

// smat_cvec_mult(sparse_banded) (mtl/operation/mat_vec_mult.hpp:709-809),
// the reference's segment loop: rows [from, to) share the band range
// [begin_pos, end_pos]; each row sums those bands from 0 in ascending order.
// Leading rows with no band and rows past the lowest band's end are set to 0
// by an assignment and not touched by += / -=.  (With += / -= and only
// strictly lower bands the reference reads begin_rows[nb]; this restatement
// steps back to the last band there, as the assignment branch does.)
void orc_banded_spmv(idx_t nr, idx_t nc, idx_t nb, const idx_t* bands, const double* data, const double* v,
                     double* w, int mode)
{
    const bool init_to_zero = mode == 0;
    auto update = [&](idx_t r, double t) {
        if (mode == 0) w[r] = t;
        else if (mode == 1) w[r] += t;
        else w[r] -= t;
    };
    if (nr == 0) return;
    if (nb == 0) {
        if (init_to_zero)
            for (idx_t r = 0; r < nr; r++) w[r] = 0.0;
        return;
    }
    std::vector<idx_t> begin_rows(static_cast<size_t>(nb)), end_rows(static_cast<size_t>(nb));
    for (idx_t b = 0; b < nb; b++) {
        begin_rows[size_t(b)] = std::max<idx_t>(0, -bands[b]);
        end_rows[size_t(b)] = std::min<idx_t>(nr, nc - bands[b]);
    }
    idx_t begin_pos = 0, end_pos = nb - 1;
    while (begin_pos < nb && begin_rows[size_t(begin_pos)] > 0) begin_pos++;
    if (begin_pos == nb) {
        --begin_pos;
        if (init_to_zero)
            for (idx_t r = 0; r < begin_rows[size_t(begin_pos)]; r++) w[r] = 0.0;
    }
    idx_t from = begin_rows[size_t(begin_pos)];
    while (begin_pos > 0 && begin_rows[size_t(begin_pos - 1)] == from) begin_pos--;
    for (bool active = true; active;) {
        const idx_t to = begin_pos > 0 && begin_rows[size_t(begin_pos - 1)] <= end_rows[size_t(end_pos)]
                             ? begin_rows[size_t(begin_pos - 1)]
                             : end_rows[size_t(end_pos)];
        for (idx_t r = from; r < to; r++) {
            double tmp = 0.0;
            const double* Ap = data + r * nb + begin_pos;
            for (idx_t b = begin_pos; b <= end_pos; ++b, ++Ap) tmp += *Ap * v[r + bands[b]];
            update(r, tmp);
        }
        if (begin_pos > 0) {
            if (begin_rows[size_t(begin_pos - 1)] == to) begin_pos--;
            if (end_rows[size_t(end_pos)] == to) end_pos--;
        } else {
            if (end_rows[0] == to) active = false;
            else end_pos--;
        }
        from = to;
    }
    if (end_rows[0] < nr && init_to_zero)
        for (idx_t r = end_rows[0]; r < nr; r++) w[r] = 0.0;
}

// mtl::seed<double>::counter (mtl/operation/random.hpp:29-38): every
// random(vector) call reseeds the C library generator with 17 * ++counter and
// fills the vector with rand() values.
static int orc_seed_counter = 0;
void orc_set_random_counter(int c) { orc_seed_counter = c; }
int orc_get_random_counter() { return orc_seed_counter; }

// itl::idr_s (idr_s.hpp:35-103); both preconditioners are ignored as in the
// reference.  P's columns are random(P) (column j: srand(17 * ++counter),
// rand()), column 0 replaced by r, then orth(P) (orth.hpp:35-54).  A singular
// M in lu_solve returns -2 (the reference throws matrix_singular).
int orc_idr_s(idx_t n, const idx_t* S, const idx_t* I, const double* D, double* x, const double* b,
              orc_iteration* it, int s)
{
    if (s < 1) s = 1;
    std::vector<double> v(n), t(n), q(n), r(n), tmp(n);
    std::vector<std::vector<double>> dR(s, std::vector<double>(n, 0.0)), dX(s, std::vector<double>(n, 0.0)),
        P(s, std::vector<double>(n, 0.0));
    std::vector<double> m(s), c(s), dm(s), M(size_t(s) * s, 0.0);
    orc_residual(n, S, I, D, x, b, r.data());
    for (int j = 0; j < s; j++) {
        srand(unsigned(17 * ++orc_seed_counter));
        for (idx_t e = 0; e < n; e++) P[j][e] = double(rand());
    }
    P[0] = r;
    for (int j = 0; j < s; j++) {
        for (int i = 0; i < j; i++) {
            const double d = dot4(n, P[i].data(), P[j].data());
            for (idx_t e = 0; e < n; e++) P[j][e] -= d * P[i][e];
        }
        const double nrm = two_norm(n, P[j].data());
        for (idx_t e = 0; e < n; e++) P[j][e] /= nrm;
    }
    double omega = 0.0;
    for (int k = 0; k < s; k++) {
        orc_spmv(n, S, I, D, r.data(), v.data(), 0);
        omega = dot4(n, v.data(), r.data()) / dot4(n, v.data(), v.data());
        for (idx_t e = 0; e < n; e++) dX[k][e] = omega * r[e];
        for (idx_t e = 0; e < n; e++) dR[k][e] = -omega * v[e];
        for (idx_t e = 0; e < n; e++) x[e] += dX[k][e];
        for (idx_t e = 0; e < n; e++) r[e] += dR[k][e];
        ++it->i;
        if (orc_finished(it, two_norm(n, r.data()))) return it->error;
        for (int i = 0; i < s; i++) M[size_t(i) * s + k] = dot4(n, P[i].data(), dR[k].data());
    }
    int oldest = 0;
    for (int i = 0; i < s; i++) m[i] = dot4(n, P[i].data(), r.data());
    while (!orc_finished(it, two_norm(n, r.data()))) {
        for (int k = 0; k < s; k++) {
            if (!dense_lu_solve(s, M, m.data(), c.data())) return -2;
            std::fill(q.begin(), q.end(), 0.0);
            for (int j = 0; j < s; j++)
                for (idx_t e = 0; e < n; e++) q[e] += dR[j][e] * -c[j];
            for (idx_t e = 0; e < n; e++) v[e] = r[e] + q[e];
            std::fill(tmp.begin(), tmp.end(), 0.0);   // dX * c
            for (int j = 0; j < s; j++)
                for (idx_t e = 0; e < n; e++) tmp[e] += dX[j][e] * c[j];
            if (k == 0) {
                orc_spmv(n, S, I, D, v.data(), t.data(), 0);
                omega = dot4(n, t.data(), v.data()) / dot4(n, t.data(), t.data());
                for (idx_t e = 0; e < n; e++) dR[oldest][e] = q[e] - omega * t[e];
                for (idx_t e = 0; e < n; e++) dX[oldest][e] = omega * v[e] - tmp[e];
            } else {
                for (idx_t e = 0; e < n; e++) dX[oldest][e] = omega * v[e] - tmp[e];
                for (idx_t e = 0; e < n; e++) tmp[e] = -dX[oldest][e];
                orc_spmv(n, S, I, D, tmp.data(), dR[oldest].data(), 0);
            }
            for (idx_t e = 0; e < n; e++) r[e] += dR[oldest][e];
            for (idx_t e = 0; e < n; e++) x[e] += dX[oldest][e];
            ++it->i;
            if (orc_finished(it, two_norm(n, r.data()))) return it->error;
            for (int i = 0; i < s; i++) dm[i] = dot4(n, P[i].data(), dR[oldest].data());
            for (int i = 0; i < s; i++) M[size_t(i) * s + oldest] = dm[i];
            for (int i = 0; i < s; i++) m[i] += dm[i];
            oldest = (oldest + 1) % s;
        }
    }
    return it->error;
}

}  // extern "C"
