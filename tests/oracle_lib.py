"""ctypes view of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: used by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg only.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# M4G_ORACLE_LIBRARY selects another build of the oracle (e.g. a sanitized one)
LIB_PATH = os.environ.get("M4G_ORACLE_LIBRARY") or os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64

PC_IDENTITY, PC_DIAGONAL, PC_IC0, PC_ILU0 = 0, 1, 2, 3
OK, MAX_ITER, ERR_NOT_SQUARE, ERR_MISSING_DIAGONAL, ERR_ORTHOGONALITY = 0, 1, 2, 3, 4


class Iteration(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("norm_r0", ctypes.c_double), ("i", ctypes.c_int), ("error", ctypes.c_int),
                ("is_finished", ctypes.c_int), ("resid", ctypes.c_double),
                ("history", ctypes.POINTER(ctypes.c_double)), ("history_cap", ctypes.c_int)]


class PC(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("n", ctypes.c_int64), ("inv_diag", ctypes.c_void_p),
                ("LS", ctypes.c_void_p), ("LI", ctypes.c_void_p), ("LD", ctypes.c_void_p),
                ("US", ctypes.c_void_p), ("UI", ctypes.c_void_p), ("UD", ctypes.c_void_p)]


def _load():
    if not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "orc_laplacian_nnz": (_i64, [ctypes.c_int, ctypes.c_int]),
        "orc_laplacian_setup": (None, [ctypes.c_int, ctypes.c_int, _i64p, _i64p, _f64p]),
        "orc_spmv": (None, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.c_int]),
        "orc_residual": (None, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, _f64p]),
        "orc_dot": (ctypes.c_double, [_i64, _f64p, _f64p]),
        "orc_two_norm": (ctypes.c_double, [_i64, _f64p]),
        "orc_unary_dot": (ctypes.c_double, [_i64, _f64p]),
        "orc_fused_spmv_dot": (ctypes.c_double, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p]),
        "orc_fused_axpy_unary_dot": (ctypes.c_double, [_i64, ctypes.c_double, _f64p, _f64p]),
        "orc_upper_nnz": (_i64, [_i64, _i64p, _i64p]),
        "orc_strict_lower_nnz": (_i64, [_i64, _i64p, _i64p]),
        "orc_ic0_factor": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p]),
        "orc_csr_transpose": (None, [_i64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p]),
        "orc_ic0_solve": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p, _f64p, _f64p]),
        "orc_ic0_solve_f32": (ctypes.c_int, [_i64, _i64p, _i64p, _f32p, _i64p, _i64p, _f32p, _f64p, _f64p]),
        "orc_ilu0_factor": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p]),
        "orc_ilu0_solve": (None, [_i64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p, _f64p, _f64p]),
        "orc_ilu0_solve_f32": (None, [_i64, _i64p, _i64p, _f32p, _i64p, _i64p, _f32p, _f64p, _f64p]),
        "orc_ilu0_adjoint_solve": (None, [_i64, _i64p, _i64p, _f64p, _i64p, _i64p, _f64p, _f64p, _f64p]),
        "orc_diagonal_setup": (None, [_i64, _i64p, _i64p, _f64p, _f64p]),
        "orc_cg": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC), ctypes.POINTER(Iteration)]),
        "orc_bicgstab": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC), ctypes.POINTER(Iteration)]),
        "orc_cgs": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC), ctypes.POINTER(Iteration)]),
        "orc_bicg": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC), ctypes.POINTER(Iteration)]),
        "orc_tfqmr": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC), ctypes.POINTER(PC),
                                     ctypes.POINTER(Iteration)]),
        "orc_qmr": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC), ctypes.POINTER(PC),
                                   ctypes.POINTER(Iteration)]),
        "orc_gmres": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC), ctypes.POINTER(PC),
                                     ctypes.POINTER(Iteration), ctypes.c_int]),
        "orc_gmres_full": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC),
                                          ctypes.POINTER(PC), ctypes.POINTER(Iteration)]),
        "orc_bicgstab_ell": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(PC),
                                            ctypes.POINTER(PC), ctypes.POINTER(Iteration), ctypes.c_int]),
        "orc_idr_s": (ctypes.c_int, [_i64, _i64p, _i64p, _f64p, _f64p, _f64p, ctypes.POINTER(Iteration), ctypes.c_int]),
        "orc_set_random_counter": (None, [ctypes.c_int]),
        "orc_assemble": (_i64, [_i64, _i64, _i64p, _i64p, _f64p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
        "orc_poisson2d_dirichlet_mult": (None, [ctypes.c_int, ctypes.c_int, _f64p, _f64p, ctypes.c_int]),
        "orc_ell_layout": (None, [_i64, _i64p, _i64p, _f64p, ctypes.POINTER(ctypes.c_int64),
                                  ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_void_p]),
        "orc_ell_spmv": (None, [_i64, _i64, _i64, _i64p, _f64p, _f64p, _f64p, ctypes.c_int]),
        "orc_banded_layout": (None, [_i64, _i64p, _i64p, _f64p, ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p,
                                     ctypes.c_void_p]),
        "orc_banded_spmv": (None, [_i64, _i64, _i64, _i64p, _f64p, _f64p, _f64p, ctypes.c_int]),
        "orc_get_random_counter": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def omp_ic0_pcg(A, U, L, b, rtol, max_iter, threads):
    """IC(0)-PCG with MTL_WITH_OPENMP semantics (oracle/cpu_openmp.cpp), a CPU
    timing baseline: returns (x, iterations)."""
    path = os.path.join(ROOT, "oracle", "_build", "liboracle_omp.so")
    if not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    olib = ctypes.CDLL(path)
    f = olib.ompc_ic0_pcg
    f.restype = ctypes.c_int
    f.argtypes = [_i64] + [_i64p, _i64p, _f64p] * 3 + [_f64p, _f64p, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                      ctypes.POINTER(ctypes.c_int)]
    x = np.zeros(A.n)
    its = ctypes.c_int(0)
    f(A.n, A.starts, A.indices, A.data, U.starts, U.indices, U.data, L.starts, L.indices, L.data,
      np.ascontiguousarray(b, np.float64), x, rtol, int(max_iter), int(threads), ctypes.byref(its))
    return x, its.value


class CSR:
    """Host CSR triple with int64 indices (MTL4's default size_type)."""

    def __init__(self, starts, indices, data):
        self.starts = np.ascontiguousarray(starts, dtype=np.int64)
        self.indices = np.ascontiguousarray(indices, dtype=np.int64)
        self.data = np.ascontiguousarray(data)
        self.n = len(self.starts) - 1

    @property
    def nnz(self):
        return int(self.starts[-1])

    def dense(self):
        A = np.zeros((self.n, self.n), dtype=np.float64)
        for r in range(self.n):
            for j in range(self.starts[r], self.starts[r + 1]):
                A[r, self.indices[j]] = self.data[j]
        return A


def laplacian_setup(m, n):
    nnz = lib.orc_laplacian_nnz(m, n)
    S = np.zeros(m * n + 1, np.int64)
    I = np.zeros(nnz, np.int64)
    D = np.zeros(nnz, np.float64)
    lib.orc_laplacian_setup(m, n, S, I, D)
    return CSR(S, I, D)


def spmv(A, v, w=None, mode=0):
    v = np.ascontiguousarray(v, np.float64)
    w = np.zeros(A.n) if w is None else np.array(w, np.float64)
    lib.orc_spmv(A.n, A.starts, A.indices, A.data, v, w, mode)
    return w


def residual(A, x, b):
    r = np.zeros(A.n)
    lib.orc_residual(A.n, A.starts, A.indices, A.data, np.ascontiguousarray(x, np.float64),
                     np.ascontiguousarray(b, np.float64), r)
    return r


def dot(a, b):
    a = np.ascontiguousarray(a, np.float64)
    return lib.orc_dot(len(a), a, np.ascontiguousarray(b, np.float64))


def two_norm(a):
    a = np.ascontiguousarray(a, np.float64)
    return lib.orc_two_norm(len(a), a)


def unary_dot(a):
    a = np.ascontiguousarray(a, np.float64)
    return lib.orc_unary_dot(len(a), a)


def fused_spmv_dot(A, p):
    q = np.zeros(A.n)
    d = lib.orc_fused_spmv_dot(A.n, A.starts, A.indices, A.data, np.ascontiguousarray(p, np.float64), q)
    return q, d


def fused_axpy_unary_dot(alpha, q, r):
    r = np.array(r, np.float64)
    d = lib.orc_fused_axpy_unary_dot(len(r), alpha, np.ascontiguousarray(q, np.float64), r)
    return r, d


def csr_transpose(A):
    TS = np.zeros(A.n + 1, np.int64)
    TI = np.zeros(A.nnz, np.int64)
    TD = np.zeros(A.nnz, np.float64)
    lib.orc_csr_transpose(A.n, A.starts, A.indices, np.ascontiguousarray(A.data, np.float64), TS, TI, TD)
    return CSR(TS, TI, TD)


def ic0_factor(A):
    """Returns (status, U, L) with U = IC(0) upper factor (inverted diagonal first), L = trans(U)."""
    nu = lib.orc_upper_nnz(A.n, A.starts, A.indices)
    US = np.zeros(A.n + 1, np.int64)
    UI = np.zeros(nu, np.int64)
    UD = np.zeros(nu, np.float64)
    st = lib.orc_ic0_factor(A.n, A.starts, A.indices, A.data, US, UI, UD)
    U = CSR(US, UI, UD)
    return st, U, (csr_transpose(U) if st == OK else None)


def ic0_solve(L, U, x, factor_dtype=np.float64):
    y = np.zeros(L.n)
    x = np.ascontiguousarray(x, np.float64)
    if factor_dtype == np.float32:
        st = lib.orc_ic0_solve_f32(L.n, L.starts, L.indices, L.data.astype(np.float32), U.starts, U.indices,
                                   U.data.astype(np.float32), x, y)
    else:
        st = lib.orc_ic0_solve(L.n, L.starts, L.indices, L.data, U.starts, U.indices, U.data, x, y)
    assert st == OK
    return y


def ilu0_factor(A):
    nl = lib.orc_strict_lower_nnz(A.n, A.starts, A.indices)
    nu = A.nnz - nl
    LS = np.zeros(A.n + 1, np.int64); LI = np.zeros(nl, np.int64); LD = np.zeros(nl)
    US = np.zeros(A.n + 1, np.int64); UI = np.zeros(nu, np.int64); UD = np.zeros(nu)
    st = lib.orc_ilu0_factor(A.n, A.starts, A.indices, A.data, LS, LI, LD, US, UI, UD)
    return st, CSR(LS, LI, LD), CSR(US, UI, UD)


def ilu0_solve(L, U, x, factor_dtype=np.float64):
    y = np.zeros(L.n)
    x = np.ascontiguousarray(x, np.float64)
    if factor_dtype == np.float32:
        lib.orc_ilu0_solve_f32(L.n, L.starts, L.indices, L.data.astype(np.float32), U.starts, U.indices,
                               U.data.astype(np.float32), x, y)
    else:
        lib.orc_ilu0_solve(L.n, L.starts, L.indices, L.data, U.starts, U.indices, U.data, x, y)
    return y


def ilu0_adjoint_solve(L, U, x):
    y = np.zeros(L.n)
    lib.orc_ilu0_adjoint_solve(L.n, L.starts, L.indices, L.data, U.starts, U.indices, U.data,
                               np.ascontiguousarray(x, np.float64), y)
    return y


def make_pc(A, kind):
    """Builds an oracle PC; returns (PC struct, keepalive list)."""
    P = PC()
    P.kind = kind
    P.n = A.n
    keep = []
    if kind == PC_DIAGONAL:
        inv = np.zeros(A.n)
        lib.orc_diagonal_setup(A.n, A.starts, A.indices, A.data, inv)
        keep.append(inv)
        P.inv_diag = inv.ctypes.data
    elif kind in (PC_IC0, PC_ILU0):
        if kind == PC_IC0:
            st, U, L = ic0_factor(A)
        else:
            st, L, U = ilu0_factor(A)
        assert st == OK, st
        keep += [L, U]
        P.LS, P.LI, P.LD = L.starts.ctypes.data, L.indices.ctypes.data, L.data.ctypes.data
        P.US, P.UI, P.UD = U.starts.ctypes.data, U.indices.ctypes.data, U.data.ctypes.data
    return P, keep


def solve(method, A, b, x0=None, pc=PC_IC0, max_iter=1000, rtol=1e-10, atol=0.0, norm_r0=None,
          history=True, right=None, param=None, prebuilt=None):
    """Runs itl::cg / bicgstab / cgs / bicg / tfqmr / qmr / gmres / gmres_full /
    bicgstab_2 / bicgstab_ell / idr_s on the oracle; `pc` is the (left)
    preconditioner kind, `right` the right one (None: identity), `param` the
    restart / l / s argument; `prebuilt` a make_pc() result to use instead of
    building the left preconditioner here (so a timing can leave the
    factorisation out).  Returns dict(x, status, iterations, resid, history)."""
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(A.n) if x0 is None else np.array(x0, np.float64)
    P, keep = make_pc(A, pc) if prebuilt is None else prebuilt
    it = Iteration()
    it.max_iter = max_iter
    it.rtol, it.atol = rtol, atol
    it.norm_r0 = abs(two_norm(b)) if norm_r0 is None else norm_r0
    hist = np.full(max_iter + 2, np.nan)
    if history:
        it.history = hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        it.history_cap = len(hist)
    R, keep_r = (None, []) if right is None else make_pc(A, right)
    Rp = None if R is None else ctypes.byref(R)
    args = (A.n, A.starts, A.indices, A.data, x, b)
    if method == "tfqmr":
        st = lib.orc_tfqmr(*args, ctypes.byref(P), Rp, ctypes.byref(it))
    elif method == "qmr":
        st = lib.orc_qmr(*args, ctypes.byref(P), Rp, ctypes.byref(it))
    elif method == "gmres":
        st = lib.orc_gmres(*args, ctypes.byref(P), Rp, ctypes.byref(it), 8 if param is None else param)
    elif method == "gmres_full":
        st = lib.orc_gmres_full(*args, ctypes.byref(P), Rp, ctypes.byref(it))
    elif method == "bicgstab_2":
        st = lib.orc_bicgstab_ell(*args, ctypes.byref(P), None, ctypes.byref(it), 2)
    elif method == "bicgstab_ell":
        st = lib.orc_bicgstab_ell(*args, ctypes.byref(P), Rp, ctypes.byref(it), 8 if param is None else param)
    elif method == "idr_s":
        st = lib.orc_idr_s(*args, ctypes.byref(it), 8 if param is None else param)
    else:
        f = {"cg": lib.orc_cg, "bicgstab": lib.orc_bicgstab, "cgs": lib.orc_cgs, "bicg": lib.orc_bicg}[method]
        st = f(A.n, A.starts, A.indices, A.data, x, b, ctypes.byref(P), ctypes.byref(it))
    del keep, keep_r
    return dict(x=x, status=st, iterations=it.i, resid=it.resid, history=hist[: it.i + 1])


def ell_layout(A):
    """ell_matrix_inserter's layout of CSR A: (slots, stride, indices, data)."""
    slots, stride = ctypes.c_int64(), ctypes.c_int64()
    lib.orc_ell_layout(A.n, A.starts, A.indices, A.data, ctypes.byref(slots), ctypes.byref(stride), None, None)
    total = slots.value * stride.value
    I = np.zeros(total, np.int64)
    D = np.zeros(total, np.float64)
    lib.orc_ell_layout(A.n, A.starts, A.indices, A.data, ctypes.byref(slots), ctypes.byref(stride),
                       I.ctypes.data, D.ctypes.data)
    return slots.value, stride.value, I, D


def ell_spmv(nrows, slots, stride, I, D, v, w=None, mode=0):
    w = np.zeros(nrows) if w is None else np.array(w, np.float64)
    lib.orc_ell_spmv(nrows, stride, slots, np.ascontiguousarray(I, np.int64), np.ascontiguousarray(D, np.float64),
                     np.ascontiguousarray(v, np.float64), w, mode)
    return w


def banded_layout(A):
    """sparse_banded_inserter's layout of CSR A: (bands, data[r * nb + b])."""
    nb = ctypes.c_int64()
    lib.orc_banded_layout(A.n, A.starts, A.indices, A.data, ctypes.byref(nb), None, None)
    B = np.zeros(nb.value, np.int64)
    D = np.zeros(A.n * nb.value, np.float64)
    lib.orc_banded_layout(A.n, A.starts, A.indices, A.data, ctypes.byref(nb), B.ctypes.data, D.ctypes.data)
    return B, D


def banded_spmv(nrows, ncols, B, D, v, w=None, mode=0):
    w = np.zeros(nrows) if w is None else np.array(w, np.float64)
    lib.orc_banded_spmv(nrows, ncols, len(B), np.ascontiguousarray(B, np.int64), np.ascontiguousarray(D, np.float64),
                        np.ascontiguousarray(v, np.float64), w, mode)
    return w


def assemble(nrows, rows, cols, vals, updater=0, existing=None):
    """mat::inserter<compressed2D, Updater> fed with the triplets in order,
    starting from `existing` (a CSR or None); returns a CSR."""
    rows = np.ascontiguousarray(rows, np.int64)
    cols = np.ascontiguousarray(cols, np.int64)
    vals = np.ascontiguousarray(vals, np.float64)
    ex = (None, None, None)
    if existing is not None:
        ex = (existing.starts.ctypes.data, existing.indices.ctypes.data,
              np.ascontiguousarray(existing.data, np.float64).ctypes.data)
    keep = existing
    nnz = lib.orc_assemble(nrows, len(rows), rows, cols, vals, updater, *ex, None, None, None)
    S = np.zeros(nrows + 1, np.int64)
    I = np.zeros(nnz, np.int64)
    D = np.zeros(nnz, np.float64)
    lib.orc_assemble(nrows, len(rows), rows, cols, vals, updater, *ex, S.ctypes.data, I.ctypes.data, D.ctypes.data)
    del keep
    return CSR(S, I, D)


def poisson2d_dirichlet_mult(m, n, v, w=None, mode=0):
    """poisson2D_dirichlet(m, n).mult(v, w, Assign) on the oracle."""
    w = np.zeros(m * n) if w is None else np.array(w, np.float64)
    lib.orc_poisson2d_dirichlet_mult(m, n, np.ascontiguousarray(v, np.float64), w, mode)
    return w
