"""Python binding of libmtl4gpu (the MI355X sparse iterative-solve library).

A thin ctypes layer over the C ABI in include/mtl4gpu.h, used by the tests,
the benchmark and the torch.autograd test binding.  It mirrors the
reference's objects: ``Matrix`` is a device-resident ``mtl::compressed2D``,
``Vector`` a device-resident ``mtl::dense_vector``, ``PC`` one of
``itl::pc::{identity, diagonal, ic_0, ilu_0}``, and ``cg`` / ``bicgstab`` are
``itl::cg`` / ``itl::bicgstab`` driven by an ``Iteration`` (``basic_iteration``).

Importing this module fails loudly when the HIP library has not been built;
there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# M4G_LIBRARY selects another build of the library (e.g. a diagnostic one)
LIB_PATH = os.environ.get("M4G_LIBRARY") or os.path.join(_HERE, "lib", "libmtl4gpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"libmtl4gpu.so not built ({LIB_PATH}); run __graft_entry__.build()")

# One HIP runtime per process.  PyTorch for ROCm ships its own copy of the HIP
# runtime (same soname, libamdhip64.so.7), and two copies in one process do not
# share the GPU.  Importing torch first makes the dynamic linker resolve this
# library's libamdhip64.so.7 to torch's copy, so tensors and this library's
# kernels run on one runtime.  Without torch the system runtime is used.
try:
    import torch  # noqa: F401
except ImportError:
    pass

_lib = ctypes.CDLL(LIB_PATH)

# ---- constants (include/mtl4gpu.h) ------------------------------------------
OK, MAX_ITER = 0, 1
ERR_NOT_SQUARE, ERR_MISSING_DIAGONAL, ERR_ORTHOGONALITY, ERR_LOGIC, ERR_SIZE, ERR_INVALID, ERR_INTERRUPTED = range(2, 9)
ERR_FILE, ERR_FORMAT, BREAKDOWN, ERR_SINGULAR = 9, 10, 11, 12
ERR_HIP, ERR_NO_DEVICE, ERR_TIMEOUT = -1, -2, -3
F64, F32 = 0, 1
IDX_I32, IDX_I64, IDX_U32 = 0, 1, 2
ASSIGN, PLUS, MINUS = 0, 1, 2
PC_IDENTITY, PC_DIAGONAL, PC_IC0, PC_ILU0 = 0, 1, 2, 3
UPDATE_STORE, UPDATE_PLUS, UPDATE_MINUS, UPDATE_TIMES = 0, 1, 2, 3
METHOD_CG, METHOD_BICGSTAB = 0, 1

_c_i64 = ctypes.c_int64
_c_i64p = ctypes.POINTER(ctypes.c_int64)
_c_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


class PcStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "levels_lower", "levels_upper", "tiles_lower", "tiles_upper", "steps_lower", "steps_upper",
        "grid_nx", "grid_ny", "grid_nz", "tile_nx", "tile_ny", "tile_nz")]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Iteration(ctypes.Structure):
    """m4g_iteration: the state of itl::basic_iteration (basic_iteration.hpp:24-165)."""
    _fields_ = [("max_iter", ctypes.c_int), ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("norm_r0", ctypes.c_double), ("iterations", ctypes.c_int), ("error", ctypes.c_int),
                ("resid", ctypes.c_double), ("history", _c_dp), ("history_capacity", ctypes.c_int),
                ("interrupt", ctypes.POINTER(ctypes.c_int))]


class RefineInfo(ctypes.Structure):
    """m4g_refine_info of m4g_refined_solve."""
    _fields_ = [("outer_iterations", ctypes.c_int), ("inner_iterations", ctypes.c_int),
                ("resid", ctypes.c_double), ("inner_error", ctypes.c_int)]


_SIGS = {
    "m4g_last_error": (ctypes.c_char_p, []),
    "m4g_version": (ctypes.c_char_p, []),
    "m4g_set_device": (ctypes.c_int, [ctypes.c_int]),
    "m4g_set_stream": (ctypes.c_int, [_vp]),
    "m4g_synchronize": (ctypes.c_int, []),
    "m4g_get_stream": (_vp, []),
    "m4g_pc_apply_async": (ctypes.c_int, [_vp, _vp, _vp]),
    "m4g_matrix_create": (ctypes.c_int, [ctypes.POINTER(_vp), _c_i64, _c_i64, _c_i64, _vp, _vp, _vp, ctypes.c_int,
                                         ctypes.c_int]),
    "m4g_matrix_destroy": (ctypes.c_int, [_vp]),
    "m4g_matrix_dims": (ctypes.c_int, [_vp, _c_i64p, _c_i64p, _c_i64p]),
    "m4g_matrix_download": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "m4g_matrix_set_values": (ctypes.c_int, [_vp, _vp, ctypes.c_int]),
    "m4g_vector_create": (ctypes.c_int, [ctypes.POINTER(_vp), _c_i64, ctypes.c_int]),
    "m4g_vector_wrap": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, _c_i64, ctypes.c_int]),
    "m4g_vector_destroy": (ctypes.c_int, [_vp]),
    "m4g_vector_size": (ctypes.c_int, [_vp, _c_i64p]),
    "m4g_vector_device_ptr": (_vp, [_vp]),
    "m4g_vector_upload": (ctypes.c_int, [_vp, _vp]),
    "m4g_vector_download": (ctypes.c_int, [_vp, _vp]),
    "m4g_vector_fill": (ctypes.c_int, [_vp, ctypes.c_double]),
    "m4g_vector_copy": (ctypes.c_int, [_vp, _vp]),
    "m4g_dot": (ctypes.c_int, [_vp, _vp, _c_dp]),
    "m4g_dot_noalloc": (ctypes.c_int, [_vp, _vp, _c_dp]),
    "m4g_two_norm": (ctypes.c_int, [_vp, _c_dp]),
    "m4g_unary_dot": (ctypes.c_int, [_vp, _c_dp]),
    "m4g_axpy": (ctypes.c_int, [ctypes.c_double, _vp, _vp]),
    "m4g_xpay": (ctypes.c_int, [_vp, ctypes.c_double, _vp]),
    "m4g_axpy_unary_dot": (ctypes.c_int, [ctypes.c_double, _vp, _vp, _c_dp]),
    "m4g_spmv": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "m4g_residual": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "m4g_spmv_dot": (ctypes.c_int, [_vp, _vp, _vp, _c_dp]),
    "m4g_pc_create": (ctypes.c_int, [ctypes.POINTER(_vp), _vp, ctypes.c_int, ctypes.c_int]),
    "m4g_pc_destroy": (ctypes.c_int, [_vp]),
    "m4g_pc_solve": (ctypes.c_int, [_vp, _vp, _vp]),
    "m4g_pc_adjoint_solve": (ctypes.c_int, [_vp, _vp, _vp]),
    "m4g_pc_factor_nnz": (ctypes.c_int, [_vp, ctypes.c_int, _c_i64p]),
    "m4g_pc_download_factor": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    "m4g_pc_stats_get": (ctypes.c_int, [_vp, ctypes.POINTER(PcStats)]),
    "m4g_cg": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_bicgstab": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_pc_refactor": (ctypes.c_int, [_vp, _vp]),
    "m4g_matrix_transpose": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "m4g_pattern_gradient": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "m4g_refined_solve": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.POINTER(Iteration), ctypes.c_double,
                                         ctypes.c_int, ctypes.POINTER(RefineInfo)]),
    "m4g_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp), _c_i64, _vp, _vp, _vp, _vp]),
    "m4g_batch_destroy": (ctypes.c_int, [_vp]),
    "m4g_batch_dims": (ctypes.c_int, [_vp, _c_i64p, _c_i64p, _c_i64p]),
    "m4g_batch_factor_ic0": (ctypes.c_int, [_vp]),
    "m4g_batch_download_factor": (ctypes.c_int, [_vp, _vp]),
    "m4g_batch_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp]),
    "m4g_mm_read": (ctypes.c_int, [ctypes.c_char_p, _vp, _vp, _vp, _vp, _vp, _vp]),
    "m4g_cgs": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_bicg": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_tfqmr": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_qmr": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_gmres_full": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_gmres": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Iteration), _c_i64]),
    "m4g_bicgstab_2": (ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.POINTER(Iteration)]),
    "m4g_bicgstab_ell": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Iteration), _c_i64]),
    "m4g_idr_s": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.POINTER(Iteration), _c_i64]),
    "m4g_random_seed_counter": (ctypes.c_int, []),
    "m4g_matrix_assemble": (ctypes.c_int, [ctypes.POINTER(_vp), _c_i64, _c_i64, _c_i64, _vp, _vp, _vp, ctypes.c_int,
                                           ctypes.c_int, _vp]),
    # alternative formats (mtl4gpu_formats.h)
    "m4g_ell_create": (ctypes.c_int, [ctypes.POINTER(_vp), _c_i64, _c_i64, _vp, _vp, _vp]),
    "m4g_ell_destroy": (ctypes.c_int, [_vp]),
    "m4g_ell_shape": (ctypes.c_int, [_vp, _c_i64p, _c_i64p, _c_i64p, _c_i64p, _c_i64p]),
    "m4g_ell_download": (ctypes.c_int, [_vp, _vp, _vp]),
    "m4g_ell_spmv": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "m4g_banded_create": (ctypes.c_int, [ctypes.POINTER(_vp), _c_i64, _c_i64, _c_i64, _vp, _vp]),
    "m4g_banded_from_csr": (ctypes.c_int, [ctypes.POINTER(_vp), _c_i64, _c_i64, _vp, _vp, _vp]),
    "m4g_banded_destroy": (ctypes.c_int, [_vp]),
    "m4g_banded_shape": (ctypes.c_int, [_vp, _c_i64p, _c_i64p, _c_i64p]),
    "m4g_banded_download": (ctypes.c_int, [_vp, _vp, _vp]),
    "m4g_banded_spmv": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "m4g_poisson2d_dirichlet_mult": (ctypes.c_int, [_c_i64, _c_i64, _vp, _vp, ctypes.c_int]),
    "m4g_set_random_seed_counter": (None, [ctypes.c_int]),
    "m4g_mm_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, _vp, _vp, _vp]),
    "m4g_batch_cg_ic0": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        _vp, ctypes.c_int, _vp, _vp, _vp]),
    # host-only analysis (mtl4gpu_analysis.h)
    "m4g_analysis_grid": (ctypes.c_int, [_c_i64, _vp, _vp, _c_i64p, _c_i64p, _c_i64p]),
    "m4g_analysis_pc_solve": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _c_i64, _vp, _vp, _vp, ctypes.c_int,
                                             ctypes.c_int, _vp, _vp, _c_i64, ctypes.POINTER(PcStats)]),
    # generators (mtl4gpu_gen.h)
    "m4g_gen_laplacian2d": (_c_i64, [_c_i64, _c_i64, _vp, _vp, _vp]),
    "m4g_gen_poisson3d": (_c_i64, [_c_i64, _c_i64, _c_i64, _vp, _vp, _vp]),
    "m4g_gen_convdiff27": (_c_i64, [_c_i64, _c_i64, _c_i64, ctypes.c_uint64, ctypes.c_double, _vp, _vp, _vp]),
    "m4g_gen_varcoef5": (_c_i64, [_c_i64, _c_i64, ctypes.c_uint64, _vp, _vp, _vp]),
    "m4g_gen_edge_laplacian": (_c_i64, [_c_i64, _c_i64, _vp, _c_i64p, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "m4g_rng_uniform": (None, [ctypes.c_uint64, ctypes.c_uint64, _c_i64, _c_i64, ctypes.c_double, ctypes.c_double,
                               _vp]),
    "m4g_rng_normal": (None, [ctypes.c_uint64, ctypes.c_uint64, _c_i64, _c_i64, _vp]),
    "m4g_rng_lognormal": (None, [ctypes.c_uint64, ctypes.c_uint64, _c_i64, _c_i64, _vp]),
    "m4g_rng_int": (None, [ctypes.c_uint64, ctypes.c_uint64, _c_i64, _c_i64, _c_i64, _c_i64, _vp]),
}


def _bind(name, restype, argtypes):
    f = getattr(_lib, name, None)
    if f is None:
        return None
    f.restype = restype
    f.argtypes = argtypes
    return f


_f = {name: _bind(name, *sig) for name, sig in _SIGS.items()}
globals().update({"_" + k: v for k, v in _f.items()})


class M4GError(RuntimeError):
    def __init__(self, status, where):
        msg = _f["m4g_last_error"]().decode(errors="replace")
        super().__init__(f"{where} failed with status {status}: {msg}")
        self.status = status


def _check(status, where, allow=(OK,)):
    if status not in allow:
        raise M4GError(status, where)
    return status


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _dtype_code(dt):
    dt = np.dtype(dt)
    if dt == np.float64:
        return F64
    if dt == np.float32:
        return F32
    raise TypeError(f"unsupported dtype {dt}")


def _np_dtype(code):
    return np.float64 if code == F64 else np.float32


def version():
    return _f["m4g_version"]().decode()


def synchronize():
    _check(_f["m4g_synchronize"](), "m4g_synchronize")


def get_stream():
    return _f["m4g_get_stream"]()


def set_stream(stream_ptr):
    _check(_f["m4g_set_stream"](ctypes.c_void_p(stream_ptr) if stream_ptr else None), "m4g_set_stream")


# ---------------------------------------------------------------------------
# handles
# ---------------------------------------------------------------------------
class Ell:
    """Device-resident mtl::mat::ell_matrix<double> built from a sorted CSR
    (the ell_matrix_inserter's layout)."""

    def __init__(self, starts, indices, values, ncols=None):
        S = np.ascontiguousarray(starts, np.int64)
        I = np.ascontiguousarray(indices, np.int64)
        V = np.ascontiguousarray(values, np.float64)
        self.nrows = len(S) - 1
        self.ncols = self.nrows if ncols is None else int(ncols)
        h = ctypes.c_void_p()
        _check(_f["m4g_ell_create"](ctypes.byref(h), self.nrows, self.ncols, _ptr(S), _ptr(I), _ptr(V)),
               "m4g_ell_create")
        self.h = h
        a = [ctypes.c_int64() for _ in range(5)]
        _check(_f["m4g_ell_shape"](self.h, *[ctypes.byref(x) for x in a]), "m4g_ell_shape")
        self.slots, self.stride, self.nnz = a[2].value, a[3].value, a[4].value

    def download(self):
        """(ref_minor(), ref_data()): stride * slots entries each."""
        n = self.stride * self.slots
        I = np.zeros(n, np.int64)
        D = np.zeros(n, np.float64)
        _check(_f["m4g_ell_download"](self.h, _ptr(I), _ptr(D)), "m4g_ell_download")
        return I, D

    def spmv(self, v, w, mode=ASSIGN):
        _check(_f["m4g_ell_spmv"](self.h, v.h, w.h, mode), "m4g_ell_spmv")

    def close(self):
        if getattr(self, "h", None):
            _f["m4g_ell_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Banded:
    """Device-resident mtl::mat::sparse_banded<double>: from a CSR (the
    inserter) or, with bands=..., from the reference layout data[r * nb + b]."""

    def __init__(self, starts=None, indices=None, values=None, nrows=None, ncols=None, bands=None, data=None):
        h = ctypes.c_void_p()
        if bands is not None:
            B = np.ascontiguousarray(bands, np.int64)
            D = np.ascontiguousarray(data, np.float64)
            _check(_f["m4g_banded_create"](ctypes.byref(h), int(nrows), int(ncols), len(B), _ptr(B), _ptr(D)),
                   "m4g_banded_create")
        else:
            S = np.ascontiguousarray(starts, np.int64)
            I = np.ascontiguousarray(indices, np.int64)
            V = np.ascontiguousarray(values, np.float64)
            nr = len(S) - 1
            _check(_f["m4g_banded_from_csr"](ctypes.byref(h), nr, nr if ncols is None else int(ncols), _ptr(S),
                                             _ptr(I), _ptr(V)), "m4g_banded_from_csr")
        self.h = h
        a = [ctypes.c_int64() for _ in range(3)]
        _check(_f["m4g_banded_shape"](self.h, *[ctypes.byref(x) for x in a]), "m4g_banded_shape")
        self.nrows, self.ncols, self.nbands = a[0].value, a[1].value, a[2].value

    def download(self):
        """(ref_bands(), data) in the reference's row-major layout."""
        B = np.zeros(self.nbands, np.int64)
        D = np.zeros(self.nrows * self.nbands, np.float64)
        _check(_f["m4g_banded_download"](self.h, _ptr(B), _ptr(D)), "m4g_banded_download")
        return B, D

    def spmv(self, v, w, mode=ASSIGN):
        _check(_f["m4g_banded_spmv"](self.h, v.h, w.h, mode), "m4g_banded_spmv")

    def close(self):
        if getattr(self, "h", None):
            _f["m4g_banded_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson2d_dirichlet_mult(m, n, v, w, mode=ASSIGN):
    """mtl::mat::poisson2D_dirichlet(m, n).mult(v, w, Assign): the matrix-free 5-point operator."""
    _check(_f["m4g_poisson2d_dirichlet_mult"](int(m), int(n), v.h, w.h, mode), "m4g_poisson2d_dirichlet_mult")


class Matrix:
    """Device-resident compressed2D built from host CSR arrays (copied)."""

    def __init__(self, starts, indices, values, ncols=None):
        starts = np.ascontiguousarray(starts, dtype=np.int64)
        indices = np.ascontiguousarray(indices, dtype=np.int64)
        values = np.ascontiguousarray(values)
        if values.dtype not in (np.float64, np.float32):
            values = values.astype(np.float64)
        self.nrows = len(starts) - 1
        self.ncols = self.nrows if ncols is None else ncols
        self.nnz = int(starts[-1])
        h = ctypes.c_void_p()
        _check(_f["m4g_matrix_create"](ctypes.byref(h), self.nrows, self.ncols, self.nnz, _ptr(starts),
                                       _ptr(indices), _ptr(values), IDX_I64, _dtype_code(values.dtype)),
               "m4g_matrix_create")
        self.h = h

    def download(self):
        S = np.zeros(self.nrows + 1, np.int64)
        I = np.zeros(self.nnz, np.int64)
        V = np.zeros(self.nnz, np.float64)
        _check(_f["m4g_matrix_download"](self.h, _ptr(S), _ptr(I), _ptr(V)), "m4g_matrix_download")
        return S, I, V

    def set_values(self, values):
        """values: numpy float64 array (host) or a contiguous float64 torch tensor on the GPU."""
        if hasattr(values, "data_ptr"):
            _check(_f["m4g_matrix_set_values"](self.h, ctypes.c_void_p(values.data_ptr()), 1), "m4g_matrix_set_values")
        else:
            v = np.ascontiguousarray(values, np.float64)
            _check(_f["m4g_matrix_set_values"](self.h, _ptr(v), 0), "m4g_matrix_set_values")

    def transpose(self):
        """trans(A) as a new device matrix."""
        h = ctypes.c_void_p()
        _check(_f["m4g_matrix_transpose"](self.h, ctypes.byref(h)), "m4g_matrix_transpose")
        T = Matrix.__new__(Matrix)
        T.h, T.nrows, T.ncols, T.nnz = h, self.ncols, self.nrows, self.nnz
        return T

    def close(self):
        if getattr(self, "h", None):
            _f["m4g_matrix_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def assemble(nrows, ncols, rows, cols, values, updater=UPDATE_STORE, existing=None):
    """mat::inserter<compressed2D, Updater> on the device: the triplets (numpy
    arrays, or int64/float64 torch tensors on the GPU) folded in insertion order
    into a new Matrix, starting from the entries of `existing`."""
    h = ctypes.c_void_p()
    if hasattr(values, "data_ptr"):
        count = int(values.numel())
        ptrs = [ctypes.c_void_p(t.data_ptr()) for t in (rows, cols, values)]
        on_device = 1
    else:
        arrs = (np.ascontiguousarray(rows, np.int64), np.ascontiguousarray(cols, np.int64),
                np.ascontiguousarray(values, np.float64))
        count = len(arrs[2])
        ptrs = [_ptr(a) for a in arrs]
        on_device = 0
    _check(_f["m4g_matrix_assemble"](ctypes.byref(h), int(nrows), int(ncols), count, *ptrs, on_device, int(updater),
                                     existing.h if existing is not None else None), "m4g_matrix_assemble")
    M = Matrix.__new__(Matrix)
    M.h, M.nrows, M.ncols = h, int(nrows), int(ncols)
    nnz = ctypes.c_int64()
    _check(_f["m4g_matrix_dims"](h, None, None, ctypes.byref(nnz)), "m4g_matrix_dims")
    M.nnz = nnz.value
    return M


class Vector:
    """Device-resident dense_vector; owns its memory, or borrows a torch tensor's."""

    def __init__(self, n=None, dtype=np.float64, data=None, tensor=None):
        h = ctypes.c_void_p()
        self._tensor = tensor
        if tensor is not None:
            import torch
            assert tensor.is_cuda and tensor.is_contiguous(), "tensor must be contiguous on the GPU"
            if tensor.dtype == torch.float64:
                code = F64
            elif tensor.dtype == torch.float32:
                code = F32
            else:
                raise TypeError(f"Vector(tensor=...) takes float64 or float32 tensors, not {tensor.dtype}")
            self.n, self.dtype = tensor.numel(), _np_dtype(code)
            _check(_f["m4g_vector_wrap"](ctypes.byref(h), ctypes.c_void_p(tensor.data_ptr()), self.n, code),
                   "m4g_vector_wrap")
        else:
            if data is not None:
                data = np.ascontiguousarray(data, dtype=dtype)
                n = len(data)
            self.n, self.dtype = int(n), np.dtype(dtype).type
            _check(_f["m4g_vector_create"](ctypes.byref(h), self.n, _dtype_code(dtype)), "m4g_vector_create")
        self.h = h
        if data is not None:
            self.upload(data)

    def upload(self, data):
        data = np.ascontiguousarray(data, dtype=self.dtype)
        assert len(data) == self.n
        _check(_f["m4g_vector_upload"](self.h, _ptr(data)), "m4g_vector_upload")

    def numpy(self):
        out = np.zeros(self.n, dtype=self.dtype)
        _check(_f["m4g_vector_download"](self.h, _ptr(out)), "m4g_vector_download")
        return out

    def fill(self, value):
        _check(_f["m4g_vector_fill"](self.h, float(value)), "m4g_vector_fill")

    def copy_from(self, other):
        _check(_f["m4g_vector_copy"](self.h, other.h), "m4g_vector_copy")

    def device_ptr(self):
        return _f["m4g_vector_device_ptr"](self.h)

    def close(self):
        if getattr(self, "h", None):
            _f["m4g_vector_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PC:
    """itl::pc::identity / diagonal / ic_0 / ilu_0 built from a Matrix."""

    def __init__(self, A, kind, factor_dtype=np.float64):
        h = ctypes.c_void_p()
        self.kind, self.n = kind, A.nrows
        _check(_f["m4g_pc_create"](ctypes.byref(h), A.h, kind, _dtype_code(factor_dtype)), "m4g_pc_create")
        self.h = h

    def solve(self, x, y):
        return _check(_f["m4g_pc_solve"](self.h, x.h, y.h), "m4g_pc_solve")

    def apply_async(self, x, y):
        return _check(_f["m4g_pc_apply_async"](self.h, x.h, y.h), "m4g_pc_apply_async")

    def adjoint_solve(self, x, y):
        return _check(_f["m4g_pc_adjoint_solve"](self.h, x.h, y.h), "m4g_pc_adjoint_solve")

    def factor(self, which):
        """Returns (starts, indices, values) of L (which=0) or U (which=1)."""
        nnz = ctypes.c_int64()
        _check(_f["m4g_pc_factor_nnz"](self.h, which, ctypes.byref(nnz)), "m4g_pc_factor_nnz")
        S = np.zeros(self.n + 1, np.int64)
        I = np.zeros(nnz.value, np.int64)
        V = np.zeros(nnz.value, np.float64)
        _check(_f["m4g_pc_download_factor"](self.h, which, _ptr(S), _ptr(I), _ptr(V)), "m4g_pc_download_factor")
        return S, I, V

    def refactor(self, A):
        """Numeric refactorisation for A's current values (same pattern)."""
        _check(_f["m4g_pc_refactor"](self.h, A.h), "m4g_pc_refactor")

    def stats(self):
        st = PcStats()
        _check(_f["m4g_pc_stats_get"](self.h, ctypes.byref(st)), "m4g_pc_stats_get")
        return st.as_dict()

    def close(self):
        if getattr(self, "h", None):
            _f["m4g_pc_destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------
def _scalar(fn, where, *args):
    out = ctypes.c_double()
    _check(fn(*args, ctypes.byref(out)), where)
    return out.value


def dot(x, y):
    return _scalar(_f["m4g_dot"], "m4g_dot", x.h, y.h)


def dot_noalloc(x, y):
    """dot(x, y) through the entry point that allocates no device memory."""
    return _scalar(_f["m4g_dot_noalloc"], "m4g_dot_noalloc", x.h, y.h)


def two_norm(x):
    return _scalar(_f["m4g_two_norm"], "m4g_two_norm", x.h)


def unary_dot(x):
    return _scalar(_f["m4g_unary_dot"], "m4g_unary_dot", x.h)


def axpy(alpha, x, y):
    _check(_f["m4g_axpy"](float(alpha), x.h, y.h), "m4g_axpy")


def xpay(x, beta, y):
    _check(_f["m4g_xpay"](x.h, float(beta), y.h), "m4g_xpay")


def axpy_unary_dot(alpha, q, r):
    return _scalar(_f["m4g_axpy_unary_dot"], "m4g_axpy_unary_dot", float(alpha), q.h, r.h)


def spmv(A, v, w, mode=ASSIGN):
    _check(_f["m4g_spmv"](A.h, v.h, w.h, mode), "m4g_spmv")


def residual(A, x, b, r):
    _check(_f["m4g_residual"](A.h, x.h, b.h, r.h), "m4g_residual")


def spmv_dot(A, p, q):
    return _scalar(_f["m4g_spmv_dot"], "m4g_spmv_dot", A.h, p.h, q.h)


def pattern_gradient(A, lam, x, grad):
    """grad[p] = -lam[row(p)] * x[col(p)] on A's stored pattern."""
    _check(_f["m4g_pattern_gradient"](A.h, lam.h, x.h, grad.h), "m4g_pattern_gradient")


def make_iteration(max_iter, rtol, atol=0.0, norm_r0=0.0, history=True):
    it = Iteration()
    it.max_iter, it.rtol, it.atol, it.norm_r0 = int(max_iter), float(rtol), float(atol), float(norm_r0)
    it.iterations, it.error, it.resid = 0, 0, 0.0
    hist = None
    if history:
        hist = np.full(int(max_iter) + 2, np.nan)
        it.history = hist.ctypes.data_as(_c_dp)
        it.history_capacity = len(hist)
    return it, hist


def _solve(fn, where, A, x, b, P, it):
    st = fn(A.h, x.h, b.h, P.h, ctypes.byref(it))
    return _check(st, where, allow=(OK, MAX_ITER))


def cg(A, x, b, P, it):
    """itl::cg(A, x, b, P, iter); returns the iteration's error code."""
    return _solve(_f["m4g_cg"], "m4g_cg", A, x, b, P, it)


def bicgstab(A, x, b, P, it):
    """itl::bicgstab(A, x, b, P, iter); returns the iteration's error code."""
    return _solve(_f["m4g_bicgstab"], "m4g_bicgstab", A, x, b, P, it)


def cgs(A, x, b, P, it):
    """itl::cgs(A, x, b, P, iter); returns 0, MAX_ITER or BREAKDOWN."""
    st = _f["m4g_cgs"](A.h, x.h, b.h, P.h, ctypes.byref(it))
    return _check(st, "m4g_cgs", allow=(OK, MAX_ITER, BREAKDOWN))


def bicg(A, x, b, P, it):
    """itl::bicg(A, x, b, P, iter); returns 0, MAX_ITER or BREAKDOWN."""
    st = _f["m4g_bicg"](A.h, x.h, b.h, P.h, ctypes.byref(it))
    return _check(st, "m4g_bicg", allow=(OK, MAX_ITER, BREAKDOWN))


def tfqmr(A, x, b, L, it, R=None):
    """itl::tfqmr(A, x, b, L, R, iter) (R = None: identity); returns the iteration's error."""
    st = _f["m4g_tfqmr"](A.h, x.h, b.h, L.h, R.h if R is not None else None, ctypes.byref(it))
    return _check(st, "m4g_tfqmr", allow=(OK, MAX_ITER))


def _h(P):
    return P.h if P is not None else None


_SOLVER_STATUS = (OK, MAX_ITER, BREAKDOWN)


def qmr(A, x, b, L, it, R=None):
    """itl::qmr(A, x, b, L, R, iter) (R = None: identity); returns 0, MAX_ITER or BREAKDOWN."""
    return _check(_f["m4g_qmr"](A.h, x.h, b.h, _h(L), _h(R), ctypes.byref(it)), "m4g_qmr", allow=_SOLVER_STATUS)


def gmres(A, x, b, L, it, restart=8, R=None):
    """itl::gmres(A, x, b, L, R, iter, restart)."""
    st = _f["m4g_gmres"](A.h, x.h, b.h, _h(L), _h(R), ctypes.byref(it), int(restart))
    return _check(st, "m4g_gmres", allow=_SOLVER_STATUS)


def gmres_full(A, x, b, L, it, R=None):
    """itl::gmres_full(A, x, b, L, R, iter)."""
    st = _f["m4g_gmres_full"](A.h, x.h, b.h, _h(L), _h(R), ctypes.byref(it))
    return _check(st, "m4g_gmres_full", allow=_SOLVER_STATUS)


def bicgstab_2(A, x, b, L, it):
    """itl::bicgstab_2(A, x, b, L, iter)."""
    return _check(_f["m4g_bicgstab_2"](A.h, x.h, b.h, _h(L), ctypes.byref(it)), "m4g_bicgstab_2",
                  allow=_SOLVER_STATUS)


def bicgstab_ell(A, x, b, L, it, l=8, R=None):
    """itl::bicgstab_ell(A, x, b, L, R, iter, l)."""
    st = _f["m4g_bicgstab_ell"](A.h, x.h, b.h, _h(L), _h(R), ctypes.byref(it), int(l))
    return _check(st, "m4g_bicgstab_ell", allow=_SOLVER_STATUS)


def idr_s(A, x, b, it, s=8, L=None, R=None):
    """itl::idr_s(A, x, b, L, R, iter, s) (L and R ignored, as in the reference)."""
    st = _f["m4g_idr_s"](A.h, x.h, b.h, _h(L), _h(R), ctypes.byref(it), int(s))
    return _check(st, "m4g_idr_s", allow=_SOLVER_STATUS)


def random_seed_counter(value=None):
    """mtl::seed<double>::counter: returns it, or sets it when `value` is given."""
    if value is not None:
        _f["m4g_set_random_seed_counter"](int(value))
    return _f["m4g_random_seed_counter"]()


def refined_solve(method, A, x, b, P, it, inner_rtol, inner_max_iter=10000):
    """FP32-inner / FP64-refined solve; returns (error code, RefineInfo)."""
    info = RefineInfo()
    st = _f["m4g_refined_solve"](method, A.h, x.h, b.h, P.h if P is not None else None, ctypes.byref(it),
                                 float(inner_rtol), int(inner_max_iter), ctypes.byref(info))
    _check(st, "m4g_refined_solve", allow=(OK, MAX_ITER))
    return st, info


class Batch:
    """Many independent SPD systems (concatenated CSR) for the batched IC(0)-PCG."""

    def __init__(self, systems):
        """systems: list of (starts, indices, values) per system (local columns)."""
        sizes = np.array([len(S) - 1 for S, _, _ in systems], np.int64)
        starts = np.concatenate([np.asarray(S, np.int64) for S, _, _ in systems])
        indices = np.concatenate([np.asarray(I, np.int64) for _, I, _ in systems])
        values = np.concatenate([np.asarray(V, np.float64) for _, _, V in systems])
        self.sizes = sizes
        self.row0 = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        self.n_total = int(sizes.sum())
        h = ctypes.c_void_p()
        _check(_f["m4g_batch_create"](ctypes.byref(h), len(sizes), _ptr(sizes), _ptr(starts), _ptr(indices),
                                      _ptr(values)), "m4g_batch_create")
        self.h = h
        self._keep = (sizes, starts, indices, values)
        self.u_total = sum(int(((np.asarray(I) >= np.repeat(np.arange(len(S) - 1), np.diff(S))).sum()))
                           for S, I, _ in systems)

    def __del__(self):
        if getattr(self, "h", None) and _f.get("m4g_batch_destroy"):
            _f["m4g_batch_destroy"](self.h)

    def factor_ic0(self):
        _check(_f["m4g_batch_factor_ic0"](self.h), "m4g_batch_factor_ic0")

    def factor_values(self):
        out = np.empty(self.u_total, np.float64)
        _check(_f["m4g_batch_download_factor"](self.h, _ptr(out)), "m4g_batch_download_factor")
        return out

    def autotune(self, nrhs, b, trial_iters=0):
        """Returns (rhs_per_item per system, block size, trial times in ms)."""
        rpi = np.zeros(len(self.sizes), np.int32)
        blk = ctypes.c_int()
        trials = np.full(24, np.nan)
        _check(_f["m4g_batch_autotune"](self.h, nrhs, b.h, int(trial_iters), _ptr(rpi), ctypes.byref(blk),
                                        _ptr(trials)), "m4g_batch_autotune")
        return rpi, blk.value, trials

    def cg_ic0(self, nrhs, x, b, max_iter, rtol, atol=0.0, rhs_per_item=None, block=0):
        """Solves every (system, rhs); returns (iterations, resid, error) arrays [nsys, nrhs]."""
        ns = len(self.sizes)
        it = np.zeros(ns * nrhs, np.int32)
        res = np.zeros(ns * nrhs, np.float64)
        err = np.zeros(ns * nrhs, np.int32)
        rpi = None if rhs_per_item is None else np.ascontiguousarray(rhs_per_item, np.int32)
        _check(_f["m4g_batch_cg_ic0"](self.h, nrhs, x.h, b.h, int(max_iter), float(rtol), float(atol),
                                      _ptr(rpi) if rpi is not None else None, int(block), _ptr(it), _ptr(res),
                                      _ptr(err)), "m4g_batch_cg_ic0")
        return it.reshape(ns, nrhs), res.reshape(ns, nrhs), err.reshape(ns, nrhs)


# ---------------------------------------------------------------------------
# host-only helpers: Matrix Market files, analysis and generators
# ---------------------------------------------------------------------------
def mm_read(path):
    """Reads a Matrix Market file as matrix_market_istream >> compressed2D<double>
    (matrix_market.hpp:166-217).  Returns (starts, indices, values, nrows, ncols)."""
    p = os.fsencode(path)
    nr, nc, nz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(0)
    _check(_f["m4g_mm_read"](p, ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nz), None, None, None),
           "m4g_mm_read")
    S = np.zeros(nr.value + 1, np.int64)
    I = np.zeros(max(nz.value, 1), np.int64)
    V = np.zeros(max(nz.value, 1), np.float64)
    _check(_f["m4g_mm_read"](p, ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nz), _ptr(S), _ptr(I), _ptr(V)),
           "m4g_mm_read")
    return S, I[:nz.value], V[:nz.value], nr.value, nc.value


def mm_write(path, starts, indices, values, ncols=None):
    """Writes CSR arrays as matrix_market_ostream << compressed2D<double>
    (matrix_market.hpp:235-256)."""
    S = np.ascontiguousarray(starts, np.int64)
    I = np.ascontiguousarray(indices, np.int64)
    V = np.ascontiguousarray(values, np.float64)
    n = len(S) - 1
    _check(_f["m4g_mm_write"](os.fsencode(path), n, n if ncols is None else int(ncols), _ptr(S), _ptr(I), _ptr(V)),
           "m4g_mm_write")


def analysis_grid(starts, indices):
    S = np.ascontiguousarray(starts, np.int64)
    I = np.ascontiguousarray(indices, np.int64)
    nx, ny, nz = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _f["m4g_analysis_grid"](len(S) - 1, _ptr(S), _ptr(I), ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz))
    return nx.value, ny.value, nz.value


def analysis_pc_solve(kind, starts, indices, factor, rhs, adjoint=False, factor_dtype=np.float64,
                      vector_dtype=np.float64, max_tile_rows=0):
    S = np.ascontiguousarray(starts, np.int64)
    I = np.ascontiguousarray(indices, np.int64)
    f = np.ascontiguousarray(factor, np.float64)
    rhs = np.ascontiguousarray(rhs, vector_dtype)
    out = np.zeros_like(rhs)
    st = PcStats()
    _check(_f["m4g_analysis_pc_solve"](kind, int(adjoint), len(S) - 1, _ptr(S), _ptr(I), _ptr(f),
                                       _dtype_code(factor_dtype), _dtype_code(vector_dtype), _ptr(rhs), _ptr(out),
                                       int(max_tile_rows), ctypes.byref(st)), "m4g_analysis_pc_solve")
    return out, st.as_dict()


def _gen(fn, *args):
    nnz = fn(*args, None, None, None)
    n_rows = None
    return nnz


def _csr_from(fn, n, *args):
    nnz = fn(*args, None, None, None)
    S = np.zeros(n + 1, np.int64)
    I = np.zeros(nnz, np.int64)
    V = np.zeros(nnz, np.float64)
    fn(*args, _ptr(S), _ptr(I), _ptr(V))
    return S, I, V


def laplacian2d(m, n):
    """mtl::mat::laplacian_setup(A, m, n) (laplacian_setup.hpp:25-44)."""
    return _csr_from(_f["m4g_gen_laplacian2d"], m * n, m, n)


def poisson3d(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _csr_from(_f["m4g_gen_poisson3d"], nx * ny * nz, nx, ny, nz)


def convdiff27(nx, ny=None, nz=None, seed=3, pe=0.5):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _csr_from(_f["m4g_gen_convdiff27"], nx * ny * nz, nx, ny, nz, ctypes.c_uint64(seed), float(pe))


def varcoef5(nx, ny=None, seed=4):
    ny = nx if ny is None else ny
    return _csr_from(_f["m4g_gen_varcoef5"], nx * ny, nx, ny, ctypes.c_uint64(seed))


def edge_laplacian(nx, ny, theta=None):
    """Returns dict(S, I, V, ei, ej, aptr, aidx, acoef, n_edges) for A(theta)."""
    ne = ctypes.c_int64()
    fn = _f["m4g_gen_edge_laplacian"]
    nnz = fn(nx, ny, None, ctypes.byref(ne), None, None, None, None, None, None, None, None)
    n = nx * ny
    S = np.zeros(n + 1, np.int64); I = np.zeros(nnz, np.int64); V = np.zeros(nnz, np.float64)
    ei = np.zeros(ne.value, np.int64); ej = np.zeros(ne.value, np.int64)
    aptr = np.zeros(nnz + 1, np.int64); aidx = np.zeros(4 * ne.value, np.int64); acoef = np.zeros(4 * ne.value)
    th = None if theta is None else np.ascontiguousarray(theta, np.float64)
    fn(nx, ny, _ptr(th), ctypes.byref(ne), _ptr(S), _ptr(I), _ptr(V), _ptr(ei), _ptr(ej), _ptr(aptr), _ptr(aidx),
       _ptr(acoef))
    na = int(aptr[-1])
    return dict(S=S, I=I, V=V, ei=ei, ej=ej, aptr=aptr, aidx=aidx[:na], acoef=acoef[:na], n_edges=ne.value)


def rng_uniform(seed, stream, count, lo=-1.0, hi=1.0, offset=0):
    out = np.zeros(count)
    _f["m4g_rng_uniform"](seed, stream, offset, count, lo, hi, _ptr(out))
    return out


def rng_normal(seed, stream, count, offset=0):
    out = np.zeros(count)
    _f["m4g_rng_normal"](seed, stream, offset, count, _ptr(out))
    return out


def rng_lognormal(seed, stream, count, offset=0):
    out = np.zeros(count)
    _f["m4g_rng_lognormal"](seed, stream, offset, count, _ptr(out))
    return out


def rng_int(seed, stream, count, lo, hi, offset=0):
    out = np.zeros(count, np.int64)
    _f["m4g_rng_int"](seed, stream, offset, count, lo, hi, _ptr(out))
    return out
