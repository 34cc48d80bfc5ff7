"""asgard_amd: MI355X (gfx950) kronmult operator apply and implicit-step GMRES.

Python mirror of the reference call surface for this path, bound to the C ABI
of ``libasgard_amd.so`` (declared in ``include/asgard_amd.h``):

* :class:`KronmultMatrix` -- ``asgard::kronmult_matrix<P>``
  (src/asgard_kronmult_matrix.hpp:111-682): sparse, sparse-grid and dense
  constructors, ``apply(alpha, x, beta, y)``, sizes and flops,
  ``update_stored_coefficients``;
* :func:`simple_gmres_euler` -- ``asgard::solver::simple_gmres_euler``
  (src/solver.cpp:119-135) returning ``GmresInfo(error, iterations)``
  (src/pde/pde_base.hpp:94-98);
* :func:`connect_1d` and :func:`sparse_pattern` -- the index construction of
  ``connect_1d`` (src/asgard_grid_1d.hpp) and ``check_connected`` scans
  (src/asgard_kronmult_matrix.cpp:918-938), the latter on the GPU.

Vectors are torch CUDA tensors (device pointers, ``apply<resource::device>``)
or numpy arrays (host pointers, ``apply<resource::host>``).  There is no CPU
fallback: the HIP library must be built and a GPU present for any compute
call; errors raise :class:`AkmError` (``ValueError`` for invalid arguments,
as the reference throws ``std::invalid_argument``).
"""
from __future__ import annotations

import ctypes
import os
from collections import namedtuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libasgard_amd.so")

AKM_SUCCESS = 0
AKM_ERR_INVALID = -1
AKM_ERR_UNSUPPORTED = -2
AKM_ERR_HIP = -3
AKM_ERR_MEMORY = -4
AKM_ERR_NOT_CONVERGED = 1
AKM_F64 = 0
AKM_F32 = 1

#: every symbol include/asgard_amd.h declares
EXPORTS = (
    "akm_last_error", "akm_version", "akm_device_count", "akm_set_stream",
    "akm_synchronize", "akm_connect_1d", "akm_sparse_pattern", "akm_create_sparse",
    "akm_create_sparse_grid", "akm_create_global_grid", "akm_global_info",
    "akm_global_slab_info",
    "akm_global_jacobi", "akm_create_dense", "akm_update_coefficients",
    "akm_update_operators", "akm_imex_step", "akm_dg_term",
    "akm_apply", "akm_apply_euler", "akm_matrix_info", "akm_set_tuning",
    "akm_autotune", "akm_set_kernel_family", "akm_time_apply", "akm_destroy", "akm_gmres_euler",
    "akm_gmres_euler_noalloc", "akm_gmres_dense",
    "akm_gmres_euler_mixed", "akm_mlp_create", "akm_mlp_info", "akm_mlp_set_stream",
    "akm_mlp_get_params", "akm_mlp_set_params", "akm_mlp_forward", "akm_mlp_gradients",
    "akm_mlp_adam_step", "akm_mlp_train", "akm_precond_apply", "akm_precond_destroy",
    "akm_batch_create", "akm_batch_info", "akm_batch_solve", "akm_batch_destroy",
)

GmresInfo = namedtuple("GmresInfo", ["error", "iterations"])

#: AKM_NOALLOC_SCRATCH_BYTES of include/asgard_amd.h
NOALLOC_SCRATCH_BYTES = 32 << 20

class DgPartial(ctypes.Structure):
    """akm_dg_partial: one partial term of a 1D DG operator."""
    _fields_ = [("kind", ctypes.c_int), ("flux_scale", ctypes.c_double),
                ("gq", ctypes.c_void_p), ("ge", ctypes.c_void_p),
                ("left", ctypes.c_int), ("right", ctypes.c_int)]


DG_MASS, DG_DIV, DG_GRAD, DG_PENALTY = 0, 1, 2, 3
DG_PERIODIC, DG_DIRICHLET, DG_NEUMANN = 0, 1, 2

#: akm_imex_hook and the events it runs at
IMEX_HOOK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)
IMEX_EXPLICIT_1, IMEX_IMPLICIT_1, IMEX_EXPLICIT_2, IMEX_IMPLICIT_2 = 0, 1, 2, 3


class SolveStats(ctypes.Structure):
    """akm_solve_stats: outcome of :func:`gmres_euler_mixed`."""
    _fields_ = [("rel_residual", ctypes.c_double), ("abs_residual", ctypes.c_double),
                ("rhs_norm", ctypes.c_double), ("iterations", ctypes.c_int),
                ("cycles", ctypes.c_int), ("applies_f64", ctypes.c_int),
                ("applies_f32", ctypes.c_int), ("precond_applies", ctypes.c_int),
                ("fallback", ctypes.c_int), ("fallback_iterations", ctypes.c_int),
                ("converged", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class AkmError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[akm {code}] {message}")
        self.code = code


class AkmInvalid(AkmError, ValueError):
    pass


_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_d = ctypes.c_double


def load_library(path: str = LIB_PATH):
    """Loads libasgard_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"asgard_amd: {path} is missing; build it first (__graft_entry__.build() "
            "or make -C asgard_amd/csrc)")
    L = ctypes.CDLL(path)
    L.akm_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.akm_last_error.restype = _i
    L.akm_version.restype = ctypes.c_char_p
    L.akm_device_count.argtypes = [ctypes.POINTER(_i)]
    L.akm_set_stream.argtypes = [_vp, _vp]
    L.akm_synchronize.argtypes = [_vp]
    L.akm_connect_1d.argtypes = [_i, _i, _vp, _vp, _vp, ctypes.POINTER(_i)]
    L.akm_sparse_pattern.argtypes = [_i, _i64, _vp, _vp, _vp]
    L.akm_create_sparse.argtypes = [_i, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i64, _i,
                                    ctypes.POINTER(_vp)]
    L.akm_create_sparse_grid.argtypes = [_i, _i, _i64, _vp, _i, _i, _vp, _i64, _i,
                                         ctypes.POINTER(_vp)]
    L.akm_create_global_grid.argtypes = [_i, _i, _i64, _vp, _i, _i, _vp, _i64, _vp, _vp, _i,
                                         ctypes.POINTER(_vp)]
    L.akm_global_slab_info.argtypes = [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                       ctypes.POINTER(_i)]
    L.akm_global_info.argtypes = [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i),
                                  ctypes.POINTER(_i)]
    L.akm_global_jacobi.argtypes = [_vp, _vp]
    L.akm_create_dense.argtypes = [_i, _i, _i, _i, _i, _vp, _vp, _i, _i, _i, _i, _i,
                                   ctypes.POINTER(_vp)]
    L.akm_update_coefficients.argtypes = [_vp, _vp]
    L.akm_update_operators.argtypes = [_vp, _vp, ctypes.c_int64]
    L.akm_dg_term.argtypes = [_i, _i, _d, _d, _i, ctypes.POINTER(DgPartial), _i, _vp, _vp, _vp]
    L.akm_imex_step.argtypes = [_vp, _vp, _d, _vp, _i, _i, _d, _i, IMEX_HOOK, _vp,
                                ctypes.POINTER(_d), ctypes.POINTER(_i)]
    L.akm_apply.argtypes = [_vp, _d, _vp, _d, _vp, _i]
    L.akm_apply_euler.argtypes = [_vp, _d, _d, _vp, _d, _vp]
    L.akm_matrix_info.argtypes = [_vp] + [ctypes.POINTER(_i64)] * 5
    L.akm_set_tuning.argtypes = [_vp, _i, _i, _i]
    L.akm_time_apply.argtypes = [_vp, _d, _vp, _d, _vp, _i, ctypes.POINTER(_d)]
    L.akm_autotune.argtypes = [_vp, _vp, _vp, ctypes.POINTER(_i), ctypes.POINTER(_i)]
    L.akm_set_kernel_family.argtypes = [_vp, _i]
    L.akm_destroy.argtypes = [_vp]
    L.akm_gmres_euler.argtypes = [_vp, _d, _vp, _vp, _i, _i, _d, _i, ctypes.POINTER(_d),
                                  ctypes.POINTER(_i)]
    L.akm_gmres_euler_noalloc.argtypes = [_vp, _d, _vp, _vp, _i, _i, _d, ctypes.POINTER(_d),
                                          ctypes.POINTER(_i)]
    L.akm_gmres_dense.argtypes = [_i, _vp, _vp, _vp, _vp, _i, _i, _d, _i, ctypes.POINTER(_d),
                                  ctypes.POINTER(_i)]
    L.akm_gmres_euler_mixed.argtypes = [_vp, _vp, _vp, _d, _vp, _vp, _i, _i, _d, _d,
                                        ctypes.POINTER(SolveStats)]
    L.akm_mlp_create.argtypes = [_i, _i, _i64, _vp, _i, _i, ctypes.c_uint64, ctypes.POINTER(_vp)]
    L.akm_mlp_info.argtypes = [_vp, ctypes.POINTER(_i64)] + [ctypes.POINTER(_i)] * 5
    L.akm_mlp_set_stream.argtypes = [_vp, _vp]
    L.akm_mlp_get_params.argtypes = [_vp, _vp]
    L.akm_mlp_set_params.argtypes = [_vp, _vp]
    L.akm_mlp_forward.argtypes = [_vp, _vp, _i64, _vp]
    L.akm_mlp_gradients.argtypes = [_vp, _vp, _vp, _i64, ctypes.POINTER(_d), _vp]
    L.akm_mlp_adam_step.argtypes = [_vp, _vp, _d]
    L.akm_mlp_train.argtypes = [_vp, _vp, _vp, _i64, _i, _i, _d, ctypes.c_uint64,
                                ctypes.POINTER(_d)]
    L.akm_precond_apply.argtypes = [_vp, _vp, _vp, _vp]
    L.akm_precond_destroy.argtypes = [_vp]
    L.akm_batch_create.argtypes = [_i, _vp, _vp, _i, ctypes.POINTER(_vp)]
    L.akm_batch_info.argtypes = [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), _vp]
    L.akm_batch_solve.argtypes = [_vp, _vp, _vp, _i, _d, _vp, _vp, _vp]
    L.akm_batch_destroy.argtypes = [_vp]
    for name in EXPORTS:
        if name not in ("akm_last_error", "akm_version"):
            getattr(L, name).restype = _i
    _lib = L
    return L


def last_error() -> str:
    L = load_library()
    buf = ctypes.create_string_buffer(1024)
    L.akm_last_error(buf, 1024)
    return buf.value.decode()


def _check(rc, allow_not_converged=False):
    if rc == AKM_SUCCESS or (allow_not_converged and rc == AKM_ERR_NOT_CONVERGED):
        return rc
    msg = last_error()
    if rc == AKM_ERR_INVALID:
        raise AkmInvalid(rc, msg)
    raise AkmError(rc, msg)


def device_count() -> int:
    c = _i(0)
    _check(load_library().akm_device_count(ctypes.byref(c)))
    return c.value


# ---------------------------------------------------------------- vectors ---

def _is_torch(t):
    return type(t).__module__.startswith("torch")


def _vec_ptr(v, dtype):
    """(pointer, on_device) of a torch CUDA tensor or numpy array of `dtype`."""
    if _is_torch(v):
        import torch
        want = torch.float64 if dtype == np.float64 else torch.float32
        if v.dtype != want or not v.is_contiguous():
            raise AkmInvalid(AKM_ERR_INVALID, f"vector must be a contiguous {want} tensor")
        if not v.is_cuda:
            raise AkmInvalid(AKM_ERR_INVALID, "torch vectors must live on the GPU")
        return v.data_ptr(), 1
    if not isinstance(v, np.ndarray) or v.dtype != dtype or not v.flags.c_contiguous:
        raise AkmInvalid(AKM_ERR_INVALID, f"vector must be a contiguous {np.dtype(dtype)} array")
    return v.ctypes.data, 0


def _np(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------ index build ---

def connect_1d(max_level, full=True):
    """(pntr, indx, diag) of the 1D cell hierarchy connectivity."""
    L = load_library()
    nconn = _i(0)
    _check(L.akm_connect_1d(int(max_level), int(full), None, None, None, ctypes.byref(nconn)))
    rows = 1 << int(max_level)
    pntr = np.zeros(rows + 1, np.int32)
    indx = np.zeros(nconn.value, np.int32)
    diag = np.zeros(rows, np.int32)
    _check(L.akm_connect_1d(int(max_level), int(full), pntr.ctypes.data, indx.ctypes.data,
                            diag.ctypes.data, ctypes.byref(nconn)))
    return pntr, indx, diag


def sparse_pattern(table):
    """Cell-pair CSR (pntr int64, indx int32) computed on the GPU."""
    L = load_library()
    table = _np(table, np.int32)
    cells, twod = table.shape
    pntr = np.zeros(cells + 1, np.int64)
    _check(L.akm_sparse_pattern(twod // 2, cells, table.ctypes.data, pntr.ctypes.data, None))
    indx = np.zeros(int(pntr[-1]), np.int32)
    _check(L.akm_sparse_pattern(twod // 2, cells, table.ctypes.data, pntr.ctypes.data,
                                indx.ctypes.data))
    return pntr, indx


# ----------------------------------------------------------------- matrix ---

class KronmultMatrix:
    """Device-resident ``kronmult_matrix<P>`` (P = float64 or float32)."""

    def __init__(self, handle, prec, d, n, rows, cols, num_terms):
        self._h = handle
        self.prec = prec
        self.dtype = np.float64 if prec == AKM_F64 else np.float32
        self.num_dimensions = d
        self.kron_size = n
        self.num_rows = rows
        self.num_cols = cols
        self.num_terms = num_terms

    # ---- constructors
    @classmethod
    def sparse(cls, d, n, rows, cols, num_terms, pntr, indx, iA, vA, dtype=np.float64):
        """Sparse single-call constructor (hpp:235-272, CPU CSR format)."""
        L = load_library()
        prec = AKM_F64 if np.dtype(dtype) == np.float64 else AKM_F32
        pntr = _np(pntr, np.int64)
        indx = _np(indx, np.int32)
        iA = _np(iA, np.int32)
        vA = _np(vA, dtype)
        h = _vp()
        _check(L.akm_create_sparse(d, n, rows, cols, num_terms, pntr.ctypes.data,
                                   indx.ctypes.data, iA.ctypes.data, vA.ctypes.data, vA.size,
                                   prec, ctypes.byref(h)))
        return cls(h, prec, d, n, rows, cols, num_terms)

    @classmethod
    def sparse_grid(cls, d, n, table, max_level, num_terms, vA, dtype=np.float64):
        """make_kronmult_sparse equivalent from the element table."""
        L = load_library()
        prec = AKM_F64 if np.dtype(dtype) == np.float64 else AKM_F32
        table = _np(table, np.int32)
        vA = _np(vA, dtype)
        h = _vp()
        _check(L.akm_create_sparse_grid(d, n, table.shape[0], table.ctypes.data,
                                        int(max_level), num_terms, vA.ctypes.data, vA.size,
                                        prec, ctypes.byref(h)))
        return cls(h, prec, d, n, table.shape[0], table.shape[0], num_terms)

    @classmethod
    def global_grid(cls, d, n, table, max_level, num_terms, vA, dtype=np.float64, active=None,
                    flux_dim=None):
        """make_global_kron_matrix equivalent (the reference's global_kron_matrix)
        from the same table and vA as sparse_grid; active[t][k] = 0 marks an
        identity factor, flux_dim[t] the term's flux direction."""
        L = load_library()
        prec = AKM_F64 if np.dtype(dtype) == np.float64 else AKM_F32
        table = _np(table, np.int32)
        vA = _np(vA, dtype)
        act = None if active is None else _np(np.asarray(active).reshape(-1), np.int32)
        flux = None if flux_dim is None else _np(np.asarray(flux_dim).reshape(-1), np.int32)
        if act is not None and act.size != num_terms * d:
            raise ValueError("active needs num_terms x d entries")
        if flux is not None and flux.size != num_terms:
            raise ValueError("flux_dim needs num_terms entries")
        h = _vp()
        _check(L.akm_create_global_grid(d, n, table.shape[0], table.ctypes.data,
                                        int(max_level), num_terms, vA.ctypes.data, vA.size,
                                        None if act is None else act.ctypes.data,
                                        None if flux is None else flux.ctypes.data,
                                        prec, ctypes.byref(h)))
        return cls(h, prec, d, n, table.shape[0], table.shape[0], num_terms)

    def global_info(self):
        """(padding cells, chains, stages) of a global operator."""
        p, c, st = _i64(), _i(), _i()
        _check(load_library().akm_global_info(self._h, ctypes.byref(p), ctypes.byref(c),
                                              ctypes.byref(st)))
        return p.value, c.value, st.value

    def slab_info(self):
        """(term groups on the slab kernels, chains left to the cell-layout
        stages, vectors the final sum reads) of a global operator's slab
        layout."""
        g, r, sc = _i(), _i(), _i()
        _check(load_library().akm_global_slab_info(self._h, ctypes.byref(g), ctypes.byref(r),
                                                   ctypes.byref(sc)))
        return g.value, r.value, sc.value

    def jacobi(self):
        """The Jacobi diagonal of a global operator (FP64, active DOFs)."""
        pc = np.zeros(self.output_size(), dtype=np.float64)
        _check(load_library().akm_global_jacobi(self._h, pc.ctypes.data))
        return pc

    @classmethod
    def dense(cls, d, n, rows, cols, num_terms, terms, elem, row_offset, col_offset,
              num_1d_blocks, dtype=np.float64):
        """Dense constructor (hpp:132-176)."""
        L = load_library()
        prec = AKM_F64 if np.dtype(dtype) == np.float64 else AKM_F32
        terms = [_np(t, dtype) for t in terms]
        elem = _np(elem, np.int32).reshape(-1)
        ptrs = (ctypes.c_void_p * max(1, len(terms)))(*[t.ctypes.data for t in terms])
        h = _vp()
        _check(L.akm_create_dense(d, n, rows, cols, num_terms, ptrs, elem.ctypes.data,
                                  elem.size // d, row_offset, col_offset, num_1d_blocks, prec,
                                  ctypes.byref(h)))
        return cls(h, prec, d, n, rows, cols, num_terms)

    # ---- lifetime
    def close(self):
        if self._h:
            load_library().akm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __bool__(self):
        return bool(self._h) and self.num_dimensions > 0

    # ---- sizes
    def _info(self):
        vals = [_i64() for _ in range(5)]
        _check(load_library().akm_matrix_info(self._h, *[ctypes.byref(v) for v in vals]))
        return [v.value for v in vals]

    def tensor_size(self):
        return self._info()[0]

    def input_size(self):
        return self._info()[1]

    def output_size(self):
        return self._info()[2]

    def num_pairs(self):
        return self._info()[3]

    def flops(self):
        """Flops per apply, ×2 convention (kronmult_matrix::compute_flops)."""
        return self._info()[4]

    # ---- compute
    def set_stream(self, stream_ptr):
        _check(load_library().akm_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    def synchronize(self):
        _check(load_library().akm_synchronize(self._h))

    def set_tuning(self, waves_per_row=0, lds_cap_bytes=-1, force_generic=False):
        """Wavefronts per row workgroup (0 = automatic; 1, 2, 4, 8), the LDS
        budget for the row factor tables (-1 automatic, 0 off) and the
        block-wide LDS kernel switch."""
        _check(load_library().akm_set_tuning(self._h, int(waves_per_row), int(lds_cap_bytes),
                                             int(force_generic)))

    def apply(self, alpha, x, beta, y):
        """y = alpha*A*x + beta*y (device tensors or host arrays)."""
        px, dx = _vec_ptr(x, self.dtype)
        py, dy = _vec_ptr(y, self.dtype)
        if dx != dy:
            raise AkmInvalid(AKM_ERR_INVALID, "x and y must both be host or both device")
        _check(load_library().akm_apply(self._h, float(alpha), px, float(beta), py, dx))
        return y

    def apply_euler(self, dt, alpha, x, beta, y):
        """y = alpha*(x - dt*A*x) + beta*y on device tensors."""
        px, dx = _vec_ptr(x, self.dtype)
        py, dy = _vec_ptr(y, self.dtype)
        if not (dx and dy):
            raise AkmInvalid(AKM_ERR_INVALID, "apply_euler needs device tensors")
        _check(load_library().akm_apply_euler(self._h, float(dt), float(alpha), px, float(beta),
                                              py))
        return y

    def set_kernel_family(self, family):
        """0 = register / LDS-table kernels, 1 = sum-factorised matrix-core
        kernels (FP64 sparse-grid; (d, n) = (3, 3), (4, 2), (6, 2))."""
        _check(load_library().akm_set_kernel_family(self._h, int(family)))

    def autotune(self, x, y):
        """Times the kernel shapes on this operator (y is overwritten) and
        keeps the fastest; returns (waves_per_row, kernel) with kernel 0 =
        factors from global memory, 1 = LDS factor tables, 2 = matrix cores."""
        px, dx = _vec_ptr(x, self.dtype)
        py, dy = _vec_ptr(y, self.dtype)
        if not (dx and dy):
            raise AkmInvalid(AKM_ERR_INVALID, "autotune needs device tensors")
        w, l = _i(), _i()
        _check(load_library().akm_autotune(self._h, px, py, ctypes.byref(w), ctypes.byref(l)))
        return w.value, l.value

    def time_apply(self, alpha, x, beta, y, iterations):
        px, dx = _vec_ptr(x, self.dtype)
        py, dy = _vec_ptr(y, self.dtype)
        if not (dx and dy):
            raise AkmInvalid(AKM_ERR_INVALID, "time_apply needs device tensors")
        ms = _d()
        _check(load_library().akm_time_apply(self._h, float(alpha), px, float(beta), py,
                                             int(iterations), ctypes.byref(ms)))
        return ms.value

    def update_stored_coefficients(self, values):
        if isinstance(values, (list, tuple)):
            arrs = [_np(v, self.dtype) for v in values]
            ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
            _check(load_library().akm_update_coefficients(self._h, ptrs))
        else:
            arr = _np(values, self.dtype)
            _check(load_library().akm_update_coefficients(self._h, arr.ctypes.data))

    def update_operators(self, ops):
        """update_kronmult_coefficients on the GPU: rebuilds the stored
        factor blocks from num_terms * d square device tensors, ops[t*d + k]
        the operator of term t in dimension k.  Each tensor holds its
        operator's transpose in row-major order (i.e. the operator in
        column-major order, as ``torch.as_tensor(A.T)`` of a matrix A)."""
        if len(ops) != self.num_terms * self.num_dimensions:
            raise AkmInvalid(AKM_ERR_INVALID, "update_operators needs num_terms * d operators")
        import torch
        want = torch.float64 if self.dtype == np.float64 else torch.float32
        lda = None
        for op in ops:
            if not (_is_torch(op) and op.is_cuda):
                raise AkmInvalid(AKM_ERR_INVALID, "update_operators needs device tensors")
            if op.dtype != want or op.dim() != 2 or op.shape[0] != op.shape[1] \
                    or not op.is_contiguous():
                raise AkmInvalid(AKM_ERR_INVALID,
                                 "update_operators needs contiguous square tensors of the "
                                 "handle's precision")
            if lda is not None and op.shape[0] != lda:
                raise AkmInvalid(AKM_ERR_INVALID, "update_operators needs one operator size")
            lda = op.shape[0]
        ptrs = (ctypes.c_void_p * len(ops))(*[op.data_ptr() for op in ops])
        _check(load_library().akm_update_operators(self._h, ptrs, int(lda or 0)))


def imex_step(explicit: KronmultMatrix, implicit: KronmultMatrix, dt, f, restart=-1,
              max_iter=-1, tol=-1.0, collision=True, hook=None):
    """imex_advance with device vectors: one SSP-RK2 IMEX step on the FP64
    CUDA tensor f, in place.  hook(event, f_ptr) runs before each explicit
    apply and each implicit solve (events IMEX_EXPLICIT_1 .. IMEX_IMPLICIT_2)
    and may update either operator.  Returns (GmresInfo of each solve,
    converged)."""
    pf, on_dev = _vec_ptr(f, np.float64)
    if not on_dev:
        raise AkmInvalid(AKM_ERR_INVALID, "imex_step needs a device vector")
    failure = []

    def trampoline(_user, event, ptr):
        try:
            hook(event, ptr)
            return 0
        except BaseException as e:  # reported after the step returns
            failure.append(e)
            return 1

    cb = IMEX_HOOK(trampoline) if hook is not None else IMEX_HOOK()
    errs = (_d * 2)()
    its = (_i * 2)()
    try:
        rc = _check(load_library().akm_imex_step(explicit._h, implicit._h, float(dt), pf,
                                                 int(restart), int(max_iter), float(tol),
                                                 int(bool(collision)), cb, None, errs, its),
                    allow_not_converged=True)
    except AkmError:
        if failure:
            raise failure[0]
        raise
    return [GmresInfo(errs[i], its[i]) for i in range(2)], rc == AKM_SUCCESS


def dg_quadrature(n, level, domain=(-1.0, 1.0), nq=None):
    """Points where akm_dg_term takes the coefficients: the nq Gauss-Legendre
    points of every cell ((2^level, nq), ascending within a cell) and the
    2^level + 1 cell edges.  nq defaults to max(10, n + 1)."""
    a, b = domain
    cells = 1 << level
    h = (b - a) / cells
    nq = nq or max(10, n + 1)
    x, _ = np.polynomial.legendre.leggauss(nq)
    pts = a + h * (np.arange(cells)[:, None] + 0.5 * x[None, :] + 0.5)
    return pts, a + h * np.arange(cells + 1), nq


def dg_term(n, level, partials, F, domain=(-1.0, 1.0), nq=None, stream=None):
    """generate_coefficients on the GPU for one 1D term (periodic domain).

    partials: list of (kind, flux_scale, gq, ge[, left, right]) with kind
    DG_MASS / DG_DIV / DG_GRAD / DG_PENALTY, gq and ge FP64 CUDA tensors of the
    coefficient at the points of :func:`dg_quadrature` (ge may be None for mass
    terms, gq for penalty terms) and the boundary conditions DG_PERIODIC
    (default) / DG_DIRICHLET / DG_NEUMANN at the two ends.  F: the multiwavelet
    transform as a CUDA tensor holding it column-major (``torch.as_tensor(F.T)``).
    Returns the operator as an N x N CUDA tensor holding it column-major, the
    form :meth:`KronmultMatrix.update_operators` takes."""
    import torch
    N = n << level
    nq = nq or max(10, n + 1)
    parts = (DgPartial * len(partials))()
    keep = []
    for i, part in enumerate(partials):
        kind, flux, gq, ge = part[:4]
        left, right = (part[4], part[5]) if len(part) > 4 else (DG_PERIODIC, DG_PERIODIC)
        ptr_q = ptr_e = None
        if gq is not None:
            gq = gq.contiguous()
            keep.append(gq)
            ptr_q = gq.data_ptr()
        if ge is not None:
            ge = ge.contiguous()
            keep.append(ge)
            ptr_e = ge.data_ptr()
        parts[i] = DgPartial(int(kind), float(flux), ptr_q, ptr_e, int(left), int(right))
    Fc = F.contiguous()
    if Fc.shape != (N, N) or Fc.dtype != torch.float64 or not Fc.is_cuda:
        raise AkmInvalid(AKM_ERR_INVALID, "dg_term needs an N x N FP64 CUDA transform")
    out = torch.empty((N, N), dtype=torch.float64, device=Fc.device)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    _check(load_library().akm_dg_term(int(n), int(level), float(domain[0]), float(domain[1]),
                                      int(nq), parts, len(partials), Fc.data_ptr(),
                                      out.data_ptr(), s))
    return out


def simple_gmres_euler(dt, mat: KronmultMatrix, x, b, restart=-1, max_iter=-1, tol=-1.0):
    """Solves (I - dt*A) x = b in place; returns GmresInfo(error, iterations).

    restart/max_iter/tol <= 0 select the reference defaults (parser::NO_USER_VALUE).
    Raises ValueError when restart exceeds the system size."""
    px, dx = _vec_ptr(x, mat.dtype)
    pb, db = _vec_ptr(b, mat.dtype)
    if dx != db:
        raise AkmInvalid(AKM_ERR_INVALID, "x and b must both be host or both device")
    err = _d()
    its = _i()
    _check(load_library().akm_gmres_euler(mat._h, float(dt), px, pb, int(restart),
                                          int(max_iter), float(tol), dx, ctypes.byref(err),
                                          ctypes.byref(its)), allow_not_converged=True)
    return GmresInfo(err.value, its.value)


def gmres_euler_noalloc(dt, mat: KronmultMatrix, x, b, restart=-1, max_iter=-1, tol=-1.0):
    """simple_gmres_euler on device vectors with no device allocation: the
    solver's memory is a fixed per-GPU array of AKM_NOALLOC_SCRATCH_BYTES
    (NOALLOC_SCRATCH_BYTES here).  Same results as simple_gmres_euler;
    raises AkmInvalid when the system and restart do not fit."""
    px, dx = _vec_ptr(x, mat.dtype)
    pb, db = _vec_ptr(b, mat.dtype)
    if not (dx and db):
        raise AkmInvalid(AKM_ERR_INVALID, "x and b must be device vectors")
    err = _d()
    its = _i()
    _check(load_library().akm_gmres_euler_noalloc(mat._h, float(dt), px, pb, int(restart),
                                                  int(max_iter), float(tol), ctypes.byref(err),
                                                  ctypes.byref(its)), allow_not_converged=True)
    return GmresInfo(err.value, its.value)


def simple_gmres(A, x, b, M=None, restart=-1, max_iter=-1, tol=-1.0):
    """solver::simple_gmres on an explicit matrix (src/solver.cpp:54-74): solves
    A x = b in place for host numpy arrays (float64 or float32; A and M n x n,
    M an optional left preconditioner applied through its LU factors);
    returns GmresInfo(error, iterations).  Defaults and the absolute tolerance
    follow simple_gmres_euler; raises ValueError when restart exceeds n."""
    x_arr = x
    if not isinstance(x_arr, np.ndarray) or x_arr.dtype not in (np.float64, np.float32) \
            or not x_arr.flags.c_contiguous:
        raise AkmInvalid(AKM_ERR_INVALID, "x must be a contiguous float64/float32 numpy array")
    dtype = x_arr.dtype.type
    n = x_arr.size
    Af = np.asfortranarray(A, dtype=dtype)
    if Af.shape != (n, n):
        raise AkmInvalid(AKM_ERR_INVALID, "A must be n x n")
    bb = _np(b, dtype)
    Mf = None
    if M is not None:
        Mf = np.asfortranarray(M, dtype=dtype)
        if Mf.shape != (n, n):
            raise AkmInvalid(AKM_ERR_INVALID, "M must be n x n")
    err = _d()
    its = _i()
    _check(load_library().akm_gmres_dense(
        int(n), Af.ctypes.data, x_arr.ctypes.data, bb.ctypes.data,
        None if Mf is None else Mf.ctypes.data, int(restart), int(max_iter), float(tol),
        AKM_F64 if dtype == np.float64 else AKM_F32, ctypes.byref(err), ctypes.byref(its)),
        allow_not_converged=True)
    return GmresInfo(err.value, its.value)


def gmres_euler_mixed(dt, mat: KronmultMatrix, x, b, mat32: KronmultMatrix = None,
                      precond: "MlpPreconditioner" = None, restart=-1, max_cycles=-1,
                      tol=1e-10, inner_tol=-1.0) -> SolveStats:
    """(I - dt*A) x = b by FP64 iterative refinement around flexible GMRES
    cycles: FP64 residuals, basis, Gram-Schmidt and least squares; the FP32
    operator `mat32` inside the cycles (None = `mat`); the right
    preconditioner `precond` on every basis vector (None = none); plain FP64
    GMRES takes over when a cycle fails to reduce the FP64 residual.  `tol` is
    relative to |b|.  x (initial guess, overwritten) and b are float64 CUDA
    tensors.  Returns the :class:`SolveStats`."""
    if mat.prec != AKM_F64:
        raise AkmInvalid(AKM_ERR_INVALID, "the operator must be FP64")
    px, dx = _vec_ptr(x, np.float64)
    pb, db = _vec_ptr(b, np.float64)
    if not (dx and db):
        raise AkmInvalid(AKM_ERR_INVALID, "gmres_euler_mixed needs device tensors")
    st = SolveStats()
    _check(load_library().akm_gmres_euler_mixed(
        mat._h, mat32._h if mat32 is not None else None,
        precond._h if precond is not None else None, float(dt), px, pb, int(restart),
        int(max_cycles), float(tol), float(inner_tol), ctypes.byref(st)),
        allow_not_converged=True)
    return st


# --------------------------------------------------- learned preconditioner ---

class MlpPreconditioner:
    """Per-cell MLP right preconditioner (akm_mlp_*): z_c = r_c + |r_c| o(x_c),
    x_c = [r_c/|r_c|, cell features]; trained in FP32 on the GPU, applied with
    BF16 weights on the matrix cores."""

    def __init__(self, d, n, table, max_level, hidden=256, seed=0):
        L = load_library()
        table = _np(table, np.int32)
        h = _vp()
        _check(L.akm_mlp_create(int(d), int(n), table.shape[0], table.ctypes.data,
                                int(max_level), int(hidden), int(seed), ctypes.byref(h)))
        self._h = h
        self.d, self.n, self.cells = int(d), int(n), table.shape[0]
        np_ = _i64()
        vals = [_i() for _ in range(5)]
        _check(L.akm_mlp_info(h, ctypes.byref(np_), *[ctypes.byref(v) for v in vals]))
        self.num_params = np_.value
        (self.inputs, self.padded_inputs, self.hidden, self.outputs,
         self.padded_outputs) = [v.value for v in vals]

    def close(self):
        if self._h:
            load_library().akm_precond_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def layout(self):
        """Offsets of W1, b1, W2, b2, W3, b3 in the flat parameter array."""
        H, K1, N3 = self.hidden, self.padded_inputs, self.padded_outputs
        offs = {}
        o = 0
        for name, size in [("W1", H * K1), ("b1", H), ("W2", H * H), ("b2", H),
                           ("W3", N3 * H), ("b3", N3)]:
            offs[name] = (o, size)
            o += size
        return offs

    def set_stream(self, stream_ptr):
        _check(load_library().akm_mlp_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        _check(load_library().akm_mlp_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, params):
        arr = _np(params, np.float32)
        if arr.size != self.num_params:
            raise AkmInvalid(AKM_ERR_INVALID, "wrong parameter count")
        _check(load_library().akm_mlp_set_params(self._h, arr.ctypes.data))

    def forward(self, X, O):
        """O (rows x N3 float32 CUDA tensor) = o(X) for X (rows x K1)."""
        _check(load_library().akm_mlp_forward(self._h, X.data_ptr(), X.shape[0], O.data_ptr()))
        return O

    def gradients(self, X, T, grad=None):
        """(loss, grad) of the mean squared error of o(X) against T."""
        loss = _d()
        _check(load_library().akm_mlp_gradients(self._h, X.data_ptr(), T.data_ptr(), X.shape[0],
                                                ctypes.byref(loss),
                                                grad.data_ptr() if grad is not None else None))
        return loss.value, grad

    def adam_step(self, grad, lr):
        _check(load_library().akm_mlp_adam_step(self._h, grad.data_ptr(), float(lr)))

    def train(self, residuals, corrections, epochs=10, batch=1024, lr=1e-3, seed=0):
        """Trains on (residual, correction) pairs: float64 CUDA tensors of shape
        (pairs, cells*n^d).  Returns the mean loss of the last epoch."""
        R = residuals.contiguous()
        E = corrections.contiguous()
        if R.shape != E.shape or R.shape[-1] != self.cells * self.outputs:
            raise AkmInvalid(AKM_ERR_INVALID, "pairs must be (count, cells*n^d) tensors")
        loss = _d()
        _check(load_library().akm_mlp_train(self._h, R.data_ptr(), E.data_ptr(), R.shape[0],
                                            int(epochs), int(batch), float(lr), int(seed),
                                            ctypes.byref(loss)))
        return loss.value

    def apply(self, v, z, stream_ptr=0):
        """z = P(v) on float64 CUDA tensors (BF16 matrix-core inference)."""
        _check(load_library().akm_precond_apply(self._h, v.data_ptr(), z.data_ptr(),
                                                ctypes.c_void_p(stream_ptr)))
        return z


# ------------------------------------------------------------ batched solves ---

class BatchSolver:
    """A batch of independent implicit steps (I - dt_s A_s) x_s = b_s solved
    together (akm_batch_*): lockstep restarted GMRES, one persistent launch
    per (d, n) group per operator apply, hipGraph-captured cycles."""

    def __init__(self, mats, dts, restart=50):
        L = load_library()
        self._mats = list(mats)  # the matrices must outlive the batch
        handles = (ctypes.c_void_p * len(self._mats))(*[m._h.value for m in self._mats])
        dt = np.ascontiguousarray(dts, dtype=np.float64)
        if dt.size != len(self._mats):
            raise AkmInvalid(AKM_ERR_INVALID, "one dt per system")
        h = _vp()
        _check(L.akm_batch_create(len(self._mats), handles, dt.ctypes.data, int(restart),
                                  ctypes.byref(h)))
        self._h = h
        self.count = len(self._mats)

    def close(self):
        if self._h:
            load_library().akm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        r, g = _i(), _i()
        waves = np.zeros(64, np.int32)
        _check(load_library().akm_batch_info(self._h, ctypes.byref(r), ctypes.byref(g),
                                             waves.ctypes.data))
        return {"restart": r.value, "groups": g.value, "group_waves": waves[:g.value].tolist()}

    def solve(self, xs, bs, max_cycles=-1, tol=1e-10):
        """Solves in place (xs: float64 CUDA tensors, initial guesses); returns
        (relative residuals, operator applications, converged flags)."""
        if len(xs) != self.count or len(bs) != self.count:
            raise AkmInvalid(AKM_ERR_INVALID, "one x and one b per system")
        for v in list(xs) + list(bs):
            _vec_ptr(v, np.float64)
        px = (ctypes.c_void_p * self.count)(*[v.data_ptr() for v in xs])
        pb = (ctypes.c_void_p * self.count)(*[v.data_ptr() for v in bs])
        res = np.zeros(self.count)
        its = np.zeros(self.count, np.int32)
        conv = np.zeros(self.count, np.int32)
        _check(load_library().akm_batch_solve(self._h, px, pb, int(max_cycles), float(tol),
                                              res.ctypes.data, its.ctypes.data,
                                              conv.ctypes.data), allow_not_converged=True)
        return res, its, conv.astype(bool)
