"""CPU oracle for the kronmult apply and the implicit-step GMRES.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.
The product (``asgard_amd``) never imports, loads or calls anything here.

Thin ctypes wrapper over ``liborc.so`` (built from grid.c, kronmult.c, gmres.c
by the Makefile in this folder).  Each C function restates a reference
function; the file:line citations are in ``asgard_oracle.h``.

Parity status of each function (how it is pinned):

* ``element_table``       -- reference fixtures element_table/table_*, ids_*
* ``connect_1d``          -- known answers, src/asgard_indexset_tests.cpp:83-155
* ``sparse_pattern``      -- symmetry, diagonal, every connected pair has a 1D
                             offset, brute force over 1D supports, nnz totals
* ``kronmult_sparse``     -- explicit Kronecker product + gemv (brute force),
                             src/asgard_kronmult_matrix_tests.cpp:178-193 known answers
* ``kronmult_dense``      -- explicit Kronecker product + gemv (brute force)
* ``gmres_*``             -- 3x3 gold of src/solver_tests.cpp:141-213, dense LU
                             solve of (I - dt A) x = b
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_vp = ctypes.c_void_p
_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double


def build(force: bool = False) -> str:
    """Compile liborc.so with the Makefile in this folder."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "all", "-C", _HERE],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_num_level_tuples.restype = _c_i64
        L.orc_num_level_tuples.argtypes = [_c_int, _i32p, _c_int]
        L.orc_level_tuples.argtypes = [_c_int, _i32p, _c_int, _i32p]
        L.orc_num_cells.restype = _c_i64
        L.orc_num_cells.argtypes = [_c_int, _i32p, _c_int]
        L.orc_element_table.argtypes = [_c_int, _i32p, _c_int, _c_int, _i32p, _i64p]
        L.orc_connect_1d.argtypes = [_c_int, _c_int, _vp, _vp, _vp]
        L.orc_connect_1d_offset.argtypes = [_i32p, _i32p, _c_int, _c_int]
        L.orc_check_connected_1d.argtypes = [_c_int] * 4
        L.orc_check_connected.argtypes = [_c_int, _i32p, _i32p]
        L.orc_sparse_pattern_count.restype = _c_i64
        L.orc_sparse_pattern_count.argtypes = [_c_int, _c_i64, _i32p, _i64p]
        L.orc_sparse_pattern_fill.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_i64,
                                              _i32p, _i64p, _i32p, _vp]
        L.orc_sparse_pattern_rows.restype = _c_i64
        L.orc_sparse_pattern_rows.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_i64, _i32p,
                                              _i32p, _c_int, _i64p, _vp, _vp]
        L.orc_cell_1d_index.argtypes = [_c_int, _c_i64, _i32p, _i32p]
        L.orc_build_vA.argtypes = [_c_int, _c_int, _c_int, _c_int, _vp, _c_i64, _f64p]
        L.orc_build_dense_terms.argtypes = [_c_int, _c_int, _c_int, _c_int, _vp,
                                            _c_i64, _vp]
        L.orc_kronmult_sparse.argtypes = [_c_int, _c_int, _c_int, _i64p, _i32p, _c_int,
                                          _i32p, _f64p, _c_dbl, _f64p, _c_dbl, _f64p]
        L.orc_kronmult_sparse_f32.argtypes = [_c_int, _c_int, _c_int, _i64p, _i32p,
                                              _c_int, _i32p, _f32p, ctypes.c_float,
                                              _f32p, ctypes.c_float, _f32p]
        L.orc_kronmult_dense.argtypes = [_c_int, _c_int, _c_int, _c_int, _c_int, _i32p,
                                         _c_int, _c_int, _vp, _c_int, _c_dbl, _f64p,
                                         _c_dbl, _f64p]
        L.orc_kronecker.argtypes = [_c_int, _f64p, _c_int, _f64p, _f64p]
        L.orc_kronmult_one_bruteforce.argtypes = [_c_int, _c_int, _vp, _f64p, _f64p]
        L.orc_simple_gmres_dense.argtypes = [_c_int, _f64p, _vp, _f64p, _f64p, _c_int,
                                             _c_int, _c_dbl, ctypes.POINTER(_c_dbl),
                                             ctypes.POINTER(_c_int)]
        L.orc_gmres_euler_sparse.argtypes = [_c_int, _c_int, _c_int, _i64p, _i32p,
                                             _c_int, _i32p, _f64p, _c_dbl, _f64p, _f64p,
                                             _c_int, _c_int, _c_dbl,
                                             ctypes.POINTER(_c_dbl),
                                             ctypes.POINTER(_c_int)]
        L.orc_dense_solve.argtypes = [_c_int, _f64p, _f64p]
        L.orc_num_threads.restype = _c_int
        L.orc_set_num_threads.argtypes = [_c_int]
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr_array(arrays):
    arr = (ctypes.c_void_p * len(arrays))()
    for i, a in enumerate(arrays):
        arr[i] = a.ctypes.data
    return arr


# ------------------------------------------------------------------ grid ---

def level_tuples(levels, full_grid=False):
    levels = _i32(levels)
    d = len(levels)
    num = lib().orc_num_level_tuples(d, levels, int(full_grid))
    out = np.zeros((num, d), dtype=np.int32)
    lib().orc_level_tuples(d, levels, int(full_grid), out.reshape(-1))
    return out


def element_table(levels, max_level=None, full_grid=False):
    """Returns (table[cells, 2d], ids[cells])."""
    levels = _i32(levels)
    d = len(levels)
    if max_level is None:
        max_level = int(levels.max())
    num = lib().orc_num_cells(d, levels, int(full_grid))
    table = np.zeros((num, 2 * d), dtype=np.int32)
    ids = np.zeros(num, dtype=np.int64)
    lib().orc_element_table(d, levels, int(full_grid), int(max_level),
                            table.reshape(-1), ids)
    return table, ids


def connect_1d(max_level, full=True):
    """Returns (pntr, indx, diag) of connect_1d(max_level)."""
    mode = 1 if full else 0
    nconn = lib().orc_connect_1d(int(max_level), mode, None, None, None)
    if nconn < 0:
        raise ValueError("invalid level")
    rows = 1 << max_level
    pntr = np.zeros(rows + 1, dtype=np.int32)
    indx = np.zeros(nconn, dtype=np.int32)
    diag = np.zeros(rows, dtype=np.int32)
    lib().orc_connect_1d(int(max_level), mode, pntr.ctypes.data, indx.ctypes.data,
                         diag.ctypes.data)
    return pntr, indx, diag


def connect_1d_offset(pntr, indx, row, col):
    return lib().orc_connect_1d_offset(_i32(pntr), _i32(indx), int(row), int(col))


def check_connected_1d(L1, p1, L2, p2):
    return bool(lib().orc_check_connected_1d(L1, p1, L2, p2))


def check_connected(d, row_coords, col_coords):
    return bool(lib().orc_check_connected(d, _i32(row_coords), _i32(col_coords)))


def cell_1d_index(table):
    table = _i32(table)
    cells, twod = table.shape
    d = twod // 2
    elem = np.zeros((cells, d), dtype=np.int32)
    lib().orc_cell_1d_index(d, cells, table.reshape(-1), elem.reshape(-1))
    return elem


def sparse_pattern(table, n=2, T=1, max_level=None, with_iA=True):
    """CSR of connected cell pairs (pntr int64, indx int32) and iA (int32)."""
    table = _i32(table)
    cells, twod = table.shape
    d = twod // 2
    if max_level is None:
        max_level = int(table[:, :d].max())
    max_level = max(int(max_level), 1)
    pntr = np.zeros(cells + 1, dtype=np.int64)
    nnz = lib().orc_sparse_pattern_count(d, cells, table.reshape(-1), pntr)
    indx = np.zeros(nnz, dtype=np.int32)
    iA = np.zeros(nnz * T * d if with_iA else 0, dtype=np.int32)
    rc = lib().orc_sparse_pattern_fill(d, n, T, max_level, cells, table.reshape(-1),
                                       pntr, indx, iA.ctypes.data if with_iA else None)
    if rc != 0:
        raise RuntimeError(f"sparse_pattern_fill failed ({rc})")
    return pntr, indx, iA


def sparse_pattern_rows(table, rows, n=2, T=1, max_level=None):
    """CSR (pntr, indx) and iA of the selected row cells only."""
    table = _i32(table)
    cells, twod = table.shape
    d = twod // 2
    if max_level is None:
        max_level = int(table[:, :d].max())
    max_level = max(int(max_level), 1)
    rows = _i32(rows)
    pntr = np.zeros(len(rows) + 1, dtype=np.int64)
    nnz = lib().orc_sparse_pattern_rows(d, n, T, max_level, cells, table.reshape(-1), rows,
                                        len(rows), pntr, None, None)
    if nnz < 0:
        raise RuntimeError("pattern rows failed")
    indx = np.zeros(nnz, dtype=np.int32)
    iA = np.zeros(nnz * T * d, dtype=np.int32)
    rc = lib().orc_sparse_pattern_rows(d, n, T, max_level, cells, table.reshape(-1), rows,
                                       len(rows), pntr, indx.ctypes.data, iA.ctypes.data)
    if rc < 0:
        raise RuntimeError("pattern rows failed")
    return pntr, indx, iA


def build_vA(d, n, T, max_level, ops):
    """ops: list of T*d column-major (Fortran-order) square matrices, index t*d+k."""
    ops = [np.asfortranarray(o, dtype=np.float64) for o in ops]
    lda = ops[0].shape[0]
    nconn = lib().orc_connect_1d(int(max_level), 1, None, None, None)
    vA = np.zeros(nconn * d * T * n * n, dtype=np.float64)
    lib().orc_build_vA(d, n, T, int(max_level), _ptr_array(ops), lda, vA)
    return vA


def build_dense_terms(d, n, T, num_1d_blocks, ops):
    ops = [np.asfortranarray(o, dtype=np.float64) for o in ops]
    lda = ops[0].shape[0]
    terms = [np.zeros(d * num_1d_blocks * num_1d_blocks * n * n) for _ in range(T)]
    lib().orc_build_dense_terms(d, n, T, num_1d_blocks, _ptr_array(ops), lda,
                                _ptr_array(terms))
    return terms


# -------------------------------------------------------------- kronmult ---

def kronmult_sparse(d, n, pntr, indx, T, iA, vA, alpha, x, beta, y):
    """y <- beta*y + alpha*A*x in place (float64 or float32 by vA dtype)."""
    rows = len(pntr) - 1
    pntr = np.ascontiguousarray(pntr, dtype=np.int64)
    if vA.dtype == np.float32:
        rc = lib().orc_kronmult_sparse_f32(d, n, rows, pntr, _i32(indx), T, _i32(iA),
                                           vA, alpha, x, beta, y)
    else:
        rc = lib().orc_kronmult_sparse(d, n, rows, pntr, _i32(indx), T, _i32(iA), vA,
                                       alpha, x, beta, y)
    if rc != 0:
        raise ValueError("unsupported (d, n)")
    return y


def kronmult_dense(d, n, num_rows, num_cols, T, elem, row_offset, col_offset,
                   terms, num_1d_blocks, alpha, x, beta, y):
    rc = lib().orc_kronmult_dense(d, n, num_rows, num_cols, T, _i32(elem).reshape(-1),
                                  row_offset, col_offset, _ptr_array(terms),
                                  num_1d_blocks, alpha, x, beta, y)
    if rc != 0:
        raise ValueError("unsupported (d, n)")
    return y


def kronecker(A, B):
    """A (m x m) kron B (n x n), flat column-major in and out."""
    m = int(round(np.sqrt(len(A))))
    n = int(round(np.sqrt(len(B))))
    out = np.zeros(m * m * n * n)
    lib().orc_kronecker(m, np.ascontiguousarray(A, dtype=np.float64), n,
                        np.ascontiguousarray(B, dtype=np.float64), out)
    return out


def kronmult_one_bruteforce(factors, x, y):
    d = len(factors)
    n = int(round(np.sqrt(len(factors[0]))))
    factors = [np.ascontiguousarray(f, dtype=np.float64) for f in factors]
    lib().orc_kronmult_one_bruteforce(d, n, _ptr_array(factors), x, y)
    return y


# ----------------------------------------------------------------- GMRES ---

def gmres_dense(A, b, x0=None, M=None, restart=None, max_outer=None, tol=1e-12):
    """simple_gmres on a dense matrix; returns (x, error, iterations)."""
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    Mf = None if M is None else np.asfortranarray(M, dtype=np.float64)
    err = _c_dbl()
    its = _c_int()
    rc = lib().orc_simple_gmres_dense(n, A.reshape(-1, order="F"),
                                      None if Mf is None else Mf.ctypes.data, x,
                                      np.ascontiguousarray(b, dtype=np.float64),
                                      restart or n, max_outer or n, tol,
                                      ctypes.byref(err), ctypes.byref(its))
    if rc == -1:
        raise ValueError("restart larger than the system size")
    if rc != 0:
        raise RuntimeError(f"gmres failed ({rc})")
    return x, err.value, its.value


def gmres_euler_sparse(d, n, pntr, indx, T, iA, vA, dt, x, b, restart, max_outer,
                       tol):
    err = _c_dbl()
    its = _c_int()
    rc = lib().orc_gmres_euler_sparse(d, n, len(pntr) - 1,
                                      np.ascontiguousarray(pntr, dtype=np.int64),
                                      _i32(indx), T, _i32(iA), vA, dt, x,
                                      np.ascontiguousarray(b, dtype=np.float64),
                                      restart, max_outer, tol, ctypes.byref(err),
                                      ctypes.byref(its))
    if rc == -1:
        raise ValueError("restart larger than the system size")
    return x, err.value, its.value


def dense_solve(A, b):
    A = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    x = np.array(b, dtype=np.float64)
    flat = A.reshape(-1, order="F").copy()
    info = lib().orc_dense_solve(A.shape[0], flat, x)
    if info != 0:
        raise np.linalg.LinAlgError("singular matrix")
    return x


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    lib().orc_set_num_threads(int(n))
