"""Shared synthetic inputs for kronmult parity tests.

Two kinds of inputs:

* ``random_pointer_case`` follows the reference test generator
  make_kronmult_data (src/asgard_kronmult_tests.hpp:111-190): every row is
  connected to every column and each (pair, term, dim) picks one of
  ``num_matrices`` random n x n matrices; values U[-1, 1].
* ``sparse_grid_case`` builds the real operator: the |l|_1 <= L sparse grid,
  its cell-pair CSR and vA from per-(term, dim) dense 1D operators.

Random numbers come from numpy.random.default_rng(seed).
"""
import numpy as np

import oracle


def random_pointer_case(d, n, num_rows, T, num_matrices, seed=42):
    rng = np.random.default_rng(seed)
    N = n ** d
    pntr = np.arange(num_rows + 1, dtype=np.int64) * num_rows
    indx = np.tile(np.arange(num_rows, dtype=np.int32), num_rows)
    nnz = num_rows * num_rows
    mats = rng.integers(0, num_matrices, size=(nnz, T, d))
    iA = (mats * n * n).astype(np.int32).reshape(-1)
    vA = rng.uniform(-1, 1, size=num_matrices * n * n)
    x = rng.uniform(-1, 1, size=num_rows * N)
    y = rng.uniform(-1, 1, size=num_rows * N)
    return dict(d=d, n=n, T=T, pntr=pntr, indx=indx, iA=iA, vA=vA, x=x, y=y,
                rows=num_rows, cols=num_rows)


def sparse_grid_case(d, n, L, T, seed=0, ops=None, dist="uniform"):
    rng = np.random.default_rng(seed)
    table, _ = oracle.element_table([L] * d)
    pntr, indx, iA = oracle.sparse_pattern(table, n=n, T=T, max_level=L)
    size1d = n * (1 << L)
    if ops is None:
        ops = [np.asfortranarray(rng.uniform(-1, 1, size=(size1d, size1d)))
               for _ in range(T * d)]
    vA = oracle.build_vA(d, n, T, L, ops)
    cells = table.shape[0]
    N = n ** d
    if dist == "normal":
        x = rng.standard_normal(cells * N)
    else:
        x = rng.uniform(-1, 1, size=cells * N)
    y = rng.uniform(-1, 1, size=cells * N)
    return dict(d=d, n=n, T=T, L=L, table=table, pntr=pntr, indx=indx, iA=iA,
                vA=vA, ops=ops, x=x, y=y, rows=cells, cols=cells,
                elem=oracle.cell_1d_index(table))


def factor(case, pair, t, k):
    """The n x n factor (row-major numpy matrix) of (pair, term, dim)."""
    d, n, T = case["d"], case["n"], case["T"]
    off = case["iA"][(pair * T + t) * d + k]
    return case["vA"][off:off + n * n].reshape(n, n).T


def bruteforce(case, alpha, beta):
    """y = beta*y + alpha * sum (A_0 kron ... kron A_{d-1}) x_col, with the
    Kronecker products formed by numpy.kron."""
    d, n, T = case["d"], case["n"], case["T"]
    N = n ** d
    pntr, indx = case["pntr"], case["indx"]
    y = beta * case["y"].copy() if beta != 0 else np.zeros_like(case["y"])
    for r in range(len(pntr) - 1):
        acc = np.zeros(N)
        for j in range(pntr[r], pntr[r + 1]):
            xc = case["x"][indx[j] * N:(indx[j] + 1) * N]
            for t in range(T):
                K = np.ones((1, 1))
                for k in range(d):
                    K = np.kron(K, factor(case, j, t, k))
                acc += K @ xc
        y[r * N:(r + 1) * N] += alpha * acc
    return y


def grid_case(d, n, L, T, seed=0, refine=0.0, coupling="full", identity=()):
    return grid_case_ops(d, n, L, T, seed, refine, coupling, identity)[0]


def grid_case_ops(d, n, L, T, seed=0, refine=0.0, coupling="full", identity=()):
    """A sparse-grid case for the global operator: the |l|_1 <= L grid, or
    (refine > 0) the |l|_1 <= L-2 grid refined twice by asgard_amd.grid.refine
    up to level L; U[-1, 1] 1D operators, kept only on the volume connections
    when coupling == "volume"; the (t, k) pairs in ``identity`` get the
    identity operator and active[t][k] = 0."""
    from asgard_amd import grid

    rng = np.random.default_rng(seed)
    if refine > 0:
        table, _ = oracle.element_table([max(1, L - 2)] * d)
        table = grid.refine(table, refine, seed=seed, max_level=L)
        table = grid.refine(table, refine, seed=seed + 1, max_level=L)
    else:
        table, _ = oracle.element_table([L] * d)
    size1d = n << L
    mask = None
    if coupling == "volume":
        vp, vi, _ = oracle.connect_1d(L, full=False)
        M = 1 << L
        vm = np.zeros((M, M), bool)
        for r in range(M):
            vm[r, vi[vp[r]:vp[r + 1]]] = True
        mask = np.kron(vm, np.ones((n, n), bool))
    ops = []
    for t in range(T):
        for k in range(d):
            if (t, k) in identity:
                ops.append(np.asfortranarray(np.eye(size1d)))
                continue
            a = rng.uniform(-1, 1, size=(size1d, size1d))
            if mask is not None:
                a = np.where(mask, a, 0.0)
            ops.append(np.asfortranarray(a))
    vA = oracle.build_vA(d, n, T, L, ops)
    pntr, indx, iA = oracle.sparse_pattern(table, n=n, T=T, max_level=L)
    cells = table.shape[0]
    N = n ** d
    active = [[0 if (t, k) in identity else 1 for k in range(d)] for t in range(T)]
    return dict(d=d, n=n, T=T, L=L, table=table, pntr=pntr, indx=indx, iA=iA, vA=vA,
                x=rng.uniform(-1, 1, cells * N), y=rng.uniform(-1, 1, cells * N),
                rows=cells, cols=cells, active=active), ops
