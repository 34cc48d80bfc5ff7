"""CPU oracle for the reference's global Kronecker operator.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` (never by ``asgard_amd``).

A numpy restatement of the reference's ``global_kron_matrix`` (the default
single-node kronmult mode), small sizes only:

* ``padding``        -- ``compute_ancestry_completion`` over the volume
                        hierarchy (src/asgard_indexset.cpp:243-318)
* ``dof_indices``    -- ``complete_poly_order`` (src/asgard_indexset.cpp:333-405)
* ``permutations``   -- ``kronmult::permutes`` (src/device/asgard_kronmult.hpp:152-220)
* ``apply``          -- ``make_global_kron_matrix`` pattern (src/asgard_kronmult_matrix.cpp:
                        1086-1128, 1198-1334), ``global_cpu``/``global_cpu_one``
                        (src/device/asgard_glkronmult_cpu.cpp:108-189) and
                        ``global_kron_matrix::apply`` (src/asgard_kronmult_matrix.cpp:1528-1590)
* ``jacobi``         -- ``build_preconditioner`` (src/asgard_kronmult_matrix.cpp:1016-1067)
* ``jacobi_scale``   -- ``apply_diagonal_precond`` (src/solver.cpp:76-92)

Parity status: ``apply`` is pinned by the identity that defines it -- on every
grid the operator equals the pairwise sparse-grid operator (``oracle.kronmult_sparse``,
itself pinned by explicit Kronecker products) on the active cells, checked on
standard and refined grids, with and without identity factors
(tests/test_oracle_global.py); ``permutations`` by the closed form of the
reference's enumeration (2^(k-1) chains, upper factors before the full one,
lower after); ``padding`` by brute force (every ancestor in every dimension is
present); ``jacobi`` by the diagonal of the explicitly assembled operator.
"""
from __future__ import annotations

import numpy as np

from . import cell_1d_index, connect_1d


def _parents(i):
    """The volume-hierarchy ancestors of 1D cell i: every cell on its parent
    chain down to cell 0 (src/asgard_grid_1d.hpp:159-260, the part of a row
    before its diagonal)."""
    out = []
    while i > 0:
        i = 0 if i == 1 else i // 2
        out.append(i)
    return out


def padding(elem):
    """Cells (1D index tuples) missing from the set to make it ancestry
    complete in every dimension, sorted."""
    have = {tuple(int(v) for v in e) for e in elem}
    pad = set()
    frontier = list(have)
    while frontier:
        nxt = []
        for c in frontier:
            for k in range(len(c)):
                for a in _parents(c[k]):
                    anc = c[:k] + (a,) + c[k + 1:]
                    if anc not in have and anc not in pad:
                        pad.add(anc)
                        nxt.append(anc)
        frontier = nxt
    return sorted(pad)


def dof_indices(cells, n):
    """1D DOF indices (cell 1D index * n + digit) of every tensor entry of
    every cell, cell-major, the last dimension's digit fastest."""
    cells = np.asarray(cells, dtype=np.int64).reshape(len(cells), -1)
    d = cells.shape[1]
    N = n ** d
    e = np.arange(N)
    out = np.empty((len(cells) * N, d), dtype=np.int64)
    for k in range(d):
        digit = (e // n ** (d - 1 - k)) % n
        out[:, k] = (cells[:, k][:, None] * n + digit[None, :]).reshape(-1)
    return out


def permutations(k):
    """Chains of one term with k active dimensions: 2^(k-1) lists of
    (active-dimension position, fill), applied first to last; fill is 'U'
    (upper incl. diagonal), 'F' (full) or 'L' (strictly lower)."""
    out = []
    for p in range(2 ** (k - 1) if k > 0 else 0):
        signed = [0]
        t = p
        for dd in range(1, k):
            signed.append(dd if t % 2 == 0 else -dd)
            t //= 2
        signed.sort()
        out.append([(abs(v), "U" if v < 0 else ("L" if v > 0 else "F")) for v in signed])
    return out


def dof_operators(d, n, T, L, vA):
    """Dense 1D DOF operators op[t][k] ((n 2^L) square) holding the vA blocks
    at the connect_1d connections and zeros elsewhere."""
    pntr, indx, _ = connect_1d(L)
    M = 1 << L
    nn = n * n
    blocks = np.asarray(vA).reshape(-1, T, d, nn)
    ops = np.zeros((T, d, n * M, n * M))
    for r in range(M):
        for q in range(pntr[r], pntr[r + 1]):
            c = indx[q]
            for t in range(T):
                for k in range(d):
                    # column-major block: A[j*n + s] = A(s, j)
                    ops[t, k, r * n:(r + 1) * n, c * n:(c + 1) * n] = \
                        blocks[q, t, k].reshape(n, n).T
    return ops


def _sweep(dof, k, A, fill, v):
    """One dimension's product over the fibres (DOFs equal in all other
    dimensions), restricted to the fill."""
    if fill == "L":
        A = np.tril(A, -1)
    elif fill == "U":
        A = np.triu(A, 0)
    out = np.zeros_like(v)
    others = np.delete(dof, k, axis=1)
    if others.shape[1] == 0:  # one dimension: a single fibre
        others = np.zeros((dof.shape[0], 1), dtype=np.int64)
    order = np.lexsort(tuple(others[:, j] for j in range(others.shape[1] - 1, -1, -1)))
    keys = others[order]
    start = 0
    for i in range(1, len(order) + 1):
        if i == len(order) or np.any(keys[i] != keys[start]):
            ids = order[start:i]
            ik = dof[ids, k]
            out[ids] += A[np.ix_(ik, ik)] @ v[ids]
            start = i
    return out


def apply(d, n, table, L, T, vA, alpha, x, beta, y, active=None):
    """y = alpha * A x + beta * y with the global algorithm on the padded set.

    ``active[t][k]`` (default: all True) marks the non-identity factors; the
    factors marked inactive are taken as the identity, as the reference does
    for mass terms (src/asgard_kronmult_matrix.cpp:1175-1184, 1313-1331)."""
    elem = cell_1d_index(table)
    pad = padding(elem)
    cells = np.concatenate([elem, np.asarray(pad, dtype=np.int64).reshape(-1, d)]) \
        if pad else np.asarray(elem, dtype=np.int64)
    dof = dof_indices(cells, n)
    num_active = elem.shape[0] * n ** d
    ops = dof_operators(d, n, T, L, vA)
    xp = np.zeros(dof.shape[0])
    xp[:num_active] = x
    acc = np.zeros(dof.shape[0])
    for t in range(T):
        act = [k for k in range(d) if active is None or active[t][k]]
        if not act:
            acc += xp
            continue
        for chain in permutations(len(act)):
            w = xp
            for pos, fill in chain:
                k = act[pos]
                w = _sweep(dof, k, ops[t, k], fill, w)
            acc += w
    return beta * np.asarray(y, dtype=np.float64) + alpha * acc[:num_active]


def jacobi(d, n, table, L, T, vA):
    """pc[i] = sum over terms of the product over dimensions of the 1D
    operators' diagonal entries at DOF i."""
    elem = cell_1d_index(table)
    dof = dof_indices(elem, n)
    ops = dof_operators(d, n, T, L, vA)
    pc = np.zeros(dof.shape[0])
    for t in range(T):
        a = np.ones(dof.shape[0])
        for k in range(d):
            a *= np.diagonal(ops[t, k])[dof[:, k]]
        pc += a
    return pc


def jacobi_scale(pc, dt, x):
    """The left preconditioner of the implicit step: x / (1 - dt * pc)."""
    return np.asarray(x) / (1.0 - dt * np.asarray(pc))
