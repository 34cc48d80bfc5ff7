"""Sparse-grid inputs of the kronmult operator, built on the host.

* :func:`element_table` -- the cell list of ``elements::table``
  (src/elements.cpp:249-350): level tuples selected by
  ``permutations::get_lequal_multi`` (sparse grid: l_k <= levels[k] and
  sum(l) <= max(levels)) or ``get_max_multi`` (full grid), cells of a tuple
  enumerated with dimension 0 fastest (``get_cell_index_set``,
  src/elements.cpp:408-465).  Sparse-grid tuples are grouped by decreasing
  |l|_1 (``select_indexex`` with levels_in_reverse, src/permutations.hpp:195-219)
  and ordered inside a group the way the lower-set odometer visits them,
  which is ascending lexicographic order of the reversed tuple.
* :func:`build_vA` -- the factor-block array of ``make_kronmult_sparse``
  (src/asgard_kronmult_matrix.cpp:369-393): for every connect_1d connection
  (row, col), every term and every dimension, the n x n block of that term's
  1D operator at block (row, col), column-major.
* :func:`build_dense_terms` -- the per-term arrays of ``make_kronmult_dense``
  (src/asgard_kronmult_matrix.cpp:98-114).
* :func:`cell_1d_index` -- 0 for level 0, else 2^(l-1) + p
  (src/asgard_kronmult_matrix.cpp:121-132).
"""
from __future__ import annotations

import itertools

import numpy as np


def level_tuples(levels, full_grid=False):
    levels = [int(v) for v in levels]
    d = len(levels)
    ranges = [range(v + 1) for v in levels]
    tuples = [t for t in itertools.product(*ranges)]
    if full_grid:
        # dimension 0 fastest
        tuples.sort(key=lambda t: t[::-1])
        return np.array(tuples, dtype=np.int32).reshape(-1, d)
    limit = max(levels)
    tuples = [t for t in tuples if sum(t) <= limit]
    tuples.sort(key=lambda t: (-sum(t), t[::-1]))
    return np.array(tuples, dtype=np.int32).reshape(-1, d)


def element_table(levels, max_level=None, full_grid=False):
    """(table[cells, 2d] int32, ids[cells] int64) in elements::table order."""
    levels = [int(v) for v in levels]
    d = len(levels)
    if max_level is None:
        max_level = max(levels)
    blocks = []
    for lt in level_tuples(levels, full_grid):
        counts = [1 if l <= 1 else 1 << (l - 1) for l in lt]
        # cells of the tuple, dimension 0 fastest
        grids = np.meshgrid(*[np.arange(c) for c in counts[::-1]], indexing="ij")
        pos = np.stack([g.ravel() for g in grids[::-1]], axis=1)
        lev = np.broadcast_to(np.array(lt, dtype=np.int64), pos.shape)
        blocks.append(np.concatenate([lev, pos], axis=1))
    table = np.concatenate(blocks).astype(np.int32)
    idx1d = cell_1d_index(table).astype(np.int64)
    stride = (1 << max_level) ** np.arange(d, dtype=np.int64)
    ids = (idx1d * stride).sum(axis=1)
    return table, ids


def children_1d(level, pos):
    """Children of a 1D hierarchical cell: level 0 -> (1, 0); level 1 -> (2, 0),
    (2, 1); level l >= 2 -> (l + 1, 2p), (l + 1, 2p + 1)."""
    if level == 0:
        return [(1, 0)]
    if level == 1:
        return [(2, 0), (2, 1)]
    return [(level + 1, 2 * pos), (level + 1, 2 * pos + 1)]


def refine(table, fraction, seed=0, max_level=None):
    """Adaptive refinement of a cell table, as an input generator: the leaf
    cells (no child present in any dimension) are refined with probability
    ``fraction`` (drawn from numpy default_rng(seed)); a refined cell gets its
    children in every dimension (one level up in that dimension, the others
    unchanged), appended after the existing cells in order of discovery.
    ``max_level`` caps the level of new cells.  Returns the new table."""
    table = np.asarray(table, dtype=np.int32)
    cells, twod = table.shape
    d = twod // 2
    present = {tuple(r) for r in table.tolist()}

    def kids(cell):
        out = []
        for k in range(d):
            for (l, p) in children_1d(cell[k], cell[d + k]):
                if max_level is not None and l > max_level:
                    continue
                c = list(cell)
                c[k], c[d + k] = l, p
                out.append(tuple(c))
        return out

    rng = np.random.default_rng(seed)
    rows = table.tolist()
    leaves = [tuple(r) for r in rows if not any(c in present for c in kids(tuple(r)))]
    pick = rng.random(len(leaves)) < fraction
    new = []
    for leaf, chosen in zip(leaves, pick):
        if not chosen:
            continue
        for c in kids(leaf):
            if c not in present:
                present.add(c)
                new.append(c)
    if not new:
        return table.copy()
    return np.concatenate([table, np.array(new, dtype=np.int32).reshape(-1, twod)])


def cell_1d_index(table):
    table = np.asarray(table)
    d = table.shape[1] // 2
    lev, pos = table[:, :d].astype(np.int64), table[:, d:].astype(np.int64)
    out = np.where(lev == 0, 0, (np.int64(1) << np.maximum(lev - 1, 0)) + pos)
    return out.astype(np.int32)


def build_vA(d, n, T, max_level, ops, connectivity=None):
    """ops[t*d + k]: square 1D operator (n*2^max_level) of term t, dim k."""
    if connectivity is None:
        from . import connect_1d
        connectivity = connect_1d(max_level)
    pntr, indx, _ = connectivity
    rows = np.repeat(np.arange(len(pntr) - 1), np.diff(pntr))
    cols = np.asarray(indx)
    nconn = len(cols)
    ri = (n * rows)[:, None] + np.arange(n)[None, :]  # nconn x n
    ci = (n * cols)[:, None] + np.arange(n)[None, :]
    out = np.empty((nconn, T, d, n * n))
    for t in range(T):
        for k in range(d):
            op = np.asarray(ops[t * d + k])
            blk = op[ri[:, :, None], ci[:, None, :]]  # nconn x n(s) x n(j)
            out[:, t, k, :] = blk.transpose(0, 2, 1).reshape(nconn, n * n)  # column-major
    return out.reshape(-1)


def build_dense_terms(d, n, T, num_1d_blocks, ops):
    nb = num_1d_blocks
    terms = []
    for t in range(T):
        parts = []
        for k in range(d):
            op = np.asarray(ops[t * d + k])[:n * nb, :n * nb]
            # blocks ordered (ocol, orow), each column-major: [ocol][orow][i][s]
            b = op.reshape(nb, n, nb, n)  # orow, s, ocol, i
            parts.append(b.transpose(2, 0, 3, 1).reshape(-1))
        terms.append(np.concatenate(parts))
    return terms


def identity_mask(ops, num_terms, d):
    """active[t][k] for akm_create_global_grid: 0 where the (t, k) 1D operator
    is exactly the identity matrix, 1 elsewhere (the reference decides this
    from the term's structure: a pure mass term is the identity in the
    orthonormal basis, src/asgard_kronmult_matrix.cpp:1175-1184)."""
    out = []
    for t in range(num_terms):
        row = []
        for k in range(d):
            a = np.asarray(ops[t * d + k])
            row.append(0 if a.shape[0] == a.shape[1] and np.array_equal(a, np.eye(a.shape[0]))
                       else 1)
        out.append(row)
    return out
