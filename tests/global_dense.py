"""An independent statement of the global Kronecker operator, for tests.

The reference's global operator (global_kron_matrix,
src/asgard_kronmult_matrix.cpp:1198-1426 with global_cpu,
src/device/asgard_glkronmult_cpu.cpp:151-189) applies, per term, a sum of
chains; each stage of a chain is the 1D operator of one dimension -- its
upper part (diagonal included), all of it, or its strictly lower part --
acting along that dimension on the padded cell set S (the cells completed
with all their volume ancestors), i.e. the tensor-product operator
I x .. x A_k x .. x I restricted to S.

Here that definition is evaluated literally on the full tensor DOF grid
((n 2^L)^d entries): a vector lives on the whole grid with zeros outside S,
a stage is ``tensordot`` of the masked 1D matrix along its axis followed by
the projection onto S.  The 1D matrices are the caller's dense operators
restricted to the connect_1d block pattern, and S is the ancestry closure
computed here.  Nothing is shared with oracle/global_kron.py except the
chain enumeration ``permutations``, which its own closed-form test pins.
"""
import numpy as np

import oracle
from oracle import global_kron as gk


def volume_ancestors(i):
    out = []
    while i > 0:
        i = 0 if i == 1 else i // 2
        out.append(i)
    return out


def closure(cells):
    """Smallest superset of the cells (1D index tuples) closed under
    replacing one coordinate by one of its volume ancestors."""
    have = set(cells)
    stack = list(cells)
    while stack:
        c = stack.pop()
        for k in range(len(c)):
            for a in volume_ancestors(c[k]):
                nc = c[:k] + (a,) + c[k + 1:]
                if nc not in have:
                    have.add(nc)
                    stack.append(nc)
    return have


def masked_1d(op, n, L):
    """The dense 1D operator restricted to the connect_1d(L) block pattern."""
    pntr, indx, _ = oracle.connect_1d(L)
    M = 1 << L
    keep = np.zeros((M, M), bool)
    for r in range(M):
        keep[r, indx[pntr[r]:pntr[r + 1]]] = True
    return np.where(np.kron(keep, np.ones((n, n), bool)), op, 0.0)


def fill(A, f):
    return np.triu(A) if f == "U" else (np.tril(A, -1) if f == "L" else A)


def apply(d, n, L, T, ops, table, x, active=None, flux_dim=None):
    """A x on the active cells of `table` (x: cell-major, tensor row-major)."""
    elem = [tuple(int(v) for v in e) for e in oracle.cell_1d_index(table)]
    S = closure(elem)
    side = n << L
    mask = np.zeros((1 << L,) * d, bool)
    for c in S:
        mask[c] = True
    # DOF mask on the full grid: cell mask expanded by n in every dimension
    dmask = mask
    for k in range(d):
        dmask = np.repeat(dmask, n, axis=k)
    N = n ** d
    full = np.zeros((side,) * d)
    xs = x.reshape(len(elem), N)
    for ci, c in enumerate(elem):
        sl = tuple(slice(c[k] * n, c[k] * n + n) for k in range(d))
        full[sl] = xs[ci].reshape((n,) * d)
    acc = np.zeros_like(full)
    for t in range(T):
        act = [k for k in range(d) if active is None or active[t][k]]
        if flux_dim is not None and len(act) > 1:
            # the reference swaps positions 0 and flux_dir of the active list
            # when flux_dir is not already first
            # (src/asgard_kronmult_matrix.cpp:1321-1323)
            fd = flux_dim[t]
            if 0 <= fd < len(act) and fd != act[0]:
                act[0], act[fd] = act[fd], act[0]
        if not act:
            acc += full
            continue
        for chain in gk.permutations(len(act)):
            w = full
            for pos, f in chain:
                k = act[pos]
                A = fill(masked_1d(ops[t * d + k], n, L), f)
                w = np.moveaxis(np.tensordot(A, w, axes=([1], [k])), 0, k) * dmask
            acc += w
    out = np.zeros_like(xs)
    for ci, c in enumerate(elem):
        sl = tuple(slice(c[k] * n, c[k] * n + n) for k in range(d))
        out[ci] = acc[sl].reshape(-1)
    return out.reshape(-1)
