"""Pins the oracle's sparse-grid index construction.

* connect_1d against the known answers of the reference's
  src/asgard_indexset_tests.cpp:83-155;
* the element table (order, coordinates, ids) against the reference's
  fixtures element_table/{table,ids}_{1,2,3}d_{SG,FG}.dat
  (src/elements_tests.cpp:10-40, 245-275);
* check_connected against an independent geometric rule: two cells are
  connected iff their closed 1D supports intersect on the periodic unit
  interval in every dimension (overlap, shared edge, or shared boundary
  point through the periodic wrap);
* the cell-pair CSR: symmetric, diagonal present, every connected pair has a
  1D coefficient offset, and the nnz totals of the sparse grids used by the
  benchmarks.
"""
import itertools

import numpy as np
import pytest

import oracle
from tests.fixtures import golden, read_octave


# ------------------------------------------------------------- connect_1d ---

def test_connect_1d_full_level3():
    pntr, indx, diag = oracle.connect_1d(3, full=True)
    assert len(pntr) - 1 == 8
    assert len(indx) == 60
    assert list(np.diff(pntr)) == [8, 8, 8, 8, 7, 7, 7, 7]
    assert list(indx[pntr[4]:pntr[5]]) == [0, 1, 2, 3, 4, 5, 7]
    for r in range(8):
        assert indx[diag[r]] == r


def test_connect_1d_volume():
    pntr, indx, _ = oracle.connect_1d(1, full=False)
    assert len(pntr) - 1 == 2 and len(indx) == 4
    assert indx[0] == 0 and indx[3] == 1 and indx[1] == 1 and indx[2] == 0
    pntr, indx, _ = oracle.connect_1d(4, full=False)
    assert len(pntr) - 1 == 16 and len(indx) == 114
    assert list(indx[pntr[8]:pntr[9]]) == [0, 1, 2, 4, 8]
    assert list(indx[pntr[5]:pntr[6]]) == [0, 1, 2, 5, 10, 11]
    assert list(indx[pntr[13]:pntr[14]]) == [0, 1, 3, 6, 13]


@pytest.mark.parametrize("L", [2, 3, 4, 5, 6, 7])
def test_connect_1d_rows_sorted_and_offsets(L):
    pntr, indx, diag = oracle.connect_1d(L)
    for r in range(1 << L):
        row = indx[pntr[r]:pntr[r + 1]]
        assert np.all(np.diff(row) > 0), "columns must be strictly increasing"
        assert indx[diag[r]] == r
        for j, c in enumerate(row):
            assert oracle.connect_1d_offset(pntr, indx, r, c) == pntr[r] + j


# ---------------------------------------------------------- element table ---

ELEMENT_CASES = [((7,), "1d"), ((5, 2), "2d"), ((3, 2, 3), "3d")]


@pytest.mark.parametrize("levels,tag", ELEMENT_CASES)
@pytest.mark.parametrize("grid", ["SG", "FG"])
def test_element_table_fixture(levels, tag, grid):
    gold_table = read_octave(golden("element_table", f"table_{tag}_{grid}.dat"))
    gold_ids = read_octave(golden("element_table", f"ids_{tag}_{grid}.dat")).ravel()
    table, ids = oracle.element_table(levels, max_level=7, full_grid=(grid == "FG"))
    assert table.shape == gold_table.shape
    assert np.array_equal(table, gold_table.astype(np.int32))
    assert np.array_equal(ids, gold_ids.astype(np.int64))


def test_level_tuple_order_sparse():
    # groups of equal |l|_1 from the largest sum down, zero tuple last;
    # lexicographic (dim 0 fastest) within a group
    t = oracle.level_tuples([2, 2])
    assert [tuple(r) for r in t] == [(2, 0), (1, 1), (0, 2), (1, 0), (0, 1), (0, 0)]


def cells_closed_form(d, L):
    """Cells of the |l|_1 <= L sparse grid: sum over level tuples of
    prod 2^max(0, l-1)."""
    total = 0
    for lt in itertools.product(range(L + 1), repeat=d):
        if sum(lt) <= L:
            total += int(np.prod([1 << max(0, l - 1) for l in lt]))
    return total


@pytest.mark.parametrize("d,L", [(2, 5), (3, 6), (4, 7), (6, 7), (5, 5)])
def test_cell_counts(d, L):
    table, _ = oracle.element_table([L] * d)
    assert table.shape[0] == cells_closed_form(d, L)


# --------------------------------------------------------- connectivity ---

def support(L, p, M):
    """Closed support of 1D cell (L, p) in units of 2^-(M-1)."""
    if L <= 1:
        return 0, 1 << (M - 1)
    w = 1 << (M - L)
    return p * w, (p + 1) * w


def touches(a, b, C):
    (a0, a1), (b0, b1) = a, b
    if max(a0, b0) <= min(a1, b1):
        return True
    # periodic wrap: the point 0 == C
    return (a1 == C and b0 == 0) or (b1 == C and a0 == 0)


@pytest.mark.parametrize("M", [3, 4, 5])
def test_check_connected_1d_geometric(M):
    C = 1 << (M - 1)
    cells = [(0, 0)] + [(l, p) for l in range(1, M + 1) for p in range(1 << (l - 1))]
    for (L1, p1), (L2, p2) in itertools.product(cells, repeat=2):
        if L1 > L2:
            continue
        geo = touches(support(L1, p1, M), support(L2, p2, M), C)
        assert oracle.check_connected_1d(L1, p1, L2, p2) == geo, (L1, p1, L2, p2)


@pytest.mark.parametrize("M", [3, 5, 7])
def test_connect_1d_matches_check_connected(M):
    # the 1D connectivity used for the coefficient offsets and the
    # check_connected rule must describe the same graph
    pntr, indx, _ = oracle.connect_1d(M)
    cells = [(0, 0)] + [(l, p) for l in range(1, M + 1) for p in range(1 << (l - 1))]
    idx1d = lambda c: 0 if c[0] == 0 else (1 << (c[0] - 1)) + c[1]
    for a in cells:
        row = set(indx[pntr[idx1d(a)]:pntr[idx1d(a) + 1]].tolist())
        for b in cells:
            lo, hi = (a, b) if a[0] <= b[0] else (b, a)
            conn = oracle.check_connected_1d(lo[0], lo[1], hi[0], hi[1])
            assert (idx1d(b) in row) == conn, (a, b)


def pattern_bruteforce(table):
    d = table.shape[1] // 2
    M = max(int(table[:, :d].max()), 1)
    C = 1 << (M - 1)
    sup = [[support(int(r[k]), int(r[d + k]), M) for k in range(d)] for r in table]
    n = table.shape[0]
    dense = np.zeros((n, n), dtype=bool)
    for i in range(n):
        for j in range(n):
            dense[i, j] = all(touches(sup[i][k], sup[j][k], C) for k in range(d))
    return dense


@pytest.mark.parametrize("levels", [(3, 3), (2, 2, 2), (4, 2), (2, 2, 2, 2)])
def test_sparse_pattern_bruteforce(levels):
    table, _ = oracle.element_table(levels)
    pntr, indx, iA = oracle.sparse_pattern(table, n=2, T=2)
    dense = pattern_bruteforce(table)
    n = table.shape[0]
    for r in range(n):
        cols = indx[pntr[r]:pntr[r + 1]]
        assert np.all(np.diff(cols) > 0)
        assert np.array_equal(cols, np.nonzero(dense[r])[0])
    assert np.array_equal(dense, dense.T)
    assert np.all(np.diag(dense))
    assert np.all(iA >= 0)


@pytest.mark.parametrize("d,L,nnz", [(2, 5, 9120), (3, 6, 352768)])
def test_sparse_pattern_nnz(d, L, nnz):
    table, _ = oracle.element_table([L] * d)
    pntr, indx, _ = oracle.sparse_pattern(table, with_iA=False)
    assert pntr[-1] == nnz
    # symmetry of the full pattern
    rows = np.repeat(np.arange(len(pntr) - 1), np.diff(pntr))
    key = set(zip(rows.tolist(), indx.tolist()))
    assert all((c, r) in key for r, c in key)


def test_iA_layout():
    # iA = offset_1d * (d*T*n^2) + (t*d + k) * n^2 with offset_1d the
    # connect_1d position of (row 1D index, col 1D index) in dimension k
    d, n, T, L = 2, 3, 2, 3
    table, _ = oracle.element_table([L] * d)
    pntr, indx, iA = oracle.sparse_pattern(table, n=n, T=T)
    cp, ci, _ = oracle.connect_1d(L)
    elem = oracle.cell_1d_index(table)
    for r in range(table.shape[0]):
        for j in range(pntr[r], pntr[r + 1]):
            c = indx[j]
            for t in range(T):
                for k in range(d):
                    off = oracle.connect_1d_offset(cp, ci, elem[r, k], elem[c, k])
                    assert iA[(j * T + t) * d + k] == off * d * T * n * n + (t * d + k) * n * n
