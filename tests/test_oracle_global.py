"""The global Kronecker oracle (oracle/global_kron.py) against what fixes it.

* On standard sparse grids the one-dimension-at-a-time algorithm reproduces
  the pair-by-pair operator (oracle.kronmult_sparse, pinned by explicit
  Kronecker products) for any 1D operators; on refined grids, completed with
  their missing ancestors, it does so for hierarchical (volume) couplings --
  the property the reference's own global-kron tests check on lower sets
  (src/asgard_kronmult_matrix_tests.cpp:456-550).
* Identity factors skipped by the active mask change nothing.
* The padding is ancestry complete and minimal (brute force).
* The chain enumeration has the closed form of kronmult::permutes.
* The Jacobi diagonal is the diagonal of the explicitly assembled operator.
"""
import itertools

import numpy as np
import pytest

import oracle
from oracle import global_kron as gk
from tests.kron_cases import grid_case


def pairwise(c, alpha, beta):
    y = c["y"].copy()
    oracle.kronmult_sparse(c["d"], c["n"], c["pntr"], c["indx"], c["T"], c["iA"], c["vA"],
                           alpha, c["x"], beta, y)
    return y


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("d,n,L,T", [(2, 2, 4, 2), (3, 2, 3, 2), (2, 3, 3, 1), (3, 3, 3, 1),
                                     (4, 2, 3, 1), (1, 3, 4, 2)])
def test_standard_grid_equals_pairwise(d, n, L, T):
    c = grid_case(d, n, L, T, seed=d * 10 + n)
    g = gk.apply(d, n, c["table"], L, T, c["vA"], 1.5, c["x"], -0.5, c["y"])
    assert rel(g, pairwise(c, 1.5, -0.5)) < 1e-13


@pytest.mark.parametrize("d,n,L,T,frac", [(2, 2, 5, 1, 0.3), (3, 2, 4, 2, 0.3),
                                          (2, 3, 4, 2, 0.4), (4, 2, 4, 1, 0.2)])
def test_refined_grid_volume_coupling_equals_pairwise(d, n, L, T, frac):
    c = grid_case(d, n, L, T, seed=7 + d, refine=frac, coupling="volume")
    g = gk.apply(d, n, c["table"], L, T, c["vA"], 1.0, c["x"], 0.0, c["y"])
    assert rel(g, pairwise(c, 1.0, 0.0)) < 1e-13


def test_refined_grids_need_padding():
    pads = [len(gk.padding(oracle.cell_1d_index(grid_case(d, 2, 4, 1, seed=7 + d,
                                                          refine=0.3)["table"])))
            for d in (2, 3, 4)]
    assert max(pads) > 0


def test_identity_factors_skipped():
    ident = ((0, 1), (1, 0), (1, 2))
    c = grid_case(3, 2, 3, 2, seed=5, identity=ident)
    g = gk.apply(3, 2, c["table"], 3, 2, c["vA"], 1.0, c["x"], 0.0, c["y"], active=c["active"])
    assert rel(g, pairwise(c, 1.0, 0.0)) < 1e-13
    # a term of identities only contributes x itself
    c0 = grid_case(2, 2, 3, 1, seed=6, identity=((0, 0), (0, 1)))
    g0 = gk.apply(2, 2, c0["table"], 3, 1, c0["vA"], 2.0, c0["x"], 0.0, c0["y"],
                  active=c0["active"])
    assert np.allclose(g0, 2.0 * c0["x"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("d", [2, 3, 4])
def test_padding_complete_and_minimal(d):
    c = grid_case(d, 2, 4, 1, seed=7 + d, refine=0.3)
    elem = [tuple(int(v) for v in e) for e in oracle.cell_1d_index(c["table"])]
    pad = gk.padding(np.array(elem))
    allc = set(elem) | set(pad)
    assert not (set(elem) & set(pad))
    for cell in allc:
        for k in range(d):
            for a in gk._parents(cell[k]):
                assert cell[:k] + (a,) + cell[k + 1:] in allc
    # every padding cell is an ancestor (in one dimension at a time) of a cell
    reach = set(elem)
    frontier = list(elem)
    while frontier:
        nxt = []
        for cell in frontier:
            for k in range(d):
                for a in gk._parents(cell[k]):
                    an = cell[:k] + (a,) + cell[k + 1:]
                    if an not in reach:
                        reach.add(an)
                        nxt.append(an)
        frontier = nxt
    assert reach == allc


def test_parents_known_answers():
    assert gk._parents(0) == []
    assert gk._parents(1) == [0]
    assert gk._parents(2) == [1, 0]
    assert gk._parents(13) == [6, 3, 1, 0]


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_permutations_closed_form(k):
    chains = gk.permutations(k)
    assert len(chains) == 2 ** (k - 1)
    seen = set()
    for ch in chains:
        assert sorted(p for p, _ in ch) == list(range(k))
        fills = [f for _, f in ch]
        assert fills.count("F") == 1 and ch[fills.index("F")][0] == 0
        f0 = fills.index("F")
        assert all(f == "U" for f in fills[:f0]) and all(f == "L" for f in fills[f0 + 1:])
        # uppers in descending position, lowers ascending (sorted signed order)
        ups = [p for p, f in ch if f == "U"]
        lows = [p for p, f in ch if f == "L"]
        assert ups == sorted(ups, reverse=True) and lows == sorted(lows)
        seen.add(frozenset(ups))
    # every subset of positions 1..k-1 is the upper set of exactly one chain
    assert len(seen) == 2 ** (k - 1)


def test_jacobi_is_operator_diagonal():
    d, n, L, T = 2, 2, 3, 2
    c = grid_case(d, n, L, T, seed=11)
    size = c["rows"] * n ** d
    diag = np.zeros(size)
    for i in range(size):
        e = np.zeros(size)
        e[i] = 1.0
        y = np.zeros(size)
        oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], c["vA"], 1.0, e, 0.0, y)
        diag[i] = y[i]
    pc = gk.jacobi(d, n, c["table"], L, T, c["vA"])
    assert rel(pc, diag) < 1e-14
    assert np.allclose(gk.jacobi_scale(pc, 0.1, np.ones(size)), 1.0 / (1.0 - 0.1 * pc))


# ---- refined grids with edge couplings: the literal definition ------------

@pytest.mark.parametrize("d,n,L,T,frac,seed", [(2, 2, 5, 2, 0.3, 1), (3, 2, 4, 2, 0.3, 2),
                                               (2, 3, 4, 1, 0.4, 3), (3, 2, 4, 1, 0.2, 4)])
def test_refined_grid_edge_coupling_matches_definition(d, n, L, T, frac, seed):
    """On refined grids with full (edge) coupling the global operator is not
    the pair operator; it is fixed by its definition: per chain, the masked
    1D operators' upper / full / lower parts applied along their dimensions
    on the ancestry-closed cell set (tests/global_dense.py)."""
    from tests import global_dense
    from tests.kron_cases import grid_case_ops

    c, ops = grid_case_ops(d, n, L, T, seed=seed, refine=frac)
    got = gk.apply(d, n, c["table"], L, T, c["vA"], 1.0, c["x"], 0.0, c["y"])
    want = global_dense.apply(d, n, L, T, ops, c["table"], c["x"])
    assert rel(got, want) < 1e-13
    # and it differs from the pair operator there (edge couplings across levels)
    assert rel(got, pairwise(c, 1.0, 0.0)) > 1e-6


def test_identity_and_flux_match_definition():
    from tests import global_dense
    from tests.kron_cases import grid_case_ops

    ident = ((0, 1), (1, 0), (1, 2), (2, 0), (2, 1), (2, 2))
    c, ops = grid_case_ops(3, 2, 4, 3, seed=4, refine=0.3, identity=ident)
    got = gk.apply(3, 2, c["table"], 4, 3, c["vA"], 1.0, c["x"], 0.0, c["y"], active=c["active"])
    want = global_dense.apply(3, 2, 4, 3, ops, c["table"], c["x"], active=c["active"])
    assert rel(got, want) < 1e-13
