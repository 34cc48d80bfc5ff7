"""Pins the oracle's kronmult (cpu_sparse / cpu_dense restatements).

Checks, each chosen so a plausible slip fails it:
* the reference's Kronecker known answers (src/asgard_kronmult_matrix_tests.cpp:178-193);
* sparse kronmult against numpy.kron + matvec on the reference test sizes
  (d = 1..6, n = 1..5, src/asgard_kronmult_matrix_tests.cpp:195-303): a
  transposed factor or a factor applied to the wrong tensor index fails;
* the alpha/beta cases {0, 1, -1, other} (the reference specialises them);
* linearity in x and additivity over terms;
* the sparse-grid operator against a dense assembly of the global operator
  sum_t kron(op_{t,0}, ..., op_{t,d-1}) restricted to the grid cells.
"""
import numpy as np
import pytest

import oracle
from tests.kron_cases import bruteforce, random_pointer_case, sparse_grid_case


def test_kronecker_known_answers():
    A = [1, 2, 3, 4]
    B = [10, 20, 30, 40]
    gold = [10, 20, 20, 40, 30, 40, 60, 80, 30, 60, 40, 80, 90, 120, 120, 160]
    assert np.allclose(oracle.kronecker(A, B), gold)
    B = [1, 2, 3, 4, 5, 6, 7, 8, 9]
    gold = [1, 2, 3, 2, 4, 6, 4, 5, 6, 8, 10, 12, 7, 8, 9, 14, 16, 18,
            3, 6, 9, 4, 8, 12, 12, 15, 18, 16, 20, 24, 21, 24, 27, 28, 32, 36]
    assert np.allclose(oracle.kronecker(A, B), gold)


def test_bruteforce_matches_numpy_kron():
    rng = np.random.default_rng(1)
    for d, n in [(1, 3), (2, 2), (3, 3), (4, 2)]:
        fac = [rng.uniform(-1, 1, n * n) for _ in range(d)]
        x = rng.uniform(-1, 1, n ** d)
        y = np.zeros(n ** d)
        oracle.kronmult_one_bruteforce(fac, x, y)
        K = np.ones((1, 1))
        for f in fac:
            K = np.kron(K, f.reshape(n, n).T)
        assert np.allclose(y, K @ x, rtol=1e-13, atol=1e-13)


# reference test sizes: (d, n range, num_rows, num_terms, num_matrices)
REF_SIZES = ([(1, n, 11, 2, 7) for n in range(1, 7)] +
             [(2, n, 12, 3, 7) for n in range(1, 6)] +
             [(3, n, 12, 2, 7) for n in range(1, 6)] +
             [(4, n, 9, 2, 7) for n in range(1, 6)] +
             [(5, n, 8, 2, 7) for n in range(1, 5)] +
             [(6, n, 6, 2, 7) for n in range(1, 4)])


@pytest.mark.parametrize("d,n,rows,T,nmat", REF_SIZES)
def test_sparse_vs_bruteforce(d, n, rows, T, nmat):
    c = random_pointer_case(d, n, rows, T, nmat)
    y = c["y"].copy()
    oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], c["vA"], 1.0,
                           c["x"], 1.0, y)
    ref = bruteforce(c, 1.0, 1.0)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("alpha", [0.0, 1.0, -1.0, 0.37])
@pytest.mark.parametrize("beta", [0.0, 1.0, -1.0, -2.5])
def test_alpha_beta_cases(alpha, beta):
    c = random_pointer_case(3, 2, 5, 2, 4, seed=3)
    y = c["y"].copy()
    oracle.kronmult_sparse(3, 2, c["pntr"], c["indx"], 2, c["iA"], c["vA"], alpha,
                           c["x"], beta, y)
    ref = bruteforce(c, alpha, beta)
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13)


def test_beta_zero_ignores_nan():
    c = random_pointer_case(2, 3, 4, 1, 3)
    y = np.full_like(c["y"], np.nan)
    oracle.kronmult_sparse(2, 3, c["pntr"], c["indx"], 1, c["iA"], c["vA"], 1.0,
                           c["x"], 0.0, y)
    assert np.all(np.isfinite(y))


def test_linearity_and_terms():
    c = random_pointer_case(4, 2, 6, 3, 5, seed=11)
    rng = np.random.default_rng(5)
    x2 = rng.uniform(-1, 1, c["x"].shape)
    ap = lambda x: oracle.kronmult_sparse(4, 2, c["pntr"], c["indx"], 3, c["iA"],
                                          c["vA"], 1.0, x, 0.0, np.zeros_like(x))
    assert np.allclose(ap(2 * c["x"] - 3 * x2), 2 * ap(c["x"]) - 3 * ap(x2),
                       rtol=1e-12, atol=1e-12)
    # additivity over terms: one call with T=3 == sum of the single-term calls
    iA = c["iA"].reshape(-1, 3, 4)
    total = np.zeros_like(c["x"])
    for t in range(3):
        oracle.kronmult_sparse(4, 2, c["pntr"], c["indx"], 1,
                               np.ascontiguousarray(iA[:, t, :]).reshape(-1), c["vA"],
                               1.0, c["x"], 1.0, total)
    assert np.allclose(total, ap(c["x"]), rtol=1e-12, atol=1e-12)


def global_operator(case):
    """Dense assembly of sum_t kron(op_{t,0}, ...) restricted to the grid:
    rows/cols are (cell, tensor entry) with the entry index of dim k being
    n*index_1d + local index."""
    d, n, T = case["d"], case["n"], case["T"]
    elem = case["elem"]
    N = n ** d
    cells = elem.shape[0]
    local = np.array(np.unravel_index(np.arange(N), [n] * d)).T  # N x d
    glob = (elem[:, None, :] * n + local[None, :, :]).reshape(cells * N, d)
    A = np.zeros((cells * N, cells * N))
    for t in range(T):
        blk = np.ones((cells * N, cells * N))
        for k in range(d):
            op = case["ops"][t * d + k]
            blk *= op[np.ix_(glob[:, k], glob[:, k])]
        A += blk
    return A


@pytest.mark.parametrize("d,n,L,T", [(2, 2, 3, 2), (2, 3, 3, 1), (3, 2, 2, 2)])
def test_sparse_grid_vs_global_operator(d, n, L, T):
    c = sparse_grid_case(d, n, L, T, seed=7)
    A = global_operator(c)
    N = n ** d
    cells = c["rows"]
    # keep only connected cell pairs (sparse-mode semantics)
    mask = np.zeros((cells, cells), dtype=bool)
    for r in range(cells):
        mask[r, c["indx"][c["pntr"][r]:c["pntr"][r + 1]]] = True
    A *= np.kron(mask, np.ones((N, N)))
    y = np.zeros(cells * N)
    oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], c["vA"], 1.0,
                           c["x"], 0.0, y)
    assert np.allclose(y, A @ c["x"], rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("d,n", [(1, 2), (2, 1), (2, 3), (3, 2), (4, 2)])
def test_dense_mode_vs_global_operator(d, n):
    L, T = 2, 2
    c = sparse_grid_case(d, n, L, T, seed=9)
    nb = 1 << L
    terms = oracle.build_dense_terms(d, n, T, nb, c["ops"])
    cells = c["rows"]
    y = c["y"].copy()
    oracle.kronmult_dense(d, n, cells, cells, T, c["elem"], 0, 0, terms, nb, -0.5,
                          c["x"], 2.0, y)
    ref = 2.0 * c["y"] - 0.5 * global_operator(c) @ c["x"]
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)


def test_f32_variant():
    c = random_pointer_case(3, 3, 6, 2, 5)
    y64 = c["y"].copy()
    oracle.kronmult_sparse(3, 3, c["pntr"], c["indx"], 2, c["iA"], c["vA"], 1.0,
                           c["x"], 1.0, y64)
    y32 = c["y"].astype(np.float32)
    oracle.kronmult_sparse(3, 3, c["pntr"], c["indx"], 2, c["iA"],
                           c["vA"].astype(np.float32), 1.0, c["x"].astype(np.float32),
                           1.0, y32)
    assert np.allclose(y32, y64, rtol=1e-4, atol=1e-4 * np.abs(y64).max())
