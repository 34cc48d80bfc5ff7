"""Pins the oracle's GMRES (restatement of solver::simple_gmres).

* the 3x3 gold of src/solver_tests.cpp:141-213, with and without the point
  Jacobi (dense LU) preconditioner, tolerance = machine epsilon;
* simple_gmres_euler on a sparse-grid operator against a dense LU solve of
  (I - dt*A) x = b (the reference's own gold, src/solver_tests.cpp:55-69);
* GMRES(m) invariants: with restart == n the method is exact; the returned
  error estimate matches the true residual norm of the returned x; and
  restart > n is rejected (src/solver.cpp:175-181).
"""
import numpy as np
import pytest

import oracle
from tests.kron_cases import sparse_grid_case

A_GOLD = np.array([[3.383861628748717e+00, 1.113343240310116e-02, 2.920740795411032e+00],
                   [3.210305545769361e+00, 3.412141162288144e+00, 3.934494120167269e+00],
                   [1.723479266939425e+00, 1.710451084172946e+00, 4.450671104482062e+00]])
PRECOND = np.diag([3.383861628748717e+00, 3.412141162288144e+00, 4.450671104482062e+00])
B_GOLD = np.array([2.084406360034887e-01, 6.444769305362776e-01, 3.687335330031937e-01])
X_GOLD = np.array([4.715561567725287e-02, 1.257695999382253e-01, 1.625351700791827e-02])
B_GOLD_2 = np.array([9.789303188021963e-01, 8.085725142873675e-01, 7.370498473207234e-01])
X_GOLD_2 = np.array([1.812300946484165e-01, -7.824949213916167e-02, 1.254969087137521e-01])
EPS = np.finfo(np.float64).eps


@pytest.mark.parametrize("b,xg", [(B_GOLD, X_GOLD), (B_GOLD_2, X_GOLD_2)])
@pytest.mark.parametrize("M", [None, PRECOND])
def test_gmres_3x3_gold(b, xg, M):
    x, err, its = oracle.gmres_dense(A_GOLD, b, M=M, restart=3, max_outer=3, tol=EPS)
    assert err < EPS
    assert np.allclose(x, xg, rtol=1e-13, atol=1e-15)


def test_gmres_restart_too_large():
    with pytest.raises(ValueError):
        oracle.gmres_dense(A_GOLD, B_GOLD, restart=4, max_outer=3)


def euler_case(d, n, L, T, dt, seed=0):
    c = sparse_grid_case(d, n, L, T, seed=seed)
    N = n ** d
    size = c["rows"] * N
    # dense assembly of the sparse kronmult operator column by column
    A = np.zeros((size, size))
    e = np.zeros(size)
    col = np.zeros(size)
    for j in range(size):
        e[:] = 0
        e[j] = 1
        oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], c["vA"], 1.0, e,
                               0.0, col)
        A[:, j] = col
    return c, A, size


@pytest.mark.parametrize("d,n,L,T,dt", [(2, 2, 3, 2, 0.05), (3, 2, 2, 2, 0.01)])
def test_gmres_euler_vs_dense_solve(d, n, L, T, dt):
    c, A, size = euler_case(d, n, L, T, dt)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, size)
    gold = oracle.dense_solve(np.eye(size) - dt * A, b)
    x = np.zeros(size)
    x, err, its = oracle.gmres_euler_sparse(d, n, c["pntr"], c["indx"], T, c["iA"],
                                            c["vA"], dt, x, b, 20, 50, 1e-12)
    assert err <= 1e-12
    assert np.allclose(x, gold, rtol=1e-9, atol=1e-10)
    # the estimate is the residual of the final x (up to rounding)
    true_res = np.linalg.norm(b - (x - dt * (A @ x)))
    assert true_res < 1e-10


def test_gmres_exact_with_full_restart():
    rng = np.random.default_rng(8)
    n = 12
    A = rng.uniform(-1, 1, (n, n)) + 4 * np.eye(n)
    b = rng.uniform(-1, 1, n)
    x, err, its = oracle.gmres_dense(A, b, restart=n, max_outer=1, tol=1e-14)
    assert np.allclose(A @ x, b, rtol=1e-11, atol=1e-11)
    assert its <= n + 1
