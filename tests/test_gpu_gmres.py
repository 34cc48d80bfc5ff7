"""GPU parity of the device-resident GMRES (simple_gmres_euler) against the
oracle's restatement of solver::simple_gmres and a dense LU solve.

The iteration sequence is the reference's, so on the same inputs the device
solver takes the same number of matvecs as the oracle (up to one when the
residual estimate lands within rounding of the tolerance) and both solutions
agree with the dense solution of (I - dt*A) x = b.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
import oracle  # noqa: E402
from tests.kron_cases import sparse_grid_case  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dense_operator(c):
    d, n, T = c["d"], c["n"], c["T"]
    size = c["rows"] * n ** d
    A = np.zeros((size, size))
    e = np.zeros(size)
    col = np.zeros(size)
    for j in range(size):
        e[:] = 0
        e[j] = 1
        oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], c["vA"], 1.0, e, 0.0, col)
        A[:, j] = col
    return A


@pytest.mark.parametrize("d,n,L,T,dt,restart", [(2, 2, 3, 2, 0.05, 10), (3, 2, 3, 2, 0.02, 20),
                                                (2, 3, 3, 1, 0.1, 5)])
def test_gmres_matches_oracle_and_dense(d, n, L, T, dt, restart):
    c = sparse_grid_case(d, n, L, T, seed=1)
    size = c["rows"] * n ** d
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, size)
    tol = 1e-12
    xo, erro, itso = oracle.gmres_euler_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], c["vA"],
                                               dt, np.zeros(size), b, restart, 100, tol)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, c["vA"])
    x = dev(np.zeros(size))
    info = asgard_amd.simple_gmres_euler(dt, m, x, dev(b), restart, 100, tol)
    xg = host(x)
    assert info.error <= tol
    assert abs(info.iterations - itso) <= 1
    gold = oracle.dense_solve(np.eye(size) - dt * dense_operator(c), b)
    assert np.abs(xg - gold).max() < 1e-9 * max(1, np.abs(gold).max())
    assert np.abs(xg - xo).max() < 1e-9 * max(1, np.abs(xo).max())


def test_gmres_host_vectors_and_initial_guess():
    d, n, L, T, dt = 2, 2, 4, 2, 0.01
    c = sparse_grid_case(d, n, L, T, seed=3)
    size = c["rows"] * n ** d
    b = np.random.default_rng(4).uniform(-1, 1, size)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, c["vA"])
    x = np.zeros(size)
    info = asgard_amd.simple_gmres_euler(dt, m, x, b, 15, 50, 1e-12)
    assert info.error <= 1e-12
    # restarting from the converged solution needs only the residual check
    info2 = asgard_amd.simple_gmres_euler(dt, m, x, b, 15, 50, 1e-10)
    assert info2.iterations == 1 and info2.error <= 1e-10


def test_gmres_restart_larger_than_system():
    c = sparse_grid_case(1, 2, 1, 1, seed=0)
    size = c["rows"] * 2
    m = asgard_amd.KronmultMatrix.sparse_grid(1, 2, c["table"], 1, 1, c["vA"])
    with pytest.raises(ValueError):
        asgard_amd.simple_gmres_euler(0.1, m, dev(np.zeros(size)), dev(np.ones(size)),
                                      size + 1, 1, 1e-10)


def test_gmres_not_converged_is_reported():
    c = sparse_grid_case(3, 2, 4, 2, seed=5)
    size = c["rows"] * 8
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 2, c["table"], 4, 2, c["vA"])
    b = np.random.default_rng(6).uniform(-1, 1, size)
    info = asgard_amd.simple_gmres_euler(0.5, m, dev(np.zeros(size)), dev(b), 3, 2, 1e-14)
    assert info.error > 1e-14
    assert info.iterations == 2 * (3 + 1)


def test_gmres_fp32():
    d, n, L, T, dt = 3, 2, 3, 2, 0.01
    c = sparse_grid_case(d, n, L, T, seed=7)
    size = c["rows"] * n ** d
    b = np.random.default_rng(8).uniform(-1, 1, size)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T,
                                              c["vA"].astype(np.float32), dtype=np.float32)
    x = dev(np.zeros(size, np.float32))
    info = asgard_amd.simple_gmres_euler(dt, m, x, dev(b.astype(np.float32)), 20, 20, 1e-5)
    assert info.error <= 1e-5
    gold = oracle.dense_solve(np.eye(size) - dt * dense_operator(c), b)
    assert np.abs(host(x) - gold).max() < 1e-4


# ---- solver::simple_gmres on an explicit matrix (akm_gmres_dense) ---------

A_GOLD = np.array([[3.383861628748717e+00, 1.113343240310116e-02, 2.920740795411032e+00],
                   [3.210305545769361e+00, 3.412141162288144e+00, 3.934494120167269e+00],
                   [1.723479266939425e+00, 1.710451084172946e+00, 4.450671104482062e+00]])
PRECOND = np.diag([3.383861628748717e+00, 3.412141162288144e+00, 4.450671104482062e+00])
B_GOLD = np.array([2.084406360034887e-01, 6.444769305362776e-01, 3.687335330031937e-01])
X_GOLD = np.array([4.715561567725287e-02, 1.257695999382253e-01, 1.625351700791827e-02])
B_GOLD_2 = np.array([9.789303188021963e-01, 8.085725142873675e-01, 7.370498473207234e-01])
X_GOLD_2 = np.array([1.812300946484165e-01, -7.824949213916167e-02, 1.254969087137521e-01])


@pytest.mark.parametrize("dtype,rtol", [(np.float64, 1e-13), (np.float32, 1e-5)])
@pytest.mark.parametrize("b,xg", [(B_GOLD, X_GOLD), (B_GOLD_2, X_GOLD_2)])
@pytest.mark.parametrize("M", [None, PRECOND])
def test_dense_gmres_3x3_gold(dtype, rtol, b, xg, M):
    """The reference's gold (src/solver_tests.cpp:141-213): restart = max
    outer = 3, tolerance = machine epsilon, with and without the point
    Jacobi preconditioner."""
    eps = float(np.finfo(dtype).eps)
    x = np.zeros(3, dtype)
    info = asgard_amd.simple_gmres(A_GOLD.astype(dtype), x, b.astype(dtype),
                                   None if M is None else M.astype(dtype), 3, 3, eps)
    assert info.error < eps
    assert np.allclose(x, xg, rtol=rtol, atol=10 * eps)


def test_dense_gmres_matches_oracle():
    n = 120
    rng = np.random.default_rng(11)
    A = np.eye(n) * 4 + rng.uniform(-1, 1, (n, n)) / np.sqrt(n)
    M = np.diag(np.diag(A)) + np.tril(rng.uniform(-0.1, 0.1, (n, n)), -1)
    b = rng.uniform(-1, 1, n)
    for pre in (None, M):
        xo, erro, itso = oracle.gmres_dense(A, b, M=pre, restart=20, max_outer=50, tol=1e-12)
        x = np.zeros(n)
        info = asgard_amd.simple_gmres(A, x, b, pre, 20, 50, 1e-12)
        assert info.error <= 1e-12
        assert abs(info.iterations - itso) <= 1
        assert np.abs(x - xo).max() <= 1e-10
        assert np.abs(x - np.linalg.solve(A, b)).max() <= 1e-9


def test_dense_gmres_errors():
    with pytest.raises(ValueError):
        asgard_amd.simple_gmres(A_GOLD, np.zeros(3), B_GOLD, None, 4, 3, 1e-12)
    with pytest.raises(ValueError):
        asgard_amd.simple_gmres(A_GOLD, np.zeros(3), B_GOLD, np.zeros((3, 3)), 3, 3, 1e-12)
