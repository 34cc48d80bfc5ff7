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
