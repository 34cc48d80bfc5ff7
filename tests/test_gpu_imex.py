"""GPU IMEX step (akm_imex_step) against the IMEX oracle.

The oracle runs the same stage arithmetic (oracle/imex.py) with the kronmult
and GMRES oracles.  Bar: the two solutions agree to 1e-9 relative with both
GMRES runs at an absolute tolerance of 1e-11 (the solves may stop one
iteration apart); without the collision operator the step has no solve and
agrees to 1e-13.  The hook test updates the implicit operator inside the
step through akm_update_operators, as a time-dependent coefficient would.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
import oracle  # noqa: E402
from oracle import imex  # noqa: E402
from tests.kron_cases import sparse_grid_case  # noqa: E402

D, NN, L, T = 2, 2, 4, 2
DT = 0.01
TOL = 1e-11
RESTART, MAX_OUTER = 20, 50


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def dev_ops(ops):
    return [torch.from_numpy(np.ascontiguousarray(np.asarray(o).T)).cuda() for o in ops]


@pytest.fixture(scope="module")
def case():
    ce = sparse_grid_case(D, NN, L, T, seed=31)
    rng = np.random.default_rng(32)
    s = NN << L
    ops_b = [np.asfortranarray(rng.uniform(-1, 1, (s, s))) for _ in range(T * D)]
    vA_b = oracle.build_vA(D, NN, T, L, ops_b)
    return ce, ops_b, vA_b


def oracle_step(ce, vA_of, collision=True, hook=None):
    pntr, indx, iA = ce["pntr"], ce["indx"], ce["iA"]

    def explicit_apply(x):
        y = np.zeros_like(x)
        oracle.kronmult_sparse(D, NN, pntr, indx, T, iA, ce["vA"], 1.0, x, 0.0, y)
        return y

    def implicit_solve(c, guess, rhs):
        x = np.array(guess, dtype=np.float64)
        oracle.gmres_euler_sparse(D, NN, pntr, indx, T, iA, vA_of(), c, x, rhs, RESTART,
                                  MAX_OUTER, TOL)
        return x

    return imex.imex_step(explicit_apply, implicit_solve, DT, ce["x"], collision, hook)


def handles(ce, vA_b):
    E = asgard_amd.KronmultMatrix.sparse_grid(D, NN, ce["table"], L, T, ce["vA"])
    B = asgard_amd.KronmultMatrix.sparse_grid(D, NN, ce["table"], L, T, vA_b)
    return E, B


@pytest.mark.parametrize("collision", [True, False])
def test_imex_step_matches_oracle(case, collision):
    ce, _, vA_b = case
    E, B = handles(ce, vA_b)
    f = torch.from_numpy(ce["x"].copy()).cuda()
    infos, converged = asgard_amd.imex_step(E, B, DT, f, RESTART, MAX_OUTER, TOL,
                                            collision=collision)
    want = oracle_step(ce, lambda: vA_b, collision)
    got = f.cpu().numpy()
    assert converged
    if collision:
        assert all(i.iterations > 0 and i.error <= TOL for i in infos)
        assert rel(got, want) < 1e-9
    else:
        assert all(i.iterations == 0 for i in infos)
        assert rel(got, want) < 1e-13


@pytest.mark.parametrize("d,n,lev,terms", [(3, 2, 6, 3), (4, 3, 4, 2)])
def test_imex_step_larger_grids(d, n, lev, terms):
    # the collision solve on grids of 5,504 and 15,552 unknowns; the implicit
    # factors are scaled by 1/sqrt(size) so that dt*A stays well conditioned
    ce = sparse_grid_case(d, n, lev, terms, seed=41)
    rng = np.random.default_rng(42)
    s = n << lev
    ops_b = [np.asfortranarray(rng.uniform(-1, 1, (s, s)) / np.sqrt(s))
             for _ in range(terms * d)]
    vA_b = oracle.build_vA(d, n, terms, lev, ops_b)
    E = asgard_amd.KronmultMatrix.sparse_grid(d, n, ce["table"], lev, terms, ce["vA"])
    B = asgard_amd.KronmultMatrix.sparse_grid(d, n, ce["table"], lev, terms, vA_b)
    f = torch.from_numpy(ce["x"].copy()).cuda()
    infos, converged = asgard_amd.imex_step(E, B, DT, f, RESTART, MAX_OUTER, TOL,
                                            collision=True)
    pntr, indx, iA = ce["pntr"], ce["indx"], ce["iA"]

    def explicit_apply(x):
        y = np.zeros_like(x)
        oracle.kronmult_sparse(d, n, pntr, indx, terms, iA, ce["vA"], 1.0, x, 0.0, y)
        return y

    def implicit_solve(c, guess, rhs):
        x = np.array(guess, dtype=np.float64)
        oracle.gmres_euler_sparse(d, n, pntr, indx, terms, iA, vA_b, c, x, rhs, RESTART,
                                  MAX_OUTER, TOL)
        return x

    want = imex.imex_step(explicit_apply, implicit_solve, DT, ce["x"], True, None)
    assert converged
    assert all(i.iterations > 0 and i.error <= TOL for i in infos)
    assert rel(f.cpu().numpy(), want) < 1e-9


def test_imex_hook_updates_operator(case):
    ce, ops_b, vA_b = case
    E, B = handles(ce, vA_b)
    scaled = {1: [2.0 * o for o in ops_b], 3: [3.0 * o for o in ops_b]}
    dev_scaled = {e: dev_ops(v) for e, v in scaled.items()}
    events = []

    def gpu_hook(event, _ptr):
        events.append(event)
        if event in dev_scaled:
            B.update_operators(dev_scaled[event])

    f = torch.from_numpy(ce["x"].copy()).cuda()
    _, converged = asgard_amd.imex_step(E, B, DT, f, RESTART, MAX_OUTER, TOL, hook=gpu_hook)
    assert converged and events == [0, 1, 2, 3]

    current = {"vA": vA_b}

    def cpu_hook(event, _v):
        if event in scaled:
            current["vA"] = oracle.build_vA(D, NN, T, L, scaled[event])

    want = oracle_step(ce, lambda: current["vA"], True, cpu_hook)
    assert rel(f.cpu().numpy(), want) < 1e-9


def test_imex_hook_failure_stops_the_step(case):
    ce, _, vA_b = case
    E, B = handles(ce, vA_b)
    f = torch.from_numpy(ce["x"].copy()).cuda()

    def bad_hook(event, _ptr):
        if event == 2:
            raise RuntimeError("coefficient update failed")

    with pytest.raises(RuntimeError, match="coefficient update failed"):
        asgard_amd.imex_step(E, B, DT, f, RESTART, MAX_OUTER, TOL, hook=bad_hook)


def test_imex_step_rejects_mismatched_handles(case):
    ce, _, vA_b = case
    E, _ = handles(ce, vA_b)
    other = sparse_grid_case(D, NN, L - 1, T, seed=33)
    B = asgard_amd.KronmultMatrix.sparse_grid(D, NN, other["table"], L - 1, T, other["vA"])
    f = torch.from_numpy(ce["x"].copy()).cuda()
    with pytest.raises(asgard_amd.AkmInvalid):
        asgard_amd.imex_step(E, B, DT, f)
    B32 = asgard_amd.KronmultMatrix.sparse_grid(D, NN, ce["table"], L, T,
                                                vA_b.astype(np.float32), dtype=np.float32)
    with pytest.raises(asgard_amd.AkmInvalid):
        asgard_amd.imex_step(E, B32, DT, f)
