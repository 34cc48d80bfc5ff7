"""GPU tests of the batched implicit-step solver (akm_batch_*): systems of
mixed (d, n) and sizes on adaptively refined sparse grids solved together
must match one-at-a-time device solves and the FP64 residual tolerance, and
repeated solves (graph replays) must repeat bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
from asgard_amd import dg, grid  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def make_system(d, n, L, seed, frac=0.2):
    table, _ = grid.element_table([L] * d)
    table = grid.refine(table, frac, seed=seed)
    Lmax = int(table[:, :d].max())
    ops = dg.operators_advection_diffusion(d, n, Lmax, n_adv=max(1, d // 2), nu=0.1)
    vA = grid.build_vA(d, n, d, Lmax, ops)
    return asgard_amd.KronmultMatrix.sparse_grid(d, n, table, Lmax, d, vA)


SYSTEMS = [(2, 2, 4, 1e-3), (2, 3, 3, 2e-3), (3, 2, 3, 1e-3), (3, 3, 3, 5e-4), (2, 2, 5, 1e-3),
           (4, 2, 3, 1e-3), (3, 2, 4, 2e-3), (4, 3, 2, 1e-3)]


def rel_residual(m, dt, x, b):
    r = m.apply_euler(dt, 1.0, x, 0.0, torch.zeros_like(x))
    return float(torch.linalg.norm(b - r).item() / torch.linalg.norm(b).item())


def test_batch_matches_individual_solves():
    mats = [make_system(d, n, L, seed=i) for i, (d, n, L, _) in enumerate(SYSTEMS)]
    dts = [s[3] for s in SYSTEMS]
    rng = np.random.default_rng(3)
    bs = [dev(rng.uniform(-1, 1, m.output_size())) for m in mats]
    xs = [b.clone() for b in bs]
    batch = asgard_amd.BatchSolver(mats, dts, restart=30)
    info = batch.info()
    assert info["restart"] == 30 and info["groups"] == len({(d, n) for d, n, _, _ in SYSTEMS})
    res, its, conv = batch.solve(xs, bs, tol=1e-10)
    assert conv.all(), (res, its)
    assert (res <= 1e-10).all()
    for m, dt, x, b, r in zip(mats, dts, xs, bs, res):
        # the reported residual is the true FP64 residual of the solution
        check = rel_residual(m, dt, x, b)
        assert check <= 1.5e-10 and abs(check - r) <= 1e-12 + 1e-3 * r, (check, r)
        xp = b.clone()
        asgard_amd.simple_gmres_euler(dt, m, xp, b, 30, -1,
                                      1e-12 * float(torch.linalg.norm(b).item()))
        diff = float(torch.linalg.norm(x - xp).item() / torch.linalg.norm(xp).item())
        assert diff < 1e-8, diff


def test_batch_repeats_bitwise_and_handles_zero_rhs():
    mats = [make_system(d, n, L, seed=10 + i) for i, (d, n, L, _) in enumerate(SYSTEMS[:4])]
    dts = [s[3] for s in SYSTEMS[:4]]
    rng = np.random.default_rng(4)
    bs = [dev(rng.uniform(-1, 1, m.output_size())) for m in mats]
    bs[2].zero_()  # this system's solution is zero
    batch = asgard_amd.BatchSolver(mats, dts, restart=20)
    x1 = [torch.zeros_like(b) for b in bs]
    r1, i1, c1 = batch.solve(x1, bs)
    x2 = [torch.zeros_like(b) for b in bs]
    r2, i2, c2 = batch.solve(x2, bs)
    assert c1.all() and c2.all()
    for a, b in zip(x1, x2):
        assert np.array_equal(host(a), host(b))
    assert np.array_equal(i1, i2)
    assert not host(x1[2]).any() and i1[2] == 1


def test_batch_rejects_unsupported_systems():
    t, _ = grid.element_table([2, 2])
    vA = grid.build_vA(2, 2, 1, 2, [np.eye(8)] * 2)
    m32 = asgard_amd.KronmultMatrix.sparse_grid(2, 2, t, 2, 1, vA.astype(np.float32),
                                                dtype=np.float32)
    with pytest.raises(ValueError):
        asgard_amd.BatchSolver([m32], [0.1])
    t6, _ = grid.element_table([1] * 6)
    v6 = grid.build_vA(6, 3, 1, 1, [np.eye(6)] * 6)
    m6 = asgard_amd.KronmultMatrix.sparse_grid(6, 3, t6, 1, 1, v6)
    with pytest.raises(asgard_amd.AkmError):
        asgard_amd.BatchSolver([m6], [0.1])
