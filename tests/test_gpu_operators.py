"""GPU operator update (update_kronmult_coefficients on the device).

akm_update_operators gathers the stored factor blocks from dense 1D operator
matrices already on the GPU.  Bar: the blocks it writes are the ones the
oracle's build_vA / build_dense_terms produce from the same operators, so the
apply after a device update equals, bit for bit, the apply of a handle built
from the oracle's blocks with the same kernel shape, and matches the oracle's
apply to 1e-13 (FP64) / 1e-5 (FP32).  Global handles also rebuild their
Jacobi diagonal on the device; it must equal the host-built one exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
import oracle  # noqa: E402
from oracle import global_kron  # noqa: E402
from tests.kron_cases import sparse_grid_case  # noqa: E402

TOL64 = 1e-13
TOL32 = 1e-5


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def dev_ops(ops, dtype=np.float64):
    """Device copies of the operators in column-major order."""
    return [torch.from_numpy(np.ascontiguousarray(np.asarray(o, dtype=dtype).T)).cuda()
            for o in ops]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def new_ops(c, seed):
    rng = np.random.default_rng(seed)
    s = c["n"] << c["L"]
    return [np.asfortranarray(rng.uniform(-1, 1, (s, s))) for _ in range(c["T"] * c["d"])]


def applied(m, x, waves, lds):
    m.set_tuning(waves, lds, False)
    y = torch.zeros_like(x)
    m.apply(1.0, x, 0.0, y)
    torch.cuda.synchronize()
    return y.cpu().numpy()


@pytest.mark.parametrize("d,n,L,T", [(2, 3, 4, 2), (3, 3, 4, 3), (4, 2, 5, 2), (6, 2, 3, 2)])
def test_sparse_grid_update(d, n, L, T):
    c = sparse_grid_case(d, n, L, T, seed=21)
    ops2 = new_ops(c, seed=22)
    vA2 = oracle.build_vA(d, n, T, L, ops2)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, c["vA"])
    m.update_operators(dev_ops(ops2))
    ref_m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, vA2)
    x = dev(c["x"])
    ref = c["y"].copy()
    oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], vA2, 1.0, c["x"], 0.0, ref)
    for lds in (0, -1):  # register kernel, then the LDS-table kernel where it fits
        y = applied(m, x, 4, lds)
        assert np.array_equal(y, applied(ref_m, x, 4, lds))
        assert rel(y, ref) < TOL64


def test_sparse_grid_update_f32():
    d, n, L, T = 3, 2, 5, 2
    c = sparse_grid_case(d, n, L, T, seed=23)
    ops2 = new_ops(c, seed=24)
    vA2 = oracle.build_vA(d, n, T, L, ops2)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T,
                                              c["vA"].astype(np.float32), dtype=np.float32)
    m.update_operators(dev_ops(ops2, np.float32))
    ref_m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T,
                                                  vA2.astype(np.float32), dtype=np.float32)
    x = dev(c["x"].astype(np.float32))
    ref = np.zeros_like(c["x"])
    oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], vA2, 1.0, c["x"], 0.0, ref)
    for lds in (0, -1):
        y = applied(m, x, 2, lds)
        assert np.array_equal(y, applied(ref_m, x, 2, lds))
        assert rel(y.astype(np.float64), ref) < TOL32


@pytest.mark.parametrize("d,n", [(2, 3), (3, 2)])
def test_dense_update(d, n):
    L, T = 3, 2
    c = sparse_grid_case(d, n, L, T, seed=25)
    nb = 1 << L
    ops2 = new_ops(c, seed=26)
    cells = c["rows"]
    m = asgard_amd.KronmultMatrix.dense(d, n, cells, cells, T,
                                        oracle.build_dense_terms(d, n, T, nb, c["ops"]),
                                        c["elem"], 0, 0, nb)
    m.update_operators(dev_ops(ops2))
    terms2 = oracle.build_dense_terms(d, n, T, nb, ops2)
    ref = np.zeros_like(c["x"])
    oracle.kronmult_dense(d, n, cells, cells, T, c["elem"], 0, 0, terms2, nb, 1.0, c["x"], 0.0,
                          ref)
    y = torch.zeros(cells * n ** d, dtype=torch.float64, device="cuda")
    m.apply(1.0, dev(c["x"]), 0.0, y)
    assert rel(y.cpu().numpy(), ref) < TOL64


def test_global_update_and_jacobi():
    d, n, L, T = 3, 2, 4, 2
    c = sparse_grid_case(d, n, L, T, seed=27)
    ops2 = new_ops(c, seed=28)
    vA2 = oracle.build_vA(d, n, T, L, ops2)
    g = asgard_amd.KronmultMatrix.global_grid(d, n, c["table"], L, T, c["vA"])
    g.update_operators(dev_ops(ops2))
    ref_g = asgard_amd.KronmultMatrix.global_grid(d, n, c["table"], L, T, vA2)
    x = dev(c["x"])
    for layout in (1, 8, 2):
        g.set_tuning(waves_per_row=layout)
        ref_g.set_tuning(waves_per_row=layout)
        y, yr = torch.zeros_like(x), torch.zeros_like(x)
        g.apply(1.0, x, 0.0, y)
        ref_g.apply(1.0, x, 0.0, yr)
        assert torch.equal(y, yr)
    want = global_kron.apply(d, n, c["table"], L, T, vA2, 1.0, c["x"], 0.0,
                             np.zeros_like(c["x"]))
    assert rel(y.cpu().numpy(), want) < TOL64
    assert np.array_equal(g.jacobi(), ref_g.jacobi())
    assert rel(g.jacobi(), global_kron.jacobi(d, n, c["table"], L, T, vA2)) < 1e-15


def test_update_errors():
    c = sparse_grid_case(2, 2, 3, 1, seed=29)
    m = asgard_amd.KronmultMatrix.sparse_grid(2, 2, c["table"], 3, 1, c["vA"])
    ops = dev_ops(c["ops"])
    with pytest.raises(asgard_amd.AkmInvalid):
        m.update_operators(ops[:1])  # T*d = 2 operators needed
    small = [o[:8, :8].contiguous() for o in ops]  # lda 8 < n * 2^L = 16
    with pytest.raises(asgard_amd.AkmInvalid):
        m.update_operators(small)
    ia = asgard_amd.KronmultMatrix.sparse(2, 2, c["rows"], c["cols"], 1, c["pntr"], c["indx"],
                                          c["iA"], c["vA"])
    with pytest.raises(asgard_amd.AkmError):
        ia.update_operators(ops)


def test_global_slab_tables_follow_updates():
    # the slab layout applies compact factor tables derived from the stored
    # blocks; an operator update, a coefficient update and a stream change
    # each make the next apply rebuild them
    from tests.kron_cases import grid_case_ops

    d, n, L, T = 4, 2, 5, 4
    act = [{0}, {1, 2}, {0, 3}, {1, 2}]
    ident = tuple((t, k) for t, a in enumerate(act) for k in range(d) if k not in a)
    c, _ = grid_case_ops(d, n, L, T, seed=41, identity=ident)
    g = asgard_amd.KronmultMatrix.global_grid(d, n, c["table"], L, T, c["vA"],
                                              active=c["active"])
    g.set_tuning(waves_per_row=6)
    assert g.slab_info()[0] == 3
    x = dev(c["x"])
    y0 = torch.zeros_like(x)
    g.apply(1.0, x, 0.0, y0)
    ops2 = new_ops(c, seed=42)
    vA2 = oracle.build_vA(d, n, T, L, ops2)
    g.update_operators(dev_ops(ops2))
    ref = asgard_amd.KronmultMatrix.global_grid(d, n, c["table"], L, T, vA2,
                                                active=c["active"])
    ref.set_tuning(waves_per_row=6)
    y, yr = torch.zeros_like(x), torch.zeros_like(x)
    g.apply(1.0, x, 0.0, y)
    ref.apply(1.0, x, 0.0, yr)
    assert torch.equal(y, yr)
    want = global_kron.apply(d, n, c["table"], L, T, vA2, 1.0, c["x"], 0.0,
                             np.zeros_like(c["x"]), active=c["active"])
    assert rel(y.cpu().numpy(), want) < TOL64
    # back to the first blocks through a coefficient update
    g.update_stored_coefficients(c["vA"])
    y1 = torch.zeros_like(x)
    g.apply(1.0, x, 0.0, y1)
    assert torch.equal(y1, y0)
    # on another stream
    s = torch.cuda.Stream()
    g.set_stream(s.cuda_stream)
    g.update_stored_coefficients(vA2)
    y2 = torch.zeros_like(x)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        g.apply(1.0, x, 0.0, y2)
    s.synchronize()
    g.set_stream(0)
    assert torch.equal(y2, yr)
