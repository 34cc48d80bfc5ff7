"""akm_gmres_euler_noalloc, the solver entry point that allocates nothing,
called through the C ABI and checked against akm_gmres_euler on the same
inputs: the same kernels in the same order, so the solutions, residual
estimates and iteration counts are equal bit for bit.  Two calls in a row,
the second on a smaller system, show that the shared per-GPU scratch array
keeps nothing from one call that the next depends on.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
import oracle  # noqa: E402
from tests.kron_cases import sparse_grid_case  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def c_noalloc(mat, dt, x, b, restart, max_iter, tol):
    """akm_gmres_euler_noalloc through the C ABI: (status, err, iterations)."""
    L = asgard_amd.load_library()
    err = ctypes.c_double()
    its = ctypes.c_int()
    rc = L.akm_gmres_euler_noalloc(mat._h, float(dt), ctypes.c_void_p(x.data_ptr()),
                                   ctypes.c_void_p(b.data_ptr()), int(restart), int(max_iter),
                                   float(tol), ctypes.byref(err), ctypes.byref(its))
    return rc, err.value, its.value


def c_reference_entry(mat, dt, x, b, restart, max_iter, tol):
    """akm_gmres_euler (device vectors) through the C ABI."""
    L = asgard_amd.load_library()
    err = ctypes.c_double()
    its = ctypes.c_int()
    rc = L.akm_gmres_euler(mat._h, float(dt), ctypes.c_void_p(x.data_ptr()),
                           ctypes.c_void_p(b.data_ptr()), int(restart), int(max_iter),
                           float(tol), 1, ctypes.byref(err), ctypes.byref(its))
    return rc, err.value, its.value


def solve_both(c, dt, restart, tol, dtype=np.float64):
    m = asgard_amd.KronmultMatrix.sparse_grid(c["d"], c["n"], c["table"], c["L"], c["T"],
                                              c["vA"].astype(dtype), dtype=dtype)
    b = dev(c["x"].astype(dtype))
    x_ref = dev(c["y"].astype(dtype))
    x_new = x_ref.clone()
    ref = c_reference_entry(m, dt, x_ref, b, restart, -1, tol)
    torch.cuda.synchronize()
    got = c_noalloc(m, dt, x_new, b, restart, -1, tol)
    torch.cuda.synchronize()
    return m, ref, got, x_ref, x_new


def test_two_calls_match_the_allocating_entry_point():
    # first call: d=3 n=2 L=5 (7,168 unknowns), then a smaller d=2 system
    big = sparse_grid_case(3, 2, 5, 2, seed=11)
    small = sparse_grid_case(2, 2, 4, 2, seed=12)
    for c, dt, restart in [(big, 0.02, 20), (small, 0.05, 10)]:
        _, ref, got, x_ref, x_new = solve_both(c, dt, restart, 1e-11)
        assert ref[0] == 0 and got[0] == 0, asgard_amd.last_error()
        assert got[2] == ref[2] and got[1] == ref[1], (got, ref)
        assert torch.equal(x_new, x_ref)
        # and both agree with the oracle's GMRES restatement
        d, n, T = c["d"], c["n"], c["T"]
        xo, _, itso = oracle.gmres_euler_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], c["vA"],
                                                dt, c["y"].copy(), c["x"], restart, 1000, 1e-11)
        assert abs(got[2] - itso) <= 1
        xh = x_new.cpu().numpy()
        assert np.linalg.norm(xh - xo) / np.linalg.norm(xo) < 1e-9


def test_fp32_and_default_restart():
    c = sparse_grid_case(2, 3, 4, 2, seed=13)
    _, ref, got, x_ref, x_new = solve_both(c, 0.02, 10, 1e-5, dtype=np.float32)
    assert ref[0] == 0 and got[0] == 0, asgard_amd.last_error()
    assert got[1:] == ref[1:] and torch.equal(x_new, x_ref)
    # restart <= 0: the reference default (it fits here), the same as the
    # allocating entry point's
    _, ref, got, x_ref, x_new = solve_both(c, 0.02, -1, 1e-5, dtype=np.float32)
    assert got[1:] == ref[1:] and torch.equal(x_new, x_ref)


def test_sizes_beyond_the_scratch_are_rejected():
    # d=3 n=3 |l|_1<=6: 18,576 unknowns; restart 250 needs about 37 MB
    c = sparse_grid_case(3, 3, 6, 1, seed=14)
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 3, c["table"], 6, 1, c["vA"])
    b = dev(c["x"])
    x = torch.zeros_like(b)
    rc, _, _ = c_noalloc(m, 0.01, x, b, 250, -1, 1e-10)
    assert rc == asgard_amd.AKM_ERR_INVALID
    assert "scratch" in asgard_amd.last_error()
    # the default restart is lowered until it fits, and the solve runs
    info = asgard_amd.gmres_euler_noalloc(0.01, m, x, b, -1, 200, 1e-10)
    assert info.error <= 1e-10
    need = 18576 * 8
    assert asgard_amd.NOALLOC_SCRATCH_BYTES // need >= 2
    with pytest.raises(asgard_amd.AkmInvalid):
        asgard_amd.gmres_euler_noalloc(0.01, m, x.cpu(), b.cpu())
