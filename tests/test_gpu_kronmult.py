"""GPU parity of the kronmult apply against the CPU oracle.

Bar: FP64 results within 1e-13 relative (max |y_gpu - y_cpu| / max |y_cpu|)
of the oracle; index/connectivity construction bit-exact; FP32 within 1e-5
relative of the FP64 oracle.  Sizes: the reference's own test sizes
(src/asgard_kronmult_matrix_tests.cpp:195-303), the sparse grids of the
benchmarks (d=2..6), edge cases (empty rows, beta = 0 with NaN in y, alpha = 0,
rows cut into many chunks), and at full benchmark size through sampled rows.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
import oracle  # noqa: E402
from tests.kron_cases import random_pointer_case, sparse_grid_case  # noqa: E402

TOL64 = 1e-13
TOL32 = 1e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel(a, b):
    scale = max(np.abs(b).max(), 1e-300)
    return np.abs(a - b).max() / scale


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert asgard_amd.device_count() >= 1


def ia_matrix(c, dtype=np.float64):
    return asgard_amd.KronmultMatrix.sparse(c["d"], c["n"], c["rows"], c["cols"], c["T"],
                                            c["pntr"], c["indx"], c["iA"],
                                            c["vA"].astype(dtype), dtype=dtype)


def oracle_apply(c, alpha, beta, x=None, y=None):
    y = (c["y"] if y is None else y).copy()
    oracle.kronmult_sparse(c["d"], c["n"], c["pntr"], c["indx"], c["T"], c["iA"], c["vA"],
                           alpha, c["x"] if x is None else x, beta, y)
    return y


REF_SIZES = ([(1, n, 11, 2, 7) for n in range(1, 7)] +
             [(2, n, 12, 3, 7) for n in range(1, 6)] +
             [(3, n, 12, 2, 7) for n in range(1, 6)] +
             [(4, n, 9, 2, 7) for n in range(1, 6)] +
             [(5, n, 8, 2, 7) for n in range(1, 5)] +
             [(6, n, 6, 2, 7) for n in range(1, 5)])


@pytest.mark.parametrize("d,n,rows,T,nmat", REF_SIZES)
def test_reference_sizes(d, n, rows, T, nmat):
    c = random_pointer_case(d, n, rows, T, nmat)
    m = ia_matrix(c)
    y = dev(c["y"])
    m.apply(1.0, dev(c["x"]), 1.0, y)
    assert rel(host(y), oracle_apply(c, 1.0, 1.0)) < TOL64


@pytest.mark.parametrize("d,n", [(2, 3), (3, 2), (4, 3), (6, 2), (5, 3)])
def test_generic_kernel(d, n):
    c = random_pointer_case(d, n, 5, 2, 6, seed=4)
    m = ia_matrix(c)
    m.set_tuning(force_generic=True)
    y = dev(c["y"])
    m.apply(0.5, dev(c["x"]), -1.0, y)
    assert rel(host(y), oracle_apply(c, 0.5, -1.0)) < TOL64


@pytest.mark.parametrize("alpha", [0.0, 1.0, -1.0, 0.37])
@pytest.mark.parametrize("beta", [0.0, 1.0, -1.0, -2.5])
def test_alpha_beta(alpha, beta):
    c = random_pointer_case(3, 3, 7, 2, 5, seed=3)
    m = ia_matrix(c)
    y = dev(c["y"])
    m.apply(alpha, dev(c["x"]), beta, y)
    assert rel(host(y), oracle_apply(c, alpha, beta)) < TOL64


def test_beta_zero_never_reads_y():
    c = random_pointer_case(2, 2, 6, 2, 4)
    m = ia_matrix(c)
    y = dev(np.full_like(c["y"], np.nan))
    m.apply(1.0, dev(c["x"]), 0.0, y)
    assert np.all(np.isfinite(host(y)))


def test_empty_rows_and_no_terms():
    # rows without columns still get y = beta*y
    c = random_pointer_case(2, 2, 4, 1, 3)
    pntr = np.array([0, 4, 4, 8, 8], np.int64)
    indx = np.concatenate([c["indx"][:4], c["indx"][4:8]])
    iA = c["iA"][:8 * 2]
    m = asgard_amd.KronmultMatrix.sparse(2, 2, 4, 4, 1, pntr, indx, iA, c["vA"])
    y = dev(c["y"])
    m.apply(1.0, dev(c["x"]), 2.0, y)
    ref = c["y"].copy()
    oracle.kronmult_sparse(2, 2, pntr, indx, 1, iA, c["vA"], 1.0, c["x"], 2.0, ref)
    assert rel(host(y), ref) < TOL64
    m0 = asgard_amd.KronmultMatrix.sparse(2, 2, 4, 4, 0, pntr, indx, iA[:0], c["vA"])
    y = dev(c["y"])
    m0.apply(1.0, dev(c["x"]), -3.0, y)
    assert np.allclose(host(y), -3.0 * c["y"], rtol=0, atol=0)


@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_workgroup_shapes_and_determinism(waves):
    c = sparse_grid_case(3, 2, 5, 2, seed=5)
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 2, c["table"], 5, 2, c["vA"])
    m.set_tuning(waves_per_row=waves)
    x = dev(c["x"])
    y1 = dev(c["y"])
    m.apply(1.0, x, 1.0, y1)
    y2 = dev(c["y"])
    m.apply(1.0, x, 1.0, y2)
    a, b = host(y1), host(y2)
    assert np.array_equal(a, b), "apply must be bitwise repeatable"
    assert rel(a, oracle_apply(c, 1.0, 1.0)) < TOL64


@pytest.mark.parametrize("d,n,L,T", [(3, 3, 5, 3), (4, 2, 5, 4), (6, 2, 3, 2)])
def test_autotune_keeps_parity(d, n, L, T):
    c = sparse_grid_case(d, n, L, T, seed=11)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, c["vA"])
    x = dev(c["x"])
    scratch = dev(np.zeros_like(c["x"]))
    waves, lds = m.autotune(x, scratch)
    # 16 wavefronts per row exist for the matrix-core kernels (kernel 2) only
    assert waves in (1, 2, 4, 8) or (waves == 16 and lds == 2)
    # a second operator of the same shape reuses the cached choice
    m2 = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, c["vA"])
    assert m2.autotune(x, scratch) == (waves, lds)
    y = dev(c["y"])
    m.apply(1.0, x, 1.0, y)
    assert rel(host(y), oracle_apply(c, 1.0, 1.0)) < TOL64


@pytest.mark.parametrize("lds", [-1, 0])
@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_lds_tables_and_shapes(waves, lds):
    c = sparse_grid_case(4, 2, 5, 3, seed=12)
    m = asgard_amd.KronmultMatrix.sparse_grid(4, 2, c["table"], 5, 3, c["vA"])
    m.set_tuning(waves_per_row=waves, lds_cap_bytes=lds)
    y = dev(c["y"])
    m.apply(0.25, dev(c["x"]), -1.0, y)
    assert rel(host(y), oracle_apply(c, 0.25, -1.0)) < TOL64


def test_invalid_workgroup_shape():
    c = sparse_grid_case(2, 2, 3, 1, seed=5)
    m = asgard_amd.KronmultMatrix.sparse_grid(2, 2, c["table"], 3, 1, c["vA"])
    with pytest.raises(ValueError):
        m.set_tuning(waves_per_row=3)


def test_sparse_pattern_bit_exact_small():
    for levels in [(5, 5), (6, 6, 6), (3, 2, 3), (4, 4, 4, 4), (2, 2, 2, 2, 2, 2)]:
        table, _ = oracle.element_table(levels)
        p, i = asgard_amd.sparse_pattern(table)
        po, io, _ = oracle.sparse_pattern(table, with_iA=False)
        assert np.array_equal(p, po) and np.array_equal(i, io), levels


def test_sparse_pattern_bit_exact_d4_l7():
    table, _ = oracle.element_table([7] * 4)
    p, i = asgard_amd.sparse_pattern(table)
    po, io, _ = oracle.sparse_pattern(table, with_iA=False)
    assert p[-1] == 12258592
    assert np.array_equal(p, po) and np.array_equal(i, io)


# ------------------------------------------------ benchmark sparse grids ---

@pytest.mark.parametrize("d,n,L,T,dist", [(2, 2, 5, 2, "uniform"), (3, 3, 6, 3, "normal"),
                                          (2, 4, 4, 2, "uniform"), (5, 2, 4, 2, "uniform")])
def test_sparse_grid_parity(d, n, L, T, dist):
    c = sparse_grid_case(d, n, L, T, seed=0, dist=dist)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, c["vA"])
    assert m.num_pairs() == c["pntr"][-1]
    assert m.flops() == 2 * d * n ** (d + 1) * T * c["pntr"][-1]
    y = dev(c["y"])
    m.apply(1.0, dev(c["x"]), 0.0, y)
    assert rel(host(y), oracle_apply(c, 1.0, 0.0)) < TOL64


def test_d4_l7_full_parity():
    c = sparse_grid_case(4, 2, 7, 4, seed=0)
    m = asgard_amd.KronmultMatrix.sparse_grid(4, 2, c["table"], 7, 4, c["vA"])
    y = dev(c["y"])
    m.apply(-0.7, dev(c["x"]), 1.0, y)
    assert rel(host(y), oracle_apply(c, -0.7, 1.0)) < TOL64


def sampled_grid_parity(d, n, L, T, rows, dtype=np.float64, seed=0):
    rng = np.random.default_rng(seed)
    table, _ = oracle.element_table([L] * d)
    size1d = n * (1 << L)
    ops = [np.asfortranarray(rng.uniform(-1, 1, (size1d, size1d))) for _ in range(T * d)]
    vA = oracle.build_vA(d, n, T, L, ops)
    cells = table.shape[0]
    N = n ** d
    x = rng.uniform(-1, 1, cells * N)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, table, L, T, vA.astype(dtype), dtype=dtype)
    y = dev(np.zeros(cells * N, dtype))
    m.apply(1.0, dev(x.astype(dtype)), 0.0, y)
    yg = host(y)
    sel = np.unique(np.concatenate([[0, cells - 1], rows % cells]))
    p, i, iA = oracle.sparse_pattern_rows(table, sel, n=n, T=T, max_level=L)
    ref = np.zeros(len(sel) * N)
    oracle.kronmult_sparse(d, n, p, i, T, iA, vA, 1.0, x, 0.0, ref)
    got = np.concatenate([yg[r * N:(r + 1) * N] for r in sel])
    return m, rel(got.astype(np.float64), ref)


def test_d6_l7_sampled_rows():
    rows = np.random.default_rng(2).integers(0, 16172, 12)
    m, err = sampled_grid_parity(6, 2, 7, 6, rows)
    assert m.num_pairs() == 226043944
    assert err < TOL64


def test_d6_l7_sampled_rows_f32():
    rows = np.random.default_rng(3).integers(0, 16172, 6)
    _, err = sampled_grid_parity(6, 2, 7, 6, rows, dtype=np.float32)
    assert err < TOL32


@pytest.mark.parametrize("d,n,L,T", [(3, 3, 6, 3), (4, 2, 6, 4)])
def test_f32_vs_f64(d, n, L, T):
    c = sparse_grid_case(d, n, L, T, seed=1)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T,
                                              c["vA"].astype(np.float32), dtype=np.float32)
    y = dev(np.zeros(c["x"].shape, np.float32))
    m.apply(1.0, dev(c["x"].astype(np.float32)), 0.0, y)
    assert rel(host(y).astype(np.float64), oracle_apply(c, 1.0, 0.0)) < TOL32


# ----------------------------------------------------------- other modes ---

@pytest.mark.parametrize("d,n", [(1, 3), (2, 2), (2, 1), (3, 3), (4, 2), (6, 2)])
def test_dense_mode(d, n):
    L, T = 2, 2
    c = sparse_grid_case(d, n, L, T, seed=9)
    nb = 1 << L
    terms = oracle.build_dense_terms(d, n, T, nb, c["ops"])
    cells = c["rows"]
    ref = c["y"].copy()
    oracle.kronmult_dense(d, n, cells, cells, T, c["elem"], 0, 0, terms, nb, -0.5, c["x"], 2.0,
                          ref)
    m = asgard_amd.KronmultMatrix.dense(d, n, cells, cells, T, terms, c["elem"], 0, 0, nb)
    y = dev(c["y"])
    m.apply(-0.5, dev(c["x"]), 2.0, y)
    assert rel(host(y), ref) < TOL64


@pytest.mark.parametrize("d,n,L", [(3, 2, 5), (2, 3, 6), (4, 2, 4)])
def test_dense_mode_larger_grids(d, n, L):
    # every cell pair of the element table coupled (the reference's dense mode)
    T = 2
    c = sparse_grid_case(d, n, L, T, seed=10)
    nb = 1 << L
    terms = oracle.build_dense_terms(d, n, T, nb, c["ops"])
    cells = c["rows"]
    ref = c["y"].copy()
    oracle.kronmult_dense(d, n, cells, cells, T, c["elem"], 0, 0, terms, nb, 1.25, c["x"], 0.5,
                          ref)
    m = asgard_amd.KronmultMatrix.dense(d, n, cells, cells, T, terms, c["elem"], 0, 0, nb)
    y = dev(c["y"])
    m.apply(1.25, dev(c["x"]), 0.5, y)
    assert rel(host(y), ref) < TOL64


def test_host_pointer_apply_and_update():
    c = sparse_grid_case(2, 3, 4, 2, seed=6)
    m = asgard_amd.KronmultMatrix.sparse_grid(2, 3, c["table"], 4, 2, c["vA"])
    y = c["y"].copy()
    m.apply(1.0, c["x"], 1.0, y)  # numpy in, numpy out (apply<resource::host>)
    assert rel(y, oracle_apply(c, 1.0, 1.0)) < TOL64
    vA2 = 2.0 * c["vA"]
    m.update_stored_coefficients(vA2)
    y = np.zeros_like(c["y"])
    m.apply(1.0, c["x"], 0.0, y)
    # every one of the d = 2 factors doubles: the product scales by 2^d
    assert rel(y, 4.0 * oracle_apply(c, 1.0, 0.0)) < TOL64


def test_euler_operator():
    c = sparse_grid_case(3, 2, 4, 2, seed=8)
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 2, c["table"], 4, 2, c["vA"])
    dt, alpha, beta = 0.01, -1.0, 1.0
    y = dev(c["y"])
    m.apply_euler(dt, alpha, dev(c["x"]), beta, y)
    ref = oracle_apply(c, -dt * alpha, beta) + alpha * c["x"]
    assert rel(host(y), ref) < TOL64


def test_time_apply_reports():
    c = sparse_grid_case(3, 3, 5, 3, seed=2)
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 3, c["table"], 5, 3, c["vA"])
    x, y = dev(c["x"]), dev(c["y"])
    ms = m.time_apply(1.0, x, 0.0, y, 5)
    assert ms > 0
    assert rel(host(y), oracle_apply(c, 1.0, 0.0)) < TOL64


# ------------------------------------- lane-spread exchange through LDS ---

@pytest.mark.parametrize("d,n,L,T", [(4, 3, 4, 2), (5, 3, 3, 2), (3, 4, 3, 2), (6, 3, 2, 1),
                                     (4, 4, 2, 1)])
@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_transposed_exchange_parity(d, n, L, T, waves):
    # n >= 3 with lane-spread dimensions: the exchange is a transpose through
    # LDS and the accumulators live in the transposed layout; every kernel
    # (register, LDS tables) and both precisions must still match the oracle
    c = sparse_grid_case(d, n, L, T, seed=21)
    ref = oracle_apply(c, 1.5, -0.5)
    for dtype, tol in ((np.float64, TOL64), (np.float32, TOL32)):
        m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T,
                                                  c["vA"].astype(dtype), dtype=dtype)
        for lds in (-1, 0):
            m.set_tuning(waves_per_row=waves, lds_cap_bytes=lds)
            y = dev(c["y"].astype(dtype))
            m.apply(1.5, dev(c["x"].astype(dtype)), -0.5, y)
            assert rel(host(y), ref) < tol, (dtype, lds)
        mi = ia_matrix(c, dtype)
        mi.set_tuning(waves_per_row=waves)
        y = dev(c["y"].astype(dtype))
        mi.apply(1.5, dev(c["x"].astype(dtype)), -0.5, y)
        assert rel(host(y), ref) < tol, (dtype, "ia")


# ------------------------------ sum-factorised matrix-core kernel (d=3, n=3) ---

@pytest.mark.parametrize("L,T,seed", [(2, 1, 1), (4, 3, 2), (6, 3, 0), (5, 4, 3), (4, 5, 4),
                                      (3, 8, 5)])
@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_mfma_kernel_parity(L, T, seed, waves):
    """akm_set_kernel_family(1): dimensions 2 and 1 on v_mfma_f64_4x4x4_4b_f64
    with the group sums in the accumulator; against the oracle at every
    workgroup shape and both term bounds (T <= 4, T <= 8)."""
    c = sparse_grid_case(3, 3, L, T, seed=seed, dist="normal")
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 3, c["table"], L, T, c["vA"])
    m.set_kernel_family(1)
    m.set_tuning(waves_per_row=waves)
    y = dev(c["y"])
    m.apply(0.75, dev(c["x"]), -0.5, y)
    assert rel(host(y), oracle_apply(c, 0.75, -0.5)) < TOL64
    # the same apply again repeats bit for bit
    y2 = dev(c["y"])
    m.apply(0.75, dev(c["x"]), -0.5, y2)
    assert torch.equal(y, y2)


def test_mfma_kernel_dg_operators_and_euler():
    from asgard_amd import dg

    L, T = 5, 3
    ops = dg.operators_advection(3, 3, L)
    c = sparse_grid_case(3, 3, L, T, seed=6, ops=ops)
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 3, c["table"], L, T, c["vA"])
    m.set_kernel_family(1)
    x = dev(c["x"])
    y = m.apply_euler(0.01, 1.0, x, 0.0, dev(np.zeros_like(c["x"])))
    ax = oracle_apply(c, 1.0, 0.0)
    assert rel(host(y), c["x"] - 0.01 * ax) < TOL64


def test_mfma_kernel_refined_grid():
    from asgard_amd import grid

    table, _ = oracle.element_table([4] * 3)
    table = grid.refine(table, 0.3, seed=2)
    L = int(table[:, :3].max())
    rng = np.random.default_rng(9)
    s = 3 << L
    ops = [np.asfortranarray(rng.uniform(-1, 1, (s, s))) for _ in range(2 * 3)]
    vA = oracle.build_vA(3, 3, 2, L, ops)
    pntr, indx, iA = oracle.sparse_pattern(table, n=3, T=2, max_level=L)
    x = rng.uniform(-1, 1, table.shape[0] * 27)
    ref = np.zeros_like(x)
    oracle.kronmult_sparse(3, 3, pntr, indx, 2, iA, vA, 1.0, x, 0.0, ref)
    m = asgard_amd.KronmultMatrix.sparse_grid(3, 3, table, L, 2, vA)
    m.set_kernel_family(1)
    y = dev(np.zeros_like(x))
    m.apply(1.0, dev(x), 0.0, y)
    assert rel(host(y), ref) < TOL64


@pytest.mark.parametrize("L,T,seed", [(2, 1, 1), (3, 2, 2), (3, 5, 3), (4, 6, 4), (3, 8, 5)])
@pytest.mark.parametrize("waves", [1, 2, 4, 8, 16])
def test_mfma_d6n2_kernel_parity(L, T, seed, waves):
    """d=6, n=2: dimensions paired into dense 4 x 4 Kronecker factors, four
    terms per v_mfma_f64_16x16x4_f64 per column, then v_mfma_f64_4x4x4_4b_f64
    per group; against the oracle for every term bound and workgroup shape."""
    c = sparse_grid_case(6, 2, L, T, seed=seed, dist="normal")
    m = asgard_amd.KronmultMatrix.sparse_grid(6, 2, c["table"], L, T, c["vA"])
    m.set_kernel_family(1)
    m.set_tuning(waves_per_row=waves)
    y = dev(c["y"])
    m.apply(-1.25, dev(c["x"]), 0.5, y)
    assert rel(host(y), oracle_apply(c, -1.25, 0.5)) < TOL64
    y2 = dev(c["y"])
    m.apply(-1.25, dev(c["x"]), 0.5, y2)
    assert torch.equal(y, y2)


@pytest.mark.parametrize("L,T,seed", [(2, 1, 1), (4, 3, 2), (5, 4, 3), (4, 6, 4), (3, 8, 5)])
@pytest.mark.parametrize("waves", [1, 2, 4, 8])
def test_mfma_d4n2_kernel_parity(L, T, seed, waves):
    """d=4, n=2: the terms in the four blocks of the 4 x 4 x 4 product, both
    paired-dimension stages on the matrix cores; against the oracle."""
    c = sparse_grid_case(4, 2, L, T, seed=seed, dist="normal")
    m = asgard_amd.KronmultMatrix.sparse_grid(4, 2, c["table"], L, T, c["vA"])
    m.set_kernel_family(1)
    m.set_tuning(waves_per_row=waves)
    y = dev(c["y"])
    m.apply(0.5, dev(c["x"]), 1.0, y)
    assert rel(host(y), oracle_apply(c, 0.5, 1.0)) < TOL64


def test_mfma_d4n2_advdiff_euler():
    from asgard_amd import dg

    L, T = 5, 4
    ops = dg.operators_advection_diffusion(4, 2, L, n_adv=2, nu=0.1)
    c = sparse_grid_case(4, 2, L, T, seed=7, ops=ops)
    m = asgard_amd.KronmultMatrix.sparse_grid(4, 2, c["table"], L, T, c["vA"])
    m.set_kernel_family(1)
    y = m.apply_euler(1e-3, 1.0, dev(c["x"]), 0.0, dev(np.zeros_like(c["x"])))
    assert rel(host(y), c["x"] - 1e-3 * oracle_apply(c, 1.0, 0.0)) < TOL64


def test_mfma_d6n2_vlasov_sampled_rows():
    """The d=6 L=7 Vlasov operator through the matrix-core kernel, checked on
    sampled rows of every length against the oracle."""
    from asgard_amd import dg, grid

    d, n, L = 6, 2, 7
    table, _ = grid.element_table([L] * d)
    ops = dg.operators_vlasov(n, L)
    T = len(ops) // d
    vA = grid.build_vA(d, n, T, L, ops)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, table, L, T, vA)
    m.set_kernel_family(1)
    x = np.random.default_rng(0).uniform(-1, 1, table.shape[0] * n ** d)
    y = dev(np.zeros_like(x))
    m.apply(1.0, dev(x), 0.0, y)
    rows = np.linspace(0, table.shape[0] - 1, 12).astype(np.int32)
    pntr, indx, iA = oracle.sparse_pattern_rows(table, rows, n=n, T=T, max_level=L)
    ref = np.zeros(len(rows) * n ** d)
    oracle.kronmult_sparse(d, n, pntr, indx, T, iA, vA, 1.0, x, 0.0, ref)
    got = host(y).reshape(-1, n ** d)[rows].reshape(-1)
    assert rel(got, ref) < TOL64


@pytest.mark.parametrize("waves", [8, 16])
def test_mfma_d6n2_eight_terms_tables_past_lds(waves):
    """d=6 L=7 with 8 random terms: all six dimensions' factor tables of a
    coarse row need about 196 KB, more than a CU's LDS, so the kernel stages
    dimensions 2-5 only and reads dimensions 0, 1 from global memory; checked
    on sampled rows of every length against the oracle."""
    from asgard_amd import grid

    d, n, L, T = 6, 2, 7, 8
    table, _ = grid.element_table([L] * d)
    rng = np.random.default_rng(21)
    s = n << L
    ops = [np.asfortranarray(rng.uniform(-1, 1, (s, s))) for _ in range(T * d)]
    vA = grid.build_vA(d, n, T, L, ops)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, table, L, T, vA)
    m.set_kernel_family(1)
    m.set_tuning(waves_per_row=waves)
    x = rng.uniform(-1, 1, table.shape[0] * n ** d)
    y = dev(np.zeros_like(x))
    m.apply(1.0, dev(x), 0.0, y)
    rows = np.linspace(0, table.shape[0] - 1, 8).astype(np.int32)
    pntr, indx, iA = oracle.sparse_pattern_rows(table, rows, n=n, T=T, max_level=L)
    ref = np.zeros(len(rows) * n ** d)
    oracle.kronmult_sparse(d, n, pntr, indx, T, iA, vA, 1.0, x, 0.0, ref)
    got = host(y).reshape(-1, n ** d)[rows].reshape(-1)
    assert rel(got, ref) < TOL64


def test_mfma_kernel_family_rejected_elsewhere():
    c = sparse_grid_case(5, 2, 3, 1, seed=7)
    m = asgard_amd.KronmultMatrix.sparse_grid(5, 2, c["table"], 3, 1, c["vA"])
    with pytest.raises(asgard_amd.AkmError):
        m.set_kernel_family(1)
    c3 = sparse_grid_case(3, 3, 3, 2, seed=8)
    m32 = asgard_amd.KronmultMatrix.sparse_grid(3, 3, c3["table"], 3, 2,
                                                c3["vA"].astype(np.float32), dtype=np.float32)
    with pytest.raises(asgard_amd.AkmError):
        m32.set_kernel_family(1)
    with pytest.raises(ValueError):
        m.set_kernel_family(2)
