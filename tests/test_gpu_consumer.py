"""The compiled C++ consumer of the boundary (tests/cpp/consumer.cpp, built
against libasgard_amd.so through include/asgard_amd.hpp) run on the d=2 n=2
|l|_1<=5 two-term workload, every result checked against the oracle: the
sparse-grid and CPU-CSR constructors, apply<host> and apply<device>,
matrix_list, the empty matrix, simple_gmres_euler on host and device vectors,
its invalid_argument for restart > n, and the dense simple_gmres gold of
src/solver_tests.cpp:141-213."""
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import oracle  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONSUMER = os.path.join(ROOT, "tests", "cpp", "consumer")


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if not os.path.exists(CONSUMER):
        pytest.fail("tests/cpp/consumer is not built (run __graft_entry__.build())")


def write_input(path, d, n, L, T, dt, restart, table, vA, pntr, indx, iA, x, b):
    with open(path, "wb") as f:
        f.write(struct.pack("7i", d, n, L, T, table.shape[0], len(indx), restart))
        f.write(struct.pack("d", dt))
        f.write(np.ascontiguousarray(table, np.int32).tobytes())
        f.write(struct.pack("q", vA.size))
        f.write(np.ascontiguousarray(vA, np.float64).tobytes())
        f.write(np.ascontiguousarray(pntr, np.int32).tobytes())
        f.write(np.ascontiguousarray(indx, np.int32).tobytes())
        f.write(np.ascontiguousarray(iA, np.int32).tobytes())
        f.write(np.ascontiguousarray(x, np.float64).tobytes())
        f.write(np.ascontiguousarray(b, np.float64).tobytes())


def read_output(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            name = f.read(32)
            if not name:
                break
            (count,) = struct.unpack("q", f.read(8))
            out[name.rstrip(b"\0").decode()] = np.frombuffer(f.read(8 * count), np.float64)
    return out


def test_cpp_consumer_against_oracle(tmp_path):
    d, n, L, T, dt, restart = 2, 2, 5, 2, 0.01, 20
    rng = np.random.default_rng(0)
    table, _ = oracle.element_table([L] * d)
    size1d = n << L
    ops = [np.asfortranarray(rng.uniform(-1, 1, (size1d, size1d))) for _ in range(T * d)]
    vA = oracle.build_vA(d, n, T, L, ops)
    pntr, indx, iA = oracle.sparse_pattern(table, n=n, T=T, max_level=L)
    cells, N = table.shape[0], n ** d
    x = rng.uniform(-1, 1, cells * N)
    b = rng.uniform(-1, 1, cells * N)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(src, d, n, L, T, dt, restart, table, vA, pntr, indx, iA, x, b)
    run = subprocess.run([CONSUMER, str(src), str(dst)], capture_output=True, text=True,
                         timeout=120)
    assert run.returncode == 0, run.stderr
    out = read_output(dst)

    ref = np.zeros_like(x)
    oracle.kronmult_sparse(d, n, pntr, indx, T, iA, vA, 1.0, x, 0.0, ref)
    scale = np.abs(ref).max()
    for key in ("grid_host", "grid_device", "list_apply"):
        assert np.abs(out[key] - ref).max() <= 1e-13 * scale, key
    # every factor block doubled: the d-fold Kronecker products scale by 2^d
    assert np.abs(out["list_twice"] - 2 ** d * ref).max() <= 1e-13 * 2 ** d * scale
    assert np.abs(out["csr_axpby"] - (2 * b - 0.5 * ref)).max() <= 1e-13 * scale
    assert np.array_equal(out["empty_scale"], 3 * b)
    nnz = int(pntr[-1])
    flops = N * n * d * T * nnz
    assert out["sizes"].tolist() == [N, cells * N, cells * N, nnz, flops, 2 * flops]
    assert out["list_flags"].tolist() == [0, 1, 0]

    xo, erro, itso = oracle.gmres_euler_sparse(d, n, pntr, indx, T, iA, vA, dt,
                                               np.zeros_like(b), b, restart, cells * N, 1e-12)
    for key in ("gmres_host", "gmres_device"):
        assert np.abs(out[key] - xo).max() <= 1e-10 * np.abs(xo).max(), key
    for key in ("gmres_info", "gmres_device_info"):
        err, its = out[key]
        assert err <= 1e-12 and abs(its - itso) <= 1, key
    assert out["gmres_restart_throws"][0] == 1

    x_gold = np.array([4.715561567725287e-02, 1.257695999382253e-01, 1.625351700791827e-02])
    eps = np.finfo(np.float64).eps
    assert np.allclose(out["dense_gmres"], x_gold, rtol=1e-13, atol=1e-15)
    assert np.allclose(out["dense_gmres_jacobi"], x_gold, rtol=1e-13, atol=1e-15)
    assert (out["dense_gmres_errors"] < eps).all()
