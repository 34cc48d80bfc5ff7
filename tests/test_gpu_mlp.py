"""GPU parity of the learned preconditioner's kernels against the numpy
oracle (oracle/mlp.py): FP32 forward, backpropagated gradients, Adam, the
BF16 matrix-core preconditioner apply, and training on solver pairs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
from asgard_amd import dg, grid  # noqa: E402
from oracle import mlp as omlp  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def random_params(lay, rng, scale=0.2):
    p = rng.standard_normal(lay.num_params) * scale
    W1, b1, W2, b2, W3, b3 = lay.unpack(p)
    W1 = W1.copy()
    W1[:, lay.N + lay.d:] = 0.0
    W3 = W3.copy()
    W3[lay.N:, :] = 0.0
    b3 = b3.copy()
    b3[lay.N:] = 0.0
    return lay.pack(W1, b1, W2, b2, W3, b3).astype(np.float32)


def make(d, n, L, hidden, seed=0):
    table, _ = grid.element_table([L] * d)
    pre = asgard_amd.MlpPreconditioner(d, n, table, L, hidden=hidden, seed=seed)
    lay = omlp.Layout(n ** d, d, hidden)
    assert pre.num_params == lay.num_params
    assert (pre.padded_inputs, pre.padded_outputs) == (lay.K1, lay.N3)
    return table, pre, lay


def rows(lay, B, rng):
    X = np.zeros((B, lay.K1), np.float32)
    X[:, :lay.N + lay.d] = rng.standard_normal((B, lay.N + lay.d))
    T = np.zeros((B, lay.N3), np.float32)
    T[:, :lay.N] = rng.standard_normal((B, lay.N))
    return X, T


def test_untrained_network_is_identity():
    table, pre, lay = make(3, 2, 4, 64)
    v = np.random.default_rng(0).standard_normal(table.shape[0] * 8)
    z = torch.empty(v.size, dtype=torch.float64, device="cuda")
    pre.apply(dev(v), z)
    assert np.array_equal(host(z), v)
    # the initial parameters: zero biases, zero output layer, zero padding
    p = pre.get_params().astype(np.float64)
    W1, b1, W2, b2, W3, b3 = lay.unpack(p)
    assert not W3.any() and not b3.any() and not b1.any() and not b2.any()
    assert not W1[:, lay.N + lay.d:].any()


@pytest.mark.parametrize("d,n,hidden,B", [(2, 2, 32, 100), (3, 3, 64, 257), (6, 2, 256, 130)])
def test_forward_matches_oracle(d, n, hidden, B):
    rng = np.random.default_rng(d * 10 + n)
    _, pre, lay = make(d, n, 3, hidden)
    p = random_params(lay, rng)
    pre.set_params(p)
    assert np.array_equal(pre.get_params(), p)
    X, _ = rows(lay, B, rng)
    O = torch.empty((B, lay.N3), dtype=torch.float32, device="cuda")
    pre.forward(dev(X), O)
    ref, _ = omlp.forward(lay, p, X)
    got = host(O)
    assert np.abs(got - ref).max() < 2e-5 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("d,n,hidden,B", [(2, 2, 32, 64), (3, 3, 64, 257), (6, 2, 256, 200)])
def test_gradients_match_oracle(d, n, hidden, B):
    rng = np.random.default_rng(100 + d)
    _, pre, lay = make(d, n, 3, hidden)
    p = random_params(lay, rng)
    pre.set_params(p)
    X, T = rows(lay, B, rng)
    g = torch.zeros(lay.num_params, dtype=torch.float32, device="cuda")
    loss, _ = pre.gradients(dev(X), dev(T), g)
    lref, gref = omlp.gradients(lay, p, X, T)
    assert abs(loss - lref) < 1e-5 * lref
    got = host(g)
    for name, (o, size) in lay.offsets.items():
        a, b = got[o:o + size], gref[o:o + size]
        assert np.abs(a - b).max() < 1e-4 * max(1e-3, np.abs(b).max()), name


def test_adam_matches_oracle():
    rng = np.random.default_rng(9)
    _, pre, lay = make(2, 2, 3, 32)
    p = random_params(lay, rng)
    pre.set_params(p)
    m = np.zeros(lay.num_params)
    v = np.zeros(lay.num_params)
    pr = p.astype(np.float64)
    for t in range(1, 4):
        g = (rng.standard_normal(lay.num_params) * 0.1).astype(np.float32)
        pre.adam_step(dev(g), 1e-2)
        pr, m, v = omlp.adam(pr, g.astype(np.float64), m, v, t, 1e-2)
    got = pre.get_params()
    assert np.abs(got - pr).max() < 1e-5


@pytest.mark.parametrize("d,n,L,hidden", [(2, 2, 4, 32), (3, 3, 3, 64), (6, 2, 3, 256),
                                          (4, 2, 4, 128)])
def test_bf16_apply_matches_oracle(d, n, L, hidden):
    rng = np.random.default_rng(20 + d)
    table, pre, lay = make(d, n, L, hidden)
    p = random_params(lay, rng, scale=0.15)
    pre.set_params(p)
    cells = table.shape[0]
    v = rng.standard_normal(cells * lay.N)
    v[:lay.N] = 0.0  # a zero cell stays zero
    z = torch.empty(v.size, dtype=torch.float64, device="cuda")
    pre.apply(dev(v), z)
    got = host(z)
    ref = omlp.precondition(lay, p.astype(np.float64), v, omlp.features(table, L))
    # BF16 weights and activations: ~3 significant digits on the correction
    corr = (ref - v).reshape(cells, lay.N)
    err = (got - ref).reshape(cells, lay.N)
    assert np.all(got[:lay.N] == 0.0)
    assert np.linalg.norm(err) < 0.03 * np.linalg.norm(corr)
    assert np.abs(err).max() < 0.05 * np.abs(corr).max()


def test_training_on_solver_pairs_lowers_the_loss():
    d, n, L, dt = 2, 2, 5, 0.02
    table, _ = grid.element_table([L] * d)
    ops = dg.operators_advection_diffusion(d, n, L, n_adv=1, nu=0.1)
    vA = grid.build_vA(d, n, d, L, ops)
    mat = asgard_amd.KronmultMatrix.sparse_grid(d, n, table, L, d, vA)
    from asgard_amd import learned

    pre, info = learned.train_preconditioner(mat, table, L, dt, samples=20000, epochs=6,
                                             batch=256, lr=1e-3, hidden=64, seed=3)
    assert info["samples"] >= 20000
    assert info["loss_after"] < 0.7 * info["loss_before"], info
