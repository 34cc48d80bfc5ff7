"""The MLP preconditioner oracle (oracle/mlp.py) pinned by mathematics:
backpropagated gradients against central finite differences, Adam's first
step against its closed form, the preconditioner's identity at a zero output
layer and its positive homogeneity per cell.  CPU only."""
import numpy as np
import pytest

from oracle import mlp


def random_params(lay, rng, scale=0.3):
    p = rng.standard_normal(lay.num_params) * scale
    W1, b1, W2, b2, W3, b3 = lay.unpack(p)
    W1 = W1.copy()
    W1[:, lay.N + lay.d:] = 0.0  # padding inputs
    W3 = W3.copy()
    W3[lay.N:, :] = 0.0          # padding outputs
    b3 = b3.copy()
    b3[lay.N:] = 0.0
    return lay.pack(W1, b1, W2, b2, W3, b3)


@pytest.mark.parametrize("N,d,H", [(4, 2, 16), (8, 3, 32), (27, 3, 16)])
def test_gradients_match_finite_differences(N, d, H):
    rng = np.random.default_rng(7)
    lay = mlp.Layout(N, d, H)
    p = random_params(lay, rng)
    B = 9
    X = np.zeros((B, lay.K1))
    X[:, :N + d] = rng.standard_normal((B, N + d))
    T = np.zeros((B, lay.N3))
    T[:, :N] = rng.standard_normal((B, N))
    L, g = mlp.gradients(lay, p, X, T)
    assert abs(L - mlp.loss(lay, p, X, T)) < 1e-14
    # every bias and a random subset of weights, central differences
    idx = list(range(lay.offsets["b1"][0], lay.offsets["b1"][0] + H))
    idx += list(range(lay.offsets["b3"][0], lay.offsets["b3"][0] + N))
    idx += list(rng.choice(lay.num_params, 120, replace=False))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp.loss(lay, p + e, X, T) - mlp.loss(lay, p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    p = rng.standard_normal(50)
    g = rng.standard_normal(50)
    lr = 1e-2
    p1, m, v = mlp.adam(p, g, np.zeros(50), np.zeros(50), 1, lr)
    # bias-corrected moments are g and g^2 at t = 1
    assert np.allclose(p1, p - lr * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_preconditioner_identity_and_homogeneity():
    rng = np.random.default_rng(5)
    d, n, cells = 2, 2, 11
    lay = mlp.Layout(n ** d, d, 32)
    table = np.zeros((cells, 2 * d), np.int32)
    table[:, :d] = rng.integers(0, 5, (cells, d))
    feats = mlp.features(table, 4)
    p = random_params(lay, rng)
    W1, b1, W2, b2, W3, b3 = lay.unpack(p)
    p0 = lay.pack(W1, b1, W2, b2, np.zeros_like(W3), np.zeros_like(b3))
    v = rng.standard_normal(cells * lay.N)
    assert np.array_equal(mlp.precondition(lay, p0, v, feats), v)
    # P(c v_c) = c P(v_c) for c > 0, cell by cell; zero cells stay zero
    z = mlp.precondition(lay, p, v, feats)
    scale = np.repeat(rng.uniform(0.1, 10.0, cells), lay.N)
    zs = mlp.precondition(lay, p, v * scale, feats)
    assert np.allclose(zs, z * scale, rtol=1e-12, atol=1e-12)
    v[:lay.N] = 0.0
    assert np.all(mlp.precondition(lay, p, v, feats)[:lay.N] == 0.0)


def test_features_encode_level_and_position():
    table = np.array([[0, 0], [1, 0], [2, 1], [3, 3]], np.int32)  # d = 1
    f = mlp.features(table, 3)[:, 0]
    assert np.allclose(f, [0.5 / 4, 1.5 / 4, 2.75 / 4, 3.875 / 4])
