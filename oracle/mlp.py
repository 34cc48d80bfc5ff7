"""CPU oracle of the learned per-cell preconditioner (numpy, float64).

TEST INFRASTRUCTURE ONLY: the checker of asgard_amd's MLP kernels (forward,
backward, Adam, preconditioner apply).  Only ``tests/`` imports it.

There is no reference counterpart (ASGarD has no learned preconditioner), so
this module restates the network defined by the akm_mlp_* entry points:

* parameters in one flat array
  ``[W1 (H x K1) | b1 (H) | W2 (H x H) | b2 (H) | W3 (N3 x H) | b3 (N3)]``,
  W row-major (output x input);
* ``o(x) = W3 relu(W2 relu(W1 x + b1) + b2) + b3``;
* loss = mean over rows and the first N outputs of ``(o - t)^2``;
* Adam of Kingma & Ba (2015), beta1 0.9, beta2 0.999, eps 1e-8;
* preconditioner ``z_c = r_c + |r_c| o([r_c/|r_c|, f_c])`` per cell.

Parity status: the gradients are pinned by central finite differences of the
loss, Adam's first step by its closed form (update = lr*g/(|g|+eps)), the
preconditioner by the identity at a zero output layer and by its positive
homogeneity per cell (tests/test_oracle_mlp.py).
"""
from __future__ import annotations

import numpy as np


def round16(v):
    return (int(v) + 15) // 16 * 16


class Layout:
    def __init__(self, N, d, H):
        self.N, self.d, self.H = int(N), int(d), int(H)
        self.K1 = round16(N + d)
        self.N3 = round16(N)
        sizes = [("W1", self.H * self.K1), ("b1", self.H), ("W2", self.H * self.H),
                 ("b2", self.H), ("W3", self.N3 * self.H), ("b3", self.N3)]
        self.offsets = {}
        o = 0
        for name, size in sizes:
            self.offsets[name] = (o, size)
            o += size
        self.num_params = o

    def unpack(self, p):
        p = np.asarray(p, dtype=np.float64)
        get = lambda k: p[self.offsets[k][0]:self.offsets[k][0] + self.offsets[k][1]]
        return (get("W1").reshape(self.H, self.K1), get("b1"),
                get("W2").reshape(self.H, self.H), get("b2"),
                get("W3").reshape(self.N3, self.H), get("b3"))

    def pack(self, W1, b1, W2, b2, W3, b3):
        return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def forward(lay: Layout, p, X):
    """(O, cache) for inputs X (rows x K1)."""
    W1, b1, W2, b2, W3, b3 = lay.unpack(p)
    X = np.asarray(X, dtype=np.float64)
    A1 = X @ W1.T + b1
    H1 = np.maximum(A1, 0.0)
    A2 = H1 @ W2.T + b2
    H2 = np.maximum(A2, 0.0)
    O = H2 @ W3.T + b3
    return O, (X, H1, H2)


def loss(lay: Layout, p, X, T):
    O, _ = forward(lay, p, X)
    D = (O - np.asarray(T, dtype=np.float64))[:, :lay.N]
    return float(np.mean(D * D))


def gradients(lay: Layout, p, X, T):
    """(loss, flat gradient) by backpropagation."""
    W1, b1, W2, b2, W3, b3 = lay.unpack(p)
    O, (X, H1, H2) = forward(lay, p, X)
    B = X.shape[0]
    D = O - np.asarray(T, dtype=np.float64)
    D[:, lay.N:] = 0.0
    L = float(np.sum(D * D) / (B * lay.N))
    dO = 2.0 * D / (B * lay.N)
    gW3 = dO.T @ H2
    gb3 = dO.sum(axis=0)
    dA2 = (dO @ W3) * (H2 > 0)
    gW2 = dA2.T @ H1
    gb2 = dA2.sum(axis=0)
    dA1 = (dA2 @ W2) * (H1 > 0)
    gW1 = dA1.T @ X
    gb1 = dA1.sum(axis=0)
    return L, lay.pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step at step count t (1-based); returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def features(table, max_level):
    """Per cell and dimension: (level + (position + 1/2) / cells of the level)
    / (max_level + 1)."""
    table = np.asarray(table)
    d = table.shape[1] // 2
    lev = table[:, :d].astype(np.float64)
    pos = table[:, d:].astype(np.float64)
    count = np.where(lev <= 1, 1.0, 2.0 ** np.maximum(lev - 1, 0))
    return (lev + (pos + 0.5) / count) / (max_level + 1)


def precondition(lay: Layout, p, v, feats):
    """z = P(v) for a whole-grid vector v (cells * N)."""
    cells = feats.shape[0]
    V = np.asarray(v, dtype=np.float64).reshape(cells, lay.N)
    s = np.sqrt(np.sum(V * V, axis=1))
    X = np.zeros((cells, lay.K1))
    nz = s > 0
    X[nz, :lay.N] = V[nz] / s[nz, None]
    X[:, lay.N:lay.N + lay.d] = feats
    O, _ = forward(lay, p, X)
    return (V + s[:, None] * O[:, :lay.N]).ravel()


def samples(lay: Layout, R, E, feats):
    """(X, T) training rows of whole-grid pairs R, E (pairs x cells*N)."""
    cells = feats.shape[0]
    R = np.asarray(R, dtype=np.float64).reshape(-1, cells, lay.N)
    E = np.asarray(E, dtype=np.float64).reshape(-1, cells, lay.N)
    s = np.sqrt(np.sum(R * R, axis=2))
    inv = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 0.0)
    X = np.zeros((R.shape[0], cells, lay.K1))
    T = np.zeros((R.shape[0], cells, lay.N3))
    X[:, :, :lay.N] = R * inv[:, :, None]
    X[:, :, lay.N:lay.N + lay.d] = feats[None]
    T[:, :, :lay.N] = (E - R) * inv[:, :, None]
    return X.reshape(-1, lay.K1), T.reshape(-1, lay.N3)
