"""Training set-up of the learned preconditioner (host side).

The pairs come from the plain FP64 device solver itself: for random
right-hand sides b ~ U[-1, 1] it solves (I - dt*A) e = b from e = 0, so each
pair is a residual (b, the residual of the zero guess) and the exact
correction e that removes it.  Every cell of every pair is one training
sample of the per-cell network.
"""
from __future__ import annotations

import math

import numpy as np

from . import AKM_F64, AkmInvalid, AKM_ERR_INVALID, MlpPreconditioner, simple_gmres_euler


def training_pairs(mat, dt, count, seed=0, tol=1e-12, restart=50):
    """(R, E): float64 CUDA tensors of shape (count, n) with
    (I - dt*A) E[q] = R[q] solved to |residual| <= tol*|R[q]|."""
    import torch

    if mat.prec != AKM_F64:
        raise AkmInvalid(AKM_ERR_INVALID, "training pairs need the FP64 operator")
    n = mat.output_size()
    rng = np.random.default_rng(seed)
    R = torch.empty((count, n), dtype=torch.float64, device="cuda")
    E = torch.zeros((count, n), dtype=torch.float64, device="cuda")
    restart = min(int(restart), n)
    for q in range(count):
        b = rng.uniform(-1.0, 1.0, n)
        R[q].copy_(torch.from_numpy(b))
        torch.cuda.synchronize()
        simple_gmres_euler(dt, mat, E[q], R[q], restart, -1, tol * float(np.linalg.norm(b)))
    torch.cuda.synchronize()
    return R, E


def train_preconditioner(mat, table, max_level, dt, samples=200_000, epochs=8, batch=1024,
                         lr=1e-3, hidden=256, seed=0):
    """Builds and trains the per-cell MLP for the operator (I - dt*A) of the
    FP64 matrix `mat` on the grid `table`.  `samples` per-cell samples are
    drawn from ceil(samples / cells) solver pairs.  Returns
    (preconditioner, {"pairs", "samples", "loss_before", "loss_after"})."""
    import torch

    table = np.ascontiguousarray(table, dtype=np.int32)
    cells, twod = table.shape
    d = twod // 2
    n = mat.kron_size
    pre = MlpPreconditioner(d, n, table, max_level, hidden=hidden, seed=seed)
    pre.set_stream(torch.cuda.current_stream().cuda_stream)
    count = max(1, math.ceil(samples / cells))
    R, E = training_pairs(mat, dt, count, seed=seed + 1)
    # the untrained network is the identity: its loss is the mean square of
    # the targets (e_c - r_c)/|r_c|
    Rc = R.reshape(count, cells, -1)
    Dc = (E - R).reshape(count, cells, -1)
    s2 = (Rc * Rc).sum(dim=2).clamp_min(1e-300)
    loss0 = float(((Dc * Dc).sum(dim=2) / s2).mean().item() / Rc.shape[2])
    loss = pre.train(R, E, epochs=epochs, batch=batch, lr=lr, seed=seed + 2)
    return pre, {"pairs": count, "samples": count * cells, "loss_before": loss0,
                 "loss_after": loss}
