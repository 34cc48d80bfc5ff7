"""The IMEX oracle against closed forms of the step (CPU only).

With B = 0 the two solves are the identity and the step is SSP-RK2 of the
linear E: f0 + dt E f0 + dt^2/2 E^2 f0.  With E = 0 it is
(I - dt/2 B)^-1 (f0 + (I - dt B)^-1 f0) / 2.  The non-collision branch is the
explicit stages alone.  A dropped 1/2, the wrong stage vector or guess, or a
swapped dt fails one of these.
"""
import numpy as np

from oracle import imex


def _mats(n=12, seed=0):
    rng = np.random.default_rng(seed)
    return 0.3 * rng.standard_normal((n, n)), 0.3 * rng.standard_normal((n, n)), \
        rng.standard_normal(n)


def _solve(B):
    n = B.shape[0]
    return lambda c, guess, rhs: np.linalg.solve(np.eye(n) - c * B, rhs)


def test_b_zero_is_second_order_taylor():
    E, _, f0 = _mats()
    dt = 0.05
    got = imex.imex_step(lambda x: E @ x, lambda c, g, r: r.copy(), dt, f0)
    want = f0 + dt * E @ f0 + 0.5 * dt * dt * E @ (E @ f0)
    assert np.allclose(got, want, rtol=0, atol=1e-14 * np.abs(want).max())


def test_e_zero_is_two_solves():
    _, B, f0 = _mats(seed=1)
    dt = 0.1
    n = B.shape[0]
    got = imex.imex_step(lambda x: np.zeros_like(x), _solve(B), dt, f0)
    f1 = np.linalg.solve(np.eye(n) - dt * B, f0)
    want = np.linalg.solve(np.eye(n) - 0.5 * dt * B, 0.5 * (f0 + f1))
    assert np.allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_no_collision_is_explicit_stages():
    E, B, f0 = _mats(seed=2)
    dt = 0.02
    got = imex.imex_step(lambda x: E @ x, _solve(B), dt, f0, collision=False)
    f1 = f0 + dt * E @ f0
    want = 0.5 * (f0 + f1 + dt * E @ f1)
    assert np.allclose(got, want, rtol=0, atol=1e-15 * np.abs(want).max())


def test_hook_events_and_vectors():
    E, B, f0 = _mats(seed=3)
    dt = 0.03
    seen = []
    imex.imex_step(lambda x: E @ x, _solve(B), dt, f0, hook=lambda e, v: seen.append((e, v.copy())))
    assert [e for e, _ in seen] == [0, 1, 2, 3]
    assert np.array_equal(seen[0][1], f0)
    assert np.allclose(seen[1][1], f0 + dt * E @ f0)
