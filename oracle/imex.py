"""IMEX step oracle.

TEST INFRASTRUCTURE ONLY: the checker of ``akm_imex_step``; the product never
imports it.

``imex_step`` restates the stage arithmetic of imex_advance
(src/time_advance.cpp:760-1000): two SSP-RK2 stages, each an explicit apply
of the explicit-term operator E followed by an implicit backward-Euler solve
of the implicit-term operator B,

    f1s = f0 + dt E f0                    (:795-822, axpy(fx, f, dt))
    (I - dt B) f1 = f1s, guess f1s        (:838-868)
    f2s = (f0 + f1 + dt E f1) / 2         (:885-919, axpy, axpy, scal 0.5)
    (I - dt/2 B) f = f2s, guess f1        (:930-960)

and without the collision operator f1 = f1s and f = f2s (:869-873, :966-975).
The applies and solves are callables, so the tests plug in the kronmult and
GMRES oracles (or dense matrices).

Parity: pinned by closed forms in tests/test_oracle_imex.py -- with B = 0 the
step is the second-order Taylor polynomial of exp(dt E) applied to f0, with
E = 0 it is (I - dt/2 B)^-1 (f0 + (I - dt B)^-1 f0) / 2 -- and by the
non-collision branch reducing to the explicit SSP-RK2 stages.
"""
import numpy as np


def imex_step(explicit_apply, implicit_solve, dt, f0, collision=True, hook=None):
    """explicit_apply(x) -> E x; implicit_solve(c, guess, rhs) -> the solution
    of (I - c B) x = rhs started from `guess`; hook(event, vector) runs at
    events 0..3 (before E f0, before solve 1, before E f1, before solve 2)."""
    f0 = np.asarray(f0, dtype=np.float64)
    f = f0.copy()
    if hook is not None:
        hook(0, f)
    f = f + dt * explicit_apply(f)
    if hook is not None:
        hook(1, f)
    f1 = f.copy()
    if collision:
        f1 = implicit_solve(dt, f1, f)
    guess = f1.copy()
    if hook is not None:
        hook(2, f1)
    f2 = f1 + dt * explicit_apply(f1)
    fs = 0.5 * (f0 + f2)
    if hook is not None:
        hook(3, fs)
    if collision:
        return implicit_solve(0.5 * dt, guess, fs)
    return fs
