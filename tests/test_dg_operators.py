"""The DG operator generator (asgard_amd.dg) against the reference's own
coefficient fixtures (testing/generated-inputs/coefficients, produced by the
reference's MATLAB pipeline and read by src/coefficients_tests.cpp): cell
basis ("norotate") and multiwavelet basis, for a single partial term with a
coefficient (continuity, degree 2 and 3) and a two-partial-term diffusion
(degree 6, Neumann/Dirichlet boundaries).  Plus the algebra the rotation must
keep: F orthogonal, spectra preserved.  CPU only."""
import numpy as np
import pytest

from asgard_amd import dg
from tests.fixtures import golden, read_octave


def fixture(name):
    return read_octave(golden("coefficients", name))


def minus_one(x, t):
    return -1.0


@pytest.mark.parametrize("rot", [False, True])
def test_continuity1_div(rot):
    C = dg.partial_term("div", 2, 2, (-1.0, 1.0), g=minus_one, flux="downwind")
    if rot:
        C = dg.rotate(C, 2, 2)
    name = "continuity1_coefficients_%sl2_d2_1_1.dat" % ("" if rot else "norotate_")
    G = fixture(name)
    assert np.abs(C - G).max() < 1e-12 * max(1.0, np.abs(G).max())


@pytest.mark.parametrize("rot", [False, True])
def test_continuity2_div_and_mass(rot):
    # term 1: dim 1 div(-1) downwind periodic on [-1, 1]; dim 2 mass (no
    # coefficient, so no mass inversion) on [-2, 2]; level 4, degree 3
    div = dg.partial_term("div", 3, 4, (-1.0, 1.0), g=minus_one, flux="downwind")
    mass = dg.partial_term("mass", 3, 4, (-2.0, 2.0), g=None)
    if rot:
        div, mass = dg.rotate(div, 3, 4), dg.rotate(mass, 3, 4)
    tag = "" if rot else "norotate_"
    for C, name in [(div, f"continuity2_coefficients_{tag}l4_4_d3_1_1.dat"),
                    (mass, f"continuity2_coefficients_{tag}l4_4_d3_1_2.dat")]:
        G = fixture(name)
        assert np.abs(C - G).max() < 1e-12 * max(1.0, np.abs(G).max()), name


@pytest.mark.parametrize("rot", [False, True])
def test_diffusion1_chained_partials(rot):
    # term 1 of diffusion_1: div (no coefficient, downwind, Neumann) times
    # grad (no coefficient, upwind, Dirichlet) on [0, 1], level 5, degree 6
    C = dg.term_matrix([dict(kind="div", flux="downwind", left="neumann", right="neumann"),
                        dict(kind="grad", flux="upwind", left="dirichlet", right="dirichlet")],
                       6, 5, (0.0, 1.0), rotated=rot)
    name = "diffusion1_coefficients_%sl5_d6_1_1.dat" % ("" if rot else "norotate_")
    G = fixture(name)
    assert np.abs(C - G).max() < 1e-10 * max(1.0, np.abs(G).max())


@pytest.mark.parametrize("n,L", [(1, 3), (2, 5), (3, 4), (4, 3)])
def test_wavelet_transform_orthogonal(n, L):
    F = dg.wavelet_transform(n, L)
    assert np.abs(F @ F.T - np.eye(n << L)).max() < 1e-12


def test_rotation_preserves_spectrum():
    C = dg.partial_term("div", 2, 4, g=minus_one, flux="upwind")
    R = dg.rotate(C, 2, 4)
    ev = np.sort_complex(np.linalg.eigvals(C))
    er = np.sort_complex(np.linalg.eigvals(R))
    assert np.allclose(ev, er, atol=1e-9)


def test_rotated_blocks_follow_connect_1d():
    # in the wavelet basis the operator vanishes on 1D cell pairs that
    # connect_1d does not list, which is what the sparse kronmult relies on
    import oracle

    n, L = 2, 4
    R = dg.advection(n, L)
    pntr, indx, _ = oracle.connect_1d(L)
    M = 1 << L
    for r in range(M):
        conn = set(indx[pntr[r]:pntr[r + 1]].tolist())
        for c in range(M):
            if c not in conn:
                assert np.abs(R[n * r:n * r + n, n * c:n * c + n]).max() < 1e-12, (r, c)


# ---- penalty: pinned by the jump identity, independent of the assembly ----

def _edge_values(u, n, level, domain):
    a, b = domain
    h = (b - a) / (1 << level)
    PL = dg.legendre_onb([-1.0], n)[0][0] / np.sqrt(h)
    PR = dg.legendre_onb([1.0], n)[0][0] / np.sqrt(h)
    cells = u.reshape(-1, n)
    return cells @ PL, cells @ PR


@pytest.mark.parametrize("n,level", [(1, 3), (2, 4), (3, 3), (4, 2)])
@pytest.mark.parametrize("bc", ["periodic", "dirichlet", "neumann"])
@pytest.mark.parametrize("flux", ["upwind", "central", "downwind"])
def test_penalty_jump_identity(n, level, bc, flux):
    """u^T P u = sum over edges of s |g_e| / 2 [[u]]_e^2: interior edges
    (and the wrap edge when periodic) use the jump between the two cells'
    traces, a Dirichlet wall the trace itself, a Neumann wall nothing."""
    domain = (-1.0, 2.0)
    g = lambda x, t: 1.5 + np.cos(2 * np.pi * (x + 1.0) / 3.0)  # periodic on the domain
    s = dg.FLUX[flux]
    P = dg.cell_operator("penalty", n, level, domain, g, flux=flux, left=bc, right=bc)
    rng = np.random.default_rng(level * 10 + n)
    a, b = domain
    cells = 1 << level
    h = (b - a) / cells
    for _ in range(3):
        u = rng.standard_normal(n * cells)
        uL, uR = _edge_values(u, n, level, domain)
        want = 0.0
        for i in range(1, cells):
            want += s * abs(g(a + i * h, 0.0)) / 2 * (uL[i] - uR[i - 1]) ** 2
        if bc == "periodic":
            want += s * abs(g(a, 0.0)) / 2 * (uL[0] - uR[-1]) ** 2
        elif bc == "dirichlet":
            want += s * abs(g(a, 0.0)) / 2 * uL[0] ** 2 + s * abs(g(b, 0.0)) / 2 * uR[-1] ** 2
        got = u @ P @ u
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


@pytest.mark.parametrize("n", [2, 3])
def test_penalty_annihilates_continuous_functions(n):
    """A continuous piecewise-linear periodic function has no jumps, so the
    penalty maps it to zero (every row is a trace times a jump)."""
    level, domain = 4, (0.0, 1.0)
    cells = 1 << level
    h = 1.0 / cells
    nodes = np.random.default_rng(3).standard_normal(cells)  # values at x_i, periodic
    # cell i: linear from nodes[i] to nodes[i+1]; projected on the Legendre basis
    xq, wq = dg.gauss_legendre(8)
    Pq = dg.legendre_onb(xq, n)[0] / np.sqrt(h)
    u = np.zeros(n * cells)
    for i in range(cells):
        vals = nodes[i] + (nodes[(i + 1) % cells] - nodes[i]) * (xq + 1) / 2
        u[i * n:(i + 1) * n] = Pq.T @ (vals * wq * h / 2)
    P = dg.cell_operator("penalty", n, level, domain, lambda x, t: 2.0, flux="upwind")
    assert np.abs(P @ u).max() <= 1e-12
