"""GPU generation of 1D DG operators (akm_dg_term) against the host generator.

The host generator (asgard_amd/dg.py) is pinned against the reference's
coefficient fixtures (tests/test_dg_operators.py).  Bar: the GPU operator in
the multiwavelet basis agrees with dg.term_matrix to 1e-12 of max(1, |A|)
(dense products in another order), for mass, div (every flux), grad, penalty
and div-grad chains with constant and variable coefficients, periodic,
Dirichlet and Neumann boundaries, n = 2, 3 and levels 0-5; a
sparse-grid handle updated from GPU-generated operators applies like one
built from the host operators.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

import asgard_amd  # noqa: E402
import oracle  # noqa: E402
from asgard_amd import dg  # noqa: E402
from tests.kron_cases import sparse_grid_case  # noqa: E402

FLUX = {"downwind": -1.0, "central": 0.0, "upwind": 1.0}
KIND = {"mass": asgard_amd.DG_MASS, "div": asgard_amd.DG_DIV, "grad": asgard_amd.DG_GRAD,
        "penalty": asgard_amd.DG_PENALTY}
BC = {"periodic": asgard_amd.DG_PERIODIC, "dirichlet": asgard_amd.DG_DIRICHLET,
      "neumann": asgard_amd.DG_NEUMANN}


@pytest.fixture(scope="module", autouse=True)
def gpu_present():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def cm(a):
    """Device tensor holding a host matrix column-major."""
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).cuda()


def gpu_term(parts, n, level, domain=(-1.0, 1.0)):
    pts, edges, _ = asgard_amd.dg_quadrature(n, level, domain)
    args = []
    for p in parts:
        g = p["g"]
        gq = torch.from_numpy(np.vectorize(lambda x: g(x, 0.0))(pts).astype(np.float64)).cuda()
        ge = torch.from_numpy(np.array([g(x, 0.0) for x in edges], dtype=np.float64)).cuda()
        args.append((KIND[p["kind"]], FLUX.get(p.get("flux", "central"), 0.0), gq, ge,
                     BC[p.get("left", "periodic")], BC[p.get("right", "periodic")]))
    F = cm(dg.wavelet_transform(n, level))
    out = asgard_amd.dg_term(n, level, args, F, domain)
    torch.cuda.synchronize()
    return out.cpu().numpy().T  # back to a row-major host matrix


CASES = {
    "advection_downwind": [dict(kind="div", g=lambda x, t: -1.0, flux="downwind")],
    "advection_upwind_variable": [dict(kind="div", g=lambda x, t: np.sin(x) + 0.3,
                                       flux="upwind")],
    "central_div": [dict(kind="div", g=lambda x, t: -1.0, flux="central")],
    "diffusion": [dict(kind="div", g=lambda x, t: 1.0, flux="central"),
                  dict(kind="grad", g=lambda x, t: 0.1, flux="central")],
    "field_mass": [dict(kind="mass", g=lambda x, t: 0.5 * np.cos(np.pi * x))],
    "grad_upwind": [dict(kind="grad", g=lambda x, t: 2.0 + np.cos(x), flux="upwind")],
    "div_dirichlet": [dict(kind="div", g=lambda x, t: -1.0, flux="upwind", left="dirichlet",
                           right="dirichlet")],
    "div_neumann_mixed": [dict(kind="div", g=lambda x, t: 1.0 + 0.5 * x, flux="central",
                               left="neumann", right="dirichlet")],
    "diffusion_dirichlet": [dict(kind="div", g=lambda x, t: 1.0, flux="central",
                                 left="dirichlet", right="dirichlet"),
                            dict(kind="grad", g=lambda x, t: 0.1, flux="central",
                                 left="dirichlet", right="dirichlet")],
    "penalty_dirichlet": [dict(kind="penalty", g=lambda x, t: 1.0, flux="upwind",
                               left="dirichlet", right="neumann")],
    "penalty_periodic": [dict(kind="penalty", g=lambda x, t: 2.0, flux="upwind")],
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("n,level", [(2, 0), (2, 3), (2, 5), (3, 0), (3, 2), (3, 4)])
def test_dg_term_matches_host(name, n, level):
    parts = CASES[name]
    got = gpu_term(parts, n, level)
    want = dg.term_matrix([dict(p) for p in parts], n, level)
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_dg_term_other_domain():
    parts = [dict(kind="div", g=lambda x, t: x, flux="upwind")]
    got = gpu_term(parts, 2, 4, domain=(0.0, 3.0))
    want = dg.term_matrix([dict(p) for p in parts], 2, 4, domain=(0.0, 3.0))
    assert np.abs(got - want).max() <= 1e-12 * max(1.0, np.abs(want).max())


def test_generated_operators_drive_the_apply():
    d, n, L = 3, 2, 4
    T = d
    adv = CASES["advection_downwind"]
    dif = CASES["diffusion"]
    host_ops, dev_ops = [], []
    ident = np.eye(n << L)
    for t in range(T):
        for k in range(d):
            if k == t:
                parts = adv if t < 2 else dif
                host_ops.append(dg.term_matrix([dict(p) for p in parts], n, L))
                dev_ops.append(torch.from_numpy(np.ascontiguousarray(
                    gpu_term(parts, n, L).T)).cuda())
            else:
                host_ops.append(ident)
                dev_ops.append(cm(ident))
    c = sparse_grid_case(d, n, L, T, seed=41)
    m = asgard_amd.KronmultMatrix.sparse_grid(d, n, c["table"], L, T, c["vA"])
    m.update_operators(dev_ops)
    vA = oracle.build_vA(d, n, T, L, host_ops)
    ref = np.zeros_like(c["x"])
    oracle.kronmult_sparse(d, n, c["pntr"], c["indx"], T, c["iA"], vA, 1.0, c["x"], 0.0, ref)
    y = torch.zeros(ref.size, dtype=torch.float64, device="cuda")
    m.apply(1.0, torch.from_numpy(c["x"]).cuda(), 0.0, y)
    assert np.abs(y.cpu().numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_dg_term_rejects_bad_input():
    F = cm(dg.wavelet_transform(2, 2))
    gq = torch.zeros((4, 10), dtype=torch.float64, device="cuda")
    with pytest.raises(asgard_amd.AkmInvalid):
        asgard_amd.dg_term(2, 2, [(7, 0.0, gq, None)], F)  # unknown kind
    with pytest.raises(asgard_amd.AkmInvalid):
        asgard_amd.dg_term(2, 2, [(asgard_amd.DG_DIV, 0.0, gq, None)], F)  # div needs edges
    with pytest.raises(asgard_amd.AkmInvalid):
        asgard_amd.dg_term(2, 2, [(asgard_amd.DG_MASS, 0.0, gq, None)], cm(np.eye(4)))
