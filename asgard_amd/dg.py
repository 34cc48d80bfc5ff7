"""1D discontinuous-Galerkin operators in the hierarchical multiwavelet basis.

Host-side input generator for the kronmult operator (the role of the
reference's ``generate_coefficients`` / ``generate_all_coefficients``,
src/coefficients.cpp:18-560, with the basis of src/basis.cpp and
src/quadrature.cpp).  Everything here is setup work in numpy; the GPU path
consumes the resulting matrices through :func:`asgard_amd.grid.build_vA`.

Conventions (all as in the reference):

* a dimension is [a, b] split into 2^level cells of width h; on each cell the
  n basis functions are orthonormal Legendre polynomials of degree < n,
  P_i(xi) * sqrt((2i+1)/2) / sqrt(h) with xi in [-1, 1];
* operators are built in that cell basis with rows = test functions and
  columns = trial functions:
  - ``mass``:    int g phi_j psi_i
  - ``div``:    -int g phi_j psi_i'  + numerical fluxes on the cell edges,
                 flux = {{g u}} + flux_scale * |g|/2 [[u]]
                 (flux_scale: downwind -1, central 0, upwind +1)
  - ``grad``:    the negative transpose of ``div``
  - ``penalty``: only the jump part of the flux;
  boundary conditions ``periodic``, ``dirichlet`` (no boundary flux for
  div/grad) or ``neumann`` (outflow flux);
* every partial term is multiplied by the inverse of the unit-coefficient
  mass matrix (I/2 in this normalisation), so a coefficient-less mass term
  is the identity; ``grad`` is assembled as a ``div`` with the upwind
  direction reversed and Dirichlet/Neumann swapped, then negated and
  transposed (pinned by the reference's coefficient fixtures);
* the wavelet rotation is F M F^T with F the multiwavelet transform of
  ``operator_two_scale``; its basis functions restate
  ``generate_multi_wavelets`` (Gram-Schmidt of +-x^j on the two halves of the
  reference cell against the Legendre scalets, then among themselves,
  normalised, reversed, alternating signs).
* a term is the product of its partial-term matrices (rechain_dimension).
"""
from __future__ import annotations

import numpy as np


# --------------------------------------------------------------- basics ---

def gauss_legendre(npts, a=-1.0, b=1.0):
    x, w = np.polynomial.legendre.leggauss(npts)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def legendre_onb(xi, n):
    """Orthonormal Legendre values and derivatives on [-1, 1]:
    (len(xi) x n, len(xi) x n)."""
    xi = np.atleast_1d(np.asarray(xi, dtype=np.float64))
    P = np.zeros((len(xi), max(1, n)))
    dP = np.zeros_like(P)
    for i in range(n):
        c = np.zeros(i + 1)
        c[i] = 1.0
        s = np.sqrt((2 * i + 1) / 2.0)
        P[:, i] = np.polynomial.legendre.legval(xi, c) * s
        dP[:, i] = np.polynomial.legendre.legval(xi, np.polynomial.legendre.legder(c)) * s
    return P, dP


FLUX = {"downwind": -1.0, "central": 0.0, "upwind": 1.0}


def cell_operator(kind, n, level, domain=(-1.0, 1.0), g=None, flux="central",
                  left="periodic", right="periodic", time=0.0, nquad=None):
    """One partial term in the cell (Legendre) basis, before the mass
    inversion and the wavelet rotation."""
    a, b = domain
    cells = 1 << level
    h = (b - a) / cells
    N = n * cells
    gf = (lambda x, t: 1.0) if g is None else g
    nq = nquad or max(10, n + 1)
    qx, qw = gauss_legendre(nq)
    Pq, dPq = legendre_onb(qx, n)
    Pq = Pq / np.sqrt(h)
    dPq = dPq * (2.0 / h) / np.sqrt(h)
    PL = legendre_onb([-1.0], n)[0][0] / np.sqrt(h)
    PR = legendre_onb([1.0], n)[0][0] / np.sqrt(h)
    LtR, LtL = np.outer(PL, PR), np.outer(PL, PL)
    RtR, RtL = np.outer(PR, PR), np.outer(PR, PL)
    jac = h / 2.0
    fscale = FLUX[flux] if isinstance(flux, str) else float(flux)
    C = np.zeros((N, N))

    def blk(i, j, alpha, mat):
        C[n * i:n * i + n, n * j:n * j + n] += alpha * mat

    gdiv = kind in ("div", "grad")
    for i in range(cells):
        if kind != "penalty":
            xq = (0.5 * qx + 0.5 + i) * h + a
            c = np.array([gf(x, time) for x in xq]) * qw * jac
            tmp = c[:, None] * Pq
            if kind == "mass":
                blk(i, i, 1.0, Pq.T @ tmp)
            else:
                blk(i, i, -1.0, dPq.T @ tmp)
    if kind == "mass":
        return C

    def fluxes(xl, xr):
        fl, fr = 0.5 * gf(xl, time), 0.5 * gf(xr, time)
        fla, fra = fscale * abs(fl), fscale * abs(fr)
        if kind == "penalty":
            fl = fr = 0.0
        return fl, fr, fla, fra

    for i in range(1, cells - 1):
        xl = a + i * h
        fl, fr, fla, fra = fluxes(xl, xl + h)
        blk(i, i - 1, -fl - fla, LtR)
        blk(i, i, -fl + fla, LtL)
        blk(i, i, fr + fra, RtR)
        blk(i, i + 1, fr - fra, RtL)

    last = cells - 1
    fl, fr, fla, fra = fluxes(a, a + h)
    if left == "dirichlet":
        if kind == "penalty":
            blk(0, 0, fla, LtL)
    elif left == "neumann":
        if kind != "penalty":
            blk(0, 0, -2.0 * fl, LtL)
    else:
        blk(0, last, -fl - fla, LtR)
        blk(0, 0, -fl + fla, LtL)
    if cells > 1:
        blk(0, 0, fr + fra, RtR)
        blk(0, 1, fr - fra, RtL)
        fl, fr, fla, fra = fluxes(b - h, b)
        blk(last, last - 1, -fl - fla, LtR)
        blk(last, last, -fl + fla, LtL)
    if right == "dirichlet":
        if kind == "penalty":
            blk(last, last, fra, RtR)
    elif right == "neumann":
        if kind != "penalty":
            blk(last, last, 2.0 * fr, RtR)
    else:
        blk(last, last, fr + fra, RtR)
        blk(last, 0, fr - fra, RtL)
    if kind == "grad":
        C = -C.T
    return C


def partial_term(kind, n, level, domain=(-1.0, 1.0), g=None, **kw):
    """Cell-basis partial term with the mass inversion the reference applies
    when the term has a coefficient function (src/coefficients.cpp:64-68)."""
    if kind == "grad":
        # grad is built as a div with the upwind direction reversed and the
        # Dirichlet/Neumann roles swapped, then negated and transposed
        # (partial_term::set_flux / set_bilinear_boundary, src/pde/pde_base.hpp:240-270)
        swap = {"dirichlet": "neumann", "neumann": "dirichlet", "periodic": "periodic"}
        fl = kw.get("flux", "central")
        kw["flux"] = -(FLUX[fl] if isinstance(fl, str) else float(fl))
        kw["left"] = swap[kw.get("left", "periodic")]
        kw["right"] = swap[kw.get("right", "periodic")]
    C = cell_operator(kind, n, level, domain, g, **kw)
    M = cell_operator("mass", n, level, domain, None)
    return np.linalg.solve(M, C)


# ---------------------------------------------------------- multiwavelets ---

def _poly_onb_scalets(n):
    """Rows: monomial coefficients (highest power first) of
    sqrt(2j+1) P_j(x) on [-1, 1], j < n."""
    rows = []
    for j in range(n):
        c = np.zeros(j + 1)
        c[j] = 1.0
        mono = np.polynomial.legendre.leg2poly(c)  # lowest power first
        rows.append(np.sqrt(2 * j + 1) * mono[::-1])
    out = np.zeros((n, n))
    for j, r in enumerate(rows):
        out[j, n - len(r):] = r
    return out


def multiwavelets(n):
    """(h0, h1, g0, g1): two-scale coefficients of generate_multi_wavelets."""
    scal = _poly_onb_scalets(n)
    xl, wl = gauss_legendre(n, -1.0, 0.0)
    xr, wr = gauss_legendre(n, 0.0, 1.0)
    xa, _ = gauss_legendre(n, -1.0, 1.0)
    pv = lambda row, x: np.polyval(row, x)
    ip = lambda f1, f2, w: 0.5 * np.sum(f1 * f2 * w)
    # f_j = -x^j on (-1,0), x^j on (0,1)
    phi = np.zeros((2 * n, n))
    for i in range(n):
        phi[n - i - 1, i] = -1.0
        phi[2 * n - i - 1, i] = 1.0
    # orthogonalise against the scalets
    for r in range(n):
        proj = np.zeros(n)
        for s in range(n):
            sm = ip(pv(phi[r], xl), pv(scal[s], xl), wl) + ip(pv(phi[r + n], xr), pv(scal[s], xr), wr)
            proj = proj + scal[s] * sm
        phi[r] -= proj
        phi[r + n] -= proj
    # Gram-Schmidt among themselves (classical, lower rows first)
    for r in range(1, n):
        p1 = np.zeros(n)
        p2 = np.zeros(n)
        for s in range(r):
            e1, e2 = pv(phi[r], xl), pv(phi[s], xl)
            e3, e4 = pv(phi[r + n], xr), pv(phi[s + n], xr)
            s1 = ip(e1, e2, wl) + ip(e3, e4, wr)
            s2 = ip(e2, e2, wl) + ip(e4, e4, wr)
            p1 = p1 + phi[s] * s1 / s2
            p2 = p2 + phi[s + n] * s1 / s2
        phi[r] -= p1
        phi[r + n] -= p2
    for r in range(n):
        sm = ip(pv(phi[r], xl), pv(phi[r], xl), wl) + ip(pv(phi[r + n], xr), pv(phi[r + n], xr), wr)
        phi[r] /= np.sqrt(sm)
        phi[r + n] /= np.sqrt(sm)
    # reverse the rows of each half, alternate signs
    phi = np.concatenate([phi[:n][::-1], phi[n:][::-1]])
    sign = np.array([1.0 if i % 2 == 0 else -1.0 for i in range(n)])
    phi = phi * np.concatenate([sign, sign])[:, None]
    h0 = np.zeros((n, n))
    h1 = np.zeros((n, n))
    g0 = np.zeros((n, n))
    g1 = np.zeros((n, n))
    for r in range(n):
        for c in range(n):
            e2 = pv(scal[c], xa)
            h0[r, c] = 2.0 * ip(pv(scal[r], xl), e2, wl) / np.sqrt(2.0)
            h1[r, c] = 2.0 * ip(pv(scal[r], xr), e2, wr) / np.sqrt(2.0)
            g0[r, c] = 2.0 * ip(pv(phi[r], xl), e2, wl) / np.sqrt(2.0)
            g1[r, c] = 2.0 * ip(pv(phi[r + n], xr), e2, wr) / np.sqrt(2.0)
    for m in (h0, h1, g0, g1):
        m[np.abs(m) < 1e-12] = 0.0
    return h0, h1, g0, g1


def wavelet_transform(n, level):
    """The multiwavelet transform F (n*2^level square) of operator_two_scale."""
    if level < 1:
        return np.eye(n)
    h0, h1, g0, g1 = multiwavelets(n)
    cells = 1 << level
    N = n * cells
    hb = np.concatenate([h0, h1], axis=1)
    gb = np.concatenate([g0, g1], axis=1)
    fmwt = np.zeros((N, N))
    for i in range(cells // 2):
        fmwt[n * i:n * i + n, 2 * n * i:2 * n * i + 2 * n] = hb
        fmwt[n * (i + cells // 2):n * (i + cells // 2) + n, 2 * n * i:2 * n * i + 2 * n] = gb
    comp = np.eye(N)
    for j in range(1, level + 1):
        if j == 1:
            cf = fmwt
        else:
            cn = (1 << (level - j + 1)) * n
            cf = np.zeros((N, N))
            cf[cn:, cn:] = np.eye(N - cn)
            cf[:cn // 2, :cn] = fmwt[:cn // 2, :cn]
            cf[cn // 2:cn, :cn] = fmwt[N // 2:N // 2 + cn // 2, :cn]
        comp = cf @ comp
    comp[np.abs(comp) < 1e-12] = 0.0
    return comp


def rotate(C, n, level):
    F = wavelet_transform(n, level)
    return F @ C @ F.T


# ------------------------------------------------------ operator recipes ---

def term_matrix(partials, n, level, domain=(-1.0, 1.0), rotated=True):
    """Product of the partial-term matrices (rechain_dimension), rotated to
    the wavelet basis.  partials: list of dicts for :func:`partial_term`."""
    mat = None
    for p in partials:
        kw = dict(p)
        kind = kw.pop("kind")
        C = partial_term(kind, n, level, kw.pop("domain", domain), **kw)
        mat = C if mat is None else mat @ C
    return rotate(mat, n, level) if rotated else mat


def identity(n, level):
    return np.eye(n * (1 << level))


def advection(n, level, speed=-1.0, domain=(-1.0, 1.0)):
    """speed * d/dx as a right-hand-side operator (df/dt = A f), periodic,
    upwinded: the flux scale takes the sign of the coefficient (downwind for
    a negative coefficient, as the reference's continuity PDEs use), which
    makes the spectrum dissipative."""
    flux = "downwind" if speed < 0 else "upwind"
    return term_matrix([dict(kind="div", g=lambda x, t: speed, flux=flux)], n, level, domain)


def diffusion(n, level, coeff=1.0, domain=(-1.0, 1.0)):
    """coeff * d^2/dx^2 as grad(div): two partial terms, central fluxes."""
    return term_matrix([dict(kind="div", g=lambda x, t: 1.0, flux="central"),
                        dict(kind="grad", g=lambda x, t: coeff, flux="central")],
                       n, level, domain)


def multiply(n, level, func, domain=(-1.0, 1.0)):
    """Multiplication by func(x) (mass term with a coefficient)."""
    return term_matrix([dict(kind="mass", g=func)], n, level, domain)


def operators_advection(d, n, level, speeds=None):
    """d terms: term k = advection in dimension k, identity elsewhere."""
    speeds = speeds or [-1.0] * d
    adv = [advection(n, level, s) for s in speeds]
    I = identity(n, level)
    return [adv[t] if k == t else I for t in range(d) for k in range(d)]


def operators_advection_diffusion(d, n, level, n_adv, nu=0.1):
    """d terms: advection in dims < n_adv, diffusion (nu) in the others."""
    I = identity(n, level)
    ops = []
    for t in range(d):
        A = advection(n, level) if t < n_adv else diffusion(n, level, nu)
        ops += [A if k == t else I for k in range(d)]
    return ops


def operators_vlasov(n, level, efield=0.5):
    """3x + 3v Vlasov-type advection (d = 6, 6 terms): -v_i d/dx_i f and
    -E_i(x_i) d/dv_i f on [-1, 1]^6, periodic; the transport speeds change
    sign over the domain, so the fluxes are central (skew spectrum)."""
    I = identity(n, level)
    dx = term_matrix([dict(kind="div", g=lambda x, t: -1.0, flux="central")], n, level)
    vmul = multiply(n, level, lambda v, t: v)
    dv = term_matrix([dict(kind="div", g=lambda x, t: -1.0, flux="central")], n, level)
    emul = multiply(n, level, lambda x, t: efield * np.cos(np.pi * x))
    ops = []
    for i in range(3):  # v_i d/dx_i
        ops += [dx if k == i else (vmul if k == 3 + i else I) for k in range(6)]
    for i in range(3):  # E_i(x_i) d/dv_i
        ops += [emul if k == i else (dv if k == 3 + i else I) for k in range(6)]
    return ops
