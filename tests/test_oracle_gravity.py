"""The oracle's self-gravity (source/sources.f90:99-897, HLLC.f90:15-24) and
how it is pinned (CPU).

  * Against the reference built with -DGRAV from its sources (fixtures
    ``ref_sg_*``, oracle/ref_build.py): the whole time step with MultiGrid
    inside Godunov, its V-cycles, relaxation order, restriction,
    prolongation, residual, coarsest solve and S_Solver with
    PHIT = PHIP + PHI, and the point-mass potential (8 Msun stars): UPrim and
    PHIT bit for bit.  The serial build never applies the multipole values
    (its neighbour ranks are never negative), so these runs keep phi's ghost
    faces as set.
  * The multipole boundary values (what the MPI build applies) by what the
    mathematics fixes: the moments against sums built from numpy's Legendre
    polynomials and the closed forms of the associated functions, and the
    boundary potential against the expansion of 1/|x - x'| by the addition
    theorem (Legendre polynomials of the angle between the points only).  As
    written, the boundary formula counts the m > 0 terms twice
    (sources.f90:705-707 multiply by 2 a moment whose factor
    2 (l-m)!/(l+m)! already holds the sum over +-m); the test pins that.
  * The solver: its result satisfies the discrete Poisson equation, and for
    a Plummer sphere it converges to the analytic potential at second order.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import load_ref_sg, sg_ref_cases
from tlaloque_amd import params as P
from tlaloque_amd import problems

CASES = sg_ref_cases()


def test_fixture_set_is_complete():
    assert {"sg_sphere", "sg_mass_box"} <= set(CASES)


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_reference_with_gravity(case):
    p, u0, prims, steps, phi0, phit = load_ref_sg(case)
    o = oracle.Oracle(p)
    o.set_phi(phi0)
    o.set_cons(u0)
    sl = (slice(1, -1),) * 3
    assert np.array_equal(o.prim[sl], prims[0][sl])
    phip = _point_potential(p)
    got_phit = []
    for s in range(1, steps[-1] + 1):
        o.advance()
        got_phit.append(o.phi + phip)
        if s in steps:
            ref = prims[steps.index(s)][sl]
            assert np.array_equal(o.prim[sl], ref), (case, s, float(np.abs(o.prim[sl] - ref).max()))
    inner = (slice(1, -1),) * 3
    assert np.array_equal(got_phit[0][inner], phit[0][inner])
    assert np.array_equal(got_phit[-1][inner], phit[1][inner])


def _point_potential(p):
    x = P.xcoord(p, np.arange(0, p.nx + 2))
    y = P.ycoord(p, np.arange(0, p.ny + 2))
    z = P.zcoord(p, np.arange(0, p.nz + 2))
    phi = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2))
    for m in p.masses:   # pointsource_gravity.f90:31-34, masses in order
        r = np.sqrt((m.x - x[None, None, :]) ** 2 + (m.y - y[None, :, None]) ** 2
                    + (m.z - z[:, None, None]) ** 2)
        phi = phi - P.GR * m.mass / r
    return phi


def _mesh(n, L=P.PC, nz=None):
    nz = nz or n
    return P.HydroParams(nx=n, ny=n, nz=nz, xl=-L, xr=L, yl=-L, yr=L, zl=-L * nz / n,
                         zr=L * nz / n, self_gravity=P.SG_FIXED, bc=(P.BC_OUTFLOW,) * 6)


def _centres(p, lo=0, hi_pad=2):
    x = P.xcoord(p, np.arange(lo, p.nx + hi_pad))
    y = P.ycoord(p, np.arange(lo, p.ny + hi_pad))
    z = P.zcoord(p, np.arange(lo, p.nz + hi_pad))
    return np.meshgrid(z, y, x, indexing="ij")[::-1]   # X, Y, Z of shape (nz, ny, nx)


def test_solution_satisfies_discrete_poisson():
    p = _mesh(16, nz=8)
    o = oracle.Oracle(p)
    rng = np.random.default_rng(3)
    rho = 1e-21 * rng.uniform(0.5, 2.0, (p.nz + 2, p.ny + 2, p.nx + 2))
    X, Y, Z = _centres(p)
    o.phi[...] = -P.GR * 1e35 / np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    o.phi[1:-1, 1:-1, 1:-1] = 0.0
    cyc = o.multigrid(rho)
    assert 1 <= cyc <= 31
    f = o.phi
    lap = (f[1:-1, 1:-1, 2:] + f[1:-1, 1:-1, :-2] + f[1:-1, 2:, 1:-1] + f[1:-1, :-2, 1:-1]
           + f[2:, 1:-1, 1:-1] + f[:-2, 1:-1, 1:-1] - 6 * f[1:-1, 1:-1, 1:-1]) / p.dx ** 2
    src = 4 * P.PI * P.GR * rho[1:-1, 1:-1, 1:-1]
    assert np.abs(lap - src).max() <= 1e-9 * np.abs(src).max()


def test_plummer_sphere_converges_at_second_order():
    # rho = 3 M / (4 pi a^3) (1 + r^2/a^2)^(-5/2), Phi = -G M / sqrt(r^2 + a^2),
    # Phi given exactly on the ghost faces
    M, a = 1e3 * P.MSUN, 0.4 * P.PC
    errs = []
    for n in (16, 32):
        p = _mesh(n)
        o = oracle.Oracle(p)
        X, Y, Z = _centres(p)
        r2 = X ** 2 + Y ** 2 + Z ** 2
        rho = 3 * M / (4 * np.pi * a ** 3) * (1 + r2 / a ** 2) ** -2.5
        exact = -P.GR * M / np.sqrt(r2 + a ** 2)
        o.phi[...] = exact
        o.phi[1:-1, 1:-1, 1:-1] = 0.0
        o.multigrid(rho)
        e = o.phi[1:-1, 1:-1, 1:-1] - exact[1:-1, 1:-1, 1:-1]
        errs.append(np.abs(e).max() / np.abs(exact).max())
    assert errs[0] < 2e-2
    assert errs[1] < 0.3 * errs[0]      # 4x per halving of dx


def _lp(l, x):
    return np.polynomial.legendre.Legendre.basis(l)(x)


def _assoc(l, m, x):
    """P_l^m with the Condon-Shortley phase, from P_l by differentiation:
    (-1)^m (1-x^2)^(m/2) d^m P_l / dx^m."""
    d = np.polynomial.legendre.Legendre.basis(l).deriv(m)
    return (-1) ** m * (1 - x ** 2) ** (m / 2) * d(x)


def _blob(p, seed):
    rng = np.random.default_rng(seed)
    X, Y, Z = _centres(p)
    c = np.array([0.2, -0.3, 0.1]) * P.PC
    r2 = (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2
    rho = 1e-20 * np.exp(-r2 / (0.15 * P.PC) ** 2) * (1 + 0.3 * rng.uniform(size=X.shape))
    rho[0, :, :] = rho[-1, :, :] = rho[:, 0, :] = rho[:, -1, :] = rho[:, :, 0] = rho[:, :, -1] = 7.0
    return rho


def test_moments_against_legendre_sums():
    p = _mesh(12, nz=8)
    o = oracle.Oracle(p)
    rho = _blob(p, 5)
    c0, cm, sm = o.moments(rho)
    X, Y, Z = (a[1:-1, 1:-1, 1:-1] for a in _centres(p))
    d = rho[1:-1, 1:-1, 1:-1] * p.dx * p.dy * p.dz      # ghosts (7.0) must not count
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    ct, fi = Z / r, np.arctan2(Y, X)
    assert c0[0] == pytest.approx(d.sum(), rel=1e-12)
    for l in range(1, 5):
        assert c0[l] == pytest.approx((_lp(l, ct) * r ** l * d).sum(), rel=1e-9, abs=1e-9 * abs(c0[l]) + 1e-300)
        for m in range(1, l + 1):
            f = 2.0 * math.factorial(l - m) / math.factorial(l + m)
            w = _assoc(l, m, ct) * r ** l * d
            scale = np.abs(w).sum() * f
            assert abs(cm[l - 1, m - 1] - f * (w * np.cos(m * fi)).sum()) <= 1e-10 * scale
            assert abs(sm[l - 1, m - 1] - f * (w * np.sin(m * fi)).sum()) <= 1e-10 * scale


def test_multipole_boundary_is_the_expansion_with_m_terms_doubled():
    p = _mesh(12, nz=8)
    o = oracle.Oracle(p)
    rho = _blob(p, 6)
    c0, cm, sm = o.moments(rho)
    phi = o.phi_boundaries(p.gconst, c0, cm, sm, np.zeros((p.nz + 2, p.ny + 2, p.nx + 2)))
    X, Y, Z = _centres(p)
    Xs, Ys, Zs = (a[1:-1, 1:-1, 1:-1].ravel() for a in (X, Y, Z))
    d = (rho[1:-1, 1:-1, 1:-1] * p.dx * p.dy * p.dz).ravel()
    rs = np.sqrt(Xs ** 2 + Ys ** 2 + Zs ** 2)
    face = np.zeros(phi.shape, bool)
    face[1:-1, 1:-1, [0, -1]] = True
    face[1:-1, [0, -1], 1:-1] = True
    face[[0, -1], 1:-1, 1:-1] = True
    full, m0 = [], []
    for x, y, z in zip(X[face], Y[face], Z[face]):
        r = np.sqrt(x * x + y * y + z * z)
        cg = (x * Xs + y * Ys + z * Zs) / (r * rs)          # cos of the angle between the points
        full.append(-P.GR * sum((d * rs ** l * _lp(l, cg)).sum() / r ** (l + 1) for l in range(5)))
        m0.append(-P.GR * sum((d * rs ** l * _lp(l, Zs / rs)).sum() * _lp(l, z / r) / r ** (l + 1)
                              for l in range(5)))
    full, m0 = np.array(full), np.array(m0)
    got = phi[face]
    scale = np.abs(full).max()
    assert np.abs(got - (2 * full - m0)).max() <= 1e-10 * scale      # as written
    assert np.abs(got - full).max() > 1e-4 * scale                    # not the plain expansion
    assert np.array_equal(phi[~face], np.zeros(phi.shape)[~face])     # faces only
    # far from the mass the l <= 4 expansion is the potential itself
    c = (X[1:-1, 1:-1, 1:-1].ravel(), Y[1:-1, 1:-1, 1:-1].ravel(), Z[1:-1, 1:-1, 1:-1].ravel())
    exact = np.array([-P.GR * (d / np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)).sum()
                      for x, y, z in zip(X[face], Y[face], Z[face])])
    assert np.abs(full - exact).max() <= 2e-2 * np.abs(exact).max()


def test_boundary_modes():
    # fixed: the ghost faces stay as set; multipole: refreshed when mod(it, 10) = 0
    p = _mesh(8)
    X, Y, Z = _centres(p)
    rho = 1e-21 * np.exp(-(X ** 2 + Y ** 2 + Z ** 2) / (0.3 * P.PC) ** 2)
    init = -1e9 * np.ones((p.nz + 2, p.ny + 2, p.nx + 2))
    for mode in (P.SG_FIXED, P.SG_MULTIPOLE):
        o = oracle.Oracle(p.replace(self_gravity=mode))
        faces = []
        for it in (0, 1, 9, 10):
            o.phi[...] = init
            o.it = it
            o.multigrid(rho)
            faces.append(o.phi[0, 1:-1, 1:-1].copy())
        if mode == P.SG_FIXED:
            assert all(np.array_equal(f, init[0, 1:-1, 1:-1]) for f in faces)
        else:
            assert not np.array_equal(faces[0], init[0, 1:-1, 1:-1])
            assert np.array_equal(faces[1], init[0, 1:-1, 1:-1])
            assert np.array_equal(faces[2], init[0, 1:-1, 1:-1])
            assert np.array_equal(faces[3], faces[0])
