"""Initial states of the benchmark and test problems.

Each builder returns ``(params, u0)`` with ``u0`` the conserved state in the
reference's Fortran layout (numpy C shape (nz+4, ny+4, nx+4, 5), ghosts
included), the form ``user_initconds`` fills (user/user.f90:22-52).

  sod_planar   3D planar shock tube along x, 128x32x32, gamma 1.4, FP64, HLLC,
               outflow, CFL 0.4, t = 0.2 (exact solution: oracle/riemann.py).
  sod_diagonal the shock tube of the reference's test harness
               (tests_utilities.jl:214-224, 407-445): the cube [-0.25, 0.25]^3,
               rho, p = 4, 1 where x + y + z < 0 and 1, 0.1795 elsewhere,
               gamma 5/3, CFL 0.8, eta 0.05, outflow, t = 0.12 (exact solution
               solutions/sod_e1rpex.out along the diagonal).
  sedov        the point explosion of the harness (tests_utilities.jl:225-234,
               455-495): the cube [-5, 5]^3, rho = 1, E0 = 1e5 as uniform
               pressure in r < 0.15, p = 1.5e-20 (gamma - 1) elsewhere,
               gamma 5/3, CFL 0.8, eta 0.05, outflow, t = 0.1.
  blast        point explosion, gamma 5/3, rho = 1, p = 1e-5, E = 1 spread
               uniformly over the cells within 3 cells of the centre, outflow.
  wind         one stellar wind (radius 5 cells, 1000 km/s, 1e-6 Msun/yr, wind
               T = 1e4 K, mu = 1) at the box centre (a cell corner for even n,
               so no cell sits at r = 0) in a uniform n = 1 cm^-3, T = 1e4 K
               medium (n_amb, t_amb), cgs, tabulated cooling.
  clumpy       lognormal density (mean n = 1 cm^-3, sigma_ln rho = 1) from a
               seeded Gaussian random field with P(k) ~ k^-4, velocities of the
               same spectrum at RMS Mach 2, T = 1e4 K, a central 1e4 Msun point
               mass softened over 2 cells, periodic boundaries, cooling.
  cluster      8 winds at seeded positions within 0.2 box of the centre
               (v_w in U[500, 2000] km/s, Mdot log-uniform in [1e-7, 1e-5]
               Msun/yr), n = 10 cm^-3, T = 100 K, point-mass gravity, cooling.

Sizes are arguments so that tests run the same set-ups small.
"""
from __future__ import annotations

from typing import Tuple

import numpy as np

from . import params as P


def cons_from_prims(p: P.HydroParams, rho, vel, pres) -> np.ndarray:
    u = np.empty(p.shape + (5,), dtype=np.float64)
    u[..., 0] = rho
    u[..., 1:4] = rho[..., None] * vel
    u[..., 4] = 0.5 * rho * (vel ** 2).sum(-1) + pres / (p.gamma - 1.0)
    return u


def uniform_state(p: P.HydroParams, rho: float, pres: float) -> np.ndarray:
    """A uniform medium at rest, built in place (no temporaries of the state's
    size: the 1024^3 state alone is 43 GB)."""
    u = np.empty(p.shape + (5,), dtype=np.float64)
    u[..., 0] = rho
    u[..., 1:4] = 0.0
    u[..., 4] = pres / (p.gamma - 1.0)
    return u


def centres(p: P.HydroParams):
    """Cell-centre coordinates of indices -1..n+2 (source/HydroCore.f90:35-75)."""
    x = P.xcoord(p, np.arange(-1, p.nx + 3))
    y = P.ycoord(p, np.arange(-1, p.ny + 3))
    z = P.zcoord(p, np.arange(-1, p.nz + 3))
    return x, y, z


def sod_planar(nx=128, ny=32, nz=32, precision=P.PREC_FP64) -> Tuple[P.HydroParams, np.ndarray]:
    L = 1.0
    p = P.HydroParams(nx=nx, ny=ny, nz=nz, xl=0.0, xr=L, yl=0.0, yr=L * ny / nx, zl=0.0,
                      zr=L * nz / nx, gamma=1.4, cfl=0.4, bc=(P.BC_OUTFLOW,) * 6,
                      precision=precision)
    x, _, _ = centres(p)
    left = (x < 0.5)[None, None, :] * np.ones(p.shape, dtype=bool)
    rho = np.where(left, 1.0, 0.125)
    pres = np.where(left, 1.0, 0.1)
    vel = np.zeros(p.shape + (3,))
    return p, cons_from_prims(p, rho, vel, pres)


SOD_TFIN = 0.2


def sod_diagonal(n=150, precision=P.PREC_FP64) -> Tuple[P.HydroParams, np.ndarray]:
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-0.25, xr=0.25, yl=-0.25, yr=0.25, zl=-0.25,
                      zr=0.25, gamma=5.0 / 3.0, cfl=0.8, eta=0.05, bc=(P.BC_OUTFLOW,) * 6,
                      precision=precision)
    x, y, z = centres(p)
    left = (x[None, None, :] + y[None, :, None] + z[:, None, None]) < 0.0
    rho = np.where(left, 4.0, 1.0)
    pres = np.where(left, 1.0, 0.1795)
    return p, cons_from_prims(p, rho, np.zeros(p.shape + (3,)), pres)


SOD_DIAGONAL_TFIN = 0.12


def sedov(n=200, precision=P.PREC_FP64) -> Tuple[P.HydroParams, np.ndarray]:
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-5.0, xr=5.0, yl=-5.0, yr=5.0, zl=-5.0, zr=5.0,
                      gamma=5.0 / 3.0, cfl=0.8, eta=0.05, bc=(P.BC_OUTFLOW,) * 6,
                      precision=precision)
    x, y, z = centres(p)
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
    pin = (1e5 / ((4.0 / 3.0) * P.PI * 0.15 ** 3)) * (p.gamma - 1.0)
    pres = np.where(r < 0.15, pin, 1.5e-20 * (p.gamma - 1.0))
    return p, cons_from_prims(p, np.ones(p.shape), np.zeros(p.shape + (3,)), pres)


SEDOV_TFIN = 0.1


def blast(n=256, precision=P.PREC_FP64, solver=P.SOLVER_HLLC, energy=1.0, rho0=1.0, p0=1e-5,
          r_cells=3.0) -> Tuple[P.HydroParams, np.ndarray]:
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-0.5, xr=0.5, yl=-0.5, yr=0.5, zl=-0.5, zr=0.5,
                      gamma=5.0 / 3.0, cfl=0.4, bc=(P.BC_OUTFLOW,) * 6, precision=precision,
                      solver=solver)
    x, y, z = centres(p)
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
    inside = r < r_cells * p.dx
    ncells = int(inside[2:-2, 2:-2, 2:-2].sum())
    vol = ncells * p.dx * p.dy * p.dz
    rho = np.full(p.shape, rho0)
    pres = np.where(inside, (p.gamma - 1.0) * energy / vol, p0)
    return p, cons_from_prims(p, rho, np.zeros(p.shape + (3,)), pres)


def blast_time(radius=0.4, energy=1.0, rho0=1.0, xi0=1.151666417967) -> float:
    """Time at which the Sedov shock reaches ``radius``: R = xi0 (E t^2/rho)^(1/5)."""
    return float((radius / xi0) ** 2.5 * np.sqrt(rho0 / energy))


def _grf(n: int, slope: float, rng: np.random.Generator, dtype=np.float32) -> np.ndarray:
    """Zero-mean unit-variance periodic Gaussian random field with P(k) ~ k^slope."""
    kx = np.fft.fftfreq(n).astype(dtype)
    kz = np.fft.rfftfreq(n).astype(dtype)
    k2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    amp = k2 ** (slope / 4.0)  # sqrt(P) with P ~ k^slope
    amp[0, 0, 0] = 0.0
    noise = rng.standard_normal((n, n, n), dtype=dtype)
    try:
        import scipy.fft as sfft
        f = sfft.rfftn(noise, workers=-1)
        f *= amp
        g = sfft.irfftn(f, s=(n, n, n), workers=-1)
    except ImportError:  # pragma: no cover
        f = np.fft.rfftn(noise) * amp
        g = np.fft.irfftn(f, s=(n, n, n))
    g = g.astype(dtype)
    g -= g.mean()
    g /= g.std()
    return g


def _pad_periodic(a: np.ndarray) -> np.ndarray:
    """Interior field (n, n, n) -> (n+4)^3 with two periodic ghost layers."""
    return np.pad(a, 2, mode="wrap")


def clumpy(n=512, precision=P.PREC_FP32, seed=2024, box_pc=10.0, cooling=P.COOL_EXACT,
           mach=2.0, sigma=1.0, temp=1e4, mass_msun=1e4, soft_cells=2.0):
    L = box_pc * P.PC
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-L / 2, xr=L / 2, yl=-L / 2, yr=L / 2, zl=-L / 2,
                      zr=L / 2, gamma=5.0 / 3.0, cfl=0.4, bc=(P.BC_PERIODIC,) * 6,
                      precision=precision, cooling=cooling,
                      masses=[P.PointMass(0.0, 0.0, 0.0, mass_msun * P.MSUN)],
                      softening=soft_cells * L / n)
    rng = np.random.default_rng(seed)
    g = _grf(n, -4.0, rng)
    nd = np.exp(sigma * g.astype(np.float64))
    nd /= nd.mean()
    cs = np.sqrt(p.gamma * P.KB * temp / (p.mu0 * P.AMU))
    vel = np.empty((n, n, n, 3), dtype=np.float32)
    for c in range(3):
        vel[..., c] = _grf(n, -4.0, rng)
    vel *= np.float32(mach * cs / np.sqrt((vel.astype(np.float64) ** 2).sum(-1).mean()))
    rho = _pad_periodic(nd * p.mu0 * P.AMU)
    pres = _pad_periodic(nd * P.KB * temp)
    v = np.stack([_pad_periodic(vel[..., c].astype(np.float64)) for c in range(3)], axis=-1)
    return p, cons_from_prims(p, rho, v, pres)


def wind(n=256, precision=P.PREC_FP32, box_pc=1.0, cooling=P.COOL_EXACT, temp_wind=1e4,
         n_amb=1.0, t_amb=1e4):
    L = box_pc * P.PC
    dx = L / n
    # the wind state is singular at r = 0 (winds.f90:72): the centre must not
    # be a cell centre -- the box centre is a cell corner for even n
    c = 0.0 if n % 2 == 0 else 0.5 * dx
    w = P.Wind(xc=c, yc=c, zc=c, vx=0.0, vy=0.0, vz=0.0, radius=5.0 * dx,
               mdot=1e-6 * P.MSUN / P.YEAR, vinf=1000.0 * P.KMS, temp=temp_wind, mu=1.0)
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-L / 2, xr=L / 2, yl=-L / 2, yr=L / 2, zl=-L / 2,
                      zr=L / 2, gamma=5.0 / 3.0, cfl=0.4, bc=(P.BC_OUTFLOW,) * 6,
                      precision=precision, cooling=cooling, winds=[w])
    return p, uniform_state(p, n_amb * p.mu0 * P.AMU, n_amb * P.KB * t_amb)


def cluster(n=1024, precision=P.PREC_FP32, seed=7, box_pc=4.0, cooling=P.COOL_SURROGATE,
            nstars=8, star_mass_msun=8.0):
    L = box_pc * P.PC
    dx = L / n
    rng = np.random.default_rng(seed)
    winds, masses = [], []
    for _ in range(nstars):
        while True:
            pos = rng.uniform(-0.2 * L, 0.2 * L, 3)
            if np.linalg.norm(pos) <= 0.2 * L:
                break
        pos = (np.floor(pos / dx) + 0.5) * dx + 0.25 * dx  # off cell centres (r > 0)
        v = rng.uniform(500.0, 2000.0) * P.KMS
        mdot = 10.0 ** rng.uniform(-7.0, -5.0) * P.MSUN / P.YEAR
        winds.append(P.Wind(xc=pos[0], yc=pos[1], zc=pos[2], vx=0.0, vy=0.0, vz=0.0,
                            radius=5.0 * dx, mdot=mdot, vinf=v, temp=1e4, mu=1.0))
        masses.append(P.PointMass(pos[0], pos[1], pos[2], star_mass_msun * P.MSUN))
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-L / 2, xr=L / 2, yl=-L / 2, yr=L / 2, zl=-L / 2,
                      zr=L / 2, gamma=5.0 / 3.0, cfl=0.4, bc=(P.BC_OUTFLOW,) * 6,
                      precision=precision, cooling=cooling, winds=winds, masses=masses,
                      softening=2.0 * dx)
    return p, uniform_state(p, 10.0 * p.mu0 * P.AMU, 10.0 * P.KB * 100.0)


# ---------------------------------------------------------------------------
# The self-gravity problems of the reference's harness (tests_utilities.jl),
# at a chosen resolution.  Each returns (params, U, PHI0); PHI0 has shape
# (nz+2, ny+2, nx+2).

def _centres3(p: P.HydroParams, lo: int = -1, pad: int = 3):
    x = P.xcoord(p, np.arange(lo, p.nx + pad))
    y = P.ycoord(p, np.arange(lo, p.ny + pad))
    z = P.zcoord(p, np.arange(lo, p.nz + pad))
    return np.meshgrid(z, y, x, indexing="ij")[::-1]


def grav_sphere(n=150, precision=P.PREC_FP64, mode=P.SG_FIXED):
    """"Gravitational Potential Accuracy Test (Multigrid)" (tests_utilities.jl:
    247-256, 548-596): rho = 1 inside r < 0.25 of the unit box, 0 outside,
    p = 1.5e-5 (gamma-1), outflow, cgs G.  PHI0: zero inside, the exact
    potential of the mesh's mass, -G M / r, on the ghost layer (the harness
    starts PHI from pointmass_potential, which with no stars leaves it 0)."""
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-0.5, xr=0.5, yl=-0.5, yr=0.5, zl=-0.5, zr=0.5,
                      gamma=5.0 / 3.0, cfl=0.8, eta=0.5e-2, bc=(P.BC_OUTFLOW,) * 6,
                      precision=precision, self_gravity=mode)
    X, Y, Z = _centres3(p)
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    rho = np.where(r > 0.25, 0.0, 1.0)
    pres = np.full(r.shape, 1.5e-5 * (p.gamma - 1.0))
    u = cons_from_prims(p, rho, np.zeros(r.shape + (3,)), pres)
    mass = rho[2:-2, 2:-2, 2:-2].sum() * p.dx * p.dy * p.dz
    phi0 = -P.GR * mass / r[1:-1, 1:-1, 1:-1]
    phi0[1:-1, 1:-1, 1:-1] = 0.0
    return p, u, phi0


def evrard(n=128, precision=P.PREC_FP64, mode=P.SG_MULTIPOLE):
    """"Evrard Collapse" (tests_utilities.jl:258-267, 597-646): G = 1,
    M = R = 1, rho = M / (2 pi R^2 r) inside R with specific internal
    energy 0.05 (U(5) = 0.05 rho), 1e-4 outside, at rest, reflective box
    [-2, 2]^3, PHI0 = 0.  (The harness's serial build would keep PHI's faces
    at 0; ``mode`` defaults to the MPI build's multipole faces.)"""
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-2.0, xr=2.0, yl=-2.0, yr=2.0, zl=-2.0, zr=2.0,
                      gamma=5.0 / 3.0, cfl=0.8, eta=0.5e-2, bc=(P.BC_REFLECTIVE,) * 6,
                      precision=precision, self_gravity=mode, gconst=1.0)
    X, Y, Z = _centres3(p)
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    inside = r <= 1.0
    u = np.zeros(p.shape + (5,))
    u[..., 0] = np.where(inside, 1.0 / (2.0 * np.pi * r), 1e-4)
    u[..., 4] = np.where(inside, 0.05 / (2.0 * np.pi * r), 1e-4 * 0.05)
    return p, u, np.zeros((p.nz + 2, p.ny + 2, p.nx + 2))


def truelove(n=128, precision=P.PREC_FP64, mode=P.SG_MULTIPOLE):
    """"Truelove Collapse Test" (tests_utilities.jl:269-278, 648-697): a
    sphere of rho = 1e-15 g/cm^3 and radius 7.8e15 cm in gas 100x thinner at
    the same pressure rho_out cs^2, cs = 0.167 km/s, gamma = 1.001, outflow
    box of +-535 AU, cgs G, PHI0 = 0."""
    L = 535.0 * P.AU
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-L, xr=L, yl=-L, yr=L, zl=-L, zr=L, gamma=1.001,
                      cfl=0.8, eta=0.5e-2, bc=(P.BC_OUTFLOW,) * 6, precision=precision,
                      self_gravity=mode)
    X, Y, Z = _centres3(p)
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    rho_out, cs = 0.01 * 1.0e-15, 0.167 * P.KMS
    pres = rho_out * cs ** 2
    u = np.zeros(p.shape + (5,))
    u[..., 0] = np.where(r > 7.8e15, rho_out, 1.0e-15)
    u[..., 4] = pres / (p.gamma - 1.0)
    return p, u, np.zeros((p.nz + 2, p.ny + 2, p.nx + 2))
