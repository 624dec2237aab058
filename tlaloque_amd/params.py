"""Run-time problem parameters of the HydroCore path.

The reference fixes these at compile time in ``user/parameters.f90:10-49``
(boundary ids :10-15, limiter :18, mesh :23-25, domain :37-42, gamma :43,
mu0 :44, cfl :46, eta :47, tfin :48, cooling_limit :49).  Here they are a
run-time record passed through the C-ABI (``include/tlq.h``: ``tlq_params``),
so one build serves every configuration.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

# source/constants.f90:6-15 (cgs)
AMU = 1.660538782e-24
KB = 1.380650400e-16
GR = 6.67430e-8
PC = 3.085677588e18
AU = 1.495978707e13
YEAR = 3.155673600e7
KYR = 3.155673600e10
MSUN = 1.988920000e33
KMS = 1.0e5
PI = 3.14159265358979

# source/constants.f90:20-23
BC_OUTFLOW, BC_REFLECTIVE, BC_PERIODIC, BC_INFLOW = 1, 2, 3, 4
# source/constants.f90:26-31
LIMITER_NONE, LIMITER_VANLEER, LIMITER_MINMOD, LIMITER_UMIST, LIMITER_WOODWARD, LIMITER_SUPERBEE = (
    1, 2, 3, 4, 5, 6)

SOLVER_HLLC, SOLVER_HLL = 0, 1
PREC_FP64, PREC_FP32 = 0, 1          # FP32: FP32 storage and sweeps, FP64 guards
COOL_OFF, COOL_EXACT, COOL_SURROGATE, COOL_SUBCYCLED, COOL_INTEGRATED = 0, 1, 2, 3, 4
# Self-gravity (source/sources.f90:99-205): off; phi's ghost faces kept as set
# (the serial build); multipole values on them every 10 steps (the MPI build
# at the domain edge).
SG_OFF, SG_FIXED, SG_MULTIPOLE = 0, 1, 2


@dataclass
class Wind:
    """``spherical_wind_type`` (source/winds.f90:5-13)."""
    xc: float
    yc: float
    zc: float
    vx: float
    vy: float
    vz: float
    radius: float
    mdot: float
    vinf: float
    temp: float
    mu: float

    def as_tuple(self) -> Tuple[float, ...]:
        return dataclasses.astuple(self)


@dataclass
class PointMass:
    """A point mass of the point-source potential (source/pointsource_gravity.f90:12-39)."""
    x: float
    y: float
    z: float
    mass: float


@dataclass
class HydroParams:
    nx: int
    ny: int
    nz: int
    xl: float = -1.0
    xr: float = 1.0
    yl: float = -1.0
    yr: float = 1.0
    zl: float = -1.0
    zr: float = 1.0
    gamma: float = 5.0 / 3.0
    mu0: float = 1.0
    cfl: float = 0.4
    eta: float = 0.5e-2
    cooling_limit: float = 0.5
    limiter: int = LIMITER_VANLEER
    # LB, RB, DB, TB, FB, BB (x low, x high, y low, y high, z low, z high)
    bc: Tuple[int, int, int, int, int, int] = (BC_REFLECTIVE,) * 6
    solver: int = SOLVER_HLLC
    precision: int = PREC_FP64
    cooling: int = COOL_OFF
    winds: List[Wind] = field(default_factory=list)
    masses: List[PointMass] = field(default_factory=list)
    softening: float = 0.0
    self_gravity: int = SG_OFF
    mg_levels: int = 0          # lvlm (user/parameters.f90:94); 0: as many as the mesh halves
    gconst: float = GR          # Gconst (user/parameters.f90:45)
    mg_tol: float = 0.0         # V-cycles while Dif > mg_tol; 0: the reference's 10 (sources.f90:140)

    @property
    def dx(self) -> float:
        return (self.xr - self.xl) / self.nx

    @property
    def dy(self) -> float:
        return (self.yr - self.yl) / self.ny

    @property
    def dz(self) -> float:
        return (self.zr - self.zl) / self.nz

    @property
    def shape(self) -> Tuple[int, int, int]:
        """(nz+4, ny+4, nx+4): numpy shape of one field with two ghost layers."""
        return (self.nz + 4, self.ny + 4, self.nx + 4)

    @property
    def zones(self) -> int:
        return self.nx * self.ny * self.nz

    def replace(self, **kw) -> "HydroParams":
        return dataclasses.replace(self, **kw)


def mg_levels(p: HydroParams) -> int:
    """Multigrid levels below the mesh: p.mg_levels, or the most halvings
    that keep nx, ny, nz whole (nx = mpix 2^size with lvlm = size,
    user/parameters.f90:26,94)."""
    if p.mg_levels > 0:
        return int(p.mg_levels)
    L = 0
    while all(n % (2 << L) == 0 for n in (p.nx, p.ny, p.nz)):
        L += 1
    return L


def xcoord(p: HydroParams, i):
    """Cell-centre x of index i, source/HydroCore.f90:35-47 (single rank)."""
    return (p.xl + p.dx * i) - p.dx / 2.0


def ycoord(p: HydroParams, j):
    return (p.yl + p.dy * j) - p.dy / 2.0


def zcoord(p: HydroParams, k):
    return (p.zl + p.dz * k) - p.dz / 2.0


def load_cooling_table(path: str) -> Tuple[Sequence[float], Sequence[float]]:
    """``loaddata`` (source/cooling.f90:21-45): first line the row count, then
    rows ``T  -log10(Lambda)``; returns (T, log10 Lambda)."""
    with open(path) as fh:
        n = int(fh.readline().split()[0])
        temps, logl = [], []
        for _ in range(n):
            a, b = fh.readline().split()[:2]
            temps.append(float(a))
            logl.append(-float(b))
    return temps, logl
