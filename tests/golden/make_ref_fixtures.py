"""Write the reference-run fixtures ``tests/golden/ref_<case>.npz``.

Each case builds the reference from its own sources with flang
(``oracle/ref_build.py``, outputs under ``oracle/_ref/``), runs it from a
seeded initial state for a few steps, and stores

  * ``u0``     -- the initial conserved state, C shape (nz+4, ny+4, nx+4, 5)
                  in the reference's Fortran (e, i, j, k) order;
  * ``prims``  -- UPrim after 0, 1 and N steps (the reference's BIN output);
  * ``steps``  -- the step counts of those snapshots;
  * ``params`` -- the problem as JSON (fields of tlaloque_amd.params.HydroParams);
  * for the self-gravity cases (``ref_sg_*``, built with -DGRAV) also
    ``phi0`` -- the initial PHI, C shape (nz+2, ny+2, nx+2), whose ghost
                faces are the boundary values the serial build keeps;
    ``phit`` -- PHIT = PHIP + PHI after 1 and N steps (HydroCore.f90:284-290).

Run in the build container only (needs /root/reference and flang):
    python tests/golden/make_ref_fixtures.py
"""
from __future__ import annotations

import dataclasses
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from oracle import ref_build  # noqa: E402
from tlaloque_amd import params as P  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def cons_from_prims(p, rho, vel, pres):
    u = np.zeros(p.shape + (5,))
    u[..., 0] = rho
    u[..., 1:4] = rho[..., None] * vel
    u[..., 4] = 0.5 * rho * (vel ** 2).sum(-1) + pres / (p.gamma - 1.0)
    return u


def random_state(p, seed, vscale=1.0, rho_range=(0.5, 2.0), p_range=(0.3, 2.0)):
    rng = np.random.default_rng(seed)
    shp = p.shape
    rho = rng.uniform(*rho_range, shp)
    vel = vscale * rng.standard_normal(shp + (3,))
    pres = rng.uniform(*p_range, shp)
    return cons_from_prims(p, rho, vel, pres)


def cgs_state(p, seed, n_range, logT_range, vmax):
    rng = np.random.default_rng(seed)
    shp = p.shape
    n = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), shp))
    T = 10.0 ** rng.uniform(*logT_range, shp)
    rho = n * p.mu0 * P.AMU
    pres = n * P.KB * T
    vel = vmax * rng.uniform(-1, 1, shp + (3,))
    return cons_from_prims(p, rho, vel, pres)


def cases():
    out = []
    # Plain hydro, outflow, van Leer, supersonic random state (every HLLC branch).
    p = P.HydroParams(nx=10, ny=8, nz=6, bc=(P.BC_OUTFLOW,) * 6, eta=0.05)
    out.append(("hydro_outflow", p, random_state(p, 11), 5, None))
    # Mixed boundary types.
    p = P.HydroParams(nx=9, ny=7, nz=6, bc=(P.BC_REFLECTIVE, P.BC_REFLECTIVE, P.BC_PERIODIC,
                                           P.BC_PERIODIC, P.BC_OUTFLOW, P.BC_REFLECTIVE),
                      xl=0.0, xr=0.9, yl=-0.35, yr=0.35, zl=0.0, zr=1.2, gamma=1.4, cfl=0.3,
                      eta=0.005)
    out.append(("bc_mixed", p, random_state(p, 12), 4, None))
    # All periodic (conservation).
    p = P.HydroParams(nx=8, ny=8, nz=6, bc=(P.BC_PERIODIC,) * 6, eta=0.02)
    out.append(("periodic", p, random_state(p, 13, vscale=0.5), 4, None))
    # Every limiter.
    for lim in (P.LIMITER_NONE, P.LIMITER_MINMOD, P.LIMITER_UMIST, P.LIMITER_WOODWARD,
                P.LIMITER_SUPERBEE):
        p = P.HydroParams(nx=8, ny=6, nz=5, bc=(P.BC_OUTFLOW, P.BC_REFLECTIVE, P.BC_PERIODIC,
                                               P.BC_PERIODIC, P.BC_REFLECTIVE, P.BC_OUTFLOW),
                          limiter=lim, eta=0.01)
        out.append((f"limiter{lim}", p, random_state(p, 20 + lim, vscale=0.6), 3, None))
    # Cooling (cgs): dense warm gas hits the 0.5 clamp, T < 1e4 K is skipped.
    L = 3.0 * P.PC
    p = P.HydroParams(nx=8, ny=8, nz=6, xl=-L, xr=L, yl=-L, yr=L, zl=-L * 0.75, zr=L * 0.75,
                      bc=(P.BC_OUTFLOW,) * 6, eta=0.05, cooling=P.COOL_EXACT)
    out.append(("cooling", p, cgs_state(p, 31, (0.3, 80.0), (3.5, 7.5), 2e6), 4, "cool"))
    # Winds (cgs): one wind crossing the x-low face into the ghost layers, one inside.
    L = 1.0 * P.PC
    dx = 2 * L / 10
    winds = [P.Wind(xc=-L + 0.7 * dx, yc=0.1 * dx, zc=0.2 * dx, vx=1e6, vy=0.0, vz=-2e6,
                    radius=2.5 * dx, mdot=1e-6 * P.MSUN / P.YEAR, vinf=1000 * P.KMS, temp=1e4,
                    mu=1.0),
             P.Wind(xc=0.3 * L, yc=-0.35 * L, zc=0.15 * L, vx=0.0, vy=0.0, vz=0.0,
                    radius=2.0 * dx, mdot=3e-7 * P.MSUN / P.YEAR, vinf=700 * P.KMS, temp=5e3,
                    mu=1.3)]
    p = P.HydroParams(nx=10, ny=10, nz=8, xl=-L, xr=L, yl=-L, yr=L, zl=-0.8 * L, zr=0.8 * L,
                      bc=(P.BC_OUTFLOW,) * 6, eta=0.05, winds=winds)
    out.append(("winds", p, cgs_state(p, 41, (0.8, 1.2), (3.9, 4.1), 1e5), 3, None))
    # Winds and cooling together (order of mainHD.f90:82-97).
    p = p.replace(cooling=P.COOL_EXACT)
    out.append(("winds_cooling", p, cgs_state(p, 42, (0.5, 30.0), (4.0, 6.5), 3e5), 3, "cool"))
    return out


def sphere_state(p, r0, rho_in, rho_out, pres):
    """A dense sphere of radius r0 at rest in uniform pressure (cgs)."""
    x = P.xcoord(p, np.arange(-1, p.nx + 3))
    y = P.ycoord(p, np.arange(-1, p.ny + 3))
    z = P.zcoord(p, np.arange(-1, p.nz + 3))
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
    rho = np.where(r < r0, rho_in, rho_out)
    return cons_from_prims(p, rho, np.zeros(r.shape + (3,)), np.full(r.shape, pres))


def point_phi0(p, u0):
    """PHI with zero interior and -G M / r on the ghost layer (M: the mesh's
    mass at the origin)."""
    mass = u0[2:-2, 2:-2, 2:-2, 0].sum() * p.dx * p.dy * p.dz
    x = P.xcoord(p, np.arange(0, p.nx + 2))
    y = P.ycoord(p, np.arange(0, p.ny + 2))
    z = P.zcoord(p, np.arange(0, p.nz + 2))
    r = np.sqrt(x[None, None, :] ** 2 + y[None, :, None] ** 2 + z[:, None, None] ** 2)
    phi = -P.GR * mass / r
    phi[1:-1, 1:-1, 1:-1] = 0.0
    return phi


def gravity_cases():
    out = []
    # Self-gravity of a dense sphere, 16^3 (four levels to 1^3), outflow.
    L = 1.0 * P.PC
    p = P.HydroParams(nx=16, ny=16, nz=16, xl=-L, xr=L, yl=-L, yr=L, zl=-L, zr=L,
                      bc=(P.BC_OUTFLOW,) * 6, eta=0.02, self_gravity=P.SG_FIXED)
    u0 = sphere_state(p, 0.5 * L, 1e-20, 1e-22, 1e-10)
    out.append(("sg_sphere", p, u0, 4, point_phi0(p, u0)))
    # A 16x16x8 box (three levels to 2x2x1), mixed boundaries, an off-centre
    # sphere and a point mass (8 Msun, as pointsource_gravity.f90:23 fixes).
    p = P.HydroParams(nx=16, ny=16, nz=8, xl=-L, xr=L, yl=-L, yr=L, zl=-0.5 * L, zr=0.5 * L,
                      bc=(P.BC_OUTFLOW, P.BC_REFLECTIVE, P.BC_PERIODIC, P.BC_PERIODIC,
                          P.BC_OUTFLOW, P.BC_OUTFLOW),
                      eta=0.02, self_gravity=P.SG_FIXED,
                      masses=[P.PointMass(0.31 * L, -0.22 * L, 0.13 * L, 8.0 * P.MSUN)])
    u0 = sphere_state(p, 0.3 * L, 3e-20, 1e-22, 3e-10)
    rng = np.random.default_rng(51)
    u0[..., 1:4] += u0[..., :1] * 1e5 * rng.uniform(-1, 1, u0.shape[:3] + (3,))
    u0[..., 4] = 3e-10 / (p.gamma - 1.0) + 0.5 * (u0[..., 1:4] ** 2).sum(-1) / u0[..., 0]
    out.append(("sg_mass_box", p, u0, 4, point_phi0(p, u0)))
    return out


def main_gravity():
    for name, p, u0, nsteps, phi0 in gravity_cases():
        o = oracle.Oracle(p)
        o.set_phi(phi0)
        o.set_cons(u0)
        ts = [0.0]
        phis = []
        for _ in range(nsteps):
            o.advance()
            ts.append(o.t)
            phis.append(o.phi.copy())
        tfin = 0.5 * (ts[-2] + ts[-1])
        r = ref_build.build_and_run(name, p, u0, tfin, gravity=True, phi0=phi0)
        prims, phit = r["prims"], r["phit"]
        if len(prims) != nsteps + 1 or len(phit) != nsteps:
            raise RuntimeError(f"{name}: reference ran {len(prims) - 1} steps, expected {nsteps}")
        keep = [0, 1, nsteps]
        np.savez_compressed(os.path.join(OUT, f"ref_{name}.npz"), u0=u0,
                            prims=np.stack([prims[s] for s in keep]), steps=np.array(keep),
                            params=np.array(params_json(p)), phi0=phi0,
                            phit=np.stack([phit[0], phit[-1]]))
        sl = (slice(2, -2),) * 3
        d = float(np.abs(o.prim[sl] - prims[-1][sl]).max())
        print(f"{name}: {nsteps} steps, oracle-vs-reference max |diff| interior = {d:.3e}")


def params_json(p) -> str:
    d = dataclasses.asdict(p)
    return json.dumps(d)


def main():
    table = P.load_cooling_table(os.path.join(ROOT, "tlaloque_amd", "data", "Z1.0.dat"))
    for name, p, u0, nsteps, cool in cases():
        o = oracle.Oracle(p, table=table if cool else None)
        o.set_cons(u0)
        ts = [0.0]
        for _ in range(nsteps):
            o.advance()
            ts.append(o.t)
        tfin = 0.5 * (ts[-2] + ts[-1])
        r = ref_build.build_and_run(name, p, u0, tfin, cooling=bool(cool))
        prims = r["prims"]
        if len(prims) != nsteps + 1:
            raise RuntimeError(f"{name}: reference ran {len(prims) - 1} steps, expected {nsteps}")
        keep = [0, 1, nsteps]
        np.savez_compressed(os.path.join(OUT, f"ref_{name}.npz"), u0=u0,
                            prims=np.stack([prims[s] for s in keep]), steps=np.array(keep),
                            params=np.array(params_json(p)))
        sl = (slice(2, -2),) * 3
        d = max(float(np.abs(o.prim[sl] - prims[-1][sl]).max()), 0.0)
        print(f"{name}: {nsteps} steps, oracle-vs-reference max |diff| interior = {d:.3e}")


if __name__ == "__main__":
    if "--gravity" not in sys.argv:
        main()
    main_gravity()
