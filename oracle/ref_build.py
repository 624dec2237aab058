"""Build and run the reference (LAlbertoA/tlaloque, Fortran 90) from its own
sources with flang, for one small problem, writing only under oracle/_ref/.

TEST INFRASTRUCTURE ONLY (pins the C oracle; see oracle/__init__.py).

The reference fixes its problem at compile time in two user files,
``user/parameters.f90`` (the parameter module) and ``user/user.f90`` (initial
conditions and boundary hooks); its own test harness regenerates both for
every test (tests_utilities.jl:286-398 and :704-795) and compiles the rest of
the tree unchanged.  This script does the same: it writes a parameters module
and a user module for the problem, compiles every other file straight from
``/root/reference/source`` (constants, globals, cooling, pointsource_gravity,
sources, winds, HydroCore, init, boundaries, HLLC, mainHD -- the list of
Makefile:15-30) with ``-O2 -cpp -DDOUBLEP -fdefault-real-8`` (Makefile:46-54),
runs the program with BIN output (source/HydroCore.f90:277-283: UPrim with
ghosts, little-endian) after every step, and reads the snapshots back.

With ``gravity=True`` the build adds ``-DGRAV`` (the self-gravity multigrid
inside Godunov, HLLC.f90:15-24, and the point-mass potential,
pointsource_gravity.f90:12-39).  The user module then sets PHI from
``phi0.bin`` (its ghost faces are the potential's boundary values), zeroes
PHIP and, for point masses, calls pointmass_potential with the masses as
the stars (8 Msun each, as hard-coded there).  The run writes PHIT per
output to ``phi`` + the BIN name (HydroCore.f90:284-290), i.e. under the
directory ``phi./DATA``, which the recipe creates.  In the serial build the
neighbour ranks stay 0, so MultiGrid never applies its multipole boundary
values (sources.f90:114-136 test ``left<0`` ...): the faces keep phi0's.

The initial state comes from a file (``ic.bin``) written here, so the
reference and the C oracle start from the same bits.  Only this container
runs this script: nothing on the GPU side reads /root/reference.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
from typing import Dict, List, Optional, Sequence

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
FLANG = shutil.which("flang") or shutil.which("amdflang") or "/opt/rocm/llvm/bin/flang"
FFLAGS = ["-O2", "-cpp", "-DDOUBLEP", "-fdefault-real-8"]

# Makefile:15-30, minus the two user files that a problem provides.
REF_SOURCES_BEFORE_USER = ["source/constants.f90", None, "source/globals.f90", "source/cooling.f90",
                           "source/pointsource_gravity.f90", "source/sources.f90",
                           "source/winds.f90"]
REF_SOURCES_AFTER_USER = ["source/HydroCore.f90", "source/init.f90", "source/boundaries.f90",
                          "source/HLLC.f90", "source/mainHD.f90"]

BC_NAME = {1: "BC_OUTFLOW", 2: "BC_REFLECTIVE", 3: "BC_PERIODIC", 4: "BC_INFLOW"}


def available() -> bool:
    return os.path.isdir(os.path.join(REF, "source")) and os.path.exists(FLANG)


def _r(x: float) -> str:
    """A real(8) literal that reads back to the same double."""
    return repr(float(x)) + "_8"


def _parameters_module(p, tfin: float, cooling_file: Optional[str], levels: int = 0,
                       num_out: int = 100000) -> str:
    lb, rb, db, tb, fb, bb = (BC_NAME[int(b)] for b in p.bc)
    lines = [
        "module parameters",
        "  use constants",
        "  implicit none",
        f"  integer, parameter :: LB = {lb}, RB = {rb}, TB = {tb}, DB = {db}, FB = {fb}, BB = {bb}",
        f"  integer, parameter :: limtr = {int(p.limiter)}",
        f"  integer, parameter :: ndim = 3, neq = 5, size = {int(levels)}",
        f"  integer, parameter :: nxtot = {p.nx}, nytot = {p.ny}, nztot = {p.nz}",
        "  integer, parameter :: choice = 2",
        f"  integer, parameter :: num_out = {int(num_out)}",
        "  integer, parameter :: outfile = BIN",
        "  character(*), parameter :: outputpath = './DATA/'",
    ]
    if cooling_file:
        lines.append(f"  character(*), parameter :: cooling_file = '{cooling_file}'")
    lines += [
        "  character(*), parameter :: posfile = 'none'",
        "  integer, parameter :: nghost = 2",
        f"  real, parameter :: xl = {_r(p.xl)}, xr = {_r(p.xr)}",
        f"  real, parameter :: yl = {_r(p.yl)}, yr = {_r(p.yr)}",
        f"  real, parameter :: zl = {_r(p.zl)}, zr = {_r(p.zr)}",
        f"  real, parameter :: gamma = {_r(p.gamma)}",
        f"  real, parameter :: mu0 = {_r(p.mu0)}",
        "  real, parameter :: Gconst = GR",
        f"  real, parameter :: cfl = {_r(p.cfl)}",
        f"  real, parameter :: eta = {_r(p.eta)}",
        f"  real, parameter :: tfin = {_r(tfin)}",
        f"  real, parameter :: cooling_limit = {_r(p.cooling_limit)}",
        "  integer, parameter :: mpix = 1, mpiy = 1, mpiz = 1",
        "  integer, parameter :: nx = nxtot, ny = nytot, nz = nztot, np = 1",
        "  integer, parameter :: mpi_real_kind = 0",
        "  real, parameter :: dtout = tfin/num_out",
        "  integer, parameter :: nxmax = nx+nghost, nymax = ny+nghost, nzmax = nz+nghost",
        "  integer, parameter :: nxmin = 1-nghost, nymin = 1-nghost, nzmin = 1-nghost",
        "  real, parameter :: dx = (xr-xl)/nxtot, dy = (yr-yl)/nytot, dz = (zr-zl)/nztot",
        "  integer, parameter :: lvlm = size",
        "  logical, parameter :: logged = .false.",
        "  integer, parameter :: logu = 12",
        "  character(*), parameter :: logfile = './log.dat'",
        "  character(*), parameter :: paramFile = 'tlaloque.dat'",
        "end module parameters",
    ]
    return "\n".join(lines) + "\n"


def _user_module(p, gravity: bool = False) -> str:
    winds = list(p.winds)
    masses = list(p.masses) if gravity else []
    if masses and winds:
        raise ValueError("the generated user module takes winds or point masses, not both")
    nstars = len(masses) if masses else len(winds)
    lines = [
        "module user",
        "  use parameters",
        "  use globals, only: posstars, Nstrs" + (", PHI, PHIP" if gravity else ""),
        "  use winds",
        "  use pgravity",
        "  implicit none",
        "  type(spherical_wind_type), dimension(:), allocatable :: spherical_wind",
        "contains",
        "  subroutine user_initconds(U)",
        "    real, dimension(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax), intent(inout) :: U",
        "    open(unit=77, file='ic.bin', access='stream', form='unformatted', status='old', &",
        "         convert='LITTLE_ENDIAN')",
        "    read(77) U",
        "    close(77)",
        f"    Nstrs = {nstars}",
        f"    allocate(spherical_wind(0:{len(winds)}-1))",
        "    allocate(posstars(0:2,0:Nstrs-1))",
    ]
    for i, w in enumerate(winds):
        for name, val in zip(("xc", "yc", "zc", "vx", "vy", "vz", "radius", "mdot", "vinf",
                              "temp", "mu"), w.as_tuple()):
            lines.append(f"    spherical_wind({i})%{name} = {_r(val)}")
        lines.append(f"    posstars(:,{i}) = (/ {_r(w.xc)}, {_r(w.yc)}, {_r(w.zc)} /)")
    for i, m in enumerate(masses):
        lines.append(f"    posstars(:,{i}) = (/ {_r(m.x)}, {_r(m.y)}, {_r(m.z)} /)")
    if gravity:
        lines += [
            "    open(unit=78, file='phi0.bin', access='stream', form='unformatted', status='old', &",
            "         convert='LITTLE_ENDIAN')",
            "    read(78) PHI",
            "    close(78)",
            "    PHIP = 0.0",
        ]
        if masses:
            lines.append("    call pointmass_potential(PHIP)")
    lines += [
        "  end subroutine user_initconds",
        "  subroutine user_boundaries(U)",
        "    real, dimension(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax), intent(inout) :: U",
        "    integer :: i",
        f"    do i = 0, {len(winds)}-1",
        "       call imposeSphericalWind(spherical_wind(i), U)",
        "    enddo",
        "  end subroutine user_boundaries",
        "end module user",
    ]
    return "\n".join(lines) + "\n"


def build_and_run(name: str, p, u0: np.ndarray, tfin: float, cooling: bool = False,
                  timeout: float = 600.0, gravity: bool = False,
                  phi0: Optional[np.ndarray] = None, num_out: int = 100000,
                  fflags: Sequence[str] = ()) -> Dict[str, object]:
    """Compile the reference for problem ``p`` and run it from ``u0`` (Fortran
    layout, C shape (nz+4, ny+4, nx+4, 5)) until t >= tfin.  With ``gravity``,
    ``phi0`` (C shape (nz+2, ny+2, nx+2)) is the initial PHI; the levels are
    ``tlaloque_amd.params.mg_levels(p)``.

    ``num_out`` is the reference's output count (an output whenever t passes
    tfin/num_out: the default writes after every step); ``fflags`` replace
    the optimisation flags.

    Returns {"prims": [UPrim after 0, 1, ... steps], "log": stdout,
    "seconds": wall time of the run} and, with gravity, "phit": [PHIT after
    1, 2, ... steps]."""
    if not available():
        raise RuntimeError("reference sources or flang not available")
    case = os.path.join(OUT, name)
    src, obj, run = (os.path.join(case, d) for d in ("src", "obj", "run"))
    if os.path.isdir(case):
        shutil.rmtree(case)
    for d in (src, obj, run, os.path.join(run, "DATA")):
        os.makedirs(d)
    cfile = os.path.join(REF, "Z1.0.dat") if cooling else None
    levels = 0
    if gravity:
        import sys
        sys.path.insert(0, os.path.dirname(HERE))
        from tlaloque_amd.params import MSUN, mg_levels
        levels = mg_levels(p)
        if any(m.mass != 8.0 * MSUN for m in p.masses) or (p.masses and p.softening != 0.0):
            raise ValueError("pointsource_gravity.f90:23,32 fixes 8 Msun per star, no softening")
    with open(os.path.join(src, "parameters.f90"), "w") as fh:
        fh.write(_parameters_module(p, tfin, cfile, levels, num_out))
    with open(os.path.join(src, "user.f90"), "w") as fh:
        fh.write(_user_module(p, gravity))
    base = [f for f in FFLAGS if not f.startswith("-O")] + list(fflags) if fflags else FFLAGS
    flags = base + (["-DCOOL"] if cooling else []) + (["-DGRAV"] if gravity else [])
    objs: List[str] = []
    order: Sequence[str] = REF_SOURCES_BEFORE_USER + ["user"] + REF_SOURCES_AFTER_USER
    for s in order:
        if s is None:
            path = os.path.join(src, "parameters.f90")
        elif s == "user":
            path = os.path.join(src, "user.f90")
        else:
            path = os.path.join(REF, s)
        o = os.path.join(obj, os.path.basename(path).replace(".f90", ".o"))
        subprocess.run([FLANG, *flags, "-module-dir", obj, "-c", path, "-o", o], check=True,
                       cwd=obj, capture_output=True)
        objs.append(o)
    exe = os.path.join(case, "tlaloque_ref")
    subprocess.run([FLANG, *flags, *objs, "-o", exe], check=True, capture_output=True)
    np.ascontiguousarray(u0, dtype="<f8").tofile(os.path.join(run, "ic.bin"))
    if gravity:
        ph = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2)) if phi0 is None else phi0
        np.ascontiguousarray(ph, dtype="<f8").tofile(os.path.join(run, "phi0.bin"))
        os.makedirs(os.path.join(run, "phi.", "DATA"))
    import time
    t0 = time.perf_counter()
    res = subprocess.run([exe], cwd=run, check=True, capture_output=True, text=True,
                         timeout=timeout)
    seconds = time.perf_counter() - t0
    files = sorted(glob.glob(os.path.join(run, "DATA", "points000N*.bin")))
    shape = u0.shape
    prims = [np.fromfile(f, dtype="<f8").reshape(shape) for f in files]
    out = {"prims": prims, "log": res.stdout, "seconds": seconds}
    if gravity:
        pf = sorted(glob.glob(os.path.join(run, "phi.", "DATA", "points000N*.bin")))[1:]
        out["phit"] = [np.fromfile(f, dtype="<f8").reshape(p.nz + 2, p.ny + 2, p.nx + 2)
                       for f in pf]
    return out


def compile_default() -> Optional[str]:
    """Compile the reference for a small default problem (the checker build
    done by ``__graft_entry__.build()``); returns the executable path, or None
    when the reference sources or flang are absent."""
    if not available():
        return None
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    from tlaloque_amd import params as P
    p = P.HydroParams(nx=8, ny=6, nz=5, bc=(P.BC_OUTFLOW,) * 6)
    case = os.path.join(OUT, "default")
    src, obj = os.path.join(case, "src"), os.path.join(case, "obj")
    os.makedirs(src, exist_ok=True)
    os.makedirs(obj, exist_ok=True)
    with open(os.path.join(src, "parameters.f90"), "w") as fh:
        fh.write(_parameters_module(p, 0.1, None))
    with open(os.path.join(src, "user.f90"), "w") as fh:
        fh.write(_user_module(p))
    objs = []
    for s in REF_SOURCES_BEFORE_USER + ["user"] + REF_SOURCES_AFTER_USER:
        path = (os.path.join(src, "parameters.f90") if s is None else
                os.path.join(src, "user.f90") if s == "user" else os.path.join(REF, s))
        o = os.path.join(obj, os.path.basename(path).replace(".f90", ".o"))
        subprocess.run([FLANG, *FFLAGS, "-module-dir", obj, "-c", path, "-o", o], check=True,
                       cwd=obj, capture_output=True)
        objs.append(o)
    exe = os.path.join(case, "tlaloque_ref")
    subprocess.run([FLANG, *FFLAGS, *objs, "-o", exe], check=True, capture_output=True)
    return exe
