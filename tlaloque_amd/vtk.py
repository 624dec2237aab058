"""Outputs in the reference's formats (source/HydroCore.f90:253-395).

``write_vtk`` / ``write_bin`` write what the reference's ``output`` writes on
one rank, so its analysis scripts read the GPU path's results unchanged:

  VTK  legacy binary STRUCTURED_POINTS, big-endian, DIMENSIONS nx ny nz,
       ORIGIN (xl, yl, zl), SPACING (dx, dy, dz), then the interior cells of
       SCALARS Density, SCALARS Pressure and VECTORS Velocity (x fastest).
       Values are always real(8) under the 'double' header (the reference's
       FP32 build writes 4-byte values under that header, which misreads).
  BIN  UPrim(5, -1:nx+2, -1:ny+2, -1:nz+2) as a little-endian real(8) stream.

With self-gravity (the -DGRAV build) the outputs also carry the potential
PHIT = PHIP + PHI: the VTK gets a "SCALARS Potential double 1" block between
Pressure and Velocity, and PHIT(0:nx+1, 0:ny+1, 0:nz+1) goes to its own
little-endian stream, named 'phi' + the BIN file's path
(HydroCore.f90:284-290, 360-372; ``phi_bin_name``).

``read_vtk`` parses such a file (the reference's harness uses VTKLegacy.jl
for this, LoadVTK in tests_analisis.jl).

Restarts: the reference's warmstart (source/warmstart.f90:1-10) was left
unfinished; ``read_restart`` does what it set out to do, for BIN and VTK
dumps.  Each dump may carry a sidecar ``<dump>.time`` with one line
``t it nout`` (written by ``write_time`` and by the Fortran driver), so a
restarted run continues the same clock and output numbering.
"""
from __future__ import annotations

import os
import re
from typing import Dict

import numpy as np

from . import params as P
from . import problems


def vtk_name(outputpath: str, nout: int, rank: int = 0) -> str:
    return os.path.join(outputpath, f"EPV{rank:03d}N{nout:03d}.vtk")


def bin_name(outputpath: str, nout: int, rank: int = 0) -> str:
    return os.path.join(outputpath, f"points{rank:03d}N{nout:03d}.bin")


def phi_bin_name(outputpath: str, nout: int, rank: int = 0) -> str:
    """'phi' // trim(outputpath) // 'points...bin', as the reference builds it:
    with outputpath './DATA/' this is 'phi./DATA/points000N001.bin'."""
    return "phi" + bin_name(outputpath, nout, rank)


def _fmt_e(v: float, width: int = 15, digits: int = 7) -> str:
    """Fortran's Ew.d edit descriptor: 0.ddddddd E+xx."""
    if v == 0.0:
        s = "0." + "0" * digits + "E+00"
    else:
        e = int(np.floor(np.log10(abs(v)))) + 1
        m = abs(v) / 10.0 ** e
        md = round(m, digits)
        if md >= 1.0:
            md /= 10.0
            e += 1
        mant = f"{md:.{digits}f}"          # '0.ddddddd'
        s = ("-" if v < 0 else "") + mant + f"E{e:+03d}"
    return s.rjust(width)


def write_vtk(path: str, p: P.HydroParams, prim: np.ndarray, phit: np.ndarray = None) -> None:
    """prim: numpy (nz+4, ny+4, nx+4, 5) primitives with ghosts; phit: the
    potential (nz+2, ny+2, nx+2) of a self-gravity run, or None."""
    w = prim[2:-2, 2:-2, 2:-2]
    hdr = ["# vtk DataFile Version 2.0", "output from Diable", "BINARY",
           "DATASET STRUCTURED_POINTS",
           f"DIMENSIONS {p.nx:6d}{p.ny:6d}{p.nz:6d}",
           "ORIGIN " + "".join(_fmt_e(v) for v in (p.xl, p.yl, p.zl)),
           "SPACING" + "".join(_fmt_e(v) for v in (p.dx, p.dy, p.dz)),
           f"POINT_DATA {p.nx * p.ny * p.nz:10d}"]
    be = ">f8"
    with open(path, "wb") as fh:
        fh.write(("\n".join(hdr) + "\n").encode())
        fh.write(b"SCALARS Density double 1\nLOOKUP_TABLE default\n")
        fh.write(np.ascontiguousarray(w[..., 0]).astype(be).tobytes())
        fh.write(b"\nSCALARS Pressure double 1\nLOOKUP_TABLE default\n")
        fh.write(np.ascontiguousarray(w[..., 4]).astype(be).tobytes())
        if phit is not None:
            fh.write(b"\nSCALARS Potential double 1\nLOOKUP_TABLE default\n")
            fh.write(np.ascontiguousarray(phit[1:-1, 1:-1, 1:-1]).astype(be).tobytes())
        fh.write(b"\nVECTORS Velocity double\n")
        fh.write(np.ascontiguousarray(w[..., 1:4]).astype(be).tobytes())
        fh.write(b"\n")


def write_bin(path: str, prim: np.ndarray) -> None:
    np.ascontiguousarray(prim).astype("<f8").tofile(path)


def read_bin(path: str, p: P.HydroParams) -> np.ndarray:
    return np.fromfile(path, dtype="<f8").reshape(p.shape + (5,))


def write_phi_bin(path: str, phit: np.ndarray) -> None:
    np.ascontiguousarray(phit).astype("<f8").tofile(path)


def read_phi_bin(path: str, p: P.HydroParams) -> np.ndarray:
    return np.fromfile(path, dtype="<f8").reshape(p.nz + 2, p.ny + 2, p.nx + 2)


def read_vtk(path: str) -> Dict[str, object]:
    """Returns dims (nx, ny, nz), origin, spacing and the fields as numpy
    arrays of shape (nz, ny, nx) (vectors (nz, ny, nx, 3))."""
    with open(path, "rb") as fh:
        data = fh.read()
    pos = 0

    def line():
        nonlocal pos
        end = data.index(b"\n", pos)
        s = data[pos:end].decode()
        pos = end + 1
        return s

    out: Dict[str, object] = {"fields": {}}
    if not line().startswith("# vtk DataFile"):
        raise ValueError("not a legacy VTK file")
    out["title"] = line()
    if line().strip() != "BINARY":
        raise ValueError("only BINARY VTK is read")
    npts = None
    while pos < len(data):
        s = line().strip()
        if not s:
            continue
        key = s.split()[0]
        if key == "DATASET":
            continue
        if key == "DIMENSIONS":
            out["dims"] = tuple(int(v) for v in s.split()[1:4])
        elif key == "ORIGIN":
            out["origin"] = tuple(float(v) for v in re.split(r"\s+", s)[1:4])
        elif key == "SPACING":
            out["spacing"] = tuple(float(v) for v in re.findall(r"[-+]?\d*\.\d+E[-+]\d+", s))
        elif key == "POINT_DATA":
            npts = int(s.split()[1])
        elif key in ("SCALARS", "VECTORS"):
            name, typ = s.split()[1], s.split()[2]
            if typ != "double":
                raise ValueError(f"field {name}: type {typ}")
            ncomp = 3 if key == "VECTORS" else 1
            if key == "SCALARS":
                lk = line()
                if not lk.startswith("LOOKUP_TABLE"):
                    raise ValueError("missing LOOKUP_TABLE")
            nb = 8 * npts * ncomp
            arr = np.frombuffer(data[pos:pos + nb], dtype=">f8").astype(np.float64)
            pos += nb
            nx, ny, nz = out["dims"]
            out["fields"][name] = arr.reshape((nz, ny, nx, 3) if ncomp == 3 else (nz, ny, nx))
        else:
            raise ValueError(f"unexpected VTK line {s!r}")
    return out


def write_time(dump_path: str, t: float, it: int, nout: int) -> None:
    """The sidecar ``<dump>.time``: t (17 significant digits), it, nout."""
    with open(dump_path + ".time", "w") as fh:
        fh.write(f"{t:25.17e} {int(it)} {int(nout)}\n")


def read_time(dump_path: str):
    """(t, it, nout) from the sidecar ``<dump>.time``."""
    with open(dump_path + ".time") as fh:
        f = fh.read().split()
    return float(f[0].replace("D", "E").replace("d", "e")), int(f[1]), int(f[2])


def read_restart(path: str, p: P.HydroParams):
    """U (reference layout, with ghosts) rebuilt from a dump, and the dump's
    (t, it, nout) when its sidecar exists (else None).

    A BIN dump holds every cell; a VTK dump holds the interior only, and the
    ghosts start as copies of the nearest interior cell (the boundary fill at
    start-up, ``HydroCore.set_state``, overwrites them).
    """
    if path.endswith(".vtk"):
        d = read_vtk(path)
        if tuple(d["dims"]) != (p.nx, p.ny, p.nz):
            raise ValueError(f"{path}: DIMENSIONS {d['dims']} != {(p.nx, p.ny, p.nz)}")
        f = d["fields"]
        w = np.empty((p.nz, p.ny, p.nx, 5))
        w[..., 0] = f["Density"]
        w[..., 1:4] = f["Velocity"]
        w[..., 4] = f["Pressure"]
        w = np.pad(w, ((2, 2), (2, 2), (2, 2), (0, 0)), mode="edge")
    else:
        w = read_bin(path, p)
    u = problems.cons_from_prims(p, w[..., 0], w[..., 1:4], w[..., 4])
    clock = read_time(path) if os.path.exists(path + ".time") else None
    return u, clock
