"""Restart from a dump: the unfinished warmstart of source/warmstart.f90:1-10
(whose autosave, #warmstart.f90#:11-13, reads an output VTK back), finished
for the BIN dump of source/HydroCore.f90:277-283 and the VTK of :286-387.

CPU:
  * the Python reader rebuilds U from BIN and VTK dumps and reads the
    ``.time`` sidecar;
  * a dry run of the Fortran driver restarted from a BIN / VTK dump writes
    back the dump's primitives (BIN: ghosts too; VTK: ghosts as copies of the
    nearest interior cell) and a sidecar numbered one past the dump's;
  * the driver refuses a VTK dump of another mesh and a dump with neither a
    sidecar nor restart_t.
GPU:
  * a run restarted from its own middle BIN output (which keeps the ghosts,
    images of UP rather than of U's interior, HydroCore.f90:195-206) ends
    where the uninterrupted run ends: same step count, t and fields within
    1e-12.  Restarted from the VTK output (interior only, ghosts refilled as
    at a fresh start, which changes the walls' ghosts by about 1% here) it
    ends where the oracle restarted from the same file ends;
  * through the Python binding (tlq_resume), FP64 and FP32, and tlq_resume's
    argument checks.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import prims_from_cons
from test_fortran import _namelist, _run, driver, needs_flang  # noqa: F401  (fixture)
from tlaloque_amd import params as P
from tlaloque_amd import problems, vtk


def _dump(tmp_path, p, u0, fmt, t=0.25, it=17, nout=3):
    prim = prims_from_cons(u0, p.gamma)
    src = tmp_path / "dump"
    src.mkdir(exist_ok=True)
    path = str(src / (f"EPV000N{nout:03d}.vtk" if fmt == "vtk" else f"points000N{nout:03d}.bin"))
    if fmt == "vtk":
        vtk.write_vtk(path, p, prim)
    else:
        vtk.write_bin(path, prim)
    vtk.write_time(path, t, it, nout)
    return path, prim


@pytest.mark.parametrize("fmt", ["bin", "vtk"])
def test_python_reader_rebuilds_u_and_clock(tmp_path, fmt):
    p, u0 = problems.sedov(n=10)
    path, prim = _dump(tmp_path, p, u0, fmt)
    u, clock = vtk.read_restart(path, p)
    assert clock == (0.25, 17, 3)
    w = prims_from_cons(u, p.gamma)
    if fmt == "vtk":
        w, prim = w[2:-2, 2:-2, 2:-2], prim[2:-2, 2:-2, 2:-2]
    assert np.allclose(w, prim, rtol=1e-14, atol=0)
    assert np.allclose(u[2:-2, 2:-2, 2:-2], u0[2:-2, 2:-2, 2:-2], rtol=1e-14, atol=1e-300)


def test_python_reader_vtk_ghosts_copy_nearest_interior(tmp_path):
    p, u0 = problems.sod_diagonal(n=8)
    path, _ = _dump(tmp_path, p, u0, "vtk")
    u, _ = vtk.read_restart(path, p)
    assert np.array_equal(u[0], u[2]) and np.array_equal(u[1], u[2])
    assert np.array_equal(u[:, :, -1], u[:, :, -3])
    p2 = p.replace(nx=9)
    with pytest.raises(ValueError, match="DIMENSIONS"):
        vtk.read_restart(path, p2)


@needs_flang
@pytest.mark.parametrize("fmt", ["bin", "vtk"])
def test_driver_dry_restart_writes_back_the_dump(tmp_path, driver, fmt):
    p, u0 = problems.sedov(n=10)
    path, prim = _dump(tmp_path, p, u0, fmt, t=0.03125, it=9, nout=2)
    out = tmp_path / "DATA"
    out.mkdir()
    nml = tmp_path / "r.nml"
    _namelist(nml, p, "restart", out, tfin=0.1, num_out=4, outfile="bin",
              extra=f", dry_run=.true., restart_file='{path}'")
    log = _run(driver, nml)
    assert "Restart from" in log
    name = vtk.bin_name(str(out), 3)
    w = vtk.read_bin(name, p)
    if fmt == "vtk":
        ref = np.pad(prim[2:-2, 2:-2, 2:-2], ((2, 2), (2, 2), (2, 2), (0, 0)), mode="edge")
    else:
        ref = prim
    assert np.allclose(w, ref, rtol=1e-13, atol=0)
    t, it, nout = vtk.read_time(name)
    assert (t, it, nout) == (0.03125, 9, 3)


@needs_flang
def test_driver_restart_namelist_overrides_sidecar(tmp_path, driver):
    p, u0 = problems.sedov(n=6)
    path, _ = _dump(tmp_path, p, u0, "bin", t=0.5, it=4, nout=1)
    os.remove(path + ".time")
    out = tmp_path / "DATA"
    out.mkdir()
    nml = tmp_path / "r.nml"
    _namelist(nml, p, "restart", out, outfile="bin",
              extra=f", dry_run=.true., restart_file='{path}'")
    r = subprocess.run([driver, str(nml)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=60)
    assert r.returncode != 0 and "no sidecar" in r.stdout
    _namelist(nml, p, "restart", out, outfile="bin",
              extra=f", dry_run=.true., restart_file='{path}', restart_t=0.75, "
                    f"restart_it=12, restart_nout=5")
    _run(driver, nml)
    assert vtk.read_time(vtk.bin_name(str(out), 6)) == (0.75, 12, 6)


@needs_flang
def test_driver_refuses_vtk_of_another_mesh(tmp_path, driver):
    p, u0 = problems.sedov(n=6)
    path, _ = _dump(tmp_path, p, u0, "vtk")
    out = tmp_path / "DATA"
    out.mkdir()
    nml = tmp_path / "r.nml"
    _namelist(nml, p.replace(nx=7), "restart", out, outfile="bin",
              extra=f", dry_run=.true., restart_file='{path}'")
    r = subprocess.run([driver, str(nml)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=60)
    assert r.returncode != 0 and "DIMENSIONS" in r.stdout


@pytest.mark.gpu
@needs_flang
@pytest.mark.parametrize("fmt", ["bin", "vtk"])
def test_driver_restart_continues_the_run(tmp_path, driver, fmt):
    p, _ = problems.sod_diagonal(n=32)
    tfin = 0.12
    full = tmp_path / "FULL"
    full.mkdir()
    nml = tmp_path / "full.nml"
    _namelist(nml, p, "sod", full, tfin=tfin, num_out=2, outfile=fmt)
    _run(driver, nml)
    mid = os.path.join(full, "EPV000N001.vtk" if fmt == "vtk" else "points000N001.bin")
    t_mid, it_mid, _ = vtk.read_time(mid)
    assert 0.0 < t_mid < tfin and it_mid > 0
    rst = tmp_path / "RST"
    rst.mkdir()
    nml = tmp_path / "rst.nml"
    _namelist(nml, p, "restart", rst, tfin=tfin, num_out=2, outfile=fmt,
              extra=f", restart_file='{mid}'")
    log = _run(driver, nml)
    assert "Restart from" in log
    assert not os.listdir(rst) or all("N002" in f for f in os.listdir(rst))
    last = "EPV000N002.vtk" if fmt == "vtk" else "points000N002.bin"
    b = os.path.join(rst, last)
    tb, itb, _ = vtk.read_time(b)
    if fmt == "bin":
        # the dump's ghosts are kept: the uninterrupted run, to rounding
        a = os.path.join(full, last)
        ta, ita, _ = vtk.read_time(a)
        wa = vtk.read_bin(a, p)[2:-2, 2:-2, 2:-2]
    else:
        # interior only, ghosts refilled as at a fresh start: the oracle
        # restarted from the same file
        import oracle
        u, _ = vtk.read_restart(mid, p)
        o = oracle.Oracle(p)
        o.set_cons(u)
        o.t, o.it = t_mid, it_mid
        while o.t < tfin:
            o.advance()
        ta, ita = o.t, o.it
        wa = o.calc_prims(o.u)[2:-2, 2:-2, 2:-2]
    assert ita == itb
    assert abs(ta - tb) <= 1e-13 * ta
    if fmt == "vtk":
        fb = vtk.read_vtk(b)["fields"]
        pairs = [(wa[..., 0], fb["Density"]), (wa[..., 4], fb["Pressure"]),
                 (wa[..., 1:4], fb["Velocity"])]
    else:
        wb = vtk.read_bin(b, p)[2:-2, 2:-2, 2:-2]
        pairs = [(wa[..., e], wb[..., e]) for e in range(5)]
    for x, y in pairs:
        assert np.abs(x - y).max() <= 1e-12 * max(np.abs(x).max(), 1e-300)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(P.PREC_FP64, 1e-12), (P.PREC_FP32, 2e-5)])
def test_binding_restart_continues_the_run(prec, tol):
    from tlaloque_amd.hydro import HydroCore
    p, u0 = problems.sedov(n=40, precision=prec)
    with HydroCore(p) as h:
        h.set_state(u0)
        h.advance(15)
        _, t_mid, it_mid = h.time()
        h.calc_prims("U")
        prim = h.download("UPrim").astype(np.float64)
        h.advance(15)
        _, t_ref, it_ref = h.time()
        h.calc_prims("U")
        ref = h.download("UPrim").astype(np.float64)[2:-2, 2:-2, 2:-2]
    u = problems.cons_from_prims(p, prim[..., 0], prim[..., 1:4], prim[..., 4])
    with HydroCore(p) as h:
        h.resume(u, t_mid, it_mid)
        assert h.time()[1:] == (t_mid, it_mid)
        h.advance(15)
        _, t, it = h.time()
        h.calc_prims("U")
        w = h.download("UPrim").astype(np.float64)[2:-2, 2:-2, 2:-2]
    assert it == it_ref == 30
    assert abs(t - t_ref) <= tol * t_ref
    for e in range(5):
        scale = max(np.abs(ref[..., e]).max(), 1e-300)
        assert np.abs(w[..., e] - ref[..., e]).max() <= tol * scale, e


@pytest.mark.gpu
def test_resume_rejects_bad_arguments():
    from tlaloque_amd.hydro import HydroCore, TlqError
    p, u0 = problems.sedov(n=8)
    with HydroCore(p) as h:
        for t, it, fill in ((-1.0, 0, 0), (float("nan"), 0, 0), (float("inf"), 0, 0),
                            (0.1, -1, 0), (0.1, 0, 2)):
            with pytest.raises(TlqError):
                h.resume(u0, t, it, fill)


@pytest.mark.gpu
@needs_flang
def test_driver_restart_with_self_gravity_continues_the_run(tmp_path, driver):
    # the BIN dump plus its PHIT dump ('phi' + the BIN path) restart a
    # self-gravity run: PHI = PHIT - PHIP comes back with its multipole faces
    # (refreshed only when mod(it, 10) = 0), and the first solve starts from it
    from tlaloque_amd.hydro import HydroCore
    L = P.PC
    p = P.HydroParams(nx=16, ny=16, nz=16, xl=-L, xr=L, yl=-L, yr=L, zl=-L, zr=L,
                      bc=(P.BC_OUTFLOW,) * 6, self_gravity=P.SG_MULTIPOLE, eta=0.02)
    x = P.xcoord(p, np.arange(-1, p.nx + 3))
    r = np.sqrt(x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2)
    rho = np.where(r < 0.4 * L, 1e-20, 1e-22)
    u0 = problems.cons_from_prims(p, rho, np.zeros(r.shape + (3,)), np.full(r.shape, 1e-10))
    with HydroCore(p) as h:      # tfin between steps 13 and 14: both runs take 14 steps
        h.set_state(u0)
        h.advance(13)
        t13 = h.time()[1]
        h.advance(1)
        tfin = 0.5 * (t13 + h.time()[1])
    init = tmp_path / "u0.bin"
    np.ascontiguousarray(u0).astype("<f8").tofile(init)
    sg = f", self_gravity={P.SG_MULTIPOLE}"

    # one run directory, as a reference run has: the dumps and their PHIT
    # files ('phi' + the relative BIN path) side by side
    for sub in ("D", "R", "phi./D", "phi./R"):
        (tmp_path / sub).mkdir(parents=True)

    def run(out, problem, extra):
        nml = tmp_path / f"{out}.nml"
        _namelist(nml, p, problem, ".", tfin=tfin, num_out=2, outfile="bin", extra=extra)
        nml.write_text(nml.read_text().replace("/\n&run", sg + " /\n&run", 1)
                       .replace("outputpath='./'", f"outputpath='./{out}/'"))
        res = subprocess.run([driver, str(nml)], cwd=tmp_path, stdout=subprocess.PIPE,
                             stderr=subprocess.STDOUT, text=True, timeout=300)
        assert res.returncode == 0, res.stdout[-2000:]
        return res.stdout

    run("D", "file", f", initfile='{init}'")
    t_mid, it_mid, _ = vtk.read_time(str(tmp_path / "D" / "points000N001.bin"))
    assert 0 < it_mid < 14 and it_mid % 10 != 0
    log = run("R", "restart", ", restart_file='./D/points000N001.bin'")
    assert "PHI from" in log
    a, b = tmp_path / "D" / "points000N002.bin", tmp_path / "R" / "points000N002.bin"
    assert vtk.read_time(str(a))[1] == vtk.read_time(str(b))[1]
    wa, wb = vtk.read_bin(str(a), p)[2:-2, 2:-2, 2:-2], vtk.read_bin(str(b), p)[2:-2, 2:-2, 2:-2]
    for e in range(5):
        scale = max(np.abs(wa[..., e]).max(), 1e-300)
        assert np.abs(wa[..., e] - wb[..., e]).max() <= 1e-9 * scale, e
    pa = vtk.read_phi_bin(str(tmp_path / vtk.phi_bin_name("./D/", 2)), p)
    pb = vtk.read_phi_bin(str(tmp_path / vtk.phi_bin_name("./R/", 2)), p)
    assert np.abs(pa - pb).max() <= 1e-10 * np.abs(pa).max()
