"""Self-gravity on the GPU (MultiGrid inside Godunov, HLLC.f90:15-24;
sources.f90:99-897) against the reference and the oracle.

The GPU relaxes red-black (the reference's Gauss-Seidel order is
sequential), so each V-cycle differs from the reference's in rounding and
in its iteration error; the converged potential is the same discrete
solution.  Every case here runs in the reference's own units (cgs,
Gconst = GR), where its stopping test Dif <= 10 means a converged potential,
so the comparisons are tight:

  * the reference runs built with -DGRAV (fixtures ``ref_sg_*``, the serial
    build: phi's ghost faces kept), fused path (one solve per step) and per
    routine (a solve in every Godunov call, as the reference): UPrim after 1
    and 4 steps within 1e-9 of the field's scale, PHIT within 1e-10;
  * one solve against the oracle's on a Plummer sphere, FP64 and FP32 (FP32
    rounds the density the source reads), with the finest level on either
    relaxation kernel (the one-pass iteration or two half sweeps), and the
    two kernels bitwise equal on a 64^3 solve and three steps after it;
  * multipole boundary values (the MPI build) against the oracle's, and
    their refresh when mod(it, 10) = 0 only;
  * an FP32 run with self-gravity and a point mass against the FP64 oracle,
    L1 of rho and p within 1e-3;
  * the self-gravity tests of the reference's harness (tests_utilities.jl):
    the potential-accuracy sphere (where the reference's absolute stopping
    test ends the solve after one V-cycle, 6% from converged: there the GPU
    and the oracle agree to that iteration error, and converge to the same
    potential when solved again), and the Evrard and Truelove collapses
    against the oracle, FP64 and FP32.
"""
import numpy as np
import pytest

import oracle
from helpers import l1_rel, load_ref_sg, max_rel, sg_ref_cases
from tlaloque_amd import params as P
from tlaloque_amd import problems

pytestmark = pytest.mark.gpu

CASES = sg_ref_cases()
TOL_PRIM = 1e-9
TOL_PHI = 1e-10


def _hc():
    from tlaloque_amd.hydro import HydroCore
    return HydroCore


def _point_potential(p):
    x = P.xcoord(p, np.arange(0, p.nx + 2))
    y = P.ycoord(p, np.arange(0, p.ny + 2))
    z = P.zcoord(p, np.arange(0, p.nz + 2))
    phi = np.zeros((p.nz + 2, p.ny + 2, p.nx + 2))
    for m in p.masses:
        r = np.sqrt((m.x - x[None, None, :]) ** 2 + (m.y - y[None, :, None]) ** 2
                    + (m.z - z[:, None, None]) ** 2 + p.softening ** 2)
        phi = phi - P.GR * m.mass / r
    return phi


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("case", CASES)
def test_fused_run_matches_reference_with_gravity(case):
    p, u0, prims, steps, phi0, phit = load_ref_sg(case)
    sl = (slice(2, -2),) * 3
    with _hc()(p) as h:
        h.upload_phi(phi0)
        h.set_state(u0)
        for s in range(1, steps[-1] + 1):
            h.advance(1)
            if s in steps:
                h.calc_prims("U")
                w = h.download("UPrim")
                err = max_rel(w[sl], prims[steps.index(s)][sl])
                print(case, s, err)
                assert err <= TOL_PRIM, (case, s, err)
            if s == 1:
                e1 = _rel(h.download_phi()[1:-1, 1:-1, 1:-1] + _point_potential(p)[1:-1, 1:-1, 1:-1],
                          phit[0][1:-1, 1:-1, 1:-1])
        phi = h.download_phi()
        lv, cyc, dif, it = h.gravity_info()
    e4 = _rel(phi[1:-1, 1:-1, 1:-1] + _point_potential(p)[1:-1, 1:-1, 1:-1], phit[1][1:-1, 1:-1, 1:-1])
    print(case, "phit", e1, e4, "levels", lv, "cycles", cyc, "dif", dif, "it", it)
    assert e1 <= TOL_PHI and e4 <= TOL_PHI
    assert lv == P.mg_levels(p) and it == steps[-1]
    # the ghost faces are kept as uploaded (serial build)
    assert np.array_equal(phi[0, 1:-1, 1:-1], phi0[0, 1:-1, 1:-1])
    assert np.array_equal(phi[1:-1, 1:-1, -1], phi0[1:-1, 1:-1, -1])


@pytest.mark.parametrize("case", CASES)
def test_per_routine_run_matches_reference_with_gravity(case):
    # the thirteen calls of mainHD.f90:74-99; each tlq_godunov solves first
    p, u0, prims, steps, phi0, _ = load_ref_sg(case)
    sl = (slice(2, -2),) * 3
    with _hc()(p) as h:
        h.upload_phi(phi0)
        h.set_state(u0)
        t, it = 0.0, 0
        for s in range(1, steps[-1] + 1):
            h.calc_prims("U")
            dt = h.time_step()
            h.fluxes("U", 1)
            h.godunov("UPP", dt / 2.0)
            h.borders("UPP", 2)
            h.calc_prims("UPP")
            h.fluxes("UPP", 2)
            h.godunov("UP", dt)
            h.borders("UP", 1)
            t, it = h.step(t, it, dt)
            if s in steps:
                h.calc_prims("U")
                err = max_rel(h.download("UPrim")[sl], prims[steps.index(s)][sl])
                print(case, s, err)
                assert err <= TOL_PRIM, (case, s, err)


def _plummer(n, prec, mode=P.SG_FIXED):
    L, M, a = P.PC, 1e3 * P.MSUN, 0.4 * P.PC
    p = P.HydroParams(nx=n, ny=n, nz=n, xl=-L, xr=L, yl=-L, yr=L, zl=-L, zr=L,
                      bc=(P.BC_OUTFLOW,) * 6, self_gravity=mode, precision=prec)
    x = P.xcoord(p, np.arange(-1, n + 3))
    r2 = x[None, None, :] ** 2 + x[None, :, None] ** 2 + x[:, None, None] ** 2
    rho = 3 * M / (4 * np.pi * a ** 3) * (1 + r2 / a ** 2) ** -2.5
    u0 = problems.cons_from_prims(p, rho, np.zeros(r2.shape + (3,)), np.full(r2.shape, 1e-9))
    exact = -P.GR * M / np.sqrt(r2 + a ** 2)[1:-1, 1:-1, 1:-1]
    phi0 = exact.copy()
    phi0[1:-1, 1:-1, 1:-1] = 0.0
    return p, u0, phi0


# The relaxation kernels: by default levels of 4M cells and more take the
# one-pass red-black iteration and smaller ones two in-place half sweeps, so
# the meshes here would run the half sweeps only.  TLQ_MG_ONEPASS_MIN_CELLS
# moves the threshold when a context is created: 32768 puts these 32^3 and
# 64^3 levels on the one-pass kernel.
ONEPASS = [None, "32768"]


def _onepass(monkeypatch, value):
    if value is None:
        monkeypatch.delenv("TLQ_MG_ONEPASS_MIN_CELLS", raising=False)
    else:
        monkeypatch.setenv("TLQ_MG_ONEPASS_MIN_CELLS", value)


@pytest.mark.parametrize("onepass", ONEPASS)
@pytest.mark.parametrize("prec,tol", [(P.PREC_FP64, 1e-10), (P.PREC_FP32, 1e-6)])
def test_solve_matches_oracle(prec, tol, onepass, monkeypatch):
    _onepass(monkeypatch, onepass)
    p, u0, phi0 = _plummer(32, prec)
    o = oracle.Oracle(p.replace(precision=P.PREC_FP64))
    o.set_phi(phi0)
    o.set_cons(u0)
    with _hc()(p) as h:
        h.upload_phi(phi0)
        h.set_state(u0)
        phi = h.download_phi()
        lv, cyc, dif, _ = h.gravity_info()
        again = h.gravity_solve()
    err = _rel(phi, o.phi)
    print(prec, "levels", lv, "cycles", cyc, "dif", dif, "oracle cycles", o.sg.cycles, "err", err,
          "again", again)
    assert lv == 5 and 1 <= cyc <= 31
    assert err <= tol
    assert again[0] <= 2        # from a converged potential one or two cycles suffice


@pytest.mark.parametrize("n", [64, 48])
@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_FP32])
def test_one_pass_iteration_equals_half_sweeps(prec, n, monkeypatch):
    """The one-pass red-black iteration and the two in-place half sweeps form
    the same updates in the same order: a Plummer solve and three fused steps
    after it give bitwise equal potentials and states with the levels of at
    least 32^3 cells on either kernel.  At 64^3 the one-pass tiles (32 x 8)
    cover the 64^3 and 32^3 levels exactly; at 48^3 the 48^3 level ends in a
    half tile in x."""
    p, u0, phi0 = _plummer(n, prec)
    out = {}
    for value in ONEPASS:
        _onepass(monkeypatch, value)
        with _hc()(p) as h:
            h.upload_phi(phi0)
            h.set_state(u0)
            phi = h.download_phi()
            cyc = h.gravity_info()[1]
            h.advance(3)
            out[value] = (phi, cyc, h.download("U"), h.download_phi())
    a, b = out[None], out["32768"]
    print(prec, "cycles", a[1], b[1])
    assert a[1] == b[1] and 1 <= a[1] <= 31
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[2], b[2]) and np.isfinite(a[2]).all()
    assert np.array_equal(a[3], b[3])


def test_multipole_faces_match_oracle_and_refresh_every_ten_steps():
    p, u0, _ = _plummer(16, P.PREC_FP64, mode=P.SG_MULTIPOLE)
    # move the sphere off centre so every multipole order appears
    u0 = np.roll(u0, (1, -2, 3), axis=(0, 1, 2))
    o = oracle.Oracle(p)
    o.set_cons(u0)
    faces = (np.s_[0, 1:-1, 1:-1], np.s_[-1, 1:-1, 1:-1], np.s_[1:-1, 0, 1:-1],
             np.s_[1:-1, -1, 1:-1], np.s_[1:-1, 1:-1, 0], np.s_[1:-1, 1:-1, -1])
    with _hc()(p) as h:
        h.set_state(u0)
        phi = h.download_phi()
        for f in faces:
            assert _rel(phi[f], o.phi[f]) <= 1e-12
        assert _rel(phi, o.phi) <= 1e-10
        seen = [phi]
        for _ in range(11):
            h.advance(1)
            seen.append(h.download_phi())
    # solves at it = 0 (start-up and step 1) .. 9 keep the faces; it = 10 refreshes them
    for s in range(1, 11):
        for f in faces:
            assert np.array_equal(seen[s][f], seen[0][f]), s
    assert any(not np.array_equal(seen[11][f], seen[10][f]) for f in faces)


def test_fp32_run_with_gravity_meets_the_l1_bar():
    p, u0, phi0 = _plummer(32, P.PREC_FP32)
    p = p.replace(masses=[P.PointMass(0.3 * P.PC, -0.2 * P.PC, 0.1 * P.PC, 50 * P.MSUN)],
                  softening=0.05 * P.PC)
    o = oracle.Oracle(p.replace(precision=P.PREC_FP64))
    o.set_phi(phi0)
    o.set_cons(u0)
    for _ in range(20):
        o.advance()
    ref = o.prim[2:-2, 2:-2, 2:-2]
    with _hc()(p) as h:
        h.upload_phi(phi0)
        h.set_state(u0)
        h.advance(20)
        h.calc_prims("U")
        w = h.download("UPrim")[2:-2, 2:-2, 2:-2]
        assert h.check_state() == 0
    l1r, l1p = l1_rel(w[..., 0], ref[..., 0]), l1_rel(w[..., 4], ref[..., 4])
    vr = np.abs(w[..., 1:4] - ref[..., 1:4]).mean() / np.abs(ref[..., 1:4]).mean()
    print("L1 rho", l1r, "p", l1p, "v", vr)
    assert l1r <= 1e-3 and l1p <= 1e-3 and vr <= 1e-3


# ---------------------------------------------------------------------------
# The self-gravity tests of the reference's harness (tests_utilities.jl:
# 247-278, 548-697), at reduced meshes.

def test_harness_potential_accuracy_sphere():
    # "Gravitational Potential Accuracy Test (Multigrid)": rho = 1 g/cm^3 in
    # r < 0.25 cm, cgs G.  PHI is of order 1e-8 there, so the absolute
    # stopping test Dif <= 10 ends the reference's solve after one V-cycle,
    # 6% from the converged potential (oracle); one red-black cycle lands at
    # the same distance by another route.  Repeated solves converge both to
    # the same discrete solution, within 1.3% of the analytic uniform-sphere
    # potential (the sphere is a staircase of cells).
    p, u0, phi0 = problems.grav_sphere(n=32)
    o = oracle.Oracle(p)
    o.set_phi(phi0)
    o.set_cons(u0)
    one_orc = o.phi.copy()
    assert o.sg.cycles == 1
    for _ in range(30):
        o.multigrid(o.dens)
    x = P.xcoord(p, np.arange(0, p.nx + 2))
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    R = 0.25
    exact = np.where(r < R, -2 * np.pi * P.GR * (R ** 2 - r ** 2 / 3), -P.GR * (4 / 3) * np.pi * R ** 3 / r)
    with _hc()(p) as h:
        h.upload_phi(phi0)
        h.set_state(u0)
        one_gpu = h.download_phi()
        assert h.gravity_info()[1] == 1
        for _ in range(30):
            h.gravity_solve()
        conv_gpu = h.download_phi()
    inner = (slice(1, -1),) * 3
    s = np.abs(exact).max()
    print("one cycle: gpu vs oracle", np.abs(one_gpu - one_orc).max() / s,
          "gpu vs exact", np.abs(one_gpu - exact)[inner].max() / s,
          "converged: gpu vs oracle", np.abs(conv_gpu - o.phi).max() / s,
          "vs exact", np.abs(conv_gpu - exact)[inner].max() / s)
    assert np.abs(one_gpu - exact)[inner].max() <= 0.10 * s
    assert np.abs(one_gpu - one_orc).max() <= 0.08 * s
    assert np.abs(conv_gpu - o.phi).max() <= 1e-10 * s
    assert np.abs(conv_gpu - exact)[inner].max() <= 0.02 * s


@pytest.mark.parametrize("name,make,steps,cfl,tol,bound64", [
    ("truelove", problems.truelove, 10, None, 0.0, 1e-6),
    ("evrard", problems.evrard, 20, 0.4, 0.0, 3e-3),
    ("evrard", problems.evrard, 20, 0.4, 1e-12, 1e-7)])
@pytest.mark.parametrize("prec", [P.PREC_FP64, P.PREC_FP32])
def test_harness_collapse_matches_oracle(name, make, steps, cfl, tol, bound64, prec):
    # "Evrard Collapse" (G = 1, reflective) and "Truelove Collapse Test" (cgs,
    # gamma = 1.001) at 32^3 with multipole faces: GPU against the FP64 oracle
    # after the same number of steps.
    # * Truelove (cgs): the reference's stopping test (Dif <= 10) means a
    #   converged potential, and the L1 errors are below 1e-6.
    # * Evrard (G = 1): PHI is of order 2, and Dif <= 10 summed over 32^3
    #   cells stops each solve short of converged, the reference's and the
    #   red-black one at different points; over 20 steps rho and p differ by
    #   L1 ~1e-3.  Dif (level 0's last sweep) also misses smooth error: one
    #   cycle from the last step's potential ends with Dif ~1e-11 while
    #   ~1e-5 away.  With mg_tol = 1e-12 every solve runs its 31 cycles and
    #   reaches the discrete solution (residual ~1e-14).  The collapse at
    #   32^3 is itself sensitive: the oracle started from U0 (1 + 1e-14 noise)
    #   ends 20 steps later ~2e-5 (L1 rho) from its unperturbed run, because
    #   limiter and wave-region branches flip.  The converged GPU run must lie
    #   within 3x that measured spread of the oracle's.
    # Evrard runs at CFL 0.4: the harness's 0.8 (set for 128^3) lets the
    # first 32^3 step (dt = 0.245; the CFL ignores gravity) drive densities
    # negative, in the oracle as in the reference's scheme.
    p, u0, phi0 = make(n=32, precision=prec)
    p = p.replace(mg_tol=tol)
    if cfl:
        p = p.replace(cfl=cfl)

    def oracle_run(u):
        o = oracle.Oracle(p.replace(precision=P.PREC_FP64))
        o.set_phi(phi0)
        o.set_cons(u)
        for _ in range(steps):
            o.advance()
        return o

    o = oracle_run(u0)
    ref = o.prim[2:-2, 2:-2, 2:-2]
    if tol > 0.0:   # the problem's own sensitivity to rounding
        rng = np.random.default_rng(0)
        q = oracle_run(u0 * (1.0 + 1e-14 * rng.standard_normal(u0.shape)))
        w2 = q.prim[2:-2, 2:-2, 2:-2]
        spread = max(l1_rel(w2[..., 0], ref[..., 0]), l1_rel(w2[..., 4], ref[..., 4]),
                     np.abs(q.phi - o.phi).max() / np.abs(o.phi).max())
        bound64 = max(bound64, 3.0 * spread)
        print(name, "oracle spread under 1e-14 noise", spread)
    with _hc()(p) as h:
        h.upload_phi(phi0)
        h.set_state(u0)
        h.advance(steps)
        _, t, it = h.time()
        h.calc_prims("U")
        w = h.download("UPrim")[2:-2, 2:-2, 2:-2]
        phi = h.download_phi()
        assert h.check_state() == 0
    l1r, l1p = l1_rel(w[..., 0], ref[..., 0]), l1_rel(w[..., 4], ref[..., 4])
    ep = np.abs(phi - o.phi).max() / np.abs(o.phi).max()
    print(name, prec, "L1 rho", l1r, "p", l1p, "max rel", max_rel(w, ref), "phi", ep,
          "t", t, o.t)
    bound = bound64 if prec == P.PREC_FP64 else max(bound64, 1e-4)
    assert it == steps and abs(t - o.t) <= max(bound, 1e-7) * o.t
    assert l1r <= bound and l1p <= bound and ep <= bound
    # the collapse has started: the central cells are denser than at t = 0
    c = p.nx // 2
    core = (slice(c - 1, c + 1),) * 3
    assert w[core + (0,)].mean() > u0[2:-2, 2:-2, 2:-2, 0][core].mean()
