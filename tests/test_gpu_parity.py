"""GPU parity of the HIP path against the reference's own runs.

Fixtures ``tests/golden/ref_*.npz`` hold UPrim of the reference (built from
its sources with flang) after 0, 1 and N steps from a seeded state; the C
oracle reproduces them bit for bit (tests/test_oracle_ref.py).  Here the GPU
runs the same problems through the C-ABI:

  * the per-routine entry points, called in the order of mainHD.f90:74-99;
  * the fused two-kernel step (tlq_advance), FP64;
  * the fused step in FP32 with FP64 guards.

Tolerances: FP64 builds keep the reference's operation order without FMA
contraction, so results agree to a few ulps except where device libm (exp,
log10, pow in the cooling) differs from glibc: max relative difference per
field <= 1e-12 (1e-10 with cooling).  FP32: L1 (sum|d|/sum|ref|) of rho and
p <= 1e-5, a hundredth of the hydro path's 1e-3 target.
"""
import numpy as np
import pytest

from helpers import (cooling_table, interior, l1_rel, load_ref, max_rel, prims_from_cons,
                     ref_cases)
from tlaloque_amd import params as P
from tlaloque_amd import problems
from tlaloque_amd.hydro import HydroCore

pytestmark = pytest.mark.gpu

CASES = ref_cases()


def _core(p):
    table = cooling_table() if p.cooling else None
    return HydroCore(p, cooling_table=table)


def _tol(p):
    return 1e-10 if p.cooling else 1e-12


@pytest.mark.parametrize("case", CASES)
def test_routines_fp64_match_reference(case):
    p, u0, prims, steps = load_ref(case)
    n = steps[-1]
    with _core(p) as h:
        h.upload("U", u0)
        h.borders("U", 1)            # mainHD.f90:38
        h.calc_prims("U")            # mainHD.f90:44
        t, it = 0.0, 0
        for s in range(n):
            dt = h.time_step()                     # :74
            h.fluxes("U", 1)                       # :76
            h.godunov("UPP", dt / 2.0)             # :78-80
            h.borders("UPP", 2)                    # :82
            h.calc_prims("UPP")                    # :84
            h.fluxes("UPP", 2)                     # :86
            h.godunov("UP", dt)                    # :88
            h.borders("UP", 1)                     # :90
            t, it = h.step(t, it, dt)              # :92
            h.calc_prims("U")                      # :94
            if p.cooling:
                h.cooling(dt)                      # :96-98
            if s + 1 == 1:
                w1 = h.download("UPrim")
                assert max_rel(interior(w1), interior(prims[1])) <= _tol(p), case
        w = h.download("UPrim")
    assert it == n
    assert max_rel(interior(w), interior(prims[-1])) <= _tol(p), case


@pytest.mark.parametrize("case", CASES)
def test_fused_fp64_matches_reference(case):
    p, u0, prims, steps = load_ref(case)
    n = steps[-1]
    with _core(p) as h:
        h.set_state(u0)
        h.advance(1)
        w1 = prims_from_cons(h.download("U"), p.gamma)
        h.advance(n - 1)
        dt, t, it = h.time()
        w = prims_from_cons(h.download("U"), p.gamma)
    assert it == n
    tol = _tol(p)
    assert max_rel(interior(w1), interior(prims[1])) <= tol, case
    assert max_rel(interior(w), interior(prims[-1])) <= tol, case


@pytest.mark.parametrize("case", CASES)
def test_fused_fp32_close_to_reference(case):
    p, u0, prims, steps = load_ref(case)
    n = steps[-1]
    with _core(p.replace(precision=P.PREC_FP32)) as h:
        h.set_state(u0)
        h.advance(n)
        w = prims_from_cons(h.download("U"), p.gamma)
        assert h.check_state() == 0
    ref = interior(prims[-1])
    got = interior(w)
    # Measured 3e-8 to 8e-8 for rho and up to 1.8e-6 for p (bc_mixed).  The
    # bound is well below the 1e-3 target so that a discrete difference
    # shows: before the FP32 contact speed kept the reference's exact S* = 0
    # at reflective walls, FP32 picked the other HLLC star state there and
    # this case measured 7e-3 (LIMITER_NONE) and 7e-5 (bc_mixed).
    tol = 1e-5
    l1r, l1p = l1_rel(got[..., 0], ref[..., 0]), l1_rel(got[..., 4], ref[..., 4])
    print(f"{case}: FP32 vs reference L1 rho {l1r:.3e} p {l1p:.3e}")
    assert l1r <= tol, case
    assert l1p <= tol, case


def _quiet_contacts(n):
    """Random density, uniform pressure, zero velocity, periodic: every face
    is a stationary contact, so momenta are pure rounding noise and reach
    subnormal values within a few steps."""
    p = P.HydroParams(nx=n, ny=n, nz=n, bc=(P.BC_PERIODIC,) * 6, xl=-0.5, xr=0.5, yl=-0.5,
                      yr=0.5, zl=-0.5, zr=0.5, precision=P.PREC_FP32)
    rng = np.random.default_rng(0)
    u = np.zeros(p.shape + (5,))
    u[..., 0] = 1.0 + 0.1 * rng.random(p.shape)
    u[..., 4] = 2.5
    return p, u


def test_fp32_subnormal_slopes_stay_finite():
    # Regression: a subnormal one-sided velocity difference once made the FP32
    # van Leer average divide by a subnormal (NaN at step 3 of this case).
    p, u = _quiet_contacts(256)
    with HydroCore(p) as h:
        h.set_state(u)
        for _ in range(6):
            h.advance(1)
            assert h.check_state() == 0


def test_fp32_fluxes_with_subnormal_differences():
    p = P.HydroParams(nx=8, ny=6, nz=5, bc=(P.BC_OUTFLOW,) * 6, precision=P.PREC_FP32)
    rng = np.random.default_rng(3)
    u = np.zeros(p.shape + (5,))
    u[..., 0] = 1.0 + 0.1 * rng.random(p.shape)
    u[..., 1:4] = rng.choice([0.0, 8e-39, -8e-39, 1e-9, -1e-9], size=p.shape + (3,))
    # non-uniform pressure keeps every contact speed well away from zero
    u[..., 4] = rng.uniform(1.0, 2.0, p.shape) / (p.gamma - 1.0)
    out = {}
    for prec in (P.PREC_FP32, P.PREC_FP64):
        with HydroCore(p.replace(precision=prec)) as h:
            h.upload("UPP", u)
            h.calc_prims("UPP")
            h.fluxes("UPP", 2)
            out[prec] = [interior(h.download(b)) for b in ("F", "G", "H")]
    for f32, f64 in zip(out[P.PREC_FP32], out[P.PREC_FP64]):
        assert np.isfinite(f32).all()
        assert max_rel(f32, f64) <= 1e-5


def _oracle_steps(p, u0, n):
    import oracle
    o = oracle.Oracle(p, table=cooling_table() if p.cooling else None)
    o.set_cons(u0)
    for _ in range(n):
        o.advance()
    return o


@pytest.mark.parametrize("case", ["hydro_outflow", "bc_mixed", "periodic", "limiter3"])
def test_fused_hll_fp64_matches_oracle(case):
    # HLL.f90 is not part of the reference build; the HLL oracle is pinned by
    # the Riemann-problem properties of tests/test_oracle_pins.py
    p, u0, _, _ = load_ref(case)
    p = p.replace(solver=P.SOLVER_HLL)
    o = _oracle_steps(p, u0, 3)
    with _core(p) as h:
        h.set_state(u0)
        h.advance(3)
        w = prims_from_cons(h.download("U"), p.gamma)
    assert max_rel(interior(w), interior(o.prim)) <= 1e-12, case


def _small_problem(shape, bc, precision, seed=11):
    nx, ny, nz = shape
    p = P.HydroParams(nx=nx, ny=ny, nz=nz, xl=0.0, xr=float(nx), yl=0.0, yr=float(ny), zl=0.0,
                      zr=float(nz), gamma=1.4, cfl=0.4, eta=0.05, bc=(bc,) * 6,
                      precision=precision)
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.5 * rng.uniform(-1, 1, p.shape)
    vel = 0.3 * rng.uniform(-1, 1, p.shape + (3,))
    pres = 1.0 + 0.5 * rng.uniform(-1, 1, p.shape)
    return p, problems.cons_from_prims(p, rho, vel, pres)


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 2, 5), (2, 9, 2)])
@pytest.mark.parametrize("bc", [P.BC_OUTFLOW, P.BC_REFLECTIVE, P.BC_PERIODIC])
def test_smallest_meshes_match_oracle(shape, bc):
    """The smallest meshes tlq_create accepts (two cells on an axis): every
    interior cell is a boundary cell, and the two ghost layers of the
    boundary fill (boundaries.f90:432-533) outnumber the interior.  FP64 fused
    path against the oracle, 3 steps; the FP32 path stays finite and within
    1e-5 of it, field by field relative to the field's largest value (measured
    up to 4.8e-7).  The reflective boxes pin the FP32 van Leer slope's
    symmetry: with a slope that differed by an ulp between lim_avg(a, b) and
    lim_avg(b, a), the face states beside a wall were no longer mirror images
    and the pressure there was off by 1.2e-4 in the 2 x 2 x 2 box."""
    p, u0 = _small_problem(shape, bc, P.PREC_FP64)
    o = _oracle_steps(p, u0, 3)
    with _core(p) as h:
        h.set_state(u0)
        h.advance(3)
        w = prims_from_cons(h.download("U"), p.gamma)
    assert max_rel(interior(w), interior(o.prim)) <= 1e-12, (shape, bc)
    p32, _ = _small_problem(shape, bc, P.PREC_FP32)
    with _core(p32) as h:
        h.set_state(u0)
        h.advance(3)
        assert h.check_state() == 0
        w32 = prims_from_cons(h.download("U"), p.gamma)
    e32 = max_rel(interior(w32), interior(o.prim))
    print(shape, bc, "FP32 max rel", e32)
    assert e32 <= 1e-5, (shape, bc)


def test_routines_hll_fp64_match_oracle():
    p, u0, _, _ = load_ref("hydro_outflow")
    p = p.replace(solver=P.SOLVER_HLL)
    o = _oracle_steps(p, u0, 2)
    with _core(p) as h:
        h.upload("U", u0)
        h.borders("U", 1)
        h.calc_prims("U")
        t, it = 0.0, 0
        for _ in range(2):
            dt = h.time_step()
            h.fluxes("U", 1)
            h.godunov("UPP", dt / 2.0)
            h.borders("UPP", 2)
            h.calc_prims("UPP")
            h.fluxes("UPP", 2)
            h.godunov("UP", dt)
            h.borders("UP", 1)
            t, it = h.step(t, it, dt)
            h.calc_prims("U")
        w = h.download("UPrim")
    assert max_rel(interior(w), interior(o.prim)) <= 1e-12


def _gravity_problem(precision):
    from tlaloque_amd import problems
    return problems.clumpy(n=24, precision=precision, cooling=P.COOL_OFF)


def test_point_mass_gravity_fp64_matches_oracle():
    p, u0 = _gravity_problem(P.PREC_FP64)
    assert p.masses and p.softening > 0
    o = _oracle_steps(p, u0, 3)
    with _core(p) as h:
        h.set_state(u0)
        h.advance(3)
        w = prims_from_cons(h.download("U"), p.gamma)
    assert max_rel(interior(w), interior(o.prim)) <= 1e-12
    # per routine: godunov adds dt S from the device's own source kernel
    with _core(p) as h:
        h.upload("U", u0)
        h.borders("U", 1)
        h.calc_prims("U")
        t, it = 0.0, 0
        for _ in range(3):
            dt = h.time_step()
            h.fluxes("U", 1)
            h.godunov("UPP", dt / 2.0)
            h.borders("UPP", 2)
            h.calc_prims("UPP")
            h.fluxes("UPP", 2)
            h.godunov("UP", dt)
            h.borders("UP", 1)
            t, it = h.step(t, it, dt)
            h.calc_prims("U")
        w2 = h.download("UPrim")
    assert max_rel(interior(w2), interior(o.prim)) <= 1e-12


def test_point_mass_gravity_fp32_close_to_oracle():
    p, u0 = _gravity_problem(P.PREC_FP32)
    o = _oracle_steps(p.replace(precision=P.PREC_FP64), u0, 3)
    with _core(p) as h:
        h.set_state(u0)
        h.advance(3)
        w = prims_from_cons(h.download("U"), p.gamma)
    ref = interior(o.prim)
    got = interior(w)
    assert l1_rel(got[..., 0], ref[..., 0]) <= 1e-3
    assert l1_rel(got[..., 4], ref[..., 4]) <= 1e-3
    for e in (1, 2, 3):
        assert l1_rel(got[..., e], ref[..., e]) <= 1e-3


def test_fp32_mass_field_follows_set_point_masses():
    # In FP32 mode tlq_set_point_masses stores the masses' acceleration per
    # cell; setting other masses (or none) must replace it: a run whose masses
    # were changed before its first step equals a run created with the final
    # masses, bit for bit.
    from tlaloque_amd import hydro
    p, u0 = _gravity_problem(P.PREC_FP32)
    assert p.masses
    moved = [P.PointMass(x=q.x + 0.3 * p.dx, y=q.y - 0.2 * p.dy, z=q.z, mass=2.0 * q.mass)
             for q in p.masses]
    lib = hydro.load_library()

    def run(first, final):
        with _core(p.replace(masses=first)) as h:
            if final is not first:
                m = np.ascontiguousarray(
                    np.array([[q.x, q.y, q.z, q.mass] for q in final], dtype=np.float64).ravel())
                ptr = hydro._dptr(m) if len(final) else None
                assert lib.tlq_set_point_masses(h._ctx, ptr, len(final)) == 0
            h.set_state(u0)
            h.advance(3)
            return h.download("U")

    assert np.array_equal(run(p.masses, moved), run(moved, moved))
    assert np.array_equal(run(p.masses, []), run([], []))
    assert not np.array_equal(run(p.masses, p.masses), run(moved, moved))


@pytest.mark.parametrize("precision", [P.PREC_FP64, P.PREC_FP32])
def test_rejected_setters_leave_the_context_unchanged(precision):
    # A cooling table whose temperatures stop increasing or that holds a NaN,
    # or a mass list that is too long or not finite, is refused with
    # TLQ_ERR_ARG; the context keeps its table and masses, so a run after the
    # refused calls (and a later setter that re-sends the problem to the
    # device) equals an undisturbed run bit for bit.
    from tlaloque_amd import hydro, problems
    TLQ_ERR_ARG = -1                                          # include/tlq.h
    p, u0 = problems.clumpy(n=24, precision=precision, temp=3e5)
    lib = hydro.load_library()
    T, L = (np.ascontiguousarray(a, dtype=np.float64) for a in cooling_table())
    bad_T, bad_L = T.copy() * 1.5, L.copy() + 1.0
    bad_T[-1] = bad_T[-2]
    nan_L = L.copy() + 1.0
    nan_L[7] = np.nan
    m = np.array([[q.x, q.y, q.z, q.mass] for q in p.masses], dtype=np.float64).ravel()
    many = np.zeros(4 * 1024, dtype=np.float64)
    nan_m = m.copy() * 2.0
    nan_m[1] = np.inf

    def run(refuse):
        with _core(p) as h:
            if refuse:
                assert lib.tlq_set_cooling_table(h._ctx, hydro._dptr(bad_T), hydro._dptr(bad_L),
                                                 len(bad_T)) == TLQ_ERR_ARG
                assert lib.tlq_set_cooling_table(h._ctx, hydro._dptr(T), hydro._dptr(nan_L),
                                                 len(T)) == TLQ_ERR_ARG
                assert lib.tlq_set_point_masses(h._ctx, hydro._dptr(many), 1024) == TLQ_ERR_ARG
                assert lib.tlq_set_point_masses(h._ctx, hydro._dptr(nan_m),
                                                len(p.masses)) == TLQ_ERR_ARG
                assert lib.tlq_set_point_masses(h._ctx, hydro._dptr(m), len(p.masses)) == 0
            h.set_state(u0)
            h.advance(3)
            return h.download("U")

    assert np.array_equal(run(True), run(False))
