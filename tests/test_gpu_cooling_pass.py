"""GPU parity of the separate cooling pass: the substepped integrator (cooling
mode TLQ_COOL_SUBCYCLED) and the exact integration (TLQ_COOL_INTEGRATED).

The fused step with mode 3 is the reference's step (mainHD.f90:74-95) followed
by cooling integrated to convergence: cells with dt <= 0.01 t_cool take one
substep -- apply_cooling's formula, cooling.f90:67-112 -- and the others the
substepped integrator.  The oracle composes the same from its own pieces: one
orc_advance with cooling off, then orc_cool_subcycled per interior cell, then
calc_prims (the next step's time_step reads the cooled state).  The seeded
state of the reference's cooling fixture spans dt/t_cool from 0 to about 5 in
the first step, so both branches run.
"""
import numpy as np
import pytest

import oracle
from helpers import cooling_table, interior, l1_rel, load_ref, max_rel, prims_from_cons
from tlaloque_amd import params as P
from tlaloque_amd.hydro import HydroCore

pytestmark = pytest.mark.gpu


def _oracle_run(p, u0, steps, exact=False):
    """Oracle composition of the step with cooling integrated to convergence
    (exact=False: orc_cool_subcycled; exact=True: cells with dt <= 0.01
    t_cool take apply_cooling's one step, the others orc_cool_exact_int)."""
    o = oracle.Oracle(p.replace(cooling=P.COOL_OFF), table=cooling_table())
    o.set_cons(u0)
    nsub, nexact = [], 0
    for _ in range(steps):
        dt = o.advance()
        u = o.u[2:-2, 2:-2, 2:-2]
        w = interior(o.prim)
        nd = w[..., 0] / (p.mu0 * P.AMU)
        T = w[..., 4] * p.mu0 * P.AMU / (w[..., 0] * P.KB)
        ek = 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2) / u[..., 0]
        for idx in zip(*np.nonzero(T >= 1e4)):
            n, T0 = float(nd[idx]), float(T[idx])
            if exact:
                eth = w[idx + (4,)] / (p.gamma - 1.0)
                cdt = o.find_coef(T0) * n * n / eth * dt
                if cdt > 0.01:
                    T1 = o.cool_exact_int(n, T0, dt)
                    nexact += 1
                else:
                    T1, _ = o.cool_subcycled(n, T0, dt, eps=1e300)
            else:
                T1, ns = o.cool_subcycled(n, T0, dt, eps=0.01)
                nsub.append(ns)
            u[idx + (4,)] = ek[idx] + w[idx + (4,)] * (T1 / T0) / (p.gamma - 1.0)
        o.prim[...] = o.calc_prims(o.u)
    return o.u, (nexact if exact else max(nsub))


@pytest.mark.parametrize("precision", [P.PREC_FP64, P.PREC_FP32])
def test_subcycled_pass_matches_oracle(precision):
    p, u0, _, _ = load_ref("cooling")
    p = p.replace(cooling=P.COOL_SUBCYCLED, precision=precision)
    steps = 3
    ref, nsub = _oracle_run(p, u0, steps)
    assert nsub > 50                      # the substepped branch is exercised
    with HydroCore(p, cooling_table=cooling_table()) as h:
        h.set_state(u0)
        h.advance(steps)
        got = h.download("U")
        assert h.check_state() == 0
    w, wr = interior(prims_from_cons(got, p.gamma)), interior(prims_from_cons(ref, p.gamma))
    if precision == P.PREC_FP64:
        # device exp/log differ from glibc in the last bits; the substep loop
        # compounds them over up to a few hundred substeps
        assert max_rel(w, wr) <= 1e-9
    else:
        # FP32 hydro: the bar of the fused FP32 step (L1 of rho and p)
        assert l1_rel(w[..., 0], wr[..., 0]) <= 1e-3
        assert l1_rel(w[..., 4], wr[..., 4]) <= 1e-3


@pytest.mark.parametrize("precision", [P.PREC_FP64, P.PREC_FP32])
def test_integrated_cooling_matches_oracle(precision):
    """Cooling mode TLQ_COOL_INTEGRATED (the exact integration in the dense
    pass) against the oracle's composition on the reference's cooling
    fixture, where dt/t_cool reaches about 5 in the first step."""
    p, u0, _, _ = load_ref("cooling")
    p = p.replace(cooling=P.COOL_INTEGRATED, precision=precision)
    steps = 3
    ref, nexact = _oracle_run(p, u0, steps, exact=True)
    assert nexact > 50                    # the exact integration is exercised
    with HydroCore(p, cooling_table=cooling_table()) as h:
        h.set_state(u0)
        h.advance(steps)
        got = h.download("U")
        assert h.check_state() == 0
    w, wr = interior(prims_from_cons(got, p.gamma)), interior(prims_from_cons(ref, p.gamma))
    print(f"integrated cooling, {precision}: {nexact} cells integrated; max rel {max_rel(w, wr):.2e}, "
          f"L1 rho {l1_rel(w[..., 0], wr[..., 0]):.2e} p {l1_rel(w[..., 4], wr[..., 4]):.2e}")
    if precision == P.PREC_FP64:
        assert max_rel(w, wr) <= 1e-9         # device pow/log/exp against glibc's
    else:
        assert l1_rel(w[..., 0], wr[..., 0]) <= 1e-3
        assert l1_rel(w[..., 4], wr[..., 4]) <= 1e-3


def _refined_table(temps, logl):
    """The table's rows plus one more inside every segment, at a fraction of
    the segment's log T that varies from 0.25 to 0.75, on the segment's power
    law: the same Lambda(T) as the table, in 1001 rows whose log T spacing is
    not uniform (the rows are then found by bisection, not direct indexing)."""
    lt, ll = np.log10(np.asarray(temps)), np.asarray(logl)
    rows_t, rows_l = [lt[0]], [ll[0]]
    for i in range(1, len(lt)):
        f = 0.25 + 0.05 * ((7 * i) % 11)
        rows_t += [lt[i - 1] + f * (lt[i] - lt[i - 1]), lt[i]]
        rows_l += [ll[i - 1] + f * (ll[i] - ll[i - 1]), ll[i]]
    return list(10.0 ** np.array(rows_t)), rows_l


@pytest.mark.parametrize("mode", [P.COOL_SUBCYCLED, P.COOL_INTEGRATED])
def test_dense_pass_large_nonuniform_table(mode):
    """The FP32 dense pass stages the table in LDS (32 B per row).  With the
    1001-row refinement of Z1.0 (31 KB of LDS, rows found by bisection) the
    cooled states equal those with Z1.0 itself (501 rows, direct indexing) up
    to the FP32 rounding of the inserted rows' log Lambda: measured L1 of rho
    and p 2.7e-9 and 2.8e-8 (mode 3), 2.5e-8 and 1.7e-7 (mode 4) after three
    steps; the bounds are about ten times those."""
    p, u0, _, _ = load_ref("cooling")
    p = p.replace(cooling=mode, precision=P.PREC_FP32)
    base = cooling_table()
    fine = _refined_table(*base)
    assert len(fine[0]) == 1001
    out = []
    for table in (base, fine):
        with HydroCore(p, cooling_table=table) as h:
            h.set_state(u0)
            h.advance(3)
            queued, _ = h.cooling_counts()
            out.append(interior(prims_from_cons(h.download("U"), p.gamma)))
            assert h.check_state() == 0
        assert queued > 50                      # the dense pass ran
    a, b = out
    l1r, l1p = l1_rel(b[..., 0], a[..., 0]), l1_rel(b[..., 4], a[..., 4])
    print(f"mode {mode}: 1001-row table vs 501 rows: L1 rho {l1r:.2e} p {l1p:.2e}, "
          f"max rel {max_rel(b, a):.2e}")
    bound_r, bound_p = {P.COOL_SUBCYCLED: (3e-8, 3e-7), P.COOL_INTEGRATED: (3e-7, 2e-6)}[mode]
    assert l1r <= bound_r and l1p <= bound_p
