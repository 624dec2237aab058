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
