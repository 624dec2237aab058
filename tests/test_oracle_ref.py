"""The C oracle against the reference's own runs (CPU).

Each fixture holds UPrim of the reference, built from its sources with
flang (oracle/ref_build.py), after 0, 1 and N steps of a seeded problem.  The
oracle restates the same routines with the same operation order, so it must
reproduce them bit for bit (cooling uses glibc exp/log10/pow as the
reference does).
"""
import numpy as np
import pytest

import oracle
from helpers import cooling_table, load_ref, ref_cases

CASES = ref_cases()


def test_fixture_set_is_complete():
    # every limiter, every boundary type, cooling, winds, winds+cooling
    for name in ("hydro_outflow", "bc_mixed", "periodic", "limiter1", "limiter3", "limiter4",
                 "limiter5", "limiter6", "cooling", "winds", "winds_cooling"):
        assert name in CASES


@pytest.mark.parametrize("case", CASES)
def test_oracle_reproduces_reference(case):
    p, u0, prims, steps = load_ref(case)
    o = oracle.Oracle(p, table=cooling_table() if p.cooling else None)
    o.set_cons(u0)
    sl = (slice(1, -1),) * 3  # interior and the first ghost layer
    assert np.array_equal(o.prim[sl], prims[0][sl])
    for s in range(1, steps[-1] + 1):
        o.advance()
        if s in steps:
            got = o.prim[sl]
            ref = prims[steps.index(s)][sl]
            assert np.array_equal(got, ref), (case, s, float(np.abs(got - ref).max()))
