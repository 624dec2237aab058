"""CPU checks of the oracle's cooling routines and of the surrogate network oracle.

What pins what:
  * find_coef reproduces Z1.0.dat at its nodes; apply_cooling is pinned by the
    reference fixtures (test_oracle_ref.py: cooling, winds_cooling).
  * The converged integrator with one substep is apply_cooling (compared with
    the fixture-pinned orc_cooling), and as eps -> 0 it converges (first
    order) to the exact solution of dT/dt = -(gamma-1) n Lambda(T)/k_B, whose
    time integral t = k_B/((gamma-1) n) int_{T1}^{T0} dT/Lambda(T) is
    evaluated independently by quadrature.
  * The exact integration (closed form on the power-law table segments) is
    pinned by the same quadrature of 1/Lambda, by the analytic solution for a
    table that is one power law, by its gate and above-table cases, and as
    the eps -> 0 limit of the converged integrator.
  * The network oracle's backpropagation equals central finite differences of
    its forward pass (FP64, without bf16 rounding).
"""
import numpy as np
import pytest
from scipy import integrate, optimize

import oracle
from helpers import cooling_table
from tlaloque_amd import params as P


def _cool_problem():
    p = P.HydroParams(nx=4, ny=4, nz=4, cooling=P.COOL_EXACT)
    return p, oracle.Oracle(p, table=cooling_table())


def test_find_coef_reproduces_table_nodes():
    _, o = _cool_problem()
    T, L = cooling_table()
    for q in range(0, len(T) - 1, 7):
        assert o.find_coef(T[q]) == pytest.approx(10.0 ** L[q], rel=1e-12)


def test_apply_cooling_bounds():
    p, o = _cool_problem()
    rng = np.random.default_rng(5)
    shape = p.shape + (5,)
    n = 10 ** rng.uniform(-2, 3, p.shape)
    T = 10 ** rng.uniform(3, 8.9, p.shape)
    prim = np.zeros(shape)
    prim[..., 0] = n * P.AMU
    prim[..., 1:4] = rng.normal(0, 1e6, p.shape + (3,))
    prim[..., 4] = n * P.KB * T
    u = np.zeros(shape)
    u[..., 0] = prim[..., 0]
    u[..., 1:4] = prim[..., :1] * prim[..., 1:4]
    u[..., 4] = 0.5 * prim[..., 0] * (prim[..., 1:4] ** 2).sum(-1) + prim[..., 4] / (p.gamma - 1)
    before = prim.copy()
    o.cooling(prim, u, 3e12)
    sl = (slice(2, -2),) * 3
    T0 = (before[..., 4] / (n * P.KB))[sl]
    T1 = (prim[..., 4] / (n * P.KB))[sl]
    cold = T0 < 1e4
    assert np.array_equal(T1[cold], T0[cold])                       # gate (cooling.f90:71)
    assert np.all(T1[~cold] <= T0[~cold] * (1 + 1e-14))
    assert np.all(T1[~cold] >= np.maximum(0.5 * T0[~cold], 10.0) * (1 - 1e-14))
    assert (T1[~cold] < 0.99 * T0[~cold]).sum() > 10                 # it does cool


def test_subcycled_with_one_substep_is_apply_cooling():
    p, o = _cool_problem()
    rng = np.random.default_rng(6)
    for _ in range(200):
        n = 10 ** rng.uniform(-3, 4)
        T = 10 ** rng.uniform(4, 8.9)
        dt = 10 ** rng.uniform(8, 14)
        T1, ns = o.cool_subcycled(n, T, dt, eps=1e300)
        assert ns == 1
        prim = np.zeros(p.shape + (5,))
        prim[..., 0] = n * p.mu0 * P.AMU
        prim[..., 4] = n * P.KB * T
        u = np.zeros_like(prim)
        u[..., 0] = prim[..., 0]
        u[..., 4] = prim[..., 4] / (p.gamma - 1)
        o.cooling(prim, u, dt)
        Tref = prim[3, 3, 3, 4] / (n * P.KB)
        assert T1 == pytest.approx(Tref, rel=1e-13)


def test_subcycled_converges_to_exact_cooling_integral():
    p, o = _cool_problem()
    lam = lambda T: o.find_coef(T)  # noqa: E731
    g1 = p.gamma - 1.0
    for n, T0, frac in [(1.0, 3e6, 0.5), (50.0, 2e5, 0.3), (0.1, 1e8, 0.8), (5.0, 3e4, 0.6)]:
        # time to cool from T0 to T0*frac, by quadrature of 1/Lambda
        Tt = T0 * frac
        grid = np.geomspace(Tt, T0, 4001)
        t_exact = P.KB / (g1 * n) * integrate.simpson(1.0 / np.array([lam(T) for T in grid]), x=grid)
        errs = []
        for eps in (0.02, 0.002):
            T1, ns = o.cool_subcycled(n, T0, t_exact, eps=eps)
            errs.append(abs(T1 / Tt - 1.0))
            assert ns > 1
        assert errs[0] < 2e-2 and errs[1] < 2e-3
        assert errs[1] < 0.3 * errs[0]                                # first-order convergence


def test_exact_integration_matches_quadrature():
    p, o = _cool_problem()
    lam = lambda T: o.find_coef(T)  # noqa: E731
    g1 = p.gamma - 1.0
    for n, T0, frac in [(1.0, 3e6, 0.5), (50.0, 2e5, 0.3), (0.1, 1e8, 0.8), (5.0, 3e4, 0.6),
                        (1e3, 1.2e4, 0.9), (1e-2, 7e8, 0.2)]:
        Tt = T0 * frac
        # quadrature on the table's own nodes (Lambda is smooth between them)
        T, _ = cooling_table()
        nodes = np.concatenate(([Tt], [t for t in T if Tt < t < T0], [T0]))
        t_exact = 0.0
        for a, b in zip(nodes[:-1], nodes[1:]):
            grid = np.geomspace(a, b, 201)
            t_exact += integrate.simpson(1.0 / np.array([lam(x) for x in grid]), x=grid)
        t_exact *= P.KB / (g1 * n)
        assert o.cool_exact_int(n, T0, t_exact) == pytest.approx(Tt, rel=1e-7)
        assert o.cool_exact_int(n, T0, 0.0) == pytest.approx(T0, rel=1e-12)


def test_exact_integration_of_one_power_law():
    # a table that is one power law, Lambda = L0 (T/T0)^alpha: the cooling
    # equation integrates to T(t) = T0 (1 - (1-alpha) t / t_c)^(1/(1-alpha)),
    # t_c = k_B T0 / ((gamma-1) n L0)
    for alpha in (-0.7, 0.5, 1.0, 1.8):
        lg = np.linspace(4.0, 8.0, 41)
        L = -22.0 + alpha * (lg - 4.0)
        p = P.HydroParams(nx=4, ny=4, nz=4, cooling=P.COOL_EXACT)
        o = oracle.Oracle(p, table=(list(10.0 ** lg), list(L)))
        n, T0 = 3.0, 10 ** 6.3
        L0 = 10.0 ** (-22.0 + alpha * (6.3 - 4.0))
        tc = P.KB * T0 / ((p.gamma - 1.0) * n * L0)
        for f in (0.01, 0.3, 0.9):
            t = f * tc
            if alpha == 1.0:
                want = T0 * np.exp(-t / tc)
            else:
                base = 1.0 - (1.0 - alpha) * t / tc
                if base <= 0.0:                 # cools to zero within t: the gate case
                    continue
                want = T0 * base ** (1.0 / (1.0 - alpha))
            if want <= 1e4:
                continue
            assert o.cool_exact_int(n, T0, t) == pytest.approx(want, rel=1e-11), (alpha, f)


def test_exact_integration_gate_and_above_table():
    p, o = _cool_problem()
    T, L = cooling_table()
    g1 = p.gamma - 1.0
    n, T0 = 10.0, 5e5
    # time to the gate, by the closed form's own complement: just above it the
    # gas stops at gate_frac * T_min, just below it stays above T_min
    lo, hi = 1e6, 1e16
    for _ in range(200):
        mid = np.sqrt(lo * hi)
        if o.cool_exact_int(n, T0, mid) <= T[0]:
            hi = mid
        else:
            lo = mid
    assert o.cool_exact_int(n, T0, hi * 1.001) == pytest.approx(0.995 * T[0], rel=1e-14)
    assert o.cool_exact_int(n, T0, lo * 0.999) > T[0]
    assert o.cool_exact_int(n, 0.5 * T[0], 1e20) == 0.5 * T[0]      # below the gate: untouched
    # above the table Lambda is the last row's: dT/dt is constant
    Ttop, lam = T[-1], 10.0 ** L[-1]
    T0 = 3.0 * Ttop
    rate = g1 * n * lam / P.KB
    dt = 0.5 * (T0 - Ttop) / rate
    assert o.cool_exact_int(n, T0, dt) == pytest.approx(T0 - rate * dt, rel=1e-12)


def test_exact_integration_is_the_converged_limit():
    p, o = _cool_problem()
    rng = np.random.default_rng(11)
    for _ in range(40):
        n = 10 ** rng.uniform(-2, 3)
        T0 = 10 ** rng.uniform(4.2, 8.4)
        Tm = 0.5 * T0
        tc = P.KB * T0 / ((p.gamma - 1.0) * n * o.find_coef(T0))
        dt = tc * 10 ** rng.uniform(-1.5, 0.7)
        ex = o.cool_exact_int(n, T0, dt)
        e1 = abs(o.cool_subcycled(n, T0, dt, eps=0.02)[0] / ex - 1.0)
        e2 = abs(o.cool_subcycled(n, T0, dt, eps=0.002)[0] / ex - 1.0)
        if ex <= 1e4:                       # gated: the integrator ends within 1% below T_min
            assert e2 <= 6e-3
            continue
        assert e2 < 1e-2, (n, T0, dt, Tm)
        assert e2 < 0.2 * e1 + 1e-9                                     # first order


def test_network_oracle_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    th = (rng.standard_normal(4481) * 0.2).astype(np.float32)
    x = np.stack([rng.uniform(-3, 4, 32), rng.uniform(4, 9, 32), rng.uniform(-2, 2, 32)], 1)
    y = rng.uniform(-1, 0, 32)
    _, _, g = oracle.mlp(th, x, y, grad=True, quant=False)
    for idx in rng.choice(4481, 40, replace=False).tolist() + [0, 192, 256, 4352, 4416, 4480]:
        h = max(1e-2 * abs(float(th[idx])), 1e-3)
        tp, tm = th.copy(), th.copy()
        tp[idx] += np.float32(h)
        tm[idx] -= np.float32(h)
        lp = oracle.mlp(tp, x, y, quant=False)[1]
        lm = oracle.mlp(tm, x, y, quant=False)[1]
        fd = (lp - lm) / (float(tp[idx]) - float(tm[idx]))
        assert g[idx] == pytest.approx(fd, rel=2e-4, abs=1e-7), idx


def test_network_oracle_quantised_forward_close_to_fp64():
    rng = np.random.default_rng(1)
    th = (rng.standard_normal(4481) * 0.2).astype(np.float32)
    x = np.stack([rng.uniform(-3, 4, 256), rng.uniform(4, 9, 256), rng.uniform(-2, 2, 256)], 1)
    a = oracle.mlp(th, x)[0]
    b = oracle.mlp(th, x, quant=False)[0]
    assert np.abs(a - b).max() <= 0.05 * np.abs(b).max()
    assert np.abs(a - b).max() > 0.0                                 # the rounding is applied
