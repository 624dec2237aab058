"""Exact solution of the 1D Riemann problem for the ideal-gas Euler equations.

TEST INFRASTRUCTURE ONLY (the analytic oracle for Sod shock tubes).

Follows the published exact solver of E. F. Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", 3rd ed., Springer 2009, chapter 4:
pressure function f_K(p) (eqs. 4.6-4.7), Newton iteration for the star
pressure p* (eq. 4.44 start value, two-rarefaction guess), star velocity
(eq. 4.9), and sampling of the wave pattern (section 4.5).

The reference's harness compares its runs with ``solutions/sod_e1rpex.out``
(tests_analisis.jl:77-89), a tabulated output of such a solver; the test
suite pins this module against that file.
"""
from __future__ import annotations

import numpy as np


def _fK(p, rho, pK, cK, g):
    """Pressure function f_K and its derivative (Toro eqs. 4.6, 4.7, 4.37)."""
    if p > pK:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pK
        s = np.sqrt(A / (p + B))
        f = (p - pK) * s
        df = s * (1.0 - 0.5 * (p - pK) / (B + p))
    else:  # rarefaction
        pr = p / pK
        f = 2.0 * cK / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rho * cK) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g, tol=1e-14, itmax=200):
    """(p*, u*) of the Riemann problem (Toro section 4.3)."""
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    if 2.0 * (cl + cr) / (g - 1.0) <= ur - ul:
        raise ValueError("vacuum is generated")
    # two-rarefaction approximation as the start value (eq. 4.46)
    z = (g - 1.0) / (2.0 * g)
    p = ((cl + cr - 0.5 * (g - 1.0) * (ur - ul)) / (cl / pl ** z + cr / pr ** z)) ** (1.0 / z)
    p = max(p, 1e-14)
    for _ in range(itmax):
        fl, dfl = _fK(p, rl, pl, cl, g)
        fr, dfr = _fK(p, rr, pr, cr, g)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-14)
        if abs(pn - p) / (0.5 * (pn + p)) < tol:
            p = pn
            break
        p = pn
    fl, _ = _fK(p, rl, pl, cl, g)
    fr, _ = _fK(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, left, right, g):
    """Density, velocity and pressure at positions x (array) at time t for a
    diaphragm at x0 (Toro section 4.5, Fig. 4.14)."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    x = np.asarray(x, dtype=np.float64)
    s = (x - x0) / t
    rho = np.empty_like(s)
    u = np.empty_like(s)
    p = np.empty_like(s)
    gm = (g - 1.0) / (g + 1.0)
    for n, S in enumerate(s):
        if S <= us:  # left of the contact
            if ps > pl:  # left shock
                SL = ul - cl * np.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
                if S <= SL:
                    rho[n], u[n], p[n] = rl, ul, pl
                else:
                    rho[n] = rl * (ps / pl + gm) / (gm * ps / pl + 1.0)
                    u[n], p[n] = us, ps
            else:  # left rarefaction
                SHL = ul - cl
                cml = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                STL = us - cml
                if S <= SHL:
                    rho[n], u[n], p[n] = rl, ul, pl
                elif S >= STL:
                    rho[n] = rl * (ps / pl) ** (1.0 / g)
                    u[n], p[n] = us, ps
                else:
                    c = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * (ul - S))
                    rho[n] = rl * (c / cl) ** (2.0 / (g - 1.0))
                    u[n] = 2.0 / (g + 1.0) * (cl + 0.5 * (g - 1.0) * ul + S)
                    p[n] = pl * (c / cl) ** (2.0 * g / (g - 1.0))
        else:  # right of the contact
            if ps > pr:  # right shock
                SR = ur + cr * np.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
                if S >= SR:
                    rho[n], u[n], p[n] = rr, ur, pr
                else:
                    rho[n] = rr * (ps / pr + gm) / (gm * ps / pr + 1.0)
                    u[n], p[n] = us, ps
            else:  # right rarefaction
                SHR = ur + cr
                cmr = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                STR = us + cmr
                if S >= SHR:
                    rho[n], u[n], p[n] = rr, ur, pr
                elif S <= STR:
                    rho[n] = rr * (ps / pr) ** (1.0 / g)
                    u[n], p[n] = us, ps
                else:
                    c = 2.0 / (g + 1.0) * (cr - 0.5 * (g - 1.0) * (ur - S))
                    rho[n] = rr * (c / cr) ** (2.0 / (g - 1.0))
                    u[n] = 2.0 / (g + 1.0) * (-cr + 0.5 * (g - 1.0) * ur + S)
                    p[n] = pr * (c / cr) ** (2.0 * g / (g - 1.0))
    return rho, u, p
