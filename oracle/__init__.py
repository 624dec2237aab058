"""FP64 CPU oracle of the tlaloque HydroCore time step (ctypes wrapper).

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package, and only as the
checker (or the timed CPU baseline); the product path in ``tlaloque_amd``
never calls it.

The C library (``oracle/tlq_oracle.c``) restates the reference routines
(source/HydroCore.f90, HLLC.f90, HLL.f90, boundaries.f90, cooling.f90,
winds.f90, pointsource_gravity.f90, sources.f90:872-897, mainHD.f90:61-100);
each C function cites the lines it follows.  Arrays are numpy float64 in the
reference's Fortran order (e, i, j, k), i.e. C shape (nz+4, ny+4, nx+4, 5).

Pinning (what checks this oracle, beyond restating it):
  * the compiled reference (``oracle/_ref``, built by ``oracle/ref_build.py``
    with flang from the reference sources) -> committed fixtures in
    ``tests/golden/ref_*.npz``: calc_prims, time_step, HLLC choice 1/2
    (including the cell-centred-U star-state quirk), Godunov, boundariesI/II,
    step, apply_cooling, imposeSphericalWind and the mainHD loop order;
  * ``solutions/sod_e1rpex.out`` through the exact Riemann solver
    (``oracle/riemann.py``) and Sod runs at increasing resolution;
  * closed forms and invariants: conservation with periodic boundaries,
    consistency F(U,U) = F(U), limiter TVD bounds, cooling bounds;
  * HLL (source/HLL.f90 does not compile as shipped: ``use HydroCore`` of a
    non-module) and the point-mass source: Toro's HLL formula and the
    analytic potential gradient only -- no reference run.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libtlq_oracle.so")
_lib = None


class OrcWind(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("xc", "yc", "zc", "vx", "vy", "vz", "radius", "mdot", "vinf", "temp", "mu")]


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xl", ctypes.c_double), ("xr", ctypes.c_double), ("yl", ctypes.c_double),
        ("yr", ctypes.c_double), ("zl", ctypes.c_double), ("zr", ctypes.c_double),
        ("gamma", ctypes.c_double), ("mu0", ctypes.c_double), ("cfl", ctypes.c_double),
        ("eta", ctypes.c_double), ("cooling_limit", ctypes.c_double),
        ("limiter", ctypes.c_int), ("bc", ctypes.c_int * 6), ("solver", ctypes.c_int),
        ("cooling", ctypes.c_int), ("n_winds", ctypes.c_int), ("winds", ctypes.POINTER(OrcWind)),
        ("n_masses", ctypes.c_int), ("masses", ctypes.POINTER(ctypes.c_double)),
        ("softening", ctypes.c_double), ("n_table", ctypes.c_int),
        ("table_T", ctypes.POINTER(ctypes.c_double)), ("table_L", ctypes.POINTER(ctypes.c_double)),
    ]


class OrcSG(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("boundary", ctypes.c_int), ("gconst", ctypes.c_double),
                ("phi", ctypes.POINTER(ctypes.c_double)), ("work", ctypes.POINTER(ctypes.c_double)),
                ("it", ctypes.c_int), ("cycles", ctypes.c_int), ("dif", ctypes.c_double),
                ("tol", ctypes.c_double)]


def build(force: bool = False) -> str:
    """Compile the C oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(OrcParams)
        L.orc_calc_prims.argtypes = [pp, dp, dp]
        L.orc_time_step.argtypes = [pp, dp]
        L.orc_time_step.restype = ctypes.c_double
        L.orc_fluxes.argtypes = [pp, dp, dp, dp, dp, dp, ctypes.c_int]
        L.orc_godunov.argtypes = [pp, dp, dp, dp, dp, dp, dp, ctypes.c_double]
        L.orc_boundaries1.argtypes = [pp, dp]
        L.orc_boundaries2.argtypes = [pp, dp]
        L.orc_winds.argtypes = [pp, dp]
        L.orc_step.argtypes = [pp, dp, dp, dp, ctypes.c_double]
        L.orc_cooling.argtypes = [pp, dp, dp, ctypes.c_double]
        L.orc_find_coef.argtypes = [pp, ctypes.c_double]
        L.orc_find_coef.restype = ctypes.c_double
        L.orc_gravity_source.argtypes = [pp, dp, dp]
        L.orc_limiter_avg.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int]
        L.orc_hllc_face.argtypes = [pp, dp, dp, dp, dp, ctypes.c_int, dp]
        L.orc_hll_face.argtypes = [pp, dp, dp, dp, dp, ctypes.c_int, dp]
        L.orc_advance.argtypes = [pp, dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_advance.restype = ctypes.c_double
        sp = ctypes.POINTER(OrcSG)
        L.orc_sg_work_len.argtypes = [pp, ctypes.c_int]
        L.orc_sg_work_len.restype = ctypes.c_size_t
        L.orc_multigrid.argtypes = [pp, sp, dp]
        L.orc_multigrid.restype = ctypes.c_int
        L.orc_moments.argtypes = [pp, dp, dp, dp, dp]
        L.orc_phi_boundaries.argtypes = [pp, ctypes.c_double, dp, dp, dp, dp]
        L.orc_density.argtypes = [pp, dp, dp]
        L.orc_gravity_source_phi.argtypes = [pp, dp, dp, dp]
        L.orc_advance_sg.argtypes = [pp, dp, dp, dp, dp, dp, dp, dp, dp, sp, dp]
        L.orc_advance_sg.restype = ctypes.c_double
        L.orc_cool_subcycled.argtypes = [pp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int)]
        L.orc_cool_subcycled.restype = ctypes.c_double
        L.orc_cool_exact_int.argtypes = [pp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double]
        L.orc_cool_exact_int.restype = ctypes.c_double
        L.orc_mlp_nparams.restype = ctypes.c_int
        L.orc_mlp.argtypes = [ctypes.POINTER(ctypes.c_float), dp, dp, ctypes.c_int, dp, dp]
        L.orc_mlp.restype = ctypes.c_double
        L.orc_mlp_set_quant.argtypes = [ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


class Oracle:
    """Oracle state for one problem: U, UP, UPP, UPrim, F, G, H (and S when
    point masses are set), mirroring the globals of source/globals.f90:5-11."""

    def __init__(self, params, table=None):
        self.p = params
        self._keep = []
        op = OrcParams()
        for n in ("nx", "ny", "nz"):
            setattr(op, n, int(getattr(params, n)))
        for n in ("xl", "xr", "yl", "yr", "zl", "zr", "gamma", "mu0", "cfl", "eta",
                  "cooling_limit", "softening"):
            setattr(op, n, float(getattr(params, n)))
        op.limiter = int(params.limiter)
        for t in range(6):
            op.bc[t] = int(params.bc[t])
        op.solver = int(params.solver)
        op.cooling = 1 if int(params.cooling) != 0 else 0
        winds = list(params.winds)
        if winds:
            arr = (OrcWind * len(winds))(*[OrcWind(*w.as_tuple()) for w in winds])
            self._keep.append(arr)
            op.winds = ctypes.cast(arr, ctypes.POINTER(OrcWind))
        op.n_winds = len(winds)
        masses = list(params.masses)
        if masses:
            m = np.array([[q.x, q.y, q.z, q.mass] for q in masses], dtype=np.float64).ravel()
            self._keep.append(m)
            op.masses = _ptr(m)
        op.n_masses = len(masses)
        if table is not None:
            T = np.ascontiguousarray(table[0], dtype=np.float64)
            L = np.ascontiguousarray(table[1], dtype=np.float64)
            self._keep += [T, L]
            op.table_T, op.table_L, op.n_table = _ptr(T), _ptr(L), len(T)
        elif op.cooling:
            raise ValueError("cooling needs a table")
        self.op = op
        shape = (params.nz + 4, params.ny + 4, params.nx + 4, 5)
        self.shape = shape
        z = lambda: np.zeros(shape, dtype=np.float64)  # noqa: E731
        self.u, self.up, self.upp, self.prim = z(), z(), z(), z()
        self.ff, self.gg, self.hh = z(), z(), z()
        self.src = z() if (masses or getattr(params, "self_gravity", 0)) else None
        self.t = 0.0
        self.it = 0
        self.dt = 0.0
        self.sg = None
        if getattr(params, "self_gravity", 0):
            from tlaloque_amd.params import mg_levels
            lv = mg_levels(params)
            n = int(lib().orc_sg_work_len(self.P, lv))
            self.phi = np.zeros((params.nz + 2, params.ny + 2, params.nx + 2))
            self.work = np.zeros(n)
            self.dens = np.zeros_like(self.phi)
            sg = OrcSG()
            sg.levels, sg.boundary, sg.gconst = lv, int(params.self_gravity), float(params.gconst)
            sg.phi, sg.work, sg.it = _ptr(self.phi), _ptr(self.work), 0
            sg.tol = float(getattr(params, "mg_tol", 0.0))
            self.sg = sg

    @property
    def P(self):
        return ctypes.byref(self.op)

    def set_cons(self, u: np.ndarray) -> None:
        """initflow + borders(U,1) + calc_prims (source/mainHD.f90:32-44), and
        with self-gravity the first MultiGrid (mainHD.f90:50-52)."""
        self.u[...] = u
        L = lib()
        L.orc_boundaries1(self.P, _ptr(self.u))
        L.orc_winds(self.P, _ptr(self.u))
        L.orc_calc_prims(self.P, _ptr(self.prim), _ptr(self.u))
        if self.sg is not None:
            L.orc_density(self.P, _ptr(self.u), _ptr(self.dens))
            L.orc_multigrid(self.P, ctypes.byref(self.sg), _ptr(self.dens))

    def set_phi(self, phi: np.ndarray) -> None:
        """PHI as the user's initial conditions set it (its ghost faces are the
        boundary values), shape (nz+2, ny+2, nx+2)."""
        self.phi[...] = phi

    def multigrid(self, dens: np.ndarray) -> int:
        """One MultiGrid(dens, PHI) call at the current step count; returns the
        V-cycle count."""
        self.sg.it = self.it
        d = np.ascontiguousarray(dens, dtype=np.float64)
        return int(lib().orc_multigrid(self.P, ctypes.byref(self.sg), _ptr(d)))

    def moments(self, rho: np.ndarray):
        """Moments (sources.f90:599-676) with the MPI build's sums:
        c0[0..4], cm[l-1, m-1], sm[l-1, m-1]."""
        c0, cm, sm = np.zeros(5), np.zeros((4, 4)), np.zeros((4, 4))
        lib().orc_moments(self.P, _ptr(np.ascontiguousarray(rho, dtype=np.float64)), _ptr(c0),
                          _ptr(cm), _ptr(sm))
        return c0, cm, sm

    def phi_boundaries(self, gconst: float, c0, cm, sm, phi: np.ndarray) -> np.ndarray:
        """phi_[xyz]boundaries (sources.f90:678-787) on the six faces, in place."""
        lib().orc_phi_boundaries(self.P, float(gconst), _ptr(np.ascontiguousarray(c0)),
                                 _ptr(np.ascontiguousarray(cm)), _ptr(np.ascontiguousarray(sm)),
                                 _ptr(phi))
        return phi

    def advance(self) -> float:
        """One iteration of source/mainHD.f90:74-99; returns dt."""
        if self.sg is not None:
            self.sg.it = self.it
            dt = lib().orc_advance_sg(self.P, _ptr(self.u), _ptr(self.up), _ptr(self.upp),
                                      _ptr(self.prim), _ptr(self.ff), _ptr(self.gg),
                                      _ptr(self.hh), _ptr(self.src), ctypes.byref(self.sg),
                                      _ptr(self.dens))
        else:
            dt = lib().orc_advance(self.P, _ptr(self.u), _ptr(self.up), _ptr(self.upp),
                                   _ptr(self.prim), _ptr(self.ff), _ptr(self.gg), _ptr(self.hh),
                                   _ptr(self.src))
        self.dt = dt
        self.t += dt
        self.it += 1
        return dt

    # single routines --------------------------------------------------------
    def calc_prims(self, uu: np.ndarray) -> np.ndarray:
        out = np.zeros_like(uu)
        lib().orc_calc_prims(self.P, _ptr(out), _ptr(np.ascontiguousarray(uu)))
        return out

    def time_step(self, prim: np.ndarray) -> float:
        return lib().orc_time_step(self.P, _ptr(np.ascontiguousarray(prim)))

    def fluxes(self, uu: np.ndarray, prim: np.ndarray, choice: int):
        f, g, h = (np.zeros_like(uu) for _ in range(3))
        lib().orc_fluxes(self.P, _ptr(np.ascontiguousarray(uu)), _ptr(np.ascontiguousarray(prim)),
                         _ptr(f), _ptr(g), _ptr(h), int(choice))
        return f, g, h

    def godunov(self, uu, f, g, h, dt, src=None, out=None):
        out = np.zeros_like(uu) if out is None else out
        lib().orc_godunov(self.P, _ptr(out), _ptr(uu), _ptr(f), _ptr(g), _ptr(h), _ptr(src),
                          float(dt))
        return out

    def boundaries(self, u: np.ndarray, order: int) -> np.ndarray:
        """borders(U, order) (source/boundaries.f90:1-26): fill + winds, in place."""
        L = lib()
        (L.orc_boundaries1 if order == 1 else L.orc_boundaries2)(self.P, _ptr(u))
        L.orc_winds(self.P, _ptr(u))
        return u

    def cooling(self, prim: np.ndarray, u: np.ndarray, dt: float) -> None:
        lib().orc_cooling(self.P, _ptr(prim), _ptr(u), float(dt))

    def find_coef(self, temp: float) -> float:
        return lib().orc_find_coef(self.P, float(temp))

    def gravity_source(self, prim: np.ndarray) -> np.ndarray:
        s = np.zeros_like(prim)
        lib().orc_gravity_source(self.P, _ptr(np.ascontiguousarray(prim)), _ptr(s))
        return s

    def cool_subcycled(self, n: float, T: float, dt: float, eps: float = 0.01,
                       max_sub: int = 100000):
        """Cooling integrated to convergence; returns (T_new, substeps)."""
        ns = ctypes.c_int()
        T1 = lib().orc_cool_subcycled(self.P, float(n), float(T), float(dt), float(eps),
                                      int(max_sub), ctypes.byref(ns))
        return T1, ns.value

    def cool_exact_int(self, n: float, T: float, dt: float, gate_frac: float = 0.995) -> float:
        """Cooling integrated exactly over dt (closed form on the power-law
        table segments); gate_frac * T_min when the gate is reached."""
        return lib().orc_cool_exact_int(self.P, float(n), float(T), float(dt), float(gate_frac))

    def hllc_face(self, wl, wr, ul, ur, dim, solver=None):
        f = np.full(5, np.nan)
        args = [np.ascontiguousarray(a, dtype=np.float64) for a in (wl, wr, ul, ur)]
        fn = lib().orc_hll_face if (solver if solver is not None else self.p.solver) == 1 \
            else lib().orc_hllc_face
        fn(self.P, *[_ptr(a) for a in args], int(dim), _ptr(f))
        return f


def limiter_avg(lim: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty_like(a)
    lib().orc_limiter_avg(int(lim), _ptr(a), _ptr(b), _ptr(out), a.size)
    return out


def mlp(theta: np.ndarray, x: np.ndarray, y=None, grad: bool = False, quant: bool = True):
    """The cooling surrogate network in FP64 (oracle of the GPU kernels).

    theta: float32 parameters; x: (B, 3) inputs (log10 n, log10 T,
    log10 dt/t_cool); y: targets.  Returns (out, loss, grad)."""
    L = lib()
    L.orc_mlp_set_quant(1 if quant else 0)
    th = np.ascontiguousarray(theta, dtype=np.float32)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
    B = xx.shape[0]
    out = np.empty(B)
    yy = None if y is None else np.ascontiguousarray(y, dtype=np.float64)
    g = np.empty(L.orc_mlp_nparams()) if grad else None
    loss = L.orc_mlp(th.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _ptr(xx), _ptr(yy), B,
                     _ptr(out), _ptr(g))
    L.orc_mlp_set_quant(1)
    return out, loss, g
