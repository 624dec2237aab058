"""Python host mirror of the reference's HydroCore interface over the C-ABI.

``HydroCore`` owns one ``tlq_ctx`` (include/tlq.h) and exposes the routines
the reference's main loop calls (source/mainHD.f90:61-100) under their
reference names:

    calc_prims(src)          source/HydroCore.f90:1-33
    time_step()              source/HydroCore.f90:113-155
    fluxes(src, choice)      source/HLLC.f90:64-496 (HLL.f90:71-297 with solver=HLL)
    godunov(dst, dt)         source/HLLC.f90:1-62
    borders(buf, order)      source/boundaries.f90:1-26 (+ user_boundaries winds)
    step(t, it, dt)          source/HydroCore.f90:170-211
    cooling(dt)              source/cooling.f90:9-19

plus ``advance`` (the fused two-kernel step) and host <-> device copies in
the reference's Fortran layout (numpy C shape (nz+4, ny+4, nx+4, 5)).

Everything runs on the GPU through ``libtlq.so``; there is no CPU fallback.
A missing or unbuilt library raises ``TlqError``.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence, Tuple

import numpy as np

from . import params as P

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libtlq.so")

BUF_U, BUF_UP, BUF_UPP, BUF_UPRIM, BUF_F, BUF_G, BUF_H, BUF_S = range(8)
BUFFERS = {"U": BUF_U, "UP": BUF_UP, "UPP": BUF_UPP, "UPrim": BUF_UPRIM, "F": BUF_F, "G": BUF_G,
           "H": BUF_H, "S": BUF_S}
ERR_NAMES = {-1: "TLQ_ERR_ARG", -2: "TLQ_ERR_HIP", -3: "TLQ_ERR_NOMEM",
             -4: "TLQ_ERR_UNSUPPORTED", -5: "TLQ_ERR_STATE", -6: "TLQ_ERR_NONFINITE"}

# Every entry point include/tlq.h declares (checked by the CPU tests).
EXPORTS = (
    "tlq_create", "tlq_destroy", "tlq_last_error", "tlq_set_winds", "tlq_set_point_masses",
    "tlq_set_cooling_table", "tlq_upload", "tlq_download", "tlq_calc_prims", "tlq_time_step",
    "tlq_fluxes", "tlq_godunov", "tlq_borders", "tlq_step", "tlq_cooling", "tlq_init_state",
    "tlq_advance", "tlq_advance_n", "tlq_get_time", "tlq_sync", "tlq_check_state",
    "tlq_fallback_fraction", "tlq_cooling_counts", "tlq_resume", "tlq_upload_phi",
    "tlq_download_phi", "tlq_gravity_solve", "tlq_gravity_info", "tlq_timer_start", "tlq_timer_stop", "tlq_kernel_timing",
    "tlq_kernel_stats", "tlq_layout", "tlq_copy_probe", "tlq_copy_probe_config", "tlq_surrogate_init",
    "tlq_surrogate_set_params", "tlq_surrogate_get_params", "tlq_surrogate_generate",
    "tlq_surrogate_samples", "tlq_surrogate_train", "tlq_surrogate_eval", "tlq_surrogate_grad",
    "tlq_surrogate_grad_noalloc", "tlq_surrogate_set_threshold",
)
MLP_NPARAMS = 4481


class TlqError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERR_NAMES.get(code, code)}: {msg}")
        self.code = code


class TlqParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xl", ctypes.c_double), ("xr", ctypes.c_double), ("yl", ctypes.c_double),
        ("yr", ctypes.c_double), ("zl", ctypes.c_double), ("zr", ctypes.c_double),
        ("gamma", ctypes.c_double), ("mu0", ctypes.c_double), ("cfl", ctypes.c_double),
        ("eta", ctypes.c_double), ("cooling_limit", ctypes.c_double),
        ("limiter", ctypes.c_int), ("bc", ctypes.c_int * 6), ("solver", ctypes.c_int),
        ("precision", ctypes.c_int), ("cooling", ctypes.c_int), ("softening", ctypes.c_double),
        ("self_gravity", ctypes.c_int), ("mg_levels", ctypes.c_int), ("gconst", ctypes.c_double),
        ("mg_tol", ctypes.c_double),
    ]


class TlqWind(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("xc", "yc", "zc", "vx", "vy", "vz", "radius", "mdot", "vinf", "temp", "mu")]


_lib = None


def load_library(path: Optional[str] = None):
    """Load libtlq.so (built by tlaloque_amd.build); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise TlqError(-5, f"{path} is not built: run python -m tlaloque_amd.build")
    L = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "tlq_create": [ctypes.POINTER(TlqParams), ctypes.POINTER(vp)],
        "tlq_destroy": [vp],
        "tlq_set_winds": [vp, ctypes.POINTER(TlqWind), ip],
        "tlq_set_point_masses": [vp, dp, ip],
        "tlq_set_cooling_table": [vp, dp, dp, ip],
        "tlq_upload": [vp, ip, dp],
        "tlq_download": [vp, ip, dp],
        "tlq_calc_prims": [vp, ip],
        "tlq_time_step": [vp, dp],
        "tlq_fluxes": [vp, ip, ip],
        "tlq_godunov": [vp, ip, ctypes.c_double],
        "tlq_borders": [vp, ip, ip],
        "tlq_step": [vp, dp, ctypes.POINTER(ip), ctypes.c_double],
        "tlq_cooling": [vp, ctypes.c_double],
        "tlq_init_state": [vp],
        "tlq_advance": [vp, dp, dp, ctypes.POINTER(ip)],
        "tlq_advance_n": [vp, ip],
        "tlq_get_time": [vp, dp, dp, ctypes.POINTER(ip)],
        "tlq_sync": [vp],
        "tlq_check_state": [vp, ctypes.POINTER(ctypes.c_longlong)],
        "tlq_fallback_fraction": [vp, dp],
        "tlq_cooling_counts": [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)],
        "tlq_resume": [vp, ctypes.c_double, ip, ip],
        "tlq_upload_phi": [vp, dp],
        "tlq_download_phi": [vp, dp],
        "tlq_gravity_solve": [vp, ctypes.POINTER(ip), dp],
        "tlq_gravity_info": [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), dp, ctypes.POINTER(ip)],
        "tlq_timer_start": [vp],
        "tlq_timer_stop": [vp, dp],
        "tlq_kernel_timing": [vp, ip],
        "tlq_kernel_stats": [vp, ip, ctypes.c_char_p, ip, dp, ctypes.POINTER(ctypes.c_longlong)],
        "tlq_layout": [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ip),
                       ctypes.POINTER(ip)],
        "tlq_copy_probe": [vp, ctypes.c_longlong, ip, dp],
        "tlq_copy_probe_config": [vp, ctypes.c_longlong, ip, ip, ip, ip, dp],
        "tlq_surrogate_init": [vp, ctypes.c_ulonglong],
        "tlq_surrogate_set_params": [vp, fp, ip],
        "tlq_surrogate_get_params": [vp, fp, ip],
        "tlq_surrogate_generate": [vp, ctypes.c_longlong, ctypes.c_ulonglong],
        "tlq_surrogate_samples": [vp, ctypes.c_longlong, ctypes.c_longlong, fp, fp],
        "tlq_surrogate_train": [vp, ip, ip, ctypes.c_double, ctypes.c_double, dp],
        "tlq_surrogate_eval": [vp, fp, ctypes.c_longlong, fp],
        "tlq_surrogate_grad": [vp, fp, fp, ip, dp, dp],
        "tlq_surrogate_set_threshold": [vp, ctypes.c_double],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    L.tlq_last_error.argtypes = []
    L.tlq_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = _lib.tlq_last_error().decode() if _lib is not None else ""
        raise TlqError(rc, msg)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def to_c_params(p: P.HydroParams) -> TlqParams:
    c = TlqParams()
    for n in ("nx", "ny", "nz", "limiter", "solver", "precision", "cooling", "self_gravity",
              "mg_levels"):
        setattr(c, n, int(getattr(p, n)))
    for n in ("xl", "xr", "yl", "yr", "zl", "zr", "gamma", "mu0", "cfl", "eta", "cooling_limit",
              "softening", "gconst", "mg_tol"):
        setattr(c, n, float(getattr(p, n)))
    for q in range(6):
        c.bc[q] = int(p.bc[q])
    return c


class HydroCore:
    """One problem on the GPU (one ``tlq_ctx``)."""

    def __init__(self, params: P.HydroParams, cooling_table: Optional[Tuple[Sequence[float],
                                                                           Sequence[float]]] = None):
        L = load_library()
        self.p = params
        self._ctx = ctypes.c_void_p()
        cp = to_c_params(params)
        _check(L.tlq_create(ctypes.byref(cp), ctypes.byref(self._ctx)))
        if params.winds:
            arr = (TlqWind * len(params.winds))(*[TlqWind(*w.as_tuple()) for w in params.winds])
            _check(L.tlq_set_winds(self._ctx, arr, len(params.winds)))
        if params.masses:
            m = np.array([[q.x, q.y, q.z, q.mass] for q in params.masses], dtype=np.float64)
            m = np.ascontiguousarray(m.ravel())
            _check(L.tlq_set_point_masses(self._ctx, _dptr(m), len(params.masses)))
        if cooling_table is not None:
            T = np.ascontiguousarray(cooling_table[0], dtype=np.float64)
            Lg = np.ascontiguousarray(cooling_table[1], dtype=np.float64)
            _check(L.tlq_set_cooling_table(self._ctx, _dptr(T), _dptr(Lg), len(T)))
        self.shape = (params.nz + 4, params.ny + 4, params.nx + 4, 5)

    # lifecycle ---------------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            _check(_lib.tlq_destroy(self._ctx))
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # copies ------------------------------------------------------------------
    def upload(self, buf, host: np.ndarray) -> None:
        b = BUFFERS[buf] if isinstance(buf, str) else int(buf)
        a = np.ascontiguousarray(host, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"expected shape {self.shape}, got {a.shape}")
        _check(_lib.tlq_upload(self._ctx, b, _dptr(a)))

    def download(self, buf) -> np.ndarray:
        b = BUFFERS[buf] if isinstance(buf, str) else int(buf)
        out = np.empty(self.shape, dtype=np.float64)
        _check(_lib.tlq_download(self._ctx, b, _dptr(out)))
        return out

    # reference routines ----------------------------------------------------------
    def calc_prims(self, src="U") -> None:
        _check(_lib.tlq_calc_prims(self._ctx, BUFFERS[src] if isinstance(src, str) else src))

    def time_step(self) -> float:
        dt = ctypes.c_double()
        _check(_lib.tlq_time_step(self._ctx, ctypes.byref(dt)))
        return dt.value

    def fluxes(self, src="U", choice: int = 1) -> None:
        _check(_lib.tlq_fluxes(self._ctx, BUFFERS[src] if isinstance(src, str) else src, choice))

    def godunov(self, dst, dt: float) -> None:
        _check(_lib.tlq_godunov(self._ctx, BUFFERS[dst] if isinstance(dst, str) else dst, dt))

    def borders(self, buf, order: int) -> None:
        _check(_lib.tlq_borders(self._ctx, BUFFERS[buf] if isinstance(buf, str) else buf, order))

    def step(self, t: float, it: int, dt: float) -> Tuple[float, int]:
        tt, ii = ctypes.c_double(t), ctypes.c_int(it)
        _check(_lib.tlq_step(self._ctx, ctypes.byref(tt), ctypes.byref(ii), dt))
        return tt.value, ii.value

    def cooling(self, dt: float) -> None:
        _check(_lib.tlq_cooling(self._ctx, dt))

    # fused path ---------------------------------------------------------------------
    def set_state(self, u0: np.ndarray) -> None:
        """Upload U and run mainHD.f90:38-44 (borders(U,1), calc_prims, first CFL)."""
        self.upload("U", u0)
        _check(_lib.tlq_init_state(self._ctx))

    def advance(self, n: int = 1) -> None:
        _check(_lib.tlq_advance_n(self._ctx, int(n)))

    def time(self) -> Tuple[float, float, int]:
        """(dt of the last step, t, it) -- synchronises."""
        dt, t, it = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        _check(_lib.tlq_get_time(self._ctx, ctypes.byref(dt), ctypes.byref(t), ctypes.byref(it)))
        return dt.value, t.value, it.value

    def resume(self, u: np.ndarray, t: float, it: int, fill_ghosts: bool = False) -> None:
        """Restart from U (reference layout, with ghosts) at time t after `it`
        steps (tlq_upload + tlq_resume).  Keep the ghosts of a BIN dump
        (fill_ghosts=False); fill them for a VTK dump, which has none."""
        self.upload("U", u)
        _check(_lib.tlq_resume(self._ctx, float(t), int(it), int(fill_ghosts)))

    # self-gravity --------------------------------------------------------------
    @property
    def phi_shape(self) -> Tuple[int, int, int]:
        return (self.p.nz + 2, self.p.ny + 2, self.p.nx + 2)

    def upload_phi(self, phi: np.ndarray) -> None:
        """PHI(0:nx+1, 0:ny+1, 0:nz+1), C shape (nz+2, ny+2, nx+2); its ghost
        faces are the boundary values."""
        a = np.ascontiguousarray(phi, dtype=np.float64)
        if a.shape != self.phi_shape:
            raise ValueError(f"expected shape {self.phi_shape}, got {a.shape}")
        _check(_lib.tlq_upload_phi(self._ctx, _dptr(a)))

    def download_phi(self) -> np.ndarray:
        out = np.empty(self.phi_shape, dtype=np.float64)
        _check(_lib.tlq_download_phi(self._ctx, _dptr(out)))
        return out

    def gravity_solve(self) -> Tuple[int, float]:
        """One MultiGrid call on the current U: (V-cycles, final Dif)."""
        cyc, dif = ctypes.c_int(), ctypes.c_double()
        _check(_lib.tlq_gravity_solve(self._ctx, ctypes.byref(cyc), ctypes.byref(dif)))
        return cyc.value, dif.value

    def gravity_info(self) -> Tuple[int, int, float, int]:
        """(levels or -1, last solve's V-cycles, its Dif, step count)."""
        lv, cyc, dif, it = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        _check(_lib.tlq_gravity_info(self._ctx, ctypes.byref(lv), ctypes.byref(cyc),
                                     ctypes.byref(dif), ctypes.byref(it)))
        return lv.value, cyc.value, dif.value, it.value

    def run_until(self, tfin: float, max_steps: int = 10 ** 9) -> int:
        """mainHD.f90:61 loop: step while t < tfin (no clamp of the last dt)."""
        _, t, it0 = self.time()
        n = 0
        while t < tfin and n < max_steps:
            self.advance(1)
            n += 1
            _, t, _ = self.time()
        return n

    def sync(self) -> None:
        _check(_lib.tlq_sync(self._ctx))

    def check_state(self) -> int:
        bad = ctypes.c_longlong()
        _check(_lib.tlq_check_state(self._ctx, ctypes.byref(bad)))
        return bad.value

    def cooling_counts(self) -> Tuple[int, int]:
        """(cells queued for the dense pass, fallbacks) of the last step in
        cooling modes 2-4: mode 2 counts the network's cells and those of them
        that took the exact integration; modes 3 and 4 the integrator's cells."""
        q, f = ctypes.c_longlong(), ctypes.c_longlong()
        _check(_lib.tlq_cooling_counts(self._ctx, ctypes.byref(q), ctypes.byref(f)))
        return q.value, f.value

    def fallback_fraction(self) -> float:
        f = ctypes.c_double()
        _check(_lib.tlq_fallback_fraction(self._ctx, ctypes.byref(f)))
        return f.value

    # timing ---------------------------------------------------------------------------
    def timer_start(self) -> None:
        _check(_lib.tlq_timer_start(self._ctx))

    def timer_stop(self) -> float:
        ms = ctypes.c_double()
        _check(_lib.tlq_timer_stop(self._ctx, ctypes.byref(ms)))
        return ms.value

    def kernel_timing(self, enable: bool) -> None:
        _check(_lib.tlq_kernel_timing(self._ctx, 1 if enable else 0))

    def kernel_stats(self) -> dict:
        out = {}
        for q in range(6):
            name = ctypes.create_string_buffer(32)
            ms, n = ctypes.c_double(), ctypes.c_longlong()
            _check(_lib.tlq_kernel_stats(self._ctx, q, name, 32, ctypes.byref(ms), ctypes.byref(n)))
            out[name.value.decode()] = (ms.value, n.value)
        return out

    def copy_probe(self, nbytes: int, iters: int = 10) -> float:
        """Mean ms of one device-to-device copy of nbytes (streaming kernel)."""
        ms = ctypes.c_double()
        _check(_lib.tlq_copy_probe(self._ctx, int(nbytes), int(iters), ctypes.byref(ms)))
        return ms.value

    def copy_probe_config(self, nbytes: int, iters: int, unroll: int, blocks: int,
                          nontemporal: bool) -> float:
        """Mean ms of one copy in a given configuration: `unroll` 16-byte loads
        in flight per thread, `blocks` blocks of 256 threads, nontemporal
        loads and stores or not."""
        ms = ctypes.c_double()
        _check(_lib.tlq_copy_probe_config(self._ctx, int(nbytes), int(iters), int(unroll),
                                          int(blocks), int(bool(nontemporal)), ctypes.byref(ms)))
        return ms.value

    # learned cooling surrogate -------------------------------------------------
    def surrogate_init(self, seed: int = 1234) -> None:
        _check(_lib.tlq_surrogate_init(self._ctx, int(seed)))

    def surrogate_params(self, theta: Optional[np.ndarray] = None) -> np.ndarray:
        """Set (theta given) or get the FP32 master parameters."""
        if theta is not None:
            t = np.ascontiguousarray(theta, dtype=np.float32)
            _check(_lib.tlq_surrogate_set_params(self._ctx, _fptr(t), t.size))
            return t
        out = np.empty(MLP_NPARAMS, dtype=np.float32)
        _check(_lib.tlq_surrogate_get_params(self._ctx, _fptr(out), out.size))
        return out

    def surrogate_threshold(self, dtc_min: float) -> None:
        """Surrogate mode: cells with 0.01 <= dt/t_cool < dtc_min take the
        exact integration instead of the network (include/tlq.h)."""
        _check(_lib.tlq_surrogate_set_threshold(self._ctx, float(dtc_min)))

    def surrogate_generate(self, n: int, seed: int = 7) -> None:
        _check(_lib.tlq_surrogate_generate(self._ctx, int(n), int(seed)))

    def surrogate_samples(self, first: int, n: int) -> Tuple[np.ndarray, np.ndarray]:
        x = np.empty((n, 3), dtype=np.float32)
        y = np.empty(n, dtype=np.float32)
        _check(_lib.tlq_surrogate_samples(self._ctx, int(first), int(n), _fptr(x), _fptr(y)))
        return x, y

    def surrogate_train(self, steps: int, batch: int = 4096, lr_max: float = 3e-3,
                        lr_min: float = 3e-5) -> float:
        loss = ctypes.c_double()
        _check(_lib.tlq_surrogate_train(self._ctx, int(steps), int(batch), lr_max, lr_min,
                                        ctypes.byref(loss)))
        return loss.value

    def surrogate_eval(self, x: np.ndarray) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        out = np.empty(xx.shape[0], dtype=np.float32)
        _check(_lib.tlq_surrogate_eval(self._ctx, _fptr(xx), xx.shape[0], _fptr(out)))
        return out

    def surrogate_grad(self, x: np.ndarray, y: np.ndarray) -> Tuple[np.ndarray, float]:
        xx = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 3)
        yy = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(MLP_NPARAMS, dtype=np.float64)
        loss = ctypes.c_double()
        _check(_lib.tlq_surrogate_grad(self._ctx, _fptr(xx), _fptr(yy), xx.shape[0], _dptr(g),
                                       ctypes.byref(loss)))
        return g, loss.value

    def layout(self) -> Tuple[int, int, int]:
        cells, px, ox = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int()
        _check(_lib.tlq_layout(self._ctx, ctypes.byref(cells), ctypes.byref(px), ctypes.byref(ox)))
        return cells.value, px.value, ox.value
