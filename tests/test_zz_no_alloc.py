"""The allocation-free entry point tlq_surrogate_grad_noalloc against
tlq_surrogate_grad, through the C-ABI.

The caller's device arrays come from hipMalloc of the HIP runtime libtlq.so
itself links; the library allocates nothing for the call (its partial
gradients live in a static device array).  Two calls in a row, the second
with fewer samples (not a multiple of 64), must each equal the allocating
entry point's gradient and loss bit for bit; sizes beyond the scratch are
refused with TLQ_ERR_ARG.
"""
import ctypes
import os

import numpy as np
import pytest

from tlaloque_amd import hydro
from tlaloque_amd import params as P

pytestmark = pytest.mark.gpu

TLQ_ERR_ARG = -1
MAX_SAMPLES = 16384          # TLQ_NOALLOC_MAX_SAMPLES (include/tlq.h)
H2D, D2H = 1, 2              # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    """The HIP runtime libtlq.so is linked against (one runtime per process)."""
    lib = hydro.load_library()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path and os.path.exists(path), "libamdhip64 not mapped"
    hip = ctypes.CDLL(path)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    lib.tlq_surrogate_grad_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    lib.tlq_surrogate_grad_noalloc.restype = ctypes.c_int
    lib.tlq_sync.argtypes = [ctypes.c_void_p]
    return hip, lib


class DeviceArray:
    def __init__(self, hip, nbytes):
        self.hip, self.nbytes = hip, nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, n, dtype=np.float32):
        out = np.empty(n, dtype=dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _samples(rng, n):
    x = np.stack([rng.uniform(-3, 4, n), rng.uniform(4, 9, n), rng.uniform(-2, 2, n)],
                 1).astype(np.float32)
    y = rng.uniform(-1.5, 0.2, n).astype(np.float32)
    return x, y


def test_noalloc_gradient_equals_allocating_entry_point():
    hip, lib = _hip()
    p = P.HydroParams(nx=8, ny=8, nz=8)
    rng = np.random.default_rng(21)
    with hydro.HydroCore(p) as h:
        h.surrogate_init(seed=5)
        npar = hydro.MLP_NPARAMS
        dx = DeviceArray(hip, 3 * 4 * MAX_SAMPLES)
        dy = DeviceArray(hip, 4 * MAX_SAMPLES)
        dg = DeviceArray(hip, 4 * npar)
        dl = DeviceArray(hip, 4)
        try:
            for n in (5000, 700):
                x, y = _samples(rng, n)
                dx.put(x)
                dy.put(y)
                rc = lib.tlq_surrogate_grad_noalloc(h._ctx, dx.ptr, dy.ptr, n, dg.ptr, dl.ptr)
                assert rc == 0, lib.tlq_last_error()
                assert lib.tlq_sync(h._ctx) == 0
                g_na, l_na = dg.get(npar), dl.get(1)[0]
                g_ref, l_ref = h.surrogate_grad(x, y)
                print(f"n={n}: loss {l_na:.6e} (allocating entry point {l_ref:.6e}), "
                      f"|grad| {np.linalg.norm(g_na):.4e}")
                assert np.isfinite(g_na).all() and np.abs(g_na).max() > 0
                assert np.array_equal(g_na, g_ref.astype(np.float32))
                assert np.float32(l_na) == np.float32(l_ref)
            x, y = _samples(rng, 64)
            dx.put(x)
            dy.put(y)
            for bad in (0, MAX_SAMPLES + 1):
                assert lib.tlq_surrogate_grad_noalloc(h._ctx, dx.ptr, dy.ptr, bad, dg.ptr,
                                                      dl.ptr) == TLQ_ERR_ARG
        finally:
            for a in (dx, dy, dg, dl):
                a.free()
