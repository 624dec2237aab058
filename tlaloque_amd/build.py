"""Build ``libtlq.so`` (the C-ABI of include/tlq.h) for gfx950 with hipcc.

The FP64 kernels are compiled with ``-ffp-contract=off`` (every multiply and
add rounded separately, as in the reference's build), the FP32 kernels with
hipcc's default contraction.  Packed FP32 instructions are disabled for all
device code.  Outputs go to ``tlaloque_amd/_build/``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from typing import List

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libtlq.so")
ARCH = os.environ.get("TLQ_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]

UNITS = [
    ("tlq_fp64.hip", ["-ffp-contract=off"]),
    ("tlq_fp32.hip", []),
    ("tlq_surrogate.hip", []),
    ("tlq_gravity.hip", ["-ffp-contract=off"]),
    ("tlq_api.hip", []),
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    cc = hipcc()
    procs = []
    objs: List[str] = []
    for src, extra in UNITS:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        cmd = [cc, *COMMON, *extra, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                            text=True)))
        objs.append(obj)
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        noise = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            errors.append(f"{src}:\n" + "\n".join(noise))
        elif verbose and noise:
            print("\n".join(noise))
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(OUT, "tlq_hydro")


def flang() -> str:
    for cand in (shutil.which("flang"), "/opt/rocm/llvm/bin/flang", "/opt/rocm/lib/llvm/bin/flang"):
        if cand and os.path.exists(cand):
            return cand
    return ""


def build_fortran(verbose: bool = False) -> str:
    """The ISO_C_BINDING module and the Fortran driver, linked to libtlq.so
    (run path $ORIGIN).  Returns the driver's path, or "" without flang."""
    fc = flang()
    if not fc:
        return ""
    mod = os.path.join(OUT, "fortran")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in ("tlq_c.f90", "tlq_hydro.f90"):
        obj = os.path.join(mod, src.replace(".f90", ".o"))
        cmd = [fc, "-O2", "-module-dir", mod, "-c", os.path.join(FORTRAN_DIR, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"flang failed on {src}:\n{r.stdout}")
        objs.append(obj)
    tmp = DRIVER + ".tmp"
    cmd = [fc, *objs, "-L" + OUT, "-ltlq", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang link failed:\n" + r.stdout)
    os.replace(tmp, DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv) or "flang not found: Fortran driver not built")
