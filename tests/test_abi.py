"""The C-ABI library loads and exports every entry point of include/tlq.h (CPU)."""
import ctypes
import os
import re

import pytest

from tlaloque_amd import hydro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "tlq.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tlq_\w+)\s*\(", src, re.M)))


def test_header_declares_the_python_binding_set():
    assert declared_functions() == sorted(hydro.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(hydro.LIB_PATH):
        pytest.skip("libtlq.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(hydro.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing


def test_params_struct_matches_header():
    # tlq_params field order/types as declared in include/tlq.h
    src = open(os.path.join(ROOT, "include", "tlq.h")).read()
    body = re.search(r"typedef struct tlq_params \{(.*?)\} tlq_params;", src, re.S).group(1)
    names = []
    for line in body.splitlines():
        line = line.split("/*")[0].strip().rstrip(";")
        if not line:
            continue
        decl = line.split(None, 1)[1]
        for n in decl.split(","):
            names.append(n.strip().split("[")[0])
    assert names == [f[0] for f in hydro.TlqParams._fields_]


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(hydro, "_lib", None)
    with pytest.raises(hydro.TlqError):
        hydro.load_library(str(tmp_path / "nope.so"))


_BAD = [
    ("mesh", dict(nx=1), -1, "nx, ny, nz"),
    ("inflow", dict(bc=(1, 4, 1, 1, 1, 1)), -4, "BC_INFLOW"),
    ("bc id", dict(bc=(1, 9, 1, 1, 1, 1)), -1, "boundary"),
    ("limiter", dict(limiter=7), -1, "limiter"),
    ("solver", dict(solver=2), -1, "solver"),
    ("precision", dict(precision=2), -1, "precision"),
    ("cooling", dict(cooling=5), -1, "cooling"),
    ("gamma", dict(gamma=1.0), -1, "physical"),
    ("domain", dict(xr=-2.0), -1, "physical"),
    ("self_gravity id", dict(self_gravity=3), -1, "self_gravity"),
    ("non-cubic cells", dict(self_gravity=1, xr=3.0), -4, "cubic"),
    ("mg_levels", dict(self_gravity=1, mg_levels=4), -1, "divisible"),
    ("coarsest level", dict(self_gravity=1, nx=34, ny=34, nz=34), -4, "4096"),
    ("gconst", dict(self_gravity=2, gconst=-1.0), -1, "gconst"),
]


@pytest.mark.parametrize("name,kw,code,text", _BAD, ids=[b[0] for b in _BAD])
def test_create_rejects_bad_parameters(name, kw, code, text):
    # validation precedes every HIP call, so this runs without a GPU
    if not os.path.exists(hydro.LIB_PATH):
        pytest.skip("libtlq.so not built")
    from tlaloque_amd import params as P
    p = P.HydroParams(nx=8, ny=8, nz=8).replace(**kw)
    with pytest.raises(hydro.TlqError) as e:
        hydro.HydroCore(p)
    assert e.value.code == code
    assert text in str(e.value)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_create_without_gpu_reports_a_hip_error():
    # the product has no CPU path: without a GPU, creating a context fails
    if not os.path.exists(hydro.LIB_PATH):
        pytest.skip("libtlq.so not built")
    from tlaloque_amd import params as P
    with pytest.raises(hydro.TlqError) as e:
        hydro.HydroCore(P.HydroParams(nx=8, ny=8, nz=8))
    assert e.value.code == -2
