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
