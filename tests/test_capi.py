"""The C-ABI library loads without a GPU and exports every entry point the
headers in include/ declare."""
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(ecm_\w+)\s*\(", text, flags=re.M):
            if "typedef" in m.group(0):
                continue
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "ecm_sim_step" in names and "ecm_sim_create" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    from ecmgenerator_amd import _native
    L = _native.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_bindings_cover_every_declared_symbol():
    from ecmgenerator_amd import _native
    missing = [n for n in declared_functions() if n not in _native.SIGNATURES]
    assert not missing, f"no ctypes signature for: {missing}"


def test_create_without_gpu_fails_cleanly():
    from ecmgenerator_amd import _native
    L = _native.lib()
    if L.ecm_device_available():
        pytest.skip("a GPU is present; the no-GPU error path is not reachable")
    import ctypes as C
    p = _native.SimParams()
    L.ecm_sim_default_params(C.byref(p))
    h = C.c_void_p()
    rc = L.ecm_sim_create(C.byref(p), -10.0, -10.0, 10.0, 10.0, C.byref(h))
    assert rc == -5 and not h
    assert b"gfx950" in L.ecm_last_error()
