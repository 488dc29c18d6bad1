"""The C-ABI library loads and exports every function include/*.h declares
(CPU: no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "openrcwa_amd", "libopenrcwa_gpu.so")


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:int|double|void|long long)\s+\*?\s*([a-z_][a-z0-9_]*)\s*\(",
                            text, re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: python -m openrcwa_amd.build"
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert "rcwa_run" in names and "orcwa_gpu_solve_batch" in names
    for n in names:
        assert hasattr(lib, n), n


def test_binding_lists_the_header():
    from openrcwa_amd import _lib
    assert set(_lib.EXPORTED) | set(_lib.EXPORTED_SURROGATE) <= set(declared_functions())
    assert "orcwa_mlp_param_count" in declared_functions()


def test_errors_without_device_are_reported():
    # orcwa_gpu_create must fail cleanly (no crash) when it cannot get a device;
    # on a GPU machine it succeeds.  Either way the call returns an error code.
    from openrcwa_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.orcwa_gpu_create(ctypes.byref(h), 999)
    assert rc != 0


def test_rcwa_run_rejects_an_invalid_setup():
    # fewer than two layers: "*** invalid rcwa setup" and return 1 before any
    # device work (the reference's bridge returns 1 on errors, sol/rcwa_bridge.cpp:145)
    from openrcwa_amd import _lib
    lib = _lib.load()
    ctypes.c_int.in_dll(lib, "NRcwaLayer").value = 1
    lib.rcwa_run.argtypes = [ctypes.c_void_p]
    assert lib.rcwa_run(None) == 1
