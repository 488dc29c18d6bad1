"""CPU checks of the C++ facade (include/openrcwa_gpu.hpp) and the package
surface: a C++ program compiles against the header and links the library, and
the reference's addLayer rule (a layer whose period or first / last two edges
differ is dropped with a message) holds without touching a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_package_star_import():
    ns = {}
    exec("from openrcwa_amd import *", ns)
    for m in ("engine", "ofd", "orcwa", "surrogate", "workloads"):
        assert m in ns


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    lib = os.path.join(ROOT, "openrcwa_amd")
    assert os.path.exists(os.path.join(lib, "libopenrcwa_gpu.so")), "build the library first"
    out = str(tmp_path_factory.mktemp("cpp") / "cpp_facade_harness")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp_facade_harness.cpp"), "-L" + lib,
                        "-lopenrcwa_gpu", "-Wl,-rpath," + lib, "-o", out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_add_layer_drops_mismatching_layers(harness):
    r = subprocess.run([harness, "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert r.stdout.strip() == "NTYPES 0 2"
    assert "The new layer does not match layers before!" in r.stderr
