"""The header-only C++ facade (include/openrcwa_gpu.hpp) from a C++ program
(tests/cpp_facade_harness.cpp, built and linked against libopenrcwa_gpu.so by
openrcwa_amd/build.py): RCWASolver against the FP64 oracle and the published
grating values, runRCWA against the Python .orcwa driver (same problem, same
CSV bytes)."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import orcwa
from openrcwa_amd.build import HARNESS_BIN

pytestmark = pytest.mark.gpu


def run(*args):
    assert os.path.exists(HARNESS_BIN), "build first: python -m openrcwa_amd.build"
    r = subprocess.run([HARNESS_BIN, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        tag, idx, *vals = line.split()
        out[(tag, int(idx))] = np.array([float(v) for v in vals])
    return out


def test_solver_facade_grating_against_oracle():
    out = run(0)
    assert out[("STATUS", 0)][0] == 0
    P, lam = 0.3, 2.99792458e8 / 6.5e14 * 1e6
    xe = [0.0, 0.5 * P, P]
    eps = [np.array([[1.0, 1.0]]), np.array([[4.0, 1.0]]), np.array([[2.25, 2.25]])]
    REF, TRN, _ = o.solve_stack(5, 0, P, 1.0, xe, [0.0, 1.0], eps, [0.0, 0.2, 0.0], lam, 0.0, 0.0,
                                pols=[(0.0, 1.0), (1.0, 0.0)])
    for p in range(2):
        assert np.abs(out[("REF", p)] - REF[p]).max() < 1e-10
        assert np.abs(out[("TRN", p)] - TRN[p]).max() < 1e-10
    # published values at N = 5, f = 6.5e14 Hz (461 nm, AGENTS.md:116-118)
    assert round(out[("REF", 0)].sum(), 6) == 0.138130
    assert round(out[("REF", 1)].sum(), 6) == 0.204841


def test_solver_facade_layers_with_own_edges():
    """Layer types with different interior edges (the reference's Layer keeps
    its own coordinates) and a repeated layer type, oblique and conical."""
    out = run(1)
    assert out[("STATUS", 0)][0] == 0
    Px, Py, lam = 0.4, 0.35, 0.55
    th, ph = 20.0, 30.0
    L0 = o.Layer([0.0, 0.1, 0.3, Px], [0.0, 0.05, 0.3, Py], np.ones((3, 3)))
    e1 = np.ones((4, 4), dtype=complex)
    e1[1, 1] = 6.0 + 0.1j
    L1 = o.Layer([0.0, 0.1, 0.22, 0.3, Px], [0.0, 0.05, 0.2, 0.3, Py], e1)
    e2 = np.full((3, 3), 2.0, dtype=complex)
    e2[0, 0] = e2[0, 1] = e2[1, 0] = 3.0
    L2 = o.Layer([0.0, 0.1, 0.3, Px], [0.0, 0.05, 0.3, Py], e2)
    L3 = o.Layer([0.0, 0.1, 0.3, Px], [0.0, 0.05, 0.3, Py], np.full((3, 3), 2.25))
    s = o.RCWASolver(3, 2)
    for l_ in (L0, L1, L2, L3):
        s.add_layer(l_)
    s.disable_pml()
    k0 = 2 * math.pi / lam
    t, p_ = math.radians(th), math.radians(ph)
    s.set_bloch_wavevector(k0 * math.sin(t) * math.cos(p_), k0 * math.sin(t) * math.sin(p_))
    s.solve(lam, [0, 1, 2, 1, 3], [0.0, 0.12, 0.08, 0.05, 0.0])
    pols = [(-math.sin(p_), math.cos(p_)), (math.cos(t) * math.cos(p_), math.cos(t) * math.sin(p_))]
    for q, (px, py) in enumerate(pols):
        c = s.generate_horizontal_plane_wave(px, py, 0)
        ref, trn = s.scatter_plane_wave(c, 0, 3, k0)
        assert np.abs(out[("REF", q)] - ref).max() < 1e-9
        assert np.abs(out[("TRN", q)] - trn).max() < 1e-9


def _driver_problem():
    p = orcwa.RCWAProblem(title="facade", xmin=0.0, xmax=0.5, ymin=0.0, ymax=0.5, zmin=0.0,
                          zmax=0.6, nHx=3, nHy=3)
    p.material_eps = [1.0, 1.0, 2.25, 5.0 + 0.05j, 3.0]
    p.material_mu = [1.0, 1.0, 1.0, 1.0, 1.5 + 0.02j]
    p.material_sigma = [0.0, 0.0, 0.0, 0.0, 2.0e4]
    p.material_mag_sigma = [0.0] * 5
    p.material_dispersion = [orcwa.MaterialDispersion() for _ in range(5)]
    B = orcwa.RCWABox
    p.boxes = [B(2, orcwa.SHAPE_BOX, 0.0, 0.5, 0.0, 0.5, 0.0, 0.2),
               B(4, orcwa.SHAPE_BOX, 0.0, 0.5, 0.0, 0.5, 0.2, 0.35),
               B(3, orcwa.SHAPE_CYL_Z, 0.1, 0.4, 0.05, 0.35, 0.2, 0.35),
               B(3, orcwa.SHAPE_BOX, 0.0, 0.2, 0.1, 0.5, 0.35, 0.45)]
    p.theta, p.phi, p.pol = 15.0, 25.0, 4
    p.lambdas = [0.45, 0.6, 0.75]
    return p


def test_run_rcwa_matches_the_python_driver(gpu_ctx, tmp_path):
    f_cpp, d_cpp = tmp_path / "field_cpp.csv", tmp_path / "device_cpp.csv"
    out = run(2, f_cpp, d_cpp)
    prob = _driver_problem()
    f_py, d_py = tmp_path / "field_py.csv", tmp_path / "device_py.csv"
    req = orcwa.FieldRequest(component="Ex", slice="xz", coord=0.2, opt="re", path=str(f_py),
                             device_path=str(d_py))
    res = orcwa.run_rcwa(prob, gpu_ctx, req)
    assert len(res) == 3
    for b, r in enumerate(res):
        tot = out[("TOT", b)]
        assert tot[0] == r.lam and out[("STATUS", b)][0] == r.status == 0
        assert np.abs(out[("REF", b)] - r.REF_orders).max() < 1e-12
        assert np.abs(out[("TRN", b)] - r.TRN_orders).max() < 1e-12
        assert abs(tot[1] - r.R) < 1e-12 and abs(tot[2] - r.T) < 1e-12
        assert 0.0 < tot[3] < 1.0      # lossy and magnetic: absorption
    assert d_cpp.read_bytes() == d_py.read_bytes()
    a = np.loadtxt(f_cpp, delimiter=",")
    b_ = np.loadtxt(f_py, delimiter=",")
    assert a.shape == b_.shape and np.abs(a - b_).max() <= 1e-7 * np.abs(b_).max()
