"""The oracle's midpoint-staircase integrator (RCWAIntegrator::compute,
core/RCWAIntegrator.cpp:20-86, with the periodic branch of the sampler,
core/VarLayerSampler.cpp:52-68 and core/LayerSampler.cpp:114-144), pinned by:

* the reference's own checks of tests/test_core_integrators.cpp: a structure
  uniform in z gives Tuu = Tdd = diag(exp(i gamma L / k0)), Rud = Rdu = 0 for
  any slab count; with an edge moving along z the result converges at second
  order in the slab count;
* a staircase structure between homogeneous media with slab boundaries on the
  layer boundaries: the per-order R/T from the integrator's S-matrix equal
  the RTCM stack solve of the same layers (a different algorithm: COD
  interfaces and the cascade of rcwa/RCWASolver.cpp:378-614).
"""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o

LAM, NHX, NHY, Z0, Z1 = 1.55, 3, 2, -0.1, 1.1


def waveguide(moving):
    """The cross-section of tests/test_core_integrators.cpp:52-71: a silicon
    core in cladding; edge x[3] moves from 2.6 to 3.7 over z in [0, 1]."""
    cl, co = 1.445 ** 2, 3.48 ** 2
    x = np.array([0.0, 0.5, 1.5, 2.6, 3.7, 3.9, 5.0, 5.5])
    y = np.array([0.0, 0.5, 1.5, 1.58, 1.8, 2.8, 3.3])
    eps = np.full((6, 7), cl, dtype=complex)
    eps[3, 2] = co

    def sample(z):
        xx = x.copy()
        if moving:
            xx[3] = 2.6 if z < 0.0 else (3.7 if z > 1.0 else (1.0 - z) * 2.6 + z * 3.7)
        return xx, y, eps
    return sample


def as_layer(sample):
    return lambda z: o.Layer(*sample(z))


@pytest.mark.parametrize("N", [1, 8])
def test_uniform_in_z_closed_form(N):
    r = o.rcwa_integrator(as_layer(waveguide(False)), LAM, NHX, NHY, Z0, Z1, N)
    E = np.diag(np.exp(1j * r["gam0"] * (Z1 - Z0) * LAM / (2 * math.pi)))
    assert np.abs(r["Tuu"] - E).max() < 1e-12 and np.abs(r["Tdd"] - E).max() < 1e-12
    assert np.abs(r["Rud"]).max() < 1e-12 and np.abs(r["Rdu"]).max() < 1e-12


def test_moving_edge_second_order_convergence():
    ref = o.integrator_harmonic_operators(
        o.rcwa_integrator(as_layer(waveguide(True)), LAM, NHX, NHY, Z0, Z1, 128))
    errs = []
    for N in (8, 16, 32):
        T = o.integrator_harmonic_operators(
            o.rcwa_integrator(as_layer(waveguide(True)), LAM, NHX, NHY, Z0, Z1, N))
        errs.append(max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(T, ref)))
    assert errs[0] > errs[1] > errs[2]
    assert min(errs[0] / errs[1], errs[1] / errs[2]) > 3.0


def staircase(seed=4):
    rng = np.random.default_rng(seed)
    xe, ye = [0.0, 0.13, 0.31, 0.5], [0.0, 0.22, 0.45]
    shp = (2, 3)
    top, bot = np.full(shp, 1.0 + 0j), np.full(shp, 2.1 + 0j)
    mids = [rng.uniform(1.5, 6, shp) + 0.02j, rng.uniform(2, 9, shp) + 0j,
            rng.uniform(1, 4, shp) + 0j]
    th = [0.08, 0.12, 0.04]               # from the top down
    z1 = sum(th)
    bounds = [z1 - th[0], z1 - th[0] - th[1]]

    def sample(z):
        if z >= z1:
            g = top
        elif z <= 0.0:
            g = bot
        else:
            g = mids[0 if z > bounds[0] else (1 if z > bounds[1] else 2)]
        return xe, ye, g
    return sample, [top] + mids + [bot], [0.0] + th + [0.0], xe, ye, z1


def rt_from_operators(T_down, R_top, kx, ky, k0, eps_ref, eps_trn, px, py, i0):
    """Per-order efficiencies of a plane wave from the harmonic-space S-matrix
    (the normalisation of rcwa/RCWASolver.cpp:1833-1907)."""
    h = kx.size
    s = np.zeros(2 * h, complex)
    s[i0], s[h + i0] = px, py
    r, t = R_top @ s, T_down @ s
    kzr = -np.sqrt(k0 * k0 * eps_ref - kx ** 2 - ky ** 2 + 0j)
    kzt = np.sqrt(k0 * k0 * eps_trn - kx ** 2 - ky ** 2 + 0j)
    rz = (kx * r[:h] + ky * r[h:]) / kzr
    tz = (kx * t[:h] + ky * t[h:]) / kzt
    inc = abs(kzr[i0].real)
    R = np.abs(kzr.real) * (np.abs(r[:h]) ** 2 + np.abs(r[h:]) ** 2 + np.abs(rz) ** 2) / inc
    T = np.abs(kzt.real) * (np.abs(t[:h]) ** 2 + np.abs(t[h:]) ** 2 + np.abs(tz) ** 2) / inc
    return R, T


def test_staircase_equals_rtcm():
    sample, eps, thick, xe, ye, z1 = staircase()
    lam, theta, phi, nX, nY = 0.62, 25.0, 30.0, 3, 2
    k0 = 2 * math.pi / lam
    kx0 = k0 * math.sin(math.radians(theta)) * math.cos(math.radians(phi))
    ky0 = k0 * math.sin(math.radians(theta)) * math.sin(math.radians(phi))
    r = o.rcwa_integrator(as_layer(sample), lam, nX, nY, 0.0, z1, 24, kx0, ky0)
    T_down, R_top, _, _ = o.integrator_harmonic_operators(r)
    REF, TRN, s = o.solve_stack(nX, nY, 0.5, 0.45, xe, ye, eps, thick, lam, theta, phi)
    i0 = (2 * nY + 1) * nX + nY
    for (px, py), Rr, Tr in zip(o.polarisation_basis(theta, phi), REF, TRN):
        R, T = rt_from_operators(T_down, R_top, s.Kx, s.Ky, k0, 1.0, 2.1, px, py, i0)
        assert np.abs(R - Rr).max() < 1e-12 and np.abs(T - Tr).max() < 1e-12


# ---- the adaptive perturbation integrator (core/RedhefferIntegrator.cpp) ----
def test_redheffer_uniform_in_z_is_one_exact_segment():
    r = o.redheffer_integrator(as_layer(waveguide(False)), LAM, NHX, NHY, Z0, Z1, 1e-3)
    assert r["N"] == 1
    E = np.diag(np.exp(1j * r["gam0"] * (Z1 - Z0) * LAM / (2 * math.pi)))
    assert np.abs(r["Tuu"] - E).max() < 1e-12 and np.abs(r["Tdd"] - E).max() < 1e-12
    assert np.abs(r["Rud"]).max() < 1e-12 and np.abs(r["Rdu"]).max() < 1e-12


def test_redheffer_agrees_with_fine_staircase():
    # tests/test_core_integrators.cpp:143-146: pe = 1e-3 within 1e-3 of a fine
    # RCWAIntegrator (both in the same bases: the modes at z0 and z1)
    ref = o.rcwa_integrator(as_layer(waveguide(True)), LAM, NHX, NHY, Z0, Z1, 128)
    scale = np.abs(ref["Tuu"]).max()
    errs = {}
    for pe in (1e-2, 1e-3):
        r = o.redheffer_integrator(as_layer(waveguide(True)), LAM, NHX, NHY, Z0, Z1, pe)
        errs[pe] = np.abs(r["Tuu"] - ref["Tuu"]).max() / scale
        assert r["N"] > 1
    assert errs[1e-3] < 1e-3 and errs[1e-3] < errs[1e-2]
