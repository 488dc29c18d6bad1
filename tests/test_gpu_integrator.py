"""The midpoint-staircase integrator on the GPU (orcwa_gpu_rcwa_integrate;
RCWAIntegrator::compute, core/RCWAIntegrator.cpp:20-86) against the closed
form of a z-uniform structure, the FP64 oracle (basis-free harmonic-space
operators), the GPU stack solve of a staircase, and second-order convergence
with a moving edge (the checks of tests/test_core_integrators.cpp)."""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import engine

from .test_oracle_integrator import (LAM, NHX, NHY, Z0, Z1, as_layer, rt_from_operators,
                                     staircase, waveguide)

pytestmark = pytest.mark.gpu


def harmonic(r):
    return o.integrator_harmonic_operators(r)


@pytest.mark.parametrize("N", [1, 8])
def test_uniform_in_z_closed_form(gpu_ctx, N):
    r = engine.rcwa_integrate(gpu_ctx, NHX, NHY, waveguide(False), LAM, Z0, Z1, N)
    assert r["status"] == 0
    E = np.diag(np.exp(1j * r["gam0"] * (Z1 - Z0) * LAM / (2 * math.pi)))
    assert np.abs(r["Tuu"] - E).max() < 1e-10 and np.abs(r["Tdd"] - E).max() < 1e-10
    assert np.abs(r["Rud"]).max() < 1e-10 and np.abs(r["Rdu"]).max() < 1e-10


@pytest.mark.parametrize("N", [5, 16])
def test_moving_edge_against_oracle(gpu_ctx, N):
    g = engine.rcwa_integrate(gpu_ctx, NHX, NHY, waveguide(True), LAM, Z0, Z1, N)
    ref = o.rcwa_integrator(as_layer(waveguide(True)), LAM, NHX, NHY, Z0, Z1, N)
    assert g["status"] == 0
    for a, b in zip(harmonic(g), harmonic(ref)):
        assert np.abs(a - b).max() < 1e-9 * np.abs(b).max()


def test_moving_edge_second_order_convergence(gpu_ctx):
    def ops(N):
        return harmonic(engine.rcwa_integrate(gpu_ctx, NHX, NHY, waveguide(True), LAM, Z0, Z1,
                                              N))
    ref = ops(256)
    errs = [max(np.abs(a - b).max() / np.abs(b).max() for a, b in zip(ops(N), ref))
            for N in (8, 16, 32, 64)]
    assert all(e0 > e1 for e0, e1 in zip(errs, errs[1:]))
    assert min(e0 / e1 for e0, e1 in zip(errs, errs[1:])) > 3.0


def test_staircase_equals_stack_solve(gpu_ctx):
    sample, eps, thick, xe, ye, z1 = staircase()
    lam, theta, phi, nX, nY = 0.62, 25.0, 30.0, 3, 2
    k0 = 2 * math.pi / lam
    kx0 = k0 * math.sin(math.radians(theta)) * math.cos(math.radians(phi))
    ky0 = k0 * math.sin(math.radians(theta)) * math.sin(math.radians(phi))
    r = engine.rcwa_integrate(gpu_ctx, nX, nY, sample, lam, 0.0, z1, 24, kx0, ky0)
    assert r["status"] == 0
    T_down, R_top, _, _ = harmonic(r)
    te, tm = o.polarisation_basis(theta, phi)
    res = engine.solve_batch(gpu_ctx, nX, nY, xe, ye, thick, [lam], [kx0], [ky0],
                             np.array(eps)[None], np.array([[te, tm]], dtype=complex))
    m = np.repeat(np.arange(2 * nX + 1) - nX, 2 * nY + 1)
    n = np.tile(np.arange(2 * nY + 1) - nY, 2 * nX + 1)
    kx = 2 * math.pi * m / 0.5 + kx0
    ky = 2 * math.pi * n / 0.45 + ky0
    i0 = (2 * nY + 1) * nX + nY
    for q, (px, py) in enumerate((te, tm)):
        R, T = rt_from_operators(T_down, R_top, kx, ky, k0, 1.0, 2.1, px, py, i0)
        assert np.abs(R - res.ref[0, q]).max() < 1e-10
        assert np.abs(T - res.trn[0, q]).max() < 1e-10


def test_redheffer_uniform_in_z(gpu_ctx):
    r = engine.redheffer_integrate(gpu_ctx, NHX, NHY, waveguide(False), LAM, Z0, Z1, 1e-3)
    assert r["status"] == 0 and r["N"] == 1
    E = np.diag(np.exp(1j * r["gam0"] * (Z1 - Z0) * LAM / (2 * math.pi)))
    assert np.abs(r["Tuu"] - E).max() < 1e-10 and np.abs(r["Tdd"] - E).max() < 1e-10
    assert np.abs(r["Rud"]).max() < 1e-10 and np.abs(r["Rdu"]).max() < 1e-10


@pytest.mark.parametrize("pe", [1e-2, 3e-3])
def test_redheffer_against_oracle(gpu_ctx, pe):
    g = engine.redheffer_integrate(gpu_ctx, NHX, NHY, waveguide(True), LAM, Z0, Z1, pe)
    ref = o.redheffer_integrator(as_layer(waveguide(True)), LAM, NHX, NHY, Z0, Z1, pe)
    assert g["status"] == 0 and g["N"] == ref["N"] > 1
    for a, b in zip(harmonic(g), harmonic(ref)):
        assert np.abs(a - b).max() < 1e-9 * np.abs(b).max()


def test_redheffer_agrees_with_fine_staircase(gpu_ctx):
    # tests/test_core_integrators.cpp:143-146
    fine = harmonic(engine.rcwa_integrate(gpu_ctx, NHX, NHY, waveguide(True), LAM, Z0, Z1, 256))
    g = engine.redheffer_integrate(gpu_ctx, NHX, NHY, waveguide(True), LAM, Z0, Z1, 1e-3)
    T = harmonic(g)
    assert np.abs(T[0] - fine[0]).max() / np.abs(fine[0]).max() < 1e-3
    assert np.abs(T[2] - fine[2]).max() / np.abs(fine[2]).max() < 1e-3


def test_redheffer_sampler_error_is_raised(gpu_ctx):
    def bad(z):
        if z > 0.5:
            raise ValueError("outside the sampled range")
        return waveguide(True)(z)
    with pytest.raises(ValueError):
        engine.redheffer_integrate(gpu_ctx, NHX, NHY, bad, LAM, Z0, Z1, 1e-3)
