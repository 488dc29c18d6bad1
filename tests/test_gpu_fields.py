"""Field images on the GPU (orcwa_gpu_field_image: the inputCoeffs form of
RCWASolver::saveFieldImage, rcwa/RCWASolver.cpp:707-1011, with the mode
amplitudes of evaluateIntermediateField, :1505-1645) against the plane-wave
closed form of a homogeneous slab and against the FP64 oracle on patterned,
lossy and magnetic stacks, for every component and slice."""
import cmath
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import _lib, engine

from .test_oracle_fields import patterned_stack, slab_closed_form

pytestmark = pytest.mark.gpu

COMPS = ["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


def bloch(lam, eps_inc, theta, phi, mu_inc=1.0):
    k0 = 2 * math.pi / lam
    n = max(cmath.sqrt(eps_inc * mu_inc).real, 1e-6)
    t, p = math.radians(theta), math.radians(phi)
    return k0 * n * math.sin(t) * math.cos(p), k0 * n * math.sin(t) * math.sin(p)


@pytest.mark.parametrize("theta,pol", [(0.0, "tm"), (30.0, "tm"), (30.0, "te"), (50.0, "te")])
def test_slab_closed_form(gpu_ctx, theta, pol):
    lam, Lx = 0.6, 0.5
    xe = ye = [-Lx / 2, Lx / 2]
    eps = np.array([[[1.0]], [[2.25]], [[1.0]]], dtype=complex)
    kx0 = 2 * math.pi / lam * math.sin(math.radians(theta))
    at, kx = slab_closed_form(theta, pol)
    nres1 = 40
    xs = -Lx / 2 + Lx / nres1 * (np.arange(nres1) + 0.5)
    for comp, cid in (("Ex", o.EX), ("Hy", o.HY), ("Ez", o.EZ)) if pol == "tm" else \
            (("Ey", o.EY), ("Hx", o.HX), ("Hz", o.HZ)):
        img, status = engine.field_image(gpu_ctx, 4, 0, xe, ye, [0.0, 0.3, 0.0], lam, kx0, 0.0,
                                         eps, (1.0, 0.0) if pol == "tm" else (0.0, 1.0),
                                         comp, "xz", 0.0, nres1)
        assert status == 0 and img.shape == (24, nres1)
        zs = 0.3 / 24 * (np.arange(24) + 0.5)
        ref = np.array([[at(z)[cid] * cmath.exp(1j * kx * x) for x in xs] for z in zs])
        assert np.abs(img - ref).max() < 1e-11 * max(1.0, np.abs(ref).max())


def oracle_image(nX, nY, xe, ye, eps, thick, lam, th, ph, pol, comp, sl, coord, nres1, mu=None):
    Lx, Ly = xe[-1] - xe[0], ye[-1] - ye[0]
    _, _, s = o.solve_stack(nX, nY, Lx, Ly, xe, ye, eps, thick, lam, th, ph, mu_layers=mu)
    c = s.generate_horizontal_plane_wave(pol[0], pol[1], 0)
    return o.field_image(s, c, list(range(len(eps))), thick, sl, coord, comp, nres1)


def compare(gpu_ctx, nX, nY, xe, ye, eps, thick, lam, th, ph, pol, sl, coord, nres1, comps,
            mu=None, tol=1e-9):
    kx0, ky0 = bloch(lam, complex(eps[0][0, 0]), th, ph,
                     1.0 if mu is None else complex(mu[0][0, 0]))
    for comp in comps:
        img, status = engine.field_image(gpu_ctx, nX, nY, xe, ye, thick, lam, kx0, ky0,
                                         np.array(eps), pol, comp, sl, coord, nres1,
                                         None if mu is None else np.array(mu))
        assert status == 0
        ref = oracle_image(nX, nY, xe, ye, eps, thick, lam, th, ph, pol, _lib.FIELD_COMPONENTS[comp],
                           _lib.SLICES[sl], coord, nres1, mu)
        assert img.shape == ref.shape
        # (relative to the component's own scale; components that vanish by
        # symmetry are compared against 1e-3 of the unit incident field)
        err = np.abs(img - ref).max() / max(np.abs(ref).max(), 1e-3)
        assert err < tol, (comp, sl, err)


@pytest.mark.parametrize("sl,coord", [("xz", 0.17), ("yz", 0.05), ("xy", 0.2)])
def test_patterned_2d_against_oracle(gpu_ctx, sl, coord):
    xe, ye, eps, thick = patterned_stack(seed=5, lossless=False)
    te, tm = o.polarisation_basis(25.0, 30.0)
    pol = (0.6 * te[0] + 0.8j * tm[0], 0.6 * te[1] + 0.8j * tm[1])
    compare(gpu_ctx, 3, 2, xe, ye, eps, thick, 0.62, 25.0, 30.0, pol, sl, coord, 48, COMPS)


def test_grating_1d_against_oracle(gpu_ctx):
    # a 1D grating with a thick high-index layer (evanescent orders decay
    # across it) and a metal-like film
    rng = np.random.default_rng(11)
    xe, ye = [0.0, 0.21, 0.34, 0.6], [0.0, 0.6]
    shp = (1, 3)
    eps = [np.full(shp, 1.0 + 0j), rng.uniform(2, 12, shp) + 0.05j, np.full(shp, -8.0 + 1.5j),
           rng.uniform(1.5, 4, shp) + 0j, np.full(shp, 2.25 + 0j)]
    thick = [0.0, 0.9, 0.02, 0.25, 0.0]
    for pol in ((1.0, 0.0), (0.0, 1.0)):
        compare(gpu_ctx, 9, 0, xe, ye, eps, thick, 0.55, 20.0, 0.0, pol, "xz", 0.0, 64, COMPS)


def test_magnetic_stack_against_oracle(gpu_ctx):
    from .test_gpu_magnetic import stack_case
    xe, ye, eps, mu, thick = stack_case(7, 2)
    te, tm = o.polarisation_basis(35.0, 20.0)
    compare(gpu_ctx, 3, 2, xe, ye, eps, thick, 0.7, 35.0, 20.0, tm, "xz", 0.3, 40, COMPS, mu=mu)


def test_errors(gpu_ctx):
    eps2 = np.array([[[1.0]], [[2.25]]], dtype=complex)
    with pytest.raises(_lib.EngineError):
        engine.field_image(gpu_ctx, 2, 0, [0, 0.5], [0, 0.5], [0.0, 0.0], 0.6, 0.0, 0.0, eps2,
                           (1.0, 0.0), "Ez", "xz", 0.0)
    eps3 = np.array([[[1.0]], [[2.25]], [[1.0]]], dtype=complex)
    with pytest.raises(_lib.EngineError):     # z of an XY slice below the interior
        engine.field_image(gpu_ctx, 2, 0, [0, 0.5], [0, 0.5], [0.0, 0.3, 0.0], 0.6, 0.0, 0.0,
                           eps3, (1.0, 0.0), "Ex", "xy", 0.5)
