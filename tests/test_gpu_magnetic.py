"""Magnetic layers on the GPU (mu != 1: rcwa/Layer.cpp:129-145, 150-164,
206-213; the half-space terms of rcwa/RCWASolver.cpp:1833-1907) against the
closed-form magnetic Fresnel coefficients, the FP64 oracle order by order,
and electromagnetic duality; then the `.orcwa` magnetic cases of the
reference's tests/test_rcwa_input.cpp (test_magnetic_material,
test_material_mu_keyword)."""
import cmath
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import engine
from openrcwa_amd import orcwa as oc

pytestmark = pytest.mark.gpu


def magnetic_fresnel_R(theta, e2, m2, pol):
    t = math.radians(theta)
    k1z, s = math.cos(t), math.sin(t)
    k2z = cmath.sqrt(e2 * m2 - s * s)
    r = (m2 * k1z - k2z) / (m2 * k1z + k2z) if pol == "te" else \
        (e2 * k1z - k2z) / (e2 * k1z + k2z)
    return abs(r) ** 2


def bloch(lam, eps_inc, mu_inc, theta, phi):
    k0 = 2 * math.pi / lam
    n = max(cmath.sqrt(eps_inc * mu_inc).real, 1e-6)
    t, p = math.radians(theta), math.radians(phi)
    return k0 * n * math.sin(t) * math.cos(p), k0 * n * math.sin(t) * math.sin(p)


@pytest.mark.parametrize("e2,m2", [(4.0, 4.0), (2.25, 3.0), (2.0 + 0.3j, 1.5 + 0.2j)])
def test_magnetic_interface(gpu_ctx, e2, m2):
    # (period 0.37 at lambda 0.6: no diffracted order is exactly grazing)
    xe, ye = [0.0, 0.37], [0.0, 0.37]
    thetas = [0.0, 30.0, 45.0, 60.0]
    lam = 0.6
    te_tm = [o.polarisation_basis(t, 0.0) for t in thetas]
    pols = np.array([[te, tm] for te, tm in te_tm], dtype=complex)
    kk = [bloch(lam, 1.0, 1.0, t, 0.0) for t in thetas]
    eps = np.array([[[[1.0]], [[e2]]]] * len(thetas), dtype=complex)
    mu = np.array([[[[1.0]], [[m2]]]] * len(thetas), dtype=complex)
    res = engine.solve_batch(gpu_ctx, 3, 1, xe, ye, [0.0, 0.0], [lam] * len(thetas),
                             [k[0] for k in kk], [k[1] for k in kk], eps, pols, mu)
    assert not res.status.any()
    for i, t in enumerate(thetas):
        assert abs(res.ref[i, 0].sum() - magnetic_fresnel_R(t, e2, m2, "te")) < 1e-12
        assert abs(res.ref[i, 1].sum() - magnetic_fresnel_R(t, e2, m2, "tm")) < 1e-12
        if isinstance(e2, float):
            assert abs(res.ref[i, 0].sum() + res.trn[i, 0].sum() - 1.0) < 1e-12


def stack_case(seed, nY):
    """air / magnetic grating / non-magnetic grating / magnetic homogeneous
    slab / magnetic grating / magnetic substrate."""
    rng = np.random.default_rng(seed)
    xe, ye = [0.0, 0.13, 0.31, 0.5], ([0.0, 0.22, 0.5] if nY else [0.0, 0.5])
    shp = (len(ye) - 1, len(xe) - 1)
    def grid(lo, hi, loss):
        return rng.uniform(lo, hi, shp) + 1j * rng.uniform(0, loss, shp)
    eps = [np.full(shp, 1.0 + 0j), grid(1.5, 6.0, 0.1), grid(2.0, 9.0, 0.05),
           np.full(shp, 2.5 + 0.02j), grid(1.0, 4.0, 0.0), np.full(shp, 1.8 + 0j)]
    mu = [np.full(shp, 1.0 + 0j), grid(1.0, 3.0, 0.05), np.ones(shp, complex),
          np.full(shp, 1.7 + 0.01j), grid(1.0, 2.0, 0.0), np.full(shp, 1.3 + 0j)]
    thick = [0.0, 0.11, 0.07, 0.05, 0.13, 0.0]
    return xe, ye, eps, mu, thick


@pytest.mark.parametrize("nY", [0, 2])
def test_magnetic_stack_against_oracle(gpu_ctx, nY):
    xe, ye, eps, mu, thick = stack_case(7 + nY, nY)
    cases = [(0.55, 0.0, 0.0), (0.7, 35.0, 20.0), (0.62, 60.0, 0.0)]
    nX = 3
    lam = [c[0] for c in cases]
    kk = [bloch(c[0], eps[0][0, 0], mu[0][0, 0], c[1], c[2]) for c in cases]
    pols = np.array([list(o.polarisation_basis(c[1], c[2])) for c in cases], dtype=complex)
    res = engine.solve_batch(gpu_ctx, nX, nY, xe, ye, thick, lam, [k[0] for k in kk],
                             [k[1] for k in kk], np.array(eps), pols, np.array(mu))
    assert not res.status.any()
    for i, (lm, th, ph) in enumerate(cases):
        REF, TRN, _ = o.solve_stack(nX, nY, 0.5, 0.5, xe, ye, eps, thick, lm, th, ph,
                                    mu_layers=mu)
        for q in range(2):
            assert np.abs(res.ref[i, q] - REF[q]).max() < 1e-10
            assert np.abs(res.trn[i, q] - TRN[q]).max() < 1e-10


def test_magnetic_duality(gpu_ctx):
    # swapping eps and mu of the grating (self-dual outer media) maps TE on TM
    xe, ye = [0.0, 0.3, 1.0], [0.0, 0.6, 1.0]
    ea = np.array([[3.0 + 0.1j, 1.0], [2.0, 1.2]])
    ma = np.array([[1.5, 1.0], [1.0, 2.5 + 0.05j]])
    one, two = np.full((2, 2), 1.0 + 0j), np.full((2, 2), 2.0 + 0j)
    k = bloch(0.8, 1.0, 1.0, 25.0, 30.0)
    pols = np.array([list(o.polarisation_basis(25.0, 30.0))], dtype=complex)
    eps = np.array([[one, ea, two], [one, ma, two]])
    mu = np.array([[one, ma, two], [one, ea, two]])
    res = engine.solve_batch(gpu_ctx, 3, 2, xe, ye, [0.0, 0.4, 0.0], [0.8, 0.8], [k[0]] * 2,
                             [k[1]] * 2, eps, np.repeat(pols, 2, axis=0), mu)
    assert not res.status.any()
    assert np.abs(res.ref[0, 0] - res.ref[1, 1]).max() < 1e-12
    assert np.abs(res.ref[0, 1] - res.ref[1, 0]).max() < 1e-12
    assert np.abs(res.trn[0, 0] - res.trn[1, 1]).max() < 1e-12


def orcwa_interface(ctx, tmp_path, eps, mu, theta, pol, name):
    text = f"""OpenRCWA 4 2
title = magnetic interface
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -5e-07 10 0.0 10 1e-06
material = 1 {eps!r} 0 {mu!r} 0
geometry = 2 1 -2.5e-07 2.5e-07 -2.5e-07 2.5e-07 -5e-07 0
planewave = {theta} 0 {pol}
pbc = 1 1 0
rcwaorder = 4 0
wavelength = 0.6
end
"""
    p = tmp_path / name
    p.write_text(text)
    return oc.run_rcwa(oc.parse_rcwa_input(str(p)), ctx)[0]


def test_orcwa_magnetic_material(gpu_ctx, tmp_path):
    r0 = orcwa_interface(gpu_ctx, tmp_path, 4.0, 4.0, 0.0, 1, "m0.orcwa")
    assert r0.R < 1e-14 and abs(r0.T - 1.0) < 1e-12
    for th in (30.0, 45.0, 60.0):
        for pol, key in ((1, "tm"), (2, "te")):
            r = orcwa_interface(gpu_ctx, tmp_path, 4.0, 4.0, th, pol, f"m{th}{pol}.orcwa")
            assert abs(r.R - magnetic_fresnel_R(th, 4.0, 4.0, key)) < 1e-12
    tm = orcwa_interface(gpu_ctx, tmp_path, 2.25, 3.0, 45.0, 1, "a1.orcwa")
    te = orcwa_interface(gpu_ctx, tmp_path, 2.25, 3.0, 45.0, 2, "a2.orcwa")
    assert abs(tm.R - magnetic_fresnel_R(45.0, 2.25, 3.0, "tm")) < 1e-12
    assert abs(te.R - magnetic_fresnel_R(45.0, 2.25, 3.0, "te")) < 1e-12
    assert abs(tm.R - te.R) > 1e-2
    swap = orcwa_interface(gpu_ctx, tmp_path, 3.0, 2.25, 45.0, 1, "s1.orcwa")
    assert abs(swap.R - te.R) < 1e-12


def test_orcwa_material_mu_and_magnetic_loss(gpu_ctx, tmp_path):
    def run(lines, name, zmesh="-5e-07 10 0.0 10 1e-06", z="-5e-07 0"):
        text = f"""OpenRCWA 4 2
title = material_mu
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = {zmesh}
{lines}geometry = 2 1 -2.5e-07 2.5e-07 -2.5e-07 2.5e-07 {z}
planewave = 0 0 1
pbc = 1 1 0
rcwaorder = 4 0
wavelength = 0.6
end
"""
        p = tmp_path / name
        p.write_text(text)
        return oc.run_rcwa(oc.parse_rcwa_input(str(p)), gpu_ctx)[0]
    a = run("material = 1 4.0 0 1.0 0\nmaterial_mu = 2 4.0 0.0\n", "kw.orcwa")
    b = run("material = 1 4.0 0 4.0 0\n", "amur.orcwa")
    assert abs(a.R - b.R) < 1e-14 and a.R < 1e-14
    slab = dict(zmesh="-1e-06 10 0.0 10 3e-07 10 7e-07", z="0 3e-07")
    lossless = run("material = 1 4.0 0 4.0 0\n", "ll.orcwa", **slab)
    lossy = run("material = 1 4.0 0 1.0 0\nmaterial_mu = 2 4.0 0.5\n", "ly.orcwa", **slab)
    assert abs(lossless.A) < 1e-12
    assert lossy.A > 0.05 and lossy.R + lossy.T < 1.0 - 1e-3
    # Airy slab with n = sqrt(eps mu), impedance Z = sqrt(mu / eps) at normal incidence
    n2, z2 = cmath.sqrt(4.0 * (4.0 + 0.5j)), cmath.sqrt((4.0 + 0.5j) / 4.0)
    r12 = (z2 - 1) / (z2 + 1)
    ph = cmath.exp(2j * 2 * math.pi / 0.6 * n2 * 0.3)
    r = (r12 - r12 * ph) / (1 - r12 * r12 * ph)
    assert abs(lossy.R - abs(r) ** 2) < 1e-12
    c = run("material = 1 2.25 0\n", "c.orcwa")
    d = run("material = 1 2.25 0 1.0 0\n", "d.orcwa")
    assert abs(c.R - d.R) < 1e-15
