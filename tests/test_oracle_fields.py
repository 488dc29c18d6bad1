"""The oracle's field reconstruction (rcwa/RCWASolver.cpp:707-1011,
1322-1487, 1505-1645) pinned by physics that does not go through it:

* the fields inside a homogeneous slab from the 2x2 boundary conditions of
  plane waves (tangential E and Z0 H continuous), for TE and TM at any angle,
  with the reference's longitudinal components Ez = i [[eps]]^-1 (Kx Hy -
  Ky Hx) = -i k0 Ez_phys and Hz = (i / k0^2)(Kx Ey - Ky Ex) = (i / k0) Z0 Hz_phys
  (Ampere and Faraday for a plane wave);
* on patterned lossless stacks, the z-flux Re sum(Ex Hy* - Ey Hx*) constant
  through every interior layer and equal to T times the incident flux; the
  tangential harmonics continuous across every interface and matching the
  incident + reflected and the transmitted fields at the two ends;
* the reference's own checks of tests/test_rcwa_input.cpp: selection rules
  (Test 13, 22), |f|^2 = Re^2 + Im^2 (23), the Bloch envelope phase (26), the
  YZ slice with nx != ny (27), no interior layer (32), the device image (36).
"""
import cmath
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o


def slab_solver(theta, nHx=4, nHy=0, eps2=2.25, lam=0.6, Lx=0.5, slab=0.3):
    s = o.RCWASolver(nHx, nHy)
    s.disable_pml()
    for e in (1.0, eps2, 1.0):
        s.add_layer(o.Layer([-Lx / 2, Lx / 2], [-Lx / 2, Lx / 2], [[e]]))
    k0 = 2 * math.pi / lam
    s.set_bloch_wavevector(k0 * math.sin(math.radians(theta)), 0.0)
    s.solve(lam, [0, 1, 2], [0.0, slab, 0.0])
    return s


def slab_closed_form(theta, pol, eps=(1.0, 2.25, 1.0), lam=0.6, d=0.3):
    """Plane waves in the slab for a unit tangential incident amplitude at the
    top face: E = a e^{i kz z} + b e^{-i kz z}, Z0 H_t = Y (a e^{..} - b e^{..})
    with Y = k0 eps / kz (TM, H_t = Hy) or -kz / k0 (TE, H_t = Hx)."""
    k0 = 2 * math.pi / lam
    kx = k0 * math.sin(math.radians(theta))
    kz = [cmath.sqrt(k0 * k0 * e - kx * kx) for e in eps]
    Y = [(k0 * e / q if pol == "tm" else -q / k0) for e, q in zip(eps, kz)]
    ph = cmath.exp(1j * kz[1] * d)
    M = np.array([[-1, 1, 1, 0], [Y[0], Y[1], -Y[1], 0], [0, ph, 1 / ph, -1],
                  [0, Y[1] * ph, -Y[1] / ph, -Y[2]]], complex)
    b1, a2, b2, a3 = np.linalg.solve(M, np.array([1, Y[0], 0, 0], complex))

    def at(z):
        dn, up = a2 * cmath.exp(1j * kz[1] * z), b2 * cmath.exp(-1j * kz[1] * z)
        E, Ht = dn + up, Y[1] * (dn - up)
        if pol == "tm":
            ez_phys = -(kx / kz[1]) * (dn - up)            # k . E = 0
            return {o.EX: E, o.HY: Ht, o.EZ: -1j * k0 * ez_phys}
        return {o.EY: E, o.HX: Ht, o.HZ: 1j * kx * E / (k0 * k0)}
    return at, kx


@pytest.mark.parametrize("theta,pol", [(0.0, "tm"), (30.0, "tm"), (65.0, "tm"),
                                       (30.0, "te"), (50.0, "te")])
def test_slab_fields_closed_form(theta, pol):
    s = slab_solver(theta)
    c = s.generate_horizontal_plane_wave(*((1.0, 0.0) if pol == "tm" else (0.0, 1.0)), 0)
    at, kx = slab_closed_form(theta, pol)
    a1, a2 = o._image_axes(s, [0, 1, 2], [0.0, 0.3, 0.0], o.SLICE_XZ, 40)
    for comp in at(0.0):
        img = o.field_image(s, c, [0, 1, 2], [0.0, 0.3, 0.0], o.SLICE_XZ, 0.0, comp, nres1=40)
        ref = np.array([[at(z)[comp] * cmath.exp(1j * kx * x) for x in a1] for z in a2])
        assert img.shape == (24, 40)
        assert np.abs(img - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def patterned_stack(seed=3, lossless=True):
    rng = np.random.default_rng(seed)
    xe, ye = [0.0, 0.13, 0.31, 0.5], [0.0, 0.22, 0.45]
    shp = (2, 3)
    loss = 0.0 if lossless else 0.2
    eps = [np.full(shp, 1.0 + 0j), rng.uniform(1.5, 6, shp) + 1j * loss,
           rng.uniform(2, 9, shp) + 0j, np.full(shp, 2.5 + 0j),
           rng.uniform(1, 4, shp) + 0j, np.full(shp, 1.8 + 0j)]
    thick = [0.0, 0.11, 0.07, 0.05, 0.13, 0.0]
    return xe, ye, eps, thick


def test_flux_and_continuity_patterned():
    xe, ye, eps, thick = patterned_stack()
    th, ph = 25.0, 30.0
    REF, TRN, s = o.solve_stack(3, 2, 0.5, 0.45, xe, ye, eps, thick, 0.62, th, ph)
    stack = list(range(6))
    h = s.nx * s.ny
    for (px, py), R, T in zip(o.polarisation_basis(th, ph), REF, TRN):
        c = s.generate_horizontal_plane_wave(px, py, 0)
        c_m, c_p = o.intermediate_field(s, c, stack, thick)

        def harm(j, zp, comp):
            return o.layer_harmonics(s, c_m, c_p, stack, j, zp, zp - thick[j], comp)

        def flux(ex, ey, hx, hy):
            return np.real(np.sum(ex * np.conj(hy) - ey * np.conj(hx)))

        L0, Lt = s.layers[0], s.layers[5]
        Ei, Hi = L0.W @ c, L0.V @ c
        s_inc = flux(Ei[:h], Ei[h:], Hi[:h], Hi[h:])
        for j in range(1, 5):
            for f in (0.0, 0.4, 1.0):
                val = flux(*(harm(j, f * thick[j], q) for q in (o.EX, o.EY, o.HX, o.HY)))
                assert abs(val / s_inc - T.sum()) < 1e-12
        # interfaces: tangential harmonics continuous
        for j in range(1, 4):
            for q in (o.EX, o.EY, o.HX, o.HY):
                a, b = harm(j, thick[j], q), harm(j + 1, 0.0, q)
                assert np.abs(a - b).max() < 1e-12
        # ends: incident + reflected above, transmitted below
        r, t = s.Rud @ c, s.Tdd @ c
        top = {o.EX: (L0.W @ (c + r))[:h], o.EY: (L0.W @ (c + r))[h:],
               o.HX: (L0.V @ (c - r))[:h], o.HY: (L0.V @ (c - r))[h:]}
        bot = {o.EX: (Lt.W @ t)[:h], o.EY: (Lt.W @ t)[h:],
               o.HX: (Lt.V @ t)[:h], o.HY: (Lt.V @ t)[h:]}
        for q in top:
            assert np.abs(harm(1, 0.0, q) - top[q]).max() < 1e-12
            assert np.abs(harm(4, thick[4], q) - bot[q]).max() < 1e-12


def test_hz_divergence_free_patterned():
    # div B = 0 per harmonic in a non-magnetic patterned layer:
    # d(Hz_ref)/dz = (Kx Hx + Ky Hy) / k0 for Hz_ref = (i/k0) Z0 Hz
    xe, ye, eps, thick = patterned_stack(seed=5, lossless=False)
    lam = 0.58
    REF, TRN, s = o.solve_stack(3, 2, 0.5, 0.45, xe, ye, eps, thick, lam, 20.0, 10.0)
    c = s.generate_horizontal_plane_wave(0.3, 0.8, 0)
    stack = list(range(6))
    c_m, c_p = o.intermediate_field(s, c, stack, thick)
    k0 = 2 * math.pi / lam
    for j in (1, 2, 4):
        z, dz = 0.5 * thick[j], 1e-5
        def hz(zp):
            return o.layer_harmonics(s, c_m, c_p, stack, j, zp, zp - thick[j], o.HZ)
        d = (hz(z + dz) - hz(z - dz)) / (2 * dz)
        hx = o.layer_harmonics(s, c_m, c_p, stack, j, z, z - thick[j], o.HX)
        hy = o.layer_harmonics(s, c_m, c_p, stack, j, z, z - thick[j], o.HY)
        rhs = (s.Kx * hx + s.Ky * hy) / k0
        assert np.abs(d - rhs).max() < 1e-7 * np.abs(rhs).max()


# ---- the reference's checks (tests/test_rcwa_input.cpp) -------------------
def uniform_image(theta, pol, comp, slice_type=o.SLICE_XZ, nHy=0, nres1=500):
    s = slab_solver(theta, nHy=nHy)
    c = s.generate_horizontal_plane_wave(*((0.0, 1.0) if pol == 2 else (1.0, 0.0)), 0)
    return o.field_image(s, c, [0, 1, 2], [0.0, 0.3, 0.0], slice_type, 0.0, comp, nres1)


def test_longitudinal_selection_rules():
    # Test 13 (tests/test_rcwa_input.cpp:567-661)
    assert np.abs(uniform_image(0.0, 1, o.EZ, nres1=100)).max() < 1e-6
    assert np.abs(uniform_image(0.0, 1, o.HZ, nres1=100)).max() < 1e-6
    assert np.abs(uniform_image(30.0, 1, o.EZ, nres1=100)).max() > 1e-3
    assert np.abs(uniform_image(30.0, 1, o.HZ, nres1=100)).max() < 1e-6
    assert np.abs(uniform_image(30.0, 2, o.HZ, nres1=100)).max() > 1e-3
    assert np.abs(uniform_image(30.0, 2, o.EZ, nres1=100)).max() < 1e-6


def test_transverse_selection_rules():
    # Test 22 (tests/test_rcwa_input.cpp:1171-1210)
    mx = {q: np.abs(uniform_image(0.0, 1, q, nres1=100)).max()
          for q in (o.EX, o.EY, o.HX, o.HY)}
    assert mx[o.EX] > 0.5 and mx[o.HY] > 1e-3 and mx[o.EY] < 1e-6 and mx[o.HX] < 1e-6
    mx = {q: np.abs(uniform_image(0.0, 2, q, nres1=100)).max()
          for q in (o.EX, o.EY, o.HX, o.HY)}
    assert mx[o.EY] > 0.5 and mx[o.HX] > 1e-3 and mx[o.EX] < 1e-6 and mx[o.HY] < 1e-6


def test_bloch_envelope_phase():
    # Test 26 (tests/test_rcwa_input.cpp:1400-1460): the first row of an XZ
    # image of a uniform stack at 30 deg advances in phase as exp(i kx0 x)
    row0 = uniform_image(0.0, 1, o.EX)[0]
    assert np.ptp(row0.real) < 1e-6
    row = uniform_image(30.0, 1, o.EX)[0]
    assert np.ptp(row.real) > 1e-2 and np.ptp(np.abs(row)) < 1e-6
    kx0 = 2 * math.pi / 0.6 * math.sin(math.radians(30.0))
    n = row.size
    dphi = math.remainder(cmath.phase(row[-1]) - cmath.phase(row[0]), 2 * math.pi)
    assert abs(dphi - kx0 * 0.5 / n * (n - 1)) < 1e-3


def test_slice_yz_non_square():
    # Test 27 (tests/test_rcwa_input.cpp:1462-1480): nx = 9, ny = 3
    img = uniform_image(0.0, 1, o.EX, slice_type=o.SLICE_YZ, nHy=1)
    assert img.shape == (300, 500)
    assert 0.5 < np.abs(img).max() < 1e100


def test_no_interior_layer_rejected():
    # Test 32 (tests/test_rcwa_input.cpp:1720-1760)
    s = o.RCWASolver(4, 0)
    s.disable_pml()
    for e in (1.0, 2.25):
        s.add_layer(o.Layer([-0.25, 0.25], [-0.25, 0.25], [[e]]))
    s.set_bloch_wavevector(0.0, 0.0)
    s.solve(0.6, [0, 1], [0.0, 0.0])
    c = s.generate_horizontal_plane_wave(1.0, 0.0, 0)
    with pytest.raises(ValueError):
        o.field_image(s, c, [0, 1], [0.0, 0.0], o.SLICE_XZ, 0.0, o.EZ)


def test_device_image_two_materials():
    # Test 36 (tests/test_rcwa_input.cpp:1937-1990): a half-filled grating
    s = o.RCWASolver(4, 0)
    s.disable_pml()
    cx, cy = [-0.25, 0.0, 0.25], [-0.25, 0.25]
    for g in ([[1.0, 1.0]], [[2.25, 1.0]], [[1.0, 1.0]], [[1.0, 1.0]]):
        s.add_layer(o.Layer(cx, cy, g))
    img = o.device_image(s, [0, 1, 2, 3], [0.0, 0.3, 0.7, 0.0], o.SLICE_XZ, 0.0)
    assert img.shape == (4000, 2000)
    assert abs(img.real.min() - 1.0) < 1e-12 and abs(img.real.max() - 2.25) < 1e-12
    # rows above z = 0.3 (1200 of 4000 over 1.0) see the grating, the rest
    # the homogeneous layer
    assert np.all(img[:1200].real[:, :1000] == 2.25) and np.all(img[:1200].real[:, 1000:] == 1.0)
    assert np.all(img[1200:].real == 1.0)
