"""Pinning of the FP64 oracle (CPU only).

Each check is fixed by something other than the oracle's own formulas: the
reference's published values (tests/golden/published_values.json), closed
forms (Fresnel, the thin-film characteristic matrix, exact Fourier integrals by
quadrature, homogeneous-layer dispersion), and the invariances the reference's
own CI asserts (ci/check_rcwa_ofd*.py, tests/test_rcwa_input.cpp).
"""
import cmath
import json
import math
import os

import numpy as np
import pytest

from oracle import rcwa_oracle as o

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "published_values.json")
C0 = 2.99792458e8


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def L(e1, e2=None, fill=0.5, d=0.0, e1i=0.0, e2i=0.0, fy=1.0):
    e2 = e1 if e2 is None else e2
    return dict(eps1=e1, eps2=e2, eps1i=e1i, eps2i=e2i, fill=fill, fill_y=fy, thickness=d)


# ---------------------------------------------------------------- closed forms
def fresnel(theta_deg, n1=1.0, n2=1.5):
    """|r_s|^2, |r_p|^2 (ci/check_rcwa_ofd_incidence.py:68-76 formulas)."""
    ti = math.radians(theta_deg)
    st = n1 * math.sin(ti) / n2
    ci, ct = math.cos(ti), math.sqrt(1 - st * st)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    return rs * rs, rp * rp


def film_rt(eps_film, lam, d, theta_deg, pol, eps_sub=2.25, eps_inc=1.0):
    """Single film between two half-spaces by the characteristic (Airy) matrix,
    written with exp(-i w t) (loss = +Im eps): admittances eta = N cos for TE,
    N / cos for TM, phase delta = k0 N cos d."""
    k0 = 2 * math.pi / lam
    n0 = cmath.sqrt(eps_inc)
    s = n0 * math.sin(math.radians(theta_deg))

    def cosang(N):
        c = cmath.sqrt(1 - (s / N) ** 2)
        if (N * c).imag < 0:       # decaying branch for exp(-i w t)
            c = -c
        return c

    def eta(N):
        c = cosang(N)
        return N * c if pol == "TE" else N / c

    N1, Ns = cmath.sqrt(eps_film), cmath.sqrt(eps_sub)
    e0, e1, es = eta(n0), eta(N1), eta(Ns)
    delta = k0 * N1 * cosang(N1) * d
    B = cmath.cos(delta) - 1j * cmath.sin(delta) / e1 * es
    C = -1j * e1 * cmath.sin(delta) + cmath.cos(delta) * es
    r = (e0 * B - C) / (e0 * B + C)
    T = 4 * e0.real * es.real / abs(e0 * B + C) ** 2
    return abs(r) ** 2, T


def drude(f, wp=1.36e16, g=3.0e13):
    w = 2 * math.pi * f
    return 1.0 + wp * wp / complex(-w * w, -g * w)


# ------------------------------------------------------------- published values
@pytest.mark.parametrize("case", golden()["grating"]["cases"], ids=lambda c: f"N{c['N']}")
def test_published_grating(case):
    g = golden()["grating"]
    r = o.ofd_solve_frequency(case["N"], g["period"], g["layers"], g["freq"])
    assert abs(r["R"][0] - case["R_TE"]) < 5.1e-7
    assert abs(r["R"][1] - case["R_TM"]) < 5.1e-7


def test_published_metal_film():
    g = golden()["metal"]
    r = o.ofd_solve_frequency(g["N"], g["period"], g["layers"], g["freq"])
    for pol in (0, 1):
        assert abs(r["R"][pol] - g["R"]) < 5.1e-7
        assert abs(r["T"][pol] - g["T"]) < 5.1e-7
        assert 1 - r["R"][pol] - r["T"][pol] > 1e-3       # absorbing


def test_published_brewster():
    g = golden()["oblique"]
    r = o.ofd_solve_frequency(g["N"], g["period"], g["layers"], g["freq"], theta=g["theta"])
    assert abs(r["R"][0] - g["R_TE"]) < 5.1e-7
    assert r["R"][1] < g["R_TM_max"]


# ------------------------------------------------------------- analytic checks
@pytest.mark.parametrize("theta,phi", [(0, 0), (20, 0), (40, 0), (60, 0), (45, 37), (70, 0)])
def test_fresnel_interface(theta, phi):
    r = o.ofd_solve_frequency(2, 4e-7, [L(1.0), L(2.25)], 5e14, theta=theta, phi=phi)
    rs, rp = fresnel(theta)
    assert abs(r["R"][0] - rs) < 1e-12
    assert abs(r["R"][1] - rp) < 1e-12
    for p in (0, 1):
        assert abs(r["R"][p] + r["T"][p] - 1) < 1e-12


@pytest.mark.parametrize("theta", [0.0, 30.0])
def test_drude_film_matches_airy(theta):
    layers = [L(1.0), L(1.0, d=3e-8), L(2.25)]
    poles = {(1, 1): [(1.0, 1.36e16, 3.0e13, 0.0)], (1, 2): [(1.0, 1.36e16, 3.0e13, 0.0)]}
    for f in np.linspace(4.0e14, 7.0e14, 4):
        r = o.ofd_solve_frequency(2, 5e-7, layers, f, theta=theta, poles=poles)
        for p, pol in enumerate(("TE", "TM")):
            R, T = film_rt(drude(f), C0 / f, 3e-8, theta, pol)
            assert abs(r["R"][p] - R) < 1e-10
            assert abs(r["T"][p] - T) < 1e-10


def test_ar_coating_minimum():
    layers = [L(1.0), L(1.5, d=1.122680e-7), L(2.25)]
    f = C0 / 550e-9
    r = o.ofd_solve_frequency(3, 3e-7, layers, f)
    R, T = film_rt(1.5, 550e-9, 1.122680e-7, 0.0, "TE")
    assert abs(r["R"][0] - R) < 1e-12 and R < 1e-10
    r2 = o.ofd_solve_frequency(3, 3e-7, [L(1.0), L(2.25)], f)
    assert abs(r2["R"][0] - 0.04) < 1e-12


def test_fourier_coefficients_quadrature():
    x = np.array([0.0, 0.13, 0.5, 0.71, 1.0]) * 3.0
    f = np.array([2 + 0.1j, 1.0, 5.5, 1 - 0.2j])
    M = 9
    F = o.fourier_1d(x, f, 3.0, M)
    # midpoint quadrature of (1/d) int f(x) exp(-i 2 pi m x / d) dx
    Nq = 400000
    xs = (np.arange(Nq) + 0.5) * (3.0 / Nq)
    fx = f[np.searchsorted(x, xs, side="right") - 1]
    for i in range(M):
        m = i - (M - 1) // 2
        ref = np.mean(fx * np.exp(-2j * np.pi * m * xs / 3.0))
        assert abs(F[i] - ref) < 2e-5 * (1 + abs(m))


def test_toeplitz_is_convolution():
    # (f g)_m = sum_k f_{m-k} g_k exactly for a trigonometric polynomial g
    x = np.array([0.0, 0.3, 0.45, 1.0])
    f = np.array([3.0, 1.0 + 0.5j, 2.0])
    N = 4
    Ff = o.fourier_1d(x, f, 1.0, 4 * N + 1)
    g = np.zeros(2 * N + 1, complex)
    g[N - 1], g[N], g[N + 2] = 0.5, 1.0, -0.25j      # orders -1, 0, +2
    T = o.toeplitz_dense(Ff)   # 4N+1 coefficients -> (2N+1) x (2N+1), T[r,c] = F_{r-c}
    Tfull = np.array([[Ff[(r - c) + 2 * N] for c in range(2 * N + 1)] for r in range(2 * N + 1)])
    assert T.shape[0] == 2 * N + 1 and np.allclose(T, Tfull)
    Nq = 200000
    xs = (np.arange(Nq) + 0.5) / Nq
    fx = f[np.searchsorted(x, xs, side="right") - 1]
    gx = sum(g[k] * np.exp(2j * np.pi * (k - N) * xs) for k in range(2 * N + 1))
    for r in range(2 * N + 1):
        m = r - N
        ref = np.mean(fx * gx * np.exp(-2j * np.pi * m * xs))
        assert abs((T @ g)[r] - ref) < 5e-5


def test_block_toeplitz_index_map():
    nx, ny = 3, 2
    f2 = np.arange((2 * nx - 1) * (2 * ny - 1)).reshape(2 * nx - 1, 2 * ny - 1) + 0j
    B = o.block_toeplitz_dense(f2)
    for m in range(nx):
        for n in range(ny):
            for j in range(nx):
                for l in range(ny):
                    assert B[m * ny + n, j * ny + l] == f2[m - j + nx - 1, n - l + ny - 1]


def test_uniform_layer_convolution_and_modes():
    xe, ye = [0, 1e-7, 4e-7], [0, 2.5e-7, 4e-7]
    eps = np.full((2, 2), 3.1 + 0.2j)
    for M in o.convolution_matrices(xe, ye, eps, 5, 3):
        assert np.abs(M - (3.1 + 0.2j) * np.eye(15)).max() < 1e-13
    lay = o.Layer(xe, ye, eps)
    s = o.RCWASolver(2, 1)
    s.disable_pml()
    s.add_layer(lay)
    lam = 500e-9
    k0 = 2 * np.pi / lam
    s.set_bloch_wavevector(0.3 * k0, -0.1 * k0)
    a, b = s.alpha_beta()
    lay.solve(np.diag(a), np.diag(b), lam, 5, 3)
    kz = np.sqrt((3.1 + 0.2j) * k0 ** 2 - a ** 2 - b ** 2 + 0j)
    expect = np.sort_complex(np.concatenate([kz, kz]))
    got = np.sort_complex(np.where(lay.gamma.imag < 0, -lay.gamma, lay.gamma))
    assert np.abs(np.sort_complex(lay.gamma ** 2) - np.sort_complex(expect ** 2)).max() < 1e-6 * k0 ** 2
    assert got.size == 30


def test_li_rules_reduce_for_one_dimensional_grids():
    # y-uniform grid: inverse rule in y is the Laurent rule (fe22 == fe33);
    # x-uniform grid: fe11 == fe33
    xe, ye = [0, 1.3e-7, 4e-7], [0, 4e-7]
    eps = np.array([[4.0 + 0.1j, 1.0]])
    f11, f22, f33 = o.convolution_matrices(xe, ye, eps, 5, 3)
    assert np.abs(f22 - f33).max() < 1e-12
    assert np.abs(f11 - f33).max() > 1e-2
    eps_t = eps.T.copy()
    g11, g22, g33 = o.convolution_matrices([0, 4e-7], [0, 1.3e-7, 4e-7], eps_t, 3, 5)
    assert np.abs(g11 - g33).max() < 1e-12


def test_gamma_sqrt_branch():
    assert o.gamma_sqrt(np.array([4.0]))[0] == 2.0
    assert o.gamma_sqrt(np.array([-4.0 + 0j]))[0] == 2j
    rng = np.random.default_rng(1)
    g = rng.standard_normal(2000) + 1j * rng.standard_normal(2000)
    v = o.gamma_sqrt(g)
    assert np.abs(v * v - g).max() < 1e-14
    ang = np.angle(v)
    assert np.all((ang > -np.pi / 4 - 1e-15) & (ang <= 3 * np.pi / 4 + 1e-15))


# ---------------------------------------------------- invariances (reference CI)
GRAT = [L(1.0), L(4.0, 1.0, fill=0.4, d=2.5e-7), L(2.25)]


def test_energy_conservation_lossless():
    for th, ph in ((0, 0), (40, 30), (60, 75)):
        r = o.ofd_solve_frequency(6, 4e-7, GRAT, 5.5e14, theta=th, phi=ph)
        for p in (0, 1):
            assert abs(r["R"][p] + r["T"][p] - 1) < 1e-9


def test_mirror_symmetry_orders():
    r = o.ofd_solve_frequency(6, 3e-7, [L(1.0), L(4.0, 1.0, d=2e-7), L(2.25)], 6.5e14)
    m = r["m"]
    for p in (0, 1):
        for k in range(len(m)):
            kk = int(np.where(m == -m[k])[0][0])
            assert abs(r["REF"][p][k] - r["REF"][p][kk]) < 1e-10
            assert abs(r["TRN"][p][k] - r["TRN"][p][kk]) < 1e-10
    a = o.ofd_solve_frequency(6, 4e-7, GRAT, 5.5e14, theta=25, phi=0)
    b = o.ofd_solve_frequency(6, 4e-7, GRAT, 5.5e14, theta=-25, phi=0)
    c = o.ofd_solve_frequency(6, 4e-7, GRAT, 5.5e14, theta=25, phi=180)
    for p in (0, 1):
        assert abs(a["R"][p] - b["R"][p]) < 1e-10 and abs(a["R"][p] - c["R"][p]) < 1e-10


def test_2d_with_uniform_y_equals_1d():
    l1 = [L(1.0), L(2.25, 1.0, fill=0.4, d=2e-7), L(2.25)]
    l2 = [L(1.0), L(2.25, 1.0, fill=0.4, d=2e-7, fy=1.0), L(2.25)]
    a = o.ofd_solve_frequency(4, 5e-7, l1, 4.0e14, theta=20, phi=30)
    b = o.ofd_solve_frequency(4, 5e-7, l2, 4.0e14, theta=20, phi=30, nY=4, period_y=5e-7)
    for p in (0, 1):
        assert abs(a["R"][p] - b["R"][p]) < 1e-10 and abs(a["T"][p] - b["T"][p]) < 1e-10


def test_c4_square_pillar_te_equals_tm():
    lay = [L(1.0), L(2.25, 1.0, fill=0.6, d=2e-7, fy=0.6), L(2.25)]
    r = o.ofd_solve_frequency(3, 5e-7, lay, 4.0e14, nY=3, period_y=5e-7)
    assert abs(r["R"][0] - r["R"][1]) < 1e-9 and abs(r["T"][0] - r["T"][1]) < 1e-9
    assert abs(r["R"][0] + r["T"][0] - 1) < 1e-9


def test_rotation_swaps_te_tm():
    ly = [L(1.0), L(2.25, 1.0, fill=0.4, d=2e-7, fy=1.0), L(2.25)]
    lx = [L(1.0), L(2.25, 1.0, fill=1.0, d=2e-7, fy=0.4), L(2.25)]
    a = o.ofd_solve_frequency(3, 5e-7, ly, 4.0e14, nY=3, period_y=5e-7)
    b = o.ofd_solve_frequency(3, 5e-7, lx, 4.0e14, nY=3, period_y=5e-7)
    assert abs(a["R"][0] - b["R"][1]) < 1e-9 and abs(a["R"][1] - b["R"][0]) < 1e-9
    assert abs(a["R"][0] - a["R"][1]) > 1e-3


def test_side_swap_translation_invariance():
    poles1 = {(1, 1): [(2.0, 3.0e15, 2.0e14, 4.0e15)]}
    poles2 = {(1, 2): [(2.0, 3.0e15, 2.0e14, 4.0e15)]}
    g1 = [L(1.0), L(1.0, 2.25, fill=0.3, d=1e-7), L(2.25)]
    g2 = [L(1.0), L(2.25, 1.0, fill=0.7, d=1e-7), L(2.25)]
    a = o.ofd_solve_frequency(6, 5e-7, g1, 5e14, theta=20, poles=poles1)
    b = o.ofd_solve_frequency(6, 5e-7, g2, 5e14, theta=20, poles=poles2)
    for p in (0, 1):
        assert abs(a["R"][p] - b["R"][p]) < 1e-9 and abs(a["T"][p] - b["T"][p]) < 1e-9
        assert 1 - a["R"][p] - a["T"][p] > 1e-4


def test_two_half_poles_equal_one_pole():
    wp, g = 1.36e16, 3.0e13
    one = {(1, 1): [(1.0, wp, g, 0.0)], (1, 2): [(1.0, wp, g, 0.0)]}
    two = {(1, 1): [(1.0, wp * math.sqrt(0.3), g, 0.0), (1.0, wp * math.sqrt(0.7), g, 0.0)],
           (1, 2): [(1.0, wp, g, 0.0)]}
    layers = [L(1.0), L(1.0, d=3e-8), L(2.25)]
    a = o.ofd_solve_frequency(2, 5e-7, layers, 5e14, poles=one)
    b = o.ofd_solve_frequency(2, 5e-7, layers, 5e14, poles=two)
    for p in (0, 1):
        assert abs(a["R"][p] - b["R"][p]) < 1e-10


def test_fill_mixing_invariance():
    same = [L(1.0, fill=0.3), L(2.25, 1.0, fill=0.3, d=2e-7), L(2.25, fill=0.3)]
    mixed = [L(1.0, fill=0.5), L(2.25, 1.0, fill=0.3, d=2e-7), L(2.25, fill=0.8)]
    a = o.ofd_solve_frequency(4, 5e-7, same, 4.5e14, theta=20, phi=30)
    b = o.ofd_solve_frequency(4, 5e-7, mixed, 4.5e14, theta=20, phi=30)
    for p in (0, 1):
        assert abs(a["R"][p] - b["R"][p]) < 1e-10


def test_grating_equation_orders():
    r = o.ofd_solve_frequency(8, 3e-7, [L(1.0), L(4.0, 1.0, d=2e-7), L(2.25)], 7.5e14, theta=30)
    refl = {int(mm) for mm, v in zip(r["m"], r["REF"][0]) if v > 0 and mm != 0}
    assert refl == {-1}
    for p in (0, 1):
        assert abs(r["REF"][p].sum() - r["R"][p]) < 1e-12


def test_eigen_backend_residual():
    st = o.EigenStats()
    lay = o.Layer([0, 1.5e-7, 4e-7], [0, 2.5e-7, 4e-7], np.array([[5 + 0.05j, 1], [1, 1]]))
    s = o.RCWASolver(3, 3)
    s.disable_pml()
    s.add_layer(lay)
    s.set_bloch_wavevector(1e6, 2e6)
    a, b = s.alpha_beta()
    lay.solve(np.diag(a), np.diag(b), 500e-9, 7, 7, st)
    assert st.solves == 1 and st.fallbacks == 0 and st.max_residual < 1e-12


def test_eigen_recovery_branch_perturbation(monkeypatch):
    # the recovery solve of rcwa/MKLEigenSolver.hpp:81-93, forced by a negative
    # acceptance tolerance: on a diagonal matrix its eigenvalues are the
    # diagonal shifted by exactly 1e-11 (i + 1) max|A|, and its residual is
    # reported against the unperturbed matrix
    monkeypatch.setattr(o, "EIG_TOL", -1.0)
    diag = np.array([3.0 + 1j, -2.0, 0.5j, 7.0, 1.0 - 4j])
    A = np.diag(diag)
    st = o.EigenStats()
    d, V = o.eig_backend(A, st)
    assert st.fallbacks == 1
    scale = np.abs(diag).max()
    want = diag + 1e-11 * scale * np.arange(1, 6)
    order = [int(np.argmin(np.abs(want - x))) for x in d]
    assert sorted(order) == list(range(5))
    assert np.abs(d - want[order]).max() < 1e-15 * scale
    # residual against A, not A + D: the perturbation itself, 1e-11 scale i
    assert abs(st.max_residual - 1e-11 * 5 * scale / scale) < 1e-13


def test_eigen_recovery_keeps_patterned_results(monkeypatch):
    # every layer of a patterned 2D stack through the recovery branch: R and T
    # per order move by far less than the 1e-8 parity bar (the reference's
    # comment puts the eigenvalue shift at ~1e-11 |A|)
    layers = [L(1.0), L(6.0, 1.0, fill=0.4, d=1.5e-7, e1i=0.05, fy=0.6), L(2.25)]
    a = o.ofd_solve_frequency(3, 4e-7, layers, 5.5e14, theta=15, phi=30)
    monkeypatch.setattr(o, "EIG_TOL", -1.0)
    b = o.ofd_solve_frequency(3, 4e-7, layers, 5.5e14, theta=15, phi=30)
    for p in (0, 1):
        assert np.abs(a["REF"][p] - b["REF"][p]).max() < 1e-8
        assert np.abs(a["TRN"][p] - b["TRN"][p]).max() < 1e-8
        assert abs(a["R"][p] - b["R"][p]) > 0.0 or abs(a["T"][p] - b["T"][p]) > 0.0


def test_slab_scatter_matrix_uniform_closed_form():
    # a slab identical to the reference medium: no reflection, pure phase
    # (tests/test_core_integrators.cpp:79-112 checks the same closed form)
    rng = np.random.default_rng(0)
    n = 6
    W = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    V = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    gamma = rng.uniform(1e6, 2e7, n) + 1j * rng.uniform(0, 1e6, n)
    lam, wz = 5e-7, 3e-7
    Rud, Tuu = o.slab_scatter_matrix(W, V, gamma, W, V, wz, lam)
    X = np.exp(1j * gamma * wz * lam / (2 * np.pi))
    assert np.abs(Rud).max() < 1e-12
    assert np.abs(Tuu - np.diag(X)).max() < 1e-12
    # star of two half slabs == one full slab (associativity of the cascade)
    R1, T1 = o.slab_scatter_matrix(W, V, gamma, W, V, wz / 2, lam)
    Z = np.zeros((n, n), complex)
    st = (Z, np.eye(n, dtype=complex), Z, np.eye(n, dtype=complex))
    st = o.redheffer_star(st, R1, T1)
    st = o.redheffer_star(st, R1, T1)
    assert np.abs(st[1] - np.diag(X)).max() < 1e-12


def test_redheffer_star_of_patterned_slabs_equals_rtcm():
    # slabs in the incidence medium's basis, composed with the star product
    # (core/RCWAIntegrator.cpp:57-67), reproduce the RTCM cascade
    # (rcwa/RCWASolver.cpp:378-492) of the same stack: reflection terms of
    # every star enter through the later ones
    lam = 5.5e-7
    xe, ye = [0.0, 1.2e-7, 2.6e-7, 4.0e-7], [0.0, 1.8e-7, 4.0e-7]

    def grid(e1, e2):
        g = np.full((2, 3), e2, dtype=complex)
        g[0, :2] = e1
        return g

    air = o.Layer(xe, ye, np.full((2, 3), 1.0 + 0j))
    slabs = [o.Layer(xe, ye, grid(6.0 + 0.1j, 1.0)), o.Layer(xe, ye, grid(2.25, 4.0 + 0.05j)),
             o.Layer(xe, ye, grid(1.0, 9.0))]
    d = [1.3e-7, 0.9e-7, 0.6e-7]
    s = o.RCWASolver(2, 2)
    s.disable_pml()
    for lay in [air] + slabs:
        s.add_layer(lay)
    k0 = 2 * np.pi / lam
    s.set_bloch_wavevector(0.3 * k0, 0.1 * k0)
    s.solve(lam, [0, 1, 2, 3, 0], [0.0] + d + [0.0])
    n = s.Rud.shape[0]
    eye, zero = np.eye(n, dtype=complex), np.zeros((n, n), complex)
    st = (zero, eye, zero, eye)
    for lay, dz in zip(slabs, d):
        R, T = o.slab_scatter_matrix(lay.W, lay.V, lay.gamma * k0, air.W, air.V, dz, lam)
        st = o.redheffer_star(st, R, T)
    # (a dropped or transposed term is an O(1) difference)
    assert np.abs(st[0] - s.Rud).max() < 1e-10
    assert np.abs(st[3] - s.Tdd).max() < 1e-10


def test_redheffer_star_identity_and_associativity():
    # the star product has the identity (0, I, 0, I) on both sides and is
    # associative; a dropped or reordered term breaks one of the two
    rng = np.random.default_rng(4)
    n = 7

    def smat():
        return tuple(0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
                     for _ in range(4))

    eye, zero = np.eye(n, dtype=complex), np.zeros((n, n), complex)
    E = (zero, eye, zero, eye)
    a, b, c = smat(), smat(), smat()
    for x, y in zip(o.redheffer_star_general(E, a), a):
        assert np.abs(x - y).max() < 1e-13
    for x, y in zip(o.redheffer_star_general(a, E), a):
        assert np.abs(x - y).max() < 1e-13
    left = o.redheffer_star_general(o.redheffer_star_general(a, b), c)
    right = o.redheffer_star_general(a, o.redheffer_star_general(b, c))
    for x, y in zip(left, right):
        assert np.abs(x - y).max() < 1e-12


# ---------------------------------------------------------------------------
# magnetic layers (rcwa/Layer.cpp:129-145, 150-164, 206-213;
# rcwa/RCWASolver.cpp:1833-1907): closed-form magnetic Fresnel, and the
# electromagnetic duality of a patterned layer (swapping the eps and mu grids
# swaps TE and TM), which only holds if mu enters P where eps enters Q
# ---------------------------------------------------------------------------
def magnetic_fresnel_R(theta, e2, m2, pol):
    t = math.radians(theta)
    k1z, s = math.cos(t), math.sin(t)
    k2z = cmath.sqrt(e2 * m2 - s * s)
    r = (m2 * k1z - k2z) / (m2 * k1z + k2z) if pol == "te" else \
        (e2 * k1z - k2z) / (e2 * k1z + k2z)
    return abs(r) ** 2


@pytest.mark.parametrize("e2,m2", [(4.0, 4.0), (2.25, 3.0), (3.0, 2.25), (2.0 + 0.3j, 1.5 + 0.2j)])
@pytest.mark.parametrize("theta", [0.0, 30.0, 45.0, 60.0])
def test_magnetic_fresnel(e2, m2, theta):
    xe, ye = [0.0, 1.0], [0.0, 1.0]
    eps = [np.array([[1.0]]), np.array([[e2]])]
    mu = [np.array([[1.0]]), np.array([[m2]])]
    REF, TRN, _ = o.solve_stack(2, 0, 1.0, 1.0, xe, ye, eps, [0.0, 0.0], 0.6, theta, 0.0,
                                mu_layers=mu)
    assert abs(REF[0].sum() - magnetic_fresnel_R(theta, e2, m2, "te")) < 1e-12
    assert abs(REF[1].sum() - magnetic_fresnel_R(theta, e2, m2, "tm")) < 1e-12
    if e2 == m2 and theta == 0.0:
        assert REF[0].sum() < 1e-14        # impedance matched
    if isinstance(e2, float) and isinstance(m2, float):
        assert abs(REF[0].sum() + TRN[0].sum() - 1.0) < 1e-12


@pytest.mark.parametrize("nY,phi", [(0, 0.0), (2, 30.0)])
def test_magnetic_duality_patterned(nY, phi):
    # air / grating / self-dual substrate (eps = mu = 2): swapping the
    # grating's eps and mu grids maps TE onto TM order by order
    xe, ye = [0.0, 0.3, 1.0], [0.0, 0.6, 1.0]
    ea = np.array([[3.0 + 0.1j, 1.0], [2.0, 1.2]])
    ma = np.array([[1.5, 1.0], [1.0, 2.5 + 0.05j]])
    def run(e, m):
        eps = [np.full((2, 2), 1.0 + 0j), e, np.full((2, 2), 2.0 + 0j)]
        mu = [np.full((2, 2), 1.0 + 0j), m, np.full((2, 2), 2.0 + 0j)]
        return o.solve_stack(3, nY, 1.0, 1.0, xe, ye, eps, [0.0, 0.4, 0.0], 0.8, 25.0, phi,
                             mu_layers=mu)
    R1, T1, _ = run(ea, ma)
    R2, T2, _ = run(ma, ea)
    for a, b in ((0, 1), (1, 0)):
        assert np.abs(R1[a] - R2[b]).max() < 1e-12
        assert np.abs(T1[a] - T2[b]).max() < 1e-12
    # and a non-magnetic grating is not self-dual (the check can fail)
    R3, _, _ = run(ea, np.ones((2, 2), complex))
    R4, _, _ = run(np.ones((2, 2), complex), ea)
    assert abs(R3[0].sum() - R4[1].sum()) < 1e-12 and abs(R3[0].sum() - R3[1].sum()) > 1e-3
