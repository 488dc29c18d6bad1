"""The `.orcwa` path on the GPU (openrcwa_amd/orcwa.py: the reference parser
rcwa/RCWAInput.cpp and driver rcwa/RCWADriver.cpp over the batched engine).

Pinned by closed forms where the physics gives one (Fresnel at any angle,
Malus, circular = mean of TM and TE on an achiral interface, the Airy formula
of a film with a Drude, Lorentz or conductive permittivity, pole addition), at
the rounding level (RCWA of homogeneous layers is exact), and on patterned
inputs (staircase cylinder, z-sliced sphere, conical 1D grating) against the
FP64 oracle solving the same layers, order by order.  The inputs are the
cases of the reference's tests/test_rcwa_input.cpp."""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import orcwa as oc

pytestmark = pytest.mark.gpu

C0_UM = 2.99792458e14
EPS0 = 8.854187817e-12


def solve(gpu_ctx, tmp_path, text, name="in.orcwa"):
    p = tmp_path / name
    p.write_text(text)
    prob = oc.parse_rcwa_input(str(p))
    return prob, oc.run_rcwa(prob, gpu_ctx)


def interface(theta, phi, pol, psi=None, order="4 2", background=None):
    bg = f"background = {background}\n" if background is not None else ""
    pw = f"{theta} {phi} {pol}" + ("" if psi is None else f" {psi}")
    return f"""OpenRCWA 4 2
title = interface
xmesh = -5e-07 10 5e-07
ymesh = -5e-07 10 5e-07
zmesh = -5e-07 10 0.0 10 1e-06
{bg}material = 1 2.25 0 1 0
geometry = 2 1 -5e-07 5e-07 -5e-07 5e-07 -5e-07 0
planewave = {pw}
pbc = 1 1 0
rcwaorder = {order}
frequency1 = 5.0e+14 5.0e+14 0
end
"""


def fresnel_R(theta_deg, n1, n2, pol):
    th = math.radians(theta_deg)
    ct = math.cos(th)
    st = n1 * math.sin(th) / n2
    ctt = math.sqrt(1.0 - st * st)
    if pol == 2:
        r = (n1 * ct - n2 * ctt) / (n1 * ct + n2 * ctt)
    else:
        r = (n2 * ct - n1 * ctt) / (n2 * ct + n1 * ctt)
    return r * r


def airy_RT(eps_film, d_um, lam_um):
    """Normal-incidence air / film / air: reflectance and transmittance."""
    n2 = np.sqrt(complex(eps_film))
    k0 = 2 * math.pi / lam_um
    r12, r23 = (1 - n2) / (1 + n2), (n2 - 1) / (n2 + 1)
    t12, t23 = 2 / (1 + n2), 2 * n2 / (n2 + 1)
    ph = np.exp(1j * k0 * n2 * d_um)
    den = 1 + r12 * r23 * ph * ph
    r = (r12 + r23 * ph * ph) / den
    t = t12 * t23 * ph / den
    return abs(r) ** 2, abs(t) ** 2


def film(mat_lines, lam_line="wavelength = 0.6"):
    return f"""OpenRCWA 4 2
title = film
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -1e-06 10 0.0 10 3e-07 10 7e-07
{mat_lines}geometry = 2 1 -2.5e-07 2.5e-07 -2.5e-07 2.5e-07 0 3e-07
planewave = 0 0 1
pbc = 1 1 0
rcwaorder = 4 0
{lam_line}
end
"""


def test_fresnel_normal_incidence_sweep(gpu_ctx, tmp_path):
    text = interface(0, 0, 1, order="6 0").replace("frequency1 = 5.0e+14 5.0e+14 0",
                                                   "frequency1 = 4.0e14 7.5e14 14")
    prob, res = solve(gpu_ctx, tmp_path, text)
    assert len(res) == 15
    for r in res:
        assert abs(r.R - 0.04) < 1e-12 and abs(r.T - 0.96) < 1e-12
        assert r.A == 1.0 - r.R - r.T
        assert r.REF_orders.shape == (13,) and r.REF_orders.sum() == pytest.approx(r.R, abs=1e-15)


@pytest.mark.parametrize("theta", [20.0, 40.0, 56.31, 60.0])
@pytest.mark.parametrize("pol", [1, 2])
def test_fresnel_oblique(gpu_ctx, tmp_path, theta, pol):
    _, res = solve(gpu_ctx, tmp_path, interface(theta, 0.0, pol))
    assert abs(res[0].R - fresnel_R(theta, 1.0, 1.5, pol)) < 1e-10
    assert abs(res[0].R + res[0].T - 1.0) < 1e-10


@pytest.mark.parametrize("phi", [0.0, 45.0, 90.0])
def test_phi_rotation_invariance(gpu_ctx, tmp_path, phi):
    for pol in (1, 2):
        _, res = solve(gpu_ctx, tmp_path, interface(56.31, phi, pol))
        assert abs(res[0].R - fresnel_R(56.31, 1.0, 1.5, pol)) < 1e-10


def test_normal_incidence_polarisation_degeneracy(gpu_ctx, tmp_path):
    _, tm = solve(gpu_ctx, tmp_path, interface(0, 0, 1))
    _, te = solve(gpu_ctx, tmp_path, interface(0, 0, 2))
    assert abs(tm[0].R - te[0].R) < 1e-13 and abs(tm[0].R - 0.04) < 1e-12


def test_linear_and_circular_polarisation(gpu_ctx, tmp_path):
    th = 56.31
    r_tm = solve(gpu_ctx, tmp_path, interface(th, 0, 1, order="4 0"))[1][0].R
    r_te = solve(gpu_ctx, tmp_path, interface(th, 0, 2, order="4 0"))[1][0].R
    assert abs(solve(gpu_ctx, tmp_path, interface(th, 0, 3, 0.0, order="4 0"))[1][0].R - r_tm) < 1e-13
    assert abs(solve(gpu_ctx, tmp_path, interface(th, 0, 3, 90.0, order="4 0"))[1][0].R - r_te) < 1e-13
    for psi in (30.0, 45.0, 60.0):
        c, s = math.cos(math.radians(psi)), math.sin(math.radians(psi))
        r = solve(gpu_ctx, tmp_path, interface(th, 0, 3, psi, order="4 0"))[1][0].R
        assert abs(r - (c * c * r_tm + s * s * r_te)) < 1e-12       # Malus
    rcp = solve(gpu_ctx, tmp_path, interface(th, 0, 4, order="4 0"))[1][0]
    lcp = solve(gpu_ctx, tmp_path, interface(th, 0, 5, order="4 0"))[1][0]
    assert abs(rcp.R - 0.5 * (r_tm + r_te)) < 1e-12
    assert abs(rcp.R - lcp.R) < 1e-13 and abs(rcp.R + rcp.T - 1.0) < 1e-10


@pytest.mark.parametrize("pol", [1, 2])
def test_low_index_incidence(gpu_ctx, tmp_path, pol):
    text = interface(30.0, 0.0, pol, order="4 0", background=0.5).replace(
        "frequency1 = 5.0e+14 5.0e+14 0", "wavelength = 0.6")
    _, res = solve(gpu_ctx, tmp_path, text)
    assert abs(res[0].R - fresnel_R(30.0, math.sqrt(0.5), 1.5, pol)) < 1e-10
    assert abs(res[0].R + res[0].T - 1.0) < 1e-10


def test_background_only(gpu_ctx, tmp_path):
    text = """OpenRCWA 4 2
title = background
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -3e-07 10 0.0 10 7e-07
background = 2.25
planewave = 0 0 1
rcwaorder = 4 0
frequency1 = 5.0e+14 5.0e+14 0
end
"""
    _, res = solve(gpu_ctx, tmp_path, text)
    assert abs(res[0].R) < 1e-14 and abs(res[0].T - 1.0) < 1e-12


def test_films_against_airy(gpu_ctx, tmp_path):
    lam = 0.6
    w = 2 * math.pi * C0_UM / lam
    # lossless glass, a conductive slab, a Drude metal, material_index n + ik
    cases = [("material = 1 2.25 0 1 0\n", 2.25),
             ("material = 1 2.25 1.0e4 1 0\n", 2.25 + 1j * 1.0e4 / (w * EPS0)),
             ("material = 1 2.25 0 1 0\nmaterial_dispersion = 2 1.0 1.37e16 2.73e13 0.0\n",
              1.0 - 1.37e16 ** 2 / (w * w + 1j * 2.73e13 * w)),
             ("material = 1 1.0 0 1 0\nmaterial_index = 2 1.5 0.1\n", (1.5 + 0.1j) ** 2),
             ("material = 1 1.0 0 1 0\nmaterial_dispersion = 2 2.0 3e15 1e14 4e15\n",
              2.0 + 3e15 ** 2 / (4e15 ** 2 - w * w - 1j * 1e14 * w))]
    for i, (lines, eps) in enumerate(cases):
        _, res = solve(gpu_ctx, tmp_path, film(lines), name=f"f{i}.orcwa")
        R, T = airy_RT(eps, 0.3, lam)
        assert abs(res[0].R - R) < 1e-10 and abs(res[0].T - T) < 1e-10, (i, res[0].R, R)
        assert res[0].A == 1.0 - res[0].R - res[0].T
    # the Drude slab: metallic reflectance, opaque, small absorption
    _, res = solve(gpu_ctx, tmp_path, film(cases[2][0]), name="drude.orcwa")
    assert res[0].R > 0.8 and res[0].T < 0.01 and 0 < res[0].A < 0.2


def test_multipole_addition(gpu_ctx, tmp_path):
    a = 0.5 * 2.0 * math.pi * C0_UM / 0.6
    one = f"material = 1 2.25 0 1 0\nmaterial_dispersion = 2 1.0 {a!r} 0.0 0.0\n"
    ah = a / math.sqrt(2.0)
    two = (f"material = 1 2.25 0 1 0\nmaterial_dispersion = 2 1.0 {ah!r} 0.0 0.0\n"
           f"material_dispersion = 2 1.0 {ah!r} 0.0 0.0\n")
    half = f"material = 1 2.25 0 1 0\nmaterial_dispersion = 2 1.0 {ah!r} 0.0 0.0\n"
    r1 = solve(gpu_ctx, tmp_path, film(one), "p1.orcwa")[1][0].R
    r2 = solve(gpu_ctx, tmp_path, film(two), "p2.orcwa")[1][0].R
    rh = solve(gpu_ctx, tmp_path, film(half), "ph.orcwa")[1][0].R
    assert 1e-4 < r1 < 0.5 and abs(r1 - r2) < 1e-12 and abs(r1 - rh) > 1e-4


PATTERNED = {
    "pillar_circular_oblique": """OpenRCWA 4 2
title = pillar
xmesh = -2.5e-07 40 2.5e-07
ymesh = -2.5e-07 40 2.5e-07
zmesh = -6e-07 10 -2e-07 10 0.0 10 6e-07
material = 1 2.25 0 1 0
geometry = 2 1 -2.5e-07 2.5e-07 -2.5e-07 2.5e-07 -6e-07 -2e-07
geometry = 2 11 -1.5e-07 1.5e-07 -1.5e-07 1.5e-07 -2e-07 0
planewave = 30 20 4
rcwaorder = 4 4
frequency1 = 3.33e14 5.0e14 8
end
""",
    "sphere_zsliced": """OpenRCWA 4 2
title = sphere
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -7e-07 10 7e-07
material = 1 12.0 0 1 0
geometry = 2 2 -2e-07 2e-07 -2e-07 2e-07 -2e-07 2e-07
planewave = 0 0 1
rcwaorder = 3 3
wavelength = 0.6
end
""",
    "conical_grating_te": """OpenRCWA 4 2
title = conical
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -7e-07 10 -3e-07 10 0.0 10 7e-07
material = 1 2.25 0 1 0
geometry = 2 1 -2.5e-07 0.0 -2.5e-07 2.5e-07 -3e-07 0
planewave = 30 30 2
rcwaorder = 4 2
frequency1 = 5.0e+14 6.0e+14 2
end
""",
    "cylinder_x_lossy_linear": """OpenRCWA 4 2
title = cyl x
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -5e-07 10 -2e-07 10 2e-07 10 5e-07
material = 1 4.0 2.0e4 1 0
geometry = 2 12 -2.5e-07 2.5e-07 -1.5e-07 1.5e-07 -1.5e-07 1.5e-07
planewave = 20 60 3 35
rcwaorder = 2 3
wavelength = 0.55 0.7 1
end
""",
}


@pytest.mark.parametrize("case", sorted(PATTERNED))
def test_patterned_against_oracle(gpu_ctx, tmp_path, case):
    prob, res = solve(gpu_ctx, tmp_path, PATTERNED[case])
    bt = oc.build_batch(prob)
    lossless = "2.0e4" not in PATTERNED[case]
    for b, lam in enumerate(bt["lam"]):
        REF, TRN, _ = o.solve_stack(prob.nHx, prob.nHy, prob.Lx, prob.Ly, bt["xgrid"],
                                    bt["ygrid"], list(bt["eps"][b]), list(bt["thickness"]),
                                    lam, prob.theta, prob.phi, pols=[tuple(bt["pol"][0])])
        assert res[b].status == 0
        assert np.abs(res[b].REF_orders - REF[0]).max() < 1e-9
        assert np.abs(res[b].TRN_orders - TRN[0]).max() < 1e-9
        if lossless:
            assert abs(res[b].R + res[b].T - 1.0) < 1e-9


def test_shapes_are_distinguished(gpu_ctx, tmp_path):
    def text(shape):
        return f"""OpenRCWA 4 2
title = shape {shape}
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -7e-07 10 7e-07
material = 1 12.0 0 1 0
geometry = 2 {shape} -2e-07 2e-07 -2e-07 2e-07 -2e-07 2e-07
planewave = 0 0 1
rcwaorder = 3 3
wavelength = 0.6
end
"""
    R = {s: solve(gpu_ctx, tmp_path, text(s), f"s{s}.orcwa")[1][0] for s in (1, 2, 11, 12)}
    for r in R.values():
        assert abs(r.R + r.T - 1.0) < 1e-9
    assert abs(R[2].R - R[11].R) > 1e-2 and abs(R[2].R - R[1].R) > 1e-2
    assert abs(R[11].R - R[1].R) > 1e-2 and abs(R[12].R - R[1].R) > 1e-2


def test_command_writes_the_csv(gpu_ctx, tmp_path, capsys):
    src = tmp_path / "cmd.orcwa"
    src.write_text(PATTERNED["conical_grating_te"])
    out = tmp_path / "out.csv"
    assert oc.main(["-v", "-o", str(out), str(src)]) == 0
    assert "=== normal end ===" in capsys.readouterr().out
    lines = out.read_text().splitlines()
    prob = oc.parse_rcwa_input(str(src))
    res = oc.run_rcwa(prob, gpu_ctx)
    n_ord = (2 * prob.nHx + 1) * (2 * prob.nHy + 1)
    assert lines[0].startswith("# lambda[um], R, T, A, R+T, R_order_0")
    assert len(lines) == 1 + len(res)
    for line, r in zip(lines[1:], res):
        v = np.array([float(x) for x in line.split(",")])
        assert v.size == 5 + 2 * n_ord
        assert np.allclose(v[:5], [r.lam, r.R, r.T, r.A, r.R + r.T], rtol=1e-7, atol=1e-15)
        assert np.allclose(v[5:5 + n_ord], r.REF_orders, rtol=1e-7, atol=1e-15)
    # the default output names
    assert oc.main([str(src)]) == 0
    assert (tmp_path / "cmd_rcwa.csv").exists()
    assert oc.main(["-field", "Ez", "-device", str(src)]) == 0
    assert (tmp_path / "cmd_field.csv").exists() and (tmp_path / "cmd_device.csv").exists()


FIELD_SLAB = """OpenRCWA 4 2
title = driver field output
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -1e-06 10 0.0 10 3e-07 10 7e-07
material = 1 2.25 0 1 0
geometry = 2 1 -2.5e-07 2.5e-07 -2.5e-07 2.5e-07 0 3e-07
planewave = 30 0 1
pbc = 1 1 0
rcwaorder = 4 0
wavelength = 0.6
end
"""


def read_csv(path):
    return np.array([[float(v) for v in line.split(",")]
                     for line in path.read_text().splitlines()])


def test_driver_field_output(gpu_ctx, tmp_path, capsys):
    # Tests 25 and 23 of tests/test_rcwa_input.cpp:1339-1398, 1213-1252: an
    # oblique TM glass slab gives a finite, non-zero Ez image, and the three
    # save options agree cell by cell
    src = tmp_path / "slab.orcwa"
    src.write_text(FIELD_SLAB)
    ez = tmp_path / "ez.csv"
    assert oc.main(["-field", "Ez", "-fieldout", str(ez), "-o", str(tmp_path / "r.csv"),
                    str(src)]) == 0
    assert f"=== field  -> {ez} ===" in capsys.readouterr().out
    v = read_csv(ez)
    assert v.shape == (300, 500) and 1e-3 < np.abs(v).max() < 1e100
    r = (tmp_path / "r.csv").read_text().splitlines()[1].split(",")
    assert abs(float(r[1]) + float(r[2]) - 1.0) < 2e-3
    imgs = {}
    for opt in ("mod", "re", "im"):
        f = tmp_path / f"ex_{opt}.csv"
        assert oc.main(["-field", "Ex", "-sopt", opt, "-fieldout", str(f),
                        "-o", str(tmp_path / "r.csv"), str(src)]) == 0
        imgs[opt] = read_csv(f)
    assert imgs["mod"].shape == imgs["re"].shape == imgs["im"].shape
    assert np.abs(imgs["mod"] ** 2 - imgs["re"] ** 2 - imgs["im"] ** 2).max() < 1e-6


def test_field_request_without_interior_layer(gpu_ctx, tmp_path, capsys):
    # Test 32 (tests/test_rcwa_input.cpp:1715-1768): an interface only
    src = tmp_path / "two.orcwa"
    src.write_text(interface(0.0, 0.0, 1))
    f = tmp_path / "nope.csv"
    assert oc.main(["-field", "Ez", "-fieldout", str(f), str(src)]) == 1
    assert not f.exists()
    assert "interior layer" in capsys.readouterr().err
    prob, res = solve(gpu_ctx, tmp_path, interface(0.0, 0.0, 1), "two2.orcwa")
    assert abs(res[0].R - 0.04) < 2e-3


def test_device_output(gpu_ctx, tmp_path):
    # Test 36 (tests/test_rcwa_input.cpp:1937-1990)
    src = tmp_path / "dev.orcwa"
    src.write_text(PATTERNED["conical_grating_te"].replace("planewave = 30 30 2",
                                                           "planewave = 0 0 1"))
    d = tmp_path / "dev.csv"
    assert oc.main(["-deviceout", str(d), "-o", str(tmp_path / "r.csv"), str(src)]) == 0
    v = read_csv(d)
    assert v.shape == (1200, 2000)
    assert abs(v.min() - 1.0) < 1e-6 and abs(v.max() - 2.25) < 1e-6


@pytest.mark.parametrize("comp,sl,coord", [("Hy", "xz", 0.1), ("Ez", "yz", -0.05),
                                           ("Ex", "xy", 0.12)])
def test_field_csv_against_oracle(gpu_ctx, tmp_path, comp, sl, coord):
    # the first wavelength of a conical grating, written with 8 digits
    src = tmp_path / "g.orcwa"
    src.write_text(PATTERNED["conical_grating_te"])
    got = {}
    for opt in ("re", "im"):
        f = tmp_path / f"{opt}.csv"
        assert oc.main(["-field", comp, "-slice", sl, "-scoord", str(coord), "-sopt", opt,
                        "-fieldout", str(f), "-o", str(tmp_path / "r.csv"), str(src)]) == 0
        got[opt] = read_csv(f)
    prob = oc.parse_rcwa_input(str(src))
    bt = oc.build_batch(prob)
    thick = list(bt["thickness"])
    _, _, s = o.solve_stack(prob.nHx, prob.nHy, prob.Lx, prob.Ly, bt["xgrid"], bt["ygrid"],
                            list(bt["eps"][0]), thick, bt["lam"][0], prob.theta, prob.phi)
    c = s.generate_horizontal_plane_wave(bt["pol"][0][0], bt["pol"][0][1], 0)
    ids = {"Ex": o.EX, "Ey": o.EY, "Ez": o.EZ, "Hx": o.HX, "Hy": o.HY, "Hz": o.HZ}
    sls = {"xy": o.SLICE_XY, "xz": o.SLICE_XZ, "yz": o.SLICE_YZ}
    ref = o.field_image(s, c, list(range(len(thick))), thick, sls[sl], coord, ids[comp])
    assert got["re"].shape == ref.shape
    scale = np.abs(ref).max()
    assert np.abs(got["re"] + 1j * got["im"] - ref).max() < 1e-7 * scale
