"""End-to-end parity of the HIP engine with the FP64 oracle (per-order R/T,
both polarisations), plus the reference's published values and closed forms
through the GPU path.  Tolerances: 1e-9 absolute on per-order efficiencies for
the small cases (FP64 everywhere; the bar of the path is 1e-6), published
values to their printed 6 decimals."""
import json
import math
import os

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import engine, ofd, workloads

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "published_values.json")
TOL = 1e-9


def L(e1, e2=None, fill=0.5, d=0.0, e1i=0.0, e2i=0.0, fy=1.0):
    e2 = e1 if e2 is None else e2
    return dict(eps1=e1, eps2=e2, eps1i=e1i, eps2i=e2i, fill=fill, fill_y=fy, thickness=d)


def compare_ofd(ctx, nX, period, layers, freqs, theta=0.0, phi=0.0, nY=0, period_y=None,
                poles=None, tol=TOL):
    gpu = ofd.solve_ofd(ctx, nX, period, layers, freqs, theta, phi, nY, period_y, poles)
    assert not gpu["status"].any(), gpu["status"]
    worst = 0.0
    for f, freq in enumerate(np.atleast_1d(freqs)):
        ref = o.ofd_solve_frequency(nX, period, layers, freq, theta, phi, nY, period_y, poles)
        for p in range(2):
            worst = max(worst, abs(gpu["R"][f, p] - ref["R"][p]), abs(gpu["T"][f, p] - ref["T"][p]),
                        np.abs(gpu["REF"][f, p] - ref["REF"][p]).max(),
                        np.abs(gpu["TRN"][f, p] - ref["TRN"][p]).max())
    assert worst < tol, worst
    return gpu


def test_grating_1d_sweep(gpu_ctx):
    """The 1D binary grating (data/sample/grating.ofd, N=5, 16 frequencies)."""
    w = workloads.grating1d(nX=5, n_freq=16)
    compare_ofd(gpu_ctx, w["nX"], w["period"], w["layers"], w["freqs"])


def test_published_values_on_gpu(gpu_ctx):
    gd = json.load(open(GOLDEN))
    gr = gd["grating"]
    for c in gr["cases"]:
        r = ofd.solve_ofd(gpu_ctx, c["N"], gr["period"], gr["layers"], [gr["freq"]])
        assert abs(r["R"][0, 0] - c["R_TE"]) < 5.1e-7
        assert abs(r["R"][0, 1] - c["R_TM"]) < 5.1e-7
    m = gd["metal"]
    r = ofd.solve_ofd(gpu_ctx, m["N"], m["period"], m["layers"], [m["freq"]])
    assert abs(r["R"][0, 0] - m["R"]) < 5.1e-7 and abs(r["T"][0, 0] - m["T"]) < 5.1e-7
    b = gd["oblique"]
    r = ofd.solve_ofd(gpu_ctx, b["N"], b["period"], b["layers"], [b["freq"]], theta=b["theta"])
    assert abs(r["R"][0, 0] - b["R_TE"]) < 5.1e-7 and r["R"][0, 1] < b["R_TM_max"]


@pytest.mark.parametrize("theta,phi", [(0, 0), (30, 0), (56.30993247, 0), (70, 0), (45, 37)])
def test_fresnel_on_gpu(gpu_ctx, theta, phi):
    r = ofd.solve_ofd(gpu_ctx, 2, 4e-7, [L(1.0), L(2.25)], [5e14], theta, phi)
    ti = math.radians(theta)
    st = math.sin(ti) / 1.5
    ci, ct = math.cos(ti), math.sqrt(1 - st * st)
    rs = ((ci - 1.5 * ct) / (ci + 1.5 * ct)) ** 2
    rp = ((1.5 * ci - ct) / (1.5 * ci + ct)) ** 2
    assert abs(r["R"][0, 0] - rs) < 1e-12 and abs(r["R"][0, 1] - rp) < 1e-12
    assert np.abs(r["R"] + r["T"] - 1).max() < 1e-12


def test_oblique_conical_grating(gpu_ctx):
    grat = [L(1.0), L(4.0, 1.0, fill=0.4, d=2.5e-7), L(2.25)]
    for th, ph in ((30, 20), (40, 30), (-25, 0), (25, 180)):
        g = compare_ofd(gpu_ctx, 8, 4e-7, grat, np.linspace(4.3e14, 7.5e14, 3), th, ph)
        assert np.abs(g["R"] + g["T"] - 1).max() < 1e-9


def test_dispersive_and_lossy(gpu_ctx):
    layers = [L(1.0), L(1.0, d=3e-8), L(2.25)]
    poles = {(1, 1): [(1.0, 1.36e16, 3.0e13, 0.0)], (1, 2): [(1.0, 1.36e16, 3.0e13, 0.0)]}
    compare_ofd(gpu_ctx, 2, 5e-7, layers, np.linspace(4e14, 7e14, 7), 30.0, 0.0, poles=poles)
    lor = {(1, 1): [(2.0, 3.0e15, 2.0e14, 4.0e15)]}
    g1 = [L(1.0), L(1.0, 2.25, fill=0.3, d=1e-7), L(2.25)]
    compare_ofd(gpu_ctx, 6, 5e-7, g1, np.linspace(4e14, 7e14, 4), 20.0, 0.0, poles=lor)


def test_2d_pillars_small(gpu_ctx):
    lay = [L(1.0), L(2.25, 1.0, fill=0.6, d=2e-7, fy=0.6), L(2.25)]
    g = compare_ofd(gpu_ctx, 3, 5e-7, lay, np.linspace(3.331e14, 4.835e14, 3), nY=3,
                    period_y=5e-7)
    assert np.abs(g["R"][:, 0] - g["R"][:, 1]).max() < 1e-9          # C4: TE = TM
    lay2 = [L(1.0), L(2.25 + 0.05j, 1.0, fill=0.55, d=1.5e-7, fy=0.35),
            L(6.0, 1.0, fill=0.3, d=0.8e-7, fy=0.7), L(2.25)]
    compare_ofd(gpu_ctx, 2, 4e-7, lay2, [5.5e14, 6.5e14], 20.0, 45.0, nY=3, period_y=4.5e-7)


def pillar_subset(nX, idx):
    w = workloads.pillar2d(nX=nX, nY=nX)
    return w, np.asarray(idx)


def gpu_pillar(ctx, w, idx):
    return engine.solve_batch(ctx, w["nX"], w["nY"], w["coord_x"], w["coord_y"], w["thickness"],
                              w["lam"][idx], w["kx0"][idx], w["ky0"][idx], w["eps"],
                              w["pols"][idx])


def oracle_pillar(w, i):
    th, ph = w["angles"][i]
    return o.solve_stack(w["nX"], w["nY"], w["period"], w["period"], w["coord_x"], w["coord_y"],
                         list(w["eps"]), w["thickness"], w["lam"][i], th, ph)


@pytest.mark.parametrize("nX", [2, 4])
def test_pillar_workload_reduced_orders(gpu_ctx, nX):
    w, idx = pillar_subset(nX, [0, 1, 3, 6, 7, 1001, 2047])
    res = gpu_pillar(gpu_ctx, w, idx)
    assert not res.status.any()
    for j, i in enumerate(idx):
        REF, TRN, _ = oracle_pillar(w, i)
        for p in range(2):
            assert np.abs(res.ref[j, p] - REF[p]).max() < TOL
            assert np.abs(res.trn[j, p] - TRN[p]).max() < TOL


def test_eigen_recovery_on_gpu(gpu_ctx, monkeypatch):
    """The reference's recovery of a failed eigen solve (rcwa/MKLEigenSolver.hpp:
    81-93, 160-176), forced on every patterned layer by a negative acceptance tolerance:
    each patterned layer is solved again on A + diag(1e-11 (i+1) max|A|) and
    reported with EIG_PERTURBED (and EIG_RESIDUAL, since the forced test fails
    again).  The per-order results match the oracle taking the same branch
    and stay within the parity bar of the unperturbed solve."""
    from openrcwa_amd import _lib
    from tests import gpu_helpers as gh
    w, idx = pillar_subset(3, [0, 9, 1500])
    base = gpu_pillar(gpu_ctx, w, idx)
    assert not base.status.any()
    gh.set_eig_tolerance(gpu_ctx, -1.0)
    try:
        res = gpu_pillar(gpu_ctx, w, idx)
    finally:
        gh.set_eig_tolerance(gpu_ctx, 1e-8)
    assert (res.status == (_lib.STATUS_EIG_PERTURBED | _lib.STATUS_EIG_RESIDUAL)).all()
    # the GPU solves homogeneous layers in closed form, so only the patterned
    # layers take the branch; the oracle is forced the same way (perturbing
    # the homogeneous layers too moves R/T by ~1e-8 on this batch)
    solve = o.Layer.solve

    def solve_forced(self, *a, **k):
        uniform = (np.all(self.eps_grid == self.eps_grid.flat[0]) and
                   np.all(self.mu_grid == self.mu_grid.flat[0]))
        saved = o.EIG_TOL
        o.EIG_TOL = saved if uniform else -1.0
        try:
            return solve(self, *a, **k)
        finally:
            o.EIG_TOL = saved
    monkeypatch.setattr(o.Layer, "solve", solve_forced)
    for j, i in enumerate(idx):
        REF, TRN, _ = oracle_pillar(w, i)
        for p in range(2):
            assert np.abs(res.ref[j, p] - REF[p]).max() < TOL
            assert np.abs(res.trn[j, p] - TRN[p]).max() < TOL
    # the perturbation (1e-11 (i + 1) max|A|, up to ~1e-9 max|A| at n = 98) moves R/T by ~2e-8
    assert np.abs(res.ref - base.ref).max() < 1e-7 and np.abs(res.trn - base.trn).max() < 1e-7
    assert np.abs(res.ref - base.ref).max() > 0.0


def test_plans_reuse_cached_buffers(gpu_ctx):
    """Plans take their device buffers from the context's cache of destroyed
    plans' buffers, which still hold the old plan's data: plans of different
    sizes built one after another on one context (a larger one, a smaller one
    in its buffer, the first again) give the same results as the first run."""
    w, idx = pillar_subset(4, [0, 1, 3, 6, 7, 1001, 2047])
    first = gpu_pillar(gpu_ctx, w, idx)
    w2, idx2 = pillar_subset(2, [5, 9])
    small = gpu_pillar(gpu_ctx, w2, idx2)
    again = gpu_pillar(gpu_ctx, w, idx)
    assert not first.status.any() and not small.status.any()
    assert np.array_equal(first.ref, again.ref) and np.array_equal(first.trn, again.trn)
    REF, TRN, _ = oracle_pillar(w2, idx2[1])
    for p in range(2):
        assert np.abs(small.ref[1, p] - REF[p]).max() < TOL
        assert np.abs(small.trn[1, p] - TRN[p]).max() < TOL


def test_pillar_workload_full_size(gpu_ctx):
    """The 2D 15x15-order pillar workload at its real size (n = 450) on a
    handful of instances; the metric's bar is max |dR|, |dT| <= 1e-6."""
    w, idx = pillar_subset(7, [0, 5, 1234])
    res = gpu_pillar(gpu_ctx, w, idx)
    assert not res.status.any()
    worst = 0.0
    for j, i in enumerate(idx):
        REF, TRN, _ = oracle_pillar(w, i)
        for p in range(2):
            worst = max(worst, np.abs(res.ref[j, p] - REF[p]).max(),
                        np.abs(res.trn[j, p] - TRN[p]).max())
    assert worst < 1e-8, worst


def test_pillar_full_size_properties(gpu_ctx):
    """Size-independent checks over a larger slice of the full workload:
    lossy stack => 0 <= R, T and R + T <= 1; sum of orders == totals."""
    w, idx = pillar_subset(7, np.arange(0, 2048, 16))
    res = gpu_pillar(gpu_ctx, w, idx)
    assert not res.status.any()
    tot = res.R + res.T
    assert (res.ref > -1e-12).all() and (res.trn > -1e-12).all()
    assert (tot < 1 + 1e-9).all() and (tot > 0.3).all()


C4_CIRCLE = """OpenRCWA 4 2
title = circles at the largest order count
xmesh = -2.5e-07 10 2.5e-07
ymesh = -2.5e-07 10 2.5e-07
zmesh = -5e-07 10 -2e-07 10 0.0 10 5e-07
material = 1 6.25 0 1 0
geometry = 2 11 -1.5e-07 1.5e-07 -1.5e-07 1.5e-07 -2e-07 0
planewave = 0 0 {pol}
pbc = 1 1 0
rcwaorder = 15 15
wavelength = 0.55 0.7 1
end
"""


def test_circle_largest_order_count_properties(gpu_ctx, tmp_path):
    """A lossless staircase cylinder at 31 x 31 orders (n = 1922, the largest
    order count supported) through the .orcwa path: energy
    conservation, TE = TM at normal incidence on the C4-symmetric lattice, and
    mirror symmetry of the orders."""
    from openrcwa_amd import orcwa as oc
    out = {}
    for pol in (1, 2):
        p = tmp_path / f"c{pol}.orcwa"
        p.write_text(C4_CIRCLE.format(pol=pol))
        out[pol] = oc.run_rcwa(oc.parse_rcwa_input(str(p)), gpu_ctx)
    for r1, r2 in zip(out[1], out[2]):
        assert r1.status == 0 and r2.status == 0
        for r in (r1, r2):
            assert abs(r.R + r.T - 1.0) < 1e-9
            g = r.REF_orders.reshape(31, 31)
            assert np.abs(g - g[::-1, :]).max() < 1e-9 and np.abs(g - g[:, ::-1]).max() < 1e-9
        assert abs(r1.R - r2.R) < 1e-9 and abs(r1.T - r2.T) < 1e-9


def test_too_many_harmonics_is_an_error(gpu_ctx):
    # 37 x 37 orders (n = 2738) exceed what the cascade keeps in LDS
    from openrcwa_amd import _lib
    eps = np.array([[[1.0]], [[2.25]], [[1.0]]], dtype=complex)
    with pytest.raises(_lib.EngineError):
        engine.solve_batch(gpu_ctx, 18, 18, [0, 0.5], [0, 0.5], [0.0, 0.1, 0.0], [0.6], [0.0],
                           [0.0], eps, np.array([[[1.0, 0.0]]], dtype=complex))


def test_large2d_against_oracle(gpu_ctx):
    """The large 2D grating (31 x 31 orders, n = 1922) with a rectangle and a
    staircase circle layer, per order against the FP64 oracle."""
    w = workloads.large2d(n_freq=2, n_layers=2)
    res = engine.solve_batch(gpu_ctx, w["nX"], w["nY"], w["coord_x"], w["coord_y"],
                             w["thickness"], w["lam"], w["kx0"], w["ky0"], w["eps"], w["pols"])
    assert not res.status.any()
    REF, TRN, _ = o.solve_stack(w["nX"], w["nY"], w["period"], w["period"], w["coord_x"],
                                w["coord_y"], list(w["eps"]), list(w["thickness"]), w["lam"][0],
                                0.0, 0.0, pols=[tuple(p) for p in w["pols"][0]])
    for q in range(2):
        assert np.abs(res.ref[0, q] - REF[q]).max() < 1e-9
        assert np.abs(res.trn[0, q] - TRN[q]).max() < 1e-9
    assert (res.R + res.T <= 1 + 1e-9).all()


def _oracle_stack(s_):
    return o.solve_stack(s_["nX"], s_["nY"], s_["period"], s_["period"], s_["coord_x"],
                         s_["coord_y"], list(s_["eps"]), list(s_["thickness"]), s_["lam"],
                         s_["theta"], s_["phi"], pols=[tuple(p) for p in s_["pol"]])


def test_uneven_stacks_against_oracle(gpu_ctx):
    """The uneven batch: stacks of different order counts, layer counts, cell
    layouts and thicknesses, with Lorentz-dispersive and metal-film layers,
    grouped into batches per order count whose instances carry their own
    edges and thicknesses (shorter stacks padded with zero-thickness
    exit-medium layers; one group holds six stacks of one shape)."""
    stacks = (workloads.uneven_stacks(8, n_range=(3, 6), layer_range=(1, 5)) +
              workloads.uneven_stacks(6, seed=7, n_range=(4, 4), layer_range=(3, 3)) +
              workloads.uneven_stacks(1, seed=11, n_range=(10, 10), layer_range=(2, 2)))
    res = engine.solve_uneven(gpu_ctx, stacks)
    for s_, (ref, trn, st) in zip(stacks, res):
        assert st == 0
        REF, TRN, _ = _oracle_stack(s_)
        for q in range(2):
            assert np.abs(ref[q] - REF[q]).max() < 1e-9
            assert np.abs(trn[q] - TRN[q]).max() < 1e-9
    # the same groups spread over three contexts (streams) driven from three
    # host threads: every group is the same batch, so the results are equal
    res3 = engine.solve_uneven(gpu_ctx, stacks, streams=3)
    for (ref, trn, st), (ref3, trn3, st3) in zip(res, res3):
        assert st3 == st
        assert np.abs(ref3 - ref).max() < 1e-13 and np.abs(trn3 - trn).max() < 1e-13
    # and again, after the contexts' cached workspaces were released
    res4 = engine.solve_uneven(gpu_ctx, stacks, streams=2)
    for (ref, trn, _), (ref4, trn4, _) in zip(res, res4):
        assert np.abs(ref4 - ref).max() < 1e-13 and np.abs(trn4 - trn).max() < 1e-13
    # one batch per (order count, layer count) instead of layer padding
    res5 = engine.solve_uneven(gpu_ctx, stacks, pad_layers=False)
    for (ref, trn, _), (ref5, trn5, st5) in zip(res, res5):
        assert st5 == 0
        assert np.abs(ref5 - ref).max() < 1e-11 and np.abs(trn5 - trn).max() < 1e-11


def test_uneven_growing_mode_is_flagged(gpu_ctx):
    """Stack 326 of the 1024-stack uneven batch holds a Lorentz inclusion at
    eps = -1.56 + 0.16i whose layer carries a mode growing by ~5e23 across its
    282 nm (Im keff < 0 under the GammaSqrtOp branch): the reference's cascade
    then loses all accuracy (the oracle's R moves by 1e-3 when a thickness
    changes by 1e-12).  It is reported with ORCWA_STATUS_ILLCOND; a
    well-conditioned stack of the same batch solved beside it is not, and
    matches the oracle."""
    from openrcwa_amd import _lib
    allst = workloads.uneven_stacks(1024)
    good = [i for i in range(300, 340) if allst[i]["nX"] == 6 and i != 326][:1]
    stacks = [allst[326]] + [allst[i] for i in good]
    res = engine.solve_uneven(gpu_ctx, stacks)
    assert res[0][2] & _lib.STATUS_ILLCOND
    for s_, (ref, trn, st) in zip(stacks[1:], res[1:]):
        assert st == 0
        REF, TRN, _ = _oracle_stack(s_)
        for q in range(2):
            assert np.abs(ref[q] - REF[q]).max() < 1e-9
            assert np.abs(trn[q] - TRN[q]).max() < 1e-9


def test_uneven_graph_replays_match_eager(gpu_ctx):
    """The uneven batch with every group resident and captured in one HIP graph:
    each replay (with eager solves of the same groups in between) equals the
    group-by-group eager solve, per order, for every stack."""
    stacks = workloads.uneven_stacks(48, n_range=(3, 6), layer_range=(1, 5))
    eager = engine.solve_uneven(gpu_ctx, stacks)
    ub = engine.UnevenBatch(gpu_ctx, stacks, streams=3)
    try:
        assert len(ub.plans) > 1
        for rep in range(3):
            got = ub.run()
            assert ub.graph.reruns == 0
            for (r1, t1, s1), (r2, t2, s2) in zip(got, eager):
                assert s1 == s2
                assert np.abs(r1 - r2).max() < 1e-13 and np.abs(t1 - t2).max() < 1e-13
            # eager work on the same context between replays
            engine.solve_uneven(gpu_ctx, stacks[:5])
    finally:
        ub.close()


FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "oracle_large.npz")


def _fixture():
    if not os.path.exists(FIXTURE):
        pytest.fail("tests/golden/oracle_large.npz is missing: "
                    "python tests/golden/make_oracle_fixtures.py")
    return np.load(FIXTURE, allow_pickle=False)


def test_pillar_full_size_against_oracle_fixture(gpu_ctx):
    """16 instances spread over the 2048-solve 15 x 15-order pillar batch
    (n = 450), solved inside the full batch, per order against the FP64 oracle
    (tests/golden/oracle_large.npz)."""
    f = _fixture()
    w = workloads.pillar2d()
    res = gpu_pillar(gpu_ctx, w, np.arange(w["lam"].size))
    assert not res.status.any()
    for j, i in enumerate(f["pillar_idx"]):
        for p in range(2):
            assert np.abs(res.ref[i, p] - f["pillar_ref"][j, p]).max() < 1e-8
            assert np.abs(res.trn[i, p] - f["pillar_trn"][j, p]).max() < 1e-8


def test_uneven_large_stacks_against_oracle_fixture(gpu_ctx):
    """Uneven-generator stacks at N = 11 and 12 (n = 1058, 1250) with ten
    layers (Lorentz, lossy metal and patterned layers), in one uneven batch
    with smaller stacks, per order against the FP64 oracle."""
    f = _fixture()
    big = [workloads.uneven_stacks(1, seed=int(sd), n_range=(int(N), int(N)),
                                   layer_range=(10, 10))[0]
           for N, sd in zip(f["uneven_n"], f["uneven_seeds"])]
    stacks = big + workloads.uneven_stacks(6, n_range=(3, 5), layer_range=(1, 4))
    res = engine.solve_uneven(gpu_ctx, stacks)
    for k in range(len(big)):
        ref, trn, st = res[k]
        assert st & ~engine._lib.STATUS_ILLCOND == 0
        assert np.abs(ref - f[f"uneven_ref{k}"]).max() < 1e-8
        assert np.abs(trn - f[f"uneven_trn{k}"]).max() < 1e-8


def test_large2d_all_layers_oblique_against_oracle_fixture(gpu_ctx):
    """The 31 x 31-order grating (n = 1922) with all six patterned layers at
    theta = 20 deg, phi = 30 deg, per order against the FP64 oracle."""
    f = _fixture()
    fr, th, ph = f["large_freq_theta_phi"]
    w = workloads.large2d()
    lam = w["lam"][int(fr)]
    k0 = 2 * math.pi / lam
    t, p_ = math.radians(th), math.radians(ph)
    te, tm = ofd.polarisation_basis(th, ph)
    res = engine.solve_batch(gpu_ctx, w["nX"], w["nY"], w["coord_x"], w["coord_y"],
                             w["thickness"], [lam], [k0 * math.sin(t) * math.cos(p_)],
                             [k0 * math.sin(t) * math.sin(p_)], w["eps"], [[te, tm]])
    assert not res.status.any()
    for q in range(2):
        assert np.abs(res.ref[0, q] - f["large_ref"][q]).max() < 1e-8
        assert np.abs(res.trn[0, q] - f["large_trn"][q]).max() < 1e-8
