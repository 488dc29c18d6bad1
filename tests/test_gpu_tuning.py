"""Kernel variants and their autotuner: every multishift-QR variant (window /
shifts / workgroups per CU) and GEMM main tile gives the oracle's per-order
R/T, and orcwa_plan_autotune keeps a valid combination, records it in the
context's tuning file and leaves the plan's results those of the choice."""
import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import engine, workloads

pytestmark = pytest.mark.gpu


def _pillar(nX=4, n_freq=2):
    w = workloads.pillar2d(n_freq=n_freq, nX=nX, nY=nX)
    return w, (w["nX"], w["nY"], w["coord_x"], w["coord_y"], w["thickness"], w["lam"], w["kx0"],
               w["ky0"], w["eps"], w["pols"])


def _arch():
    # the architecture name the library compares: gcnArchName up to ':'
    return "gfx950"


@pytest.mark.parametrize("qr,tile", [(0, 0), (1, 0), (2, 0), (0, 1), (0, 2), (2, 1)])
def test_variants_against_oracle(gpu_ctx, tmp_path, qr, tile):
    w, args = _pillar()
    n = 2 * (2 * w["nX"] + 1) * (2 * w["nY"] + 1)
    f = tmp_path / "tune.txt"
    f.write_text(f"# forced variant\n{_arch()} {n} {qr} {tile} 0.0\n")
    ctx = engine.Context(0)
    try:
        ctx.set_tuning_file(f)
        res = engine.solve_batch(ctx, *args)
    finally:
        ctx.close()
    assert not res.status.any()
    for i in (0, 5, len(w["lam"]) - 1):
        th, ph = w["angles"][i]
        REF, TRN, _ = o.solve_stack(w["nX"], w["nY"], w["period"], w["period"], w["coord_x"],
                                    w["coord_y"], list(w["eps"]), w["thickness"], w["lam"][i],
                                    th, ph)
        for p in range(2):
            assert np.abs(res.ref[i, p] - REF[p]).max() < 1e-9
            assert np.abs(res.trn[i, p] - TRN[p]).max() < 1e-9


def test_autotune_records_a_choice(gpu_ctx, tmp_path):
    w, args = _pillar(nX=3)
    f = tmp_path / "tune.txt"
    ctx = engine.Context(0)
    try:
        ctx.set_tuning_file(f, append=True)
        plan = engine.Plan(ctx, *args)
        plan.execute()
        plan.sync()
        base = plan.results()
        qr, tile = plan.autotune()
        after = plan.results()
        plan.close()
    finally:
        ctx.close()
    assert 0 <= qr < 3 and 0 <= tile < 3
    line = f.read_text().split()
    n = 2 * (2 * w["nX"] + 1) ** 2
    assert line[1:4] == [str(n), str(qr), str(tile)] and float(line[4]) > 0
    assert np.abs(after.ref - base.ref).max() < 1e-10
    assert np.abs(after.trn - base.trn).max() < 1e-10
