"""GPU tests of the learned surrogate: the HIP training and inference kernels
against the FP64 network oracle (oracle/mlp_oracle.py), graph-captured
training against kernel-by-kernel training, the GPU-generated training data
against the FP64 RCWA oracle, and the surrogate with its plain-solver fallback
against the plain solver."""
import time

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from oracle import rcwa_oracle as o
from openrcwa_amd import surrogate as sg

pytestmark = pytest.mark.gpu


def _net(ctx, n_in, width, depth, n_out, members, seed=3):
    return sg.MLP(ctx, n_in, width, depth, n_out, members, seed)


@pytest.mark.parametrize("n_in,width,depth,n_out,members,rows",
                         [(12, 64, 3, 18, 3, 300), (12, 512, 4, 18, 2, 1000), (5, 96, 1, 7, 1, 65)])
def test_forward_matches_oracle(gpu_ctx, n_in, width, depth, n_out, members, rows):
    net = _net(gpu_ctx, n_in, width, depth, n_out, members)
    try:
        rng = np.random.default_rng(rows)
        p = net.get_params().astype(np.float64)
        p += 0.02 * rng.normal(size=p.shape)            # non-zero biases
        net.set_params(p.astype(np.float32))
        p = net.get_params().astype(np.float64)
        x = rng.uniform(-1, 1, size=(rows, n_in)).astype(np.float32)
        y = net.forward(x)
        for m in range(members):
            ref, _ = mo.forward(p[m], x.astype(np.float64), n_in, width, depth, n_out)
            # FP32 products and accumulation over at most 512 terms
            assert np.abs(y[m] - ref).max() <= 2e-5 * max(1.0, np.abs(ref).max())
    finally:
        net.close()


@pytest.mark.parametrize("rows", [257, 4096])
def test_gradient_matches_oracle(gpu_ctx, rows):
    n_in, width, depth, n_out, members = 12, 128, 3, 18, 2
    net = _net(gpu_ctx, n_in, width, depth, n_out, members, seed=11)
    try:
        rng = np.random.default_rng(rows)
        p = net.get_params().astype(np.float64) + 0.02 * rng.normal(size=(members, net.P))
        net.set_params(p.astype(np.float32))
        p = net.get_params().astype(np.float64)
        x = rng.uniform(-1, 1, size=(rows, n_in)).astype(np.float32)
        t = rng.uniform(0, 1, size=(rows, n_out)).astype(np.float32)
        loss, g = net.gradient(x, t)
        for m in range(members):
            L, gr = mo.gradient(p[m], x.astype(np.float64), t.astype(np.float64), n_in, width,
                                depth, n_out)
            assert abs(loss[m] - L) <= 1e-5 * L
            assert np.abs(g[m] - gr).max() <= 1e-4 * np.abs(gr).max()
    finally:
        net.close()


def test_bf16_inference(gpu_ctx):
    net = _net(gpu_ctx, 12, 512, 4, 18, 2)
    try:
        x = np.random.default_rng(1).uniform(-1, 1, size=(3000, 12)).astype(np.float32)
        y32 = net.forward(x)
        net.set_precision(True)
        y16 = net.forward(x)
        net.set_precision(False)
        scale = np.abs(y32).max()
        d = np.abs(y16 - y32).max() / scale
        # BF16 operands (8 significant bits) with FP32 accumulation
        assert 1e-5 < d < 3e-2
        assert np.array_equal(net.forward(x), y32)
    finally:
        net.close()


def test_graph_training_equals_eager_and_learns(gpu_ctx):
    rng = np.random.default_rng(4)
    x = rng.uniform(-1, 1, size=(8192, 12)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0] + x[:, 1]), x[:, 2] * x[:, 3], np.cos(x[:, 4] - x[:, 5])],
                 1).astype(np.float32)
    out = []
    for graph in (True, False):
        net = _net(gpu_ctx, 12, 64, 2, 3, 2, seed=5)
        try:
            net.set_data(x, t)
            loss = net.train(epochs=6, batch=256, lr_max=3e-3, lr_min=1e-4, seed=9,
                             use_graph=graph)
            out.append((loss, net.get_params()))
        finally:
            net.close()
    (l_g, p_g), (l_e, p_e) = out
    # the same kernels on the same data in the same order: identical parameters
    assert np.array_equal(p_g, p_e)
    assert np.allclose(l_g, l_e, rtol=1e-5)
    assert l_g[-1] < 0.2 * l_g[0]


@pytest.fixture(scope="module")
def surrogate_data(gpu_ctx):
    t0 = time.perf_counter()
    params, targets, status = sg.make_dataset(gpu_ctx, 16384, seed=sg.SEED_TRAIN)
    gen_s = time.perf_counter() - t0
    cal = sg.sample_queries(2048, sg.SEED_CALIB)
    cal_t, _ = sg.solve_queries(gpu_ctx, cal)
    return params, targets, status, cal, cal_t, gen_s


def test_training_data_against_oracle(gpu_ctx, surrogate_data):
    params, targets, status, *_ = surrogate_data
    assert not status.any()
    n = params.shape[0] // 2
    assert np.all(targets >= -1e-12)
    assert np.all(targets[:, :9].sum(1) + targets[:, 9:].sum(1) <= 1.0 + 1e-9)
    for i in (0, 1234, n - 1):
        g = params[i, :11][None]
        a = sg.stack_arrays(g)
        REF, TRN, _ = o.solve_stack(sg.NX, sg.NY, sg.PERIOD, sg.PERIOD, a["edges"][0],
                                    a["edges"][0], list(a["eps"][0]), a["thickness"][0],
                                    a["lam"][0], g[0, 9], g[0, 10],
                                    pols=[tuple(a["pols"][0, 0]), tuple(a["pols"][0, 1])])
        for p in range(2):
            ref = np.concatenate([np.asarray(REF[p])[sg.ORDER_IDX],
                                  np.asarray(TRN[p])[sg.ORDER_IDX]])
            assert np.abs(targets[i + p * n] - ref).max() < 1e-9


def test_surrogate_with_fallback(gpu_ctx, surrogate_data):
    params, targets, status, cal, cal_t, gen_s = surrogate_data
    sur = sg.Surrogate(gpu_ctx, members=4, width=512, depth=4, seed=21)
    try:
        loss = sur.fit(params, targets, epochs=40, batch=1024, lr_max=2e-3, lr_min=2e-5)
        assert loss[-1] < 0.1 * loss[0]
        tau, cal_frac = sur.calibrate(cal, cal_t)
        q = sg.sample_queries(65536, sg.SEED_HELDOUT)
        t0 = time.perf_counter()
        plain, _ = sg.solve_queries(gpu_ctx, q, chunk=8192)
        plain_s = time.perf_counter() - t0
        ans, flagged, tm = sur.query(q)
        err = np.abs(ans - plain).max()
        net_err = np.abs(sur.net.predict(sg.features(q))[0] - plain).max(1)
        print(f"\n[surrogate] data {gen_s:.1f} s, train loss {loss[0]:.2e} -> {loss[-1]:.2e}, "
              f"spread threshold {tau:.3g}, fallback {flagged.mean():.3f} (calibration {cal_frac:.3f}), "
              f"network-only max err {net_err.max():.2e} median {np.median(net_err):.2e}, "
              f"after fallback {err:.2e}; plain {plain_s:.2f} s, surrogate {tm['total_s']:.2f} s "
              f"(network {tm['network_s'] * 1e3:.1f} ms)", flush=True)
        # every answer within the tolerance of the plain solver
        assert err <= 1e-3
        # every query the network answered itself is within the tolerance
        assert (net_err[~flagged] <= 1e-3).all()
        # the network is far cheaper than the solver per query
        assert tm["network_s"] < 0.05 * plain_s
    finally:
        sur.net.close()
