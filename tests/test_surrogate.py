"""CPU tests of the surrogate's host logic and of the network oracle
(oracle/mlp_oracle.py) the GPU kernels are checked against."""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import mlp_oracle as mo
from oracle import rcwa_oracle as o
from openrcwa_amd import surrogate as sg


def test_gelu_matches_its_integral_definition():
    # GELU(z) = z Phi(z), Phi the standard normal CDF, by quadrature
    for z in (-3.0, -0.7, 0.0, 0.4, 2.5):
        phi, _ = integrate.quad(lambda u: math.exp(-0.5 * u * u) / math.sqrt(2 * math.pi),
                                -np.inf, z)
        assert abs(mo.gelu(np.array(z)) - z * phi) < 1e-12
    # derivative by central differences
    z = np.linspace(-4, 4, 33)
    h = 1e-5
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.abs(fd - mo.gelu_grad(z)).max() < 1e-9


def test_one_hidden_layer_closed_form():
    rng = np.random.default_rng(3)
    n_in, w, n_out = 3, 5, 2
    p = mo.he_init(n_in, w, 1, n_out, rng)
    p[-n_out:] = rng.normal(size=n_out)
    x = rng.normal(size=(4, n_in))
    (W0, b0), (W1, b1) = mo.unpack(p, n_in, w, 1, n_out)
    ref = np.array([[sum(W1[j, o_] * mo.gelu(sum(x[r, i] * W0[i, j] for i in range(n_in)) + b0[j])
                         for j in range(w)) + b1[o_] for o_ in range(n_out)] for r in range(4)])
    y, _ = mo.forward(p, x, n_in, w, 1, n_out)
    assert np.abs(y - ref).max() < 1e-12
    assert mo.param_count(n_in, w, 1, n_out) == p.size == n_in * w + w + w * n_out + n_out


@pytest.mark.parametrize("depth", [1, 3])
def test_backprop_equals_finite_differences(depth):
    rng = np.random.default_rng(10 + depth)
    n_in, w, n_out = 4, 6, 3
    p = mo.he_init(n_in, w, depth, n_out, rng)
    p += 0.05 * rng.normal(size=p.size)          # non-zero biases
    x = rng.normal(size=(7, n_in))
    t = rng.normal(size=(7, n_out))
    L, g = mo.gradient(p, x, t, n_in, w, depth, n_out)
    assert abs(L - mo.loss(p, x, t, n_in, w, depth, n_out)) < 1e-14
    h = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (mo.loss(p + e, x, t, n_in, w, depth, n_out) -
                 mo.loss(p - e, x, t, n_in, w, depth, n_out)) / (2 * h)
    assert np.abs(fd - g).max() < 1e-7 * max(1.0, np.abs(g).max())


def test_a_planted_backprop_mistake_is_caught():
    # dropping the GELU derivative in the backward pass must fail the check above
    rng = np.random.default_rng(5)
    n_in, w, n_out, depth = 3, 4, 2, 2
    p = mo.he_init(n_in, w, depth, n_out, rng)
    x = rng.normal(size=(5, n_in))
    t = rng.normal(size=(5, n_out))
    saved = mo.gelu_grad
    try:
        mo.gelu_grad = lambda z: np.ones_like(z)
        _, g_bad = mo.gradient(p, x, t, n_in, w, depth, n_out)
    finally:
        mo.gelu_grad = saved
    _, g = mo.gradient(p, x, t, n_in, w, depth, n_out)
    assert np.abs(g_bad - g).max() > 1e-3


def test_parameter_sampling_and_scaling():
    q = sg.sample_queries(2000, 1)
    assert q.shape == (2000, sg.N_IN)
    for j, k in enumerate(sg.PARAMS[:-1]):
        lo, hi = sg.RANGES[k]
        assert lo <= q[:, j].min() and q[:, j].max() < hi
    assert set(np.unique(q[:, -1])) == {0.0, 1.0}
    x = sg.scale_inputs(q)
    assert x.dtype == np.float32 and x.min() >= -1.0 and x.max() <= 1.0
    # the same seed gives the same rows
    assert np.array_equal(sg.sample_queries(10, 1), q[:10])


def test_stack_arrays_layers_and_orders():
    g = np.array([[0.3, 0.6, 100e-9, 150e-9, 4.0, 0.05, 2.5, 0.01, 500e-9, 10.0, 30.0]])
    a = sg.stack_arrays(g)
    P = sg.PERIOD
    assert np.allclose(a["edges"][0], [0, 0.3 * P, 0.6 * P, P])
    e1, e2 = 4.0 + 0.05j, 2.5 + 0.01j
    # layer 1: only the corner cell [0, 0.3P]^2; layer 2: the 2 x 2 corner block
    assert a["eps"][0, 1, 0, 0] == e1 and np.count_nonzero(a["eps"][0, 1] == e1) == 1
    assert np.count_nonzero(a["eps"][0, 2] == e2) == 4 and a["eps"][0, 2, 2, 2] == 1.0
    assert np.all(a["eps"][0, 0] == 1.0) and np.all(a["eps"][0, 3] == sg.EPS_SUB)
    assert np.allclose(a["thickness"][0], [0, 100e-9, 150e-9, 0])
    k0 = 2 * math.pi / 500e-9
    th, ph = math.radians(10), math.radians(30)
    assert math.isclose(a["kx0"][0], k0 * math.sin(th) * math.cos(ph))
    assert math.isclose(a["ky0"][0], k0 * math.sin(th) * math.sin(ph))
    # the 9 orders are the harmonic indices of (m, n) in {-1, 0, 1}^2
    assert sg.ORDER_IDX[4] == sg.order_index(0, 0) == (2 * sg.NY + 1) * sg.NX + sg.NY


def test_nine_orders_carry_all_power_in_the_oracle():
    # one design point through the FP64 oracle: every order outside the nine is
    # evanescent on both sides, so R + T over the nine equals the full sums
    g = np.array([[0.45, 0.7, 80e-9, 120e-9, 5.0, 0.02, 3.0, 0.0, 410e-9, 28.0, 40.0]])
    a = sg.stack_arrays(g)
    REF, TRN, _ = o.solve_stack(sg.NX, sg.NY, sg.PERIOD, sg.PERIOD, a["edges"][0],
                                a["edges"][0], list(a["eps"][0]), a["thickness"][0],
                                a["lam"][0], 28.0, 40.0,
                                pols=[tuple(a["pols"][0, 0]), tuple(a["pols"][0, 1])])
    for p in range(2):
        ref, trn = np.asarray(REF[p]), np.asarray(TRN[p])
        assert abs(ref[sg.ORDER_IDX].sum() + trn[sg.ORDER_IDX].sum() - ref.sum() - trn.sum()) < 1e-14
        assert ref.sum() + trn.sum() <= 1.0 + 1e-9


def test_vectorised_splitmix_equals_the_scalar_sequence():
    from openrcwa_amd.workloads import SplitMix64
    a, b = SplitMix64(77), SplitMix64(77)
    seq = np.array([a.uniform() for _ in range(500)])
    assert np.array_equal(seq, np.concatenate([b.uniforms(123), b.uniforms(377)]))
    assert a.state == b.state


def test_features_layout():
    q = sg.sample_queries(64, 3)
    f = sg.features(q)
    assert f.shape == (64, sg.N_FEAT) and f.dtype == np.float32
    assert np.array_equal(f[:, :sg.N_IN], sg.scale_inputs(q))


class _StubNet:
    """Numpy stand-in for the GPU ensemble in calibration: fixed member outputs
    (members, rows, n_out) for the rows it is asked about, and the device's
    error estimate max_c kappa_c (max - min over members of output c)."""

    def __init__(self, y):
        self.y = y
        self.kappa = None

    def forward(self, x):
        assert x.shape[0] == self.y.shape[1]
        return self.y

    def set_error_scale(self, kappa):
        self.kappa = np.asarray(kappa, dtype=np.float64)

    def predict(self, x):
        assert x.shape[0] == self.y.shape[1]
        spread = self.y.max(0) - self.y.min(0)
        return self.y.mean(0), (self.kappa[None, :] * spread).max(1)


def test_calibration_flags_every_calibration_miss():
    rng = np.random.default_rng(5)
    rows, tol = 400, 1e-3
    params = sg.sample_queries(rows, 99)
    t = rng.uniform(0.0, 0.05, size=(rows, sg.N_OUT))
    s = 10.0 ** rng.uniform(-5, -2, size=rows)            # per-query member spread
    off = np.full(rows, 0.3 * tol)                         # error of the ensemble mean
    bad = rng.permutation(rows)[:30]
    off[bad] = 2.0 * tol
    unphys = 7                                             # a negative answer, tiny spread
    bad = bad[bad != unphys]
    s[unphys] = 1e-7
    d = np.array([-0.5, -0.1, 0.1, 0.5])                  # member offsets: mean 0, range 1
    y = t[None] + d[:, None, None] * s[None, :, None]
    y[:, :, 0] += off[None, :]
    y[:, unphys, 0] = -3.0 * tol + d * s[unphys]
    sur = sg.Surrogate.__new__(sg.Surrogate)
    sur.space, sur.tol, sur.kappa, sur.tau = sg.DEFAULT_SPACE, tol, None, None
    sur.net = _StubNet(y)
    tau, frac = sur.calibrate(params, t, margin=1.5, tol=tol)
    # tau: the smallest spread among the physical misses, over the margin; the
    # unphysical answer does not set it (its spread is far smaller)
    assert tau == pytest.approx(s[bad].min() / 1.5, rel=1e-12)
    mean, est = sur.net.predict(np.zeros((rows, 1)))
    flagged = sur._flags(mean, est)
    assert flagged[bad].all()
    assert flagged[unphys]
    good = np.setdiff1d(np.arange(rows), np.append(bad, unphys))
    assert not flagged[good[s[good] < 0.99 * tau]].any()
    assert flagged[good[s[good] > 1.01 * tau]].all()
    assert frac == pytest.approx(flagged.mean())
