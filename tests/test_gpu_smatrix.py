"""Slab scattering matrices and the Redheffer star product on the GPU
(orcwa_gpu_slab_smatrix: rcwa/RCWAScatterMatrix.cpp:88-147;
orcwa_gpu_redheffer_star: core/RCWAIntegrator.cpp:57-67) against the FP64
oracle, and a patterned stack built slab by slab in the reference's order
(each new slab on top of the accumulated stack) against the RTCM cascade."""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import engine

pytestmark = pytest.mark.gpu


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("n", [6, 50, 98])
def test_slab_smatrix_matches_oracle(gpu_ctx, n):
    # well-conditioned mode matrices (I + a scaled random part): the results
    # are then fixed to rounding level, independently of the LU pivoting
    rng = np.random.default_rng(n)
    B = 3
    eye = np.eye(n)[None]
    W, V, W0, V0 = (eye + 0.5 * crand(rng, B, n, n) / math.sqrt(n) for _ in range(4))
    gamma = rng.uniform(1e6, 2e7, (B, n)) + 1j * rng.uniform(0.0, 3e6, (B, n))
    lam = np.array([4.5e-7, 5.5e-7, 6.5e-7])
    wz = np.array([1.0e-7, 2.5e-7, 0.4e-7])
    Rud, Tuu, info = engine.slab_smatrix(gpu_ctx, W, V, gamma, W0, V0, lam, wz)
    assert not info.any()
    for b in range(B):
        R, T = o.slab_scatter_matrix(W[b], V[b], gamma[b], W0[b], V0[b], wz[b], lam[b])
        scale = max(np.abs(R).max(), np.abs(T).max())
        assert np.abs(Rud[b] - R).max() < 1e-11 * scale
        assert np.abs(Tuu[b] - T).max() < 1e-11 * scale


@pytest.mark.parametrize("n", [5, 64, 130])
def test_redheffer_star_matches_oracle(gpu_ctx, n):
    rng = np.random.default_rng(100 + n)
    B = 2
    top = [0.3 * crand(rng, B, n, n) / math.sqrt(n) for _ in range(4)]
    bot = [0.3 * crand(rng, B, n, n) / math.sqrt(n) for _ in range(4)]
    out, info = engine.redheffer_star(gpu_ctx, top, bot)
    assert not info.any()
    for b in range(B):
        ref = o.redheffer_star_general([x[b] for x in top], [x[b] for x in bot])
        for g, r in zip(out, ref):
            assert np.abs(g[b] - r).max() < 1e-12


def test_stack_by_star_products_equals_rtcm(gpu_ctx):
    """Three patterned slabs in the air basis, composed bottom-up on the GPU
    (acc <- slab * acc, as RCWAIntegrator::compute does), reproduce the RTCM
    Rud and Tdd of the same stack (air / slabs / air)."""
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
    Rud, Tuu, info = engine.slab_smatrix(
        gpu_ctx, np.stack([l.W for l in slabs]), np.stack([l.V for l in slabs]),
        np.stack([l.gamma * k0 for l in slabs]), np.stack([air.W] * 3), np.stack([air.V] * 3),
        lam, np.array(d))
    assert not info.any()
    eye, zero = np.eye(n, dtype=complex)[None], np.zeros((1, n, n), complex)
    acc = (zero, eye, zero, eye)
    for i in (2, 1, 0):
        slab = (Rud[i:i + 1], Tuu[i:i + 1], Rud[i:i + 1], Tuu[i:i + 1])
        acc, info = engine.redheffer_star(gpu_ctx, slab, acc)
        assert not info.any()
    assert np.abs(acc[0][0] - s.Rud).max() < 1e-10
    assert np.abs(acc[3][0] - s.Tdd).max() < 1e-10
