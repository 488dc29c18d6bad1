"""Reduced-precision star-product GEMMs (FP32, split BF16): the products
against numpy in FP64, and the error they leave in per-order R/T against the
all-FP64 run of the same batch.  The eigensolves and every LU solve stay FP64."""
import numpy as np
import pytest

from openrcwa_amd import engine, workloads
from tests import gpu_helpers as g

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(17)

# error of the product relative to max|A| max|B| K: operand rounding (FP32:
# 2^-24; two BF16 slices: about 2^-16; three: about 2^-24) with FP64 sums
BOUND = {"fp32": 2e-7, "bf16x3": 4e-5, "bf16x6": 2e-7}


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture
def prec_ctx(gpu_ctx):
    yield gpu_ctx
    gpu_ctx.set_gemm_precision("fp64")


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6"])
@pytest.mark.parametrize("M,N,K,tA,cA,tB", [(64, 64, 64, 0, 0, 0), (450, 450, 450, 0, 0, 1),
                                            (130, 70, 40, 1, 1, 0), (17, 33, 5, 0, 0, 0),
                                            (20, 150, 33, 1, 0, 1)])
def test_reduced_precision_products(prec_ctx, prec, M, N, K, tA, cA, tB):
    A = crand(3, K, M) if tA else crand(3, M, K)
    B = crand(3, N, K) if tB else crand(3, K, N)
    C = crand(3, M, N)
    al, be = 0.7 - 0.2j, 1.0 + 0.0j
    prec_ctx.set_gemm_precision(prec)
    out = g.zgemm(prec_ctx, A, B, C, al, be, tA=tA, cA=cA, tB=tB)
    prec_ctx.set_gemm_precision("fp64")
    opA = np.swapaxes(A, 1, 2) if tA else A
    opA = np.conj(opA) if cA else opA
    opB = np.swapaxes(B, 1, 2) if tB else B
    ref = al * opA @ opB + be * C
    scale = np.abs(A).max() * np.abs(B).max() * K
    err = np.abs(out - ref).max() / scale
    assert err < BOUND[prec]
    # the reduced path really ran (an FP64 product is accurate to ~1e-16)
    assert err > 1e-12


def _pillar_small(n_freq=8):
    return workloads.pillar2d(n_freq=n_freq, nX=4, nY=4)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3", "bf16x6"])
def test_cascade_precision_error_vs_fp64(prec_ctx, prec):
    """Per-order R/T of a patterned batch (Nx = Ny = 4, four layers, 64
    solves) with the cascade products reduced, against the all-FP64 run."""
    w = _pillar_small()
    args = (w["nX"], w["nY"], w["coord_x"], w["coord_y"], w["thickness"], w["lam"], w["kx0"],
            w["ky0"], w["eps"], w["pols"])
    ref = engine.solve_batch(prec_ctx, *args)
    prec_ctx.set_gemm_precision(prec)
    low = engine.solve_batch(prec_ctx, *args)
    prec_ctx.set_gemm_precision("fp64")
    d = max(np.abs(low.ref - ref.ref).max(), np.abs(low.trn - ref.trn).max())
    print(f"\n[precision] {prec}: max |dR|, |dT| vs FP64 = {d:.3e}", flush=True)
    assert not low.status.any()
    assert 1e-14 < d < {"fp32": 1e-5, "bf16x3": 1e-3, "bf16x6": 1e-5}[prec]


def test_star_product_precision(prec_ctx):
    """Redheffer star products of random S-matrices in split BF16 against FP64."""
    n, B = 96, 4
    mats = [0.3 * crand(B, n, n) / np.sqrt(n) for _ in range(8)]
    top, bottom = mats[:4], mats[4:]
    ref, _ = engine.redheffer_star(prec_ctx, top, bottom)
    prec_ctx.set_gemm_precision("bf16x6")
    low, info = engine.redheffer_star(prec_ctx, top, bottom)
    prec_ctx.set_gemm_precision("fp64")
    assert not info.any()
    d = max(np.abs(a - b).max() for a, b in zip(low, ref))
    assert 1e-14 < d < 1e-5
