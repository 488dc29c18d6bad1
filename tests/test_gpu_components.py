"""GPU component checks, one kernel family at a time, against numpy (LAPACK)
and the FP64 oracle.  Bar: FP64 round-off (relative errors ~1e-13)."""
import math

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from tests import gpu_helpers as g

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(7)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("M,N,K", [(1, 1, 1), (17, 33, 5), (64, 64, 16), (100, 70, 130),
                                   (450, 450, 450), (30, 200, 40), (130, 40, 70), (70, 20, 33),
                                   (10, 100, 12), (20, 20, 20), (129, 250, 64)])
def test_zgemm_shapes(gpu_ctx, M, N, K):
    A, B, C = crand(2, M, K), crand(2, K, N), crand(2, M, N)
    al, be = 0.7 - 0.2j, -1.1 + 0.4j
    out = g.zgemm(gpu_ctx, A, B, C, al, be)
    ref = al * A @ B + be * C
    scale = np.abs(A).max() * np.abs(B).max() * K + np.abs(C).max()
    assert np.abs(out - ref).max() / scale < 1e-14


@pytest.mark.parametrize("M,N,K,batch", [(64, 64, 16, 64), (130, 70, 40, 67), (20, 150, 33, 131),
                                         (450, 450, 450, 72)])
def test_zgemm_large_batches(gpu_ctx, M, N, K, batch):
    # batches of 64 or more products take the XCD-grouped tile order; batch
    # counts that are not a multiple of 8 leave a remainder in plain order
    A, B, C = crand(batch, M, K), crand(batch, K, N), crand(batch, M, N)
    al, be = 0.3 + 0.9j, 0.5 - 0.25j
    out = g.zgemm(gpu_ctx, A, B, C, al, be)
    ref = al * A @ B + be * C
    scale = np.abs(A).max() * np.abs(B).max() * K + np.abs(C).max()
    err = np.abs(out - ref).max(axis=(1, 2)) / scale
    assert err.max() < 1e-14, np.argmax(err)


@pytest.mark.parametrize("tA,cA,tB,cB", [(1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 1), (1, 1, 1, 0)])
def test_zgemm_ops(gpu_ctx, tA, cA, tB, cB):
    M, N, K = 37, 29, 41
    A = crand(3, K, M) if tA else crand(3, M, K)
    B = crand(3, N, K) if tB else crand(3, K, N)
    C = np.zeros((3, M, N), complex)
    out = g.zgemm(gpu_ctx, A, B, C, 1.0 + 0j, 0j, tA, cA, tB, cB)
    opA = np.transpose(A, (0, 2, 1)) if tA else A
    opB = np.transpose(B, (0, 2, 1)) if tB else B
    opA = opA.conj() if cA else opA
    opB = opB.conj() if cB else opB
    assert np.abs(out - opA @ opB).max() < 1e-12


@pytest.mark.parametrize("n,nrhs,batch", [(1, 1, 2), (5, 3, 2), (31, 31, 2), (32, 32, 2),
                                          (33, 7, 2), (100, 100, 2), (200, 48, 2), (256, 17, 2),
                                          (450, 450, 2), (450, 900, 2), (512, 33, 2),
                                          (97, 40, 3), (450, 900, 3), (513, 9, 2),
                                          (600, 20, 3), (1300, 11, 2), (2100, 5, 2)])
def test_zgesv(gpu_ctx, n, nrhs, batch):
    # even batches apply the row interchanges of the solve as one gather,
    # odd ones swap by swap (both paths of zgetrs_batched); panels of more
    # than 512 rows factor in global memory, shorter ones in registers; the
    # gather takes fewer columns per workgroup as n grows (4 at n = 513, 2 at
    # 1300, 1 at 2100)
    A, B = crand(batch, n, n), crand(batch, n, nrhs)
    X, info = g.zgesv(gpu_ctx, A, B)
    assert not info.any()
    ref = np.linalg.solve(A, B)
    # (the forward error grows with the condition number, ~n for these
    # random matrices; the backward error below does not)
    assert np.abs(X - ref).max() / np.abs(ref).max() < 1e-11 * max(1.0, n / 600)
    assert np.abs(A @ X - B).max() / (np.abs(A).max() * np.abs(X).max() * n) < 1e-14


def test_zgesv_singular_flag(gpu_ctx):
    A = crand(1, 8, 8)
    A[0, :, 3] = 0.0
    _, info = g.zgesv(gpu_ctx, A, crand(1, 8, 2))
    assert info[0] & 4


def match_eigenvalues(a, b):
    b = list(b)
    worst = 0.0
    for x in a:
        j = int(np.argmin(np.abs(np.array(b) - x)))
        worst = max(worst, abs(b[j] - x))
        b.pop(j)
    return worst


@pytest.mark.parametrize("n", [2, 3, 10, 50, 64, 65, 66, 98, 130, 200])
def test_zgeev_random(gpu_ctx, n):
    A = crand(2, n, n)
    lam, W, resid, info = g.zgeev(gpu_ctx, A)
    assert not info.any()
    for b in range(2):
        ref = np.linalg.eigvals(A[b])
        assert match_eigenvalues(lam[b], ref) < 1e-11 * np.abs(ref).max()
        assert resid[b] < 1e-12
        assert np.abs(np.linalg.norm(W[b], axis=0) - 1).max() < 1e-13
        r = np.abs(A[b] @ W[b] - W[b] * lam[b]).max() / np.abs(A[b]).max()
        assert abs(r - resid[b]) < 1e-13


def rcwa_pq(N, seed):
    r = np.random.default_rng(seed)
    Lp = 400e-9
    fx, fy = r.uniform(0.2, 0.8, 2)
    e1 = complex(r.uniform(2, 12), r.uniform(0, 0.1))
    xe, ye = [0, fx * Lp, Lp], [0, fy * Lp, Lp]
    eps = np.array([[e1, 1.0], [1.0, 1.0]])
    lay = o.Layer(xe, ye, eps)
    s = o.RCWASolver(N, N)
    s.disable_pml()
    s.add_layer(lay)
    lam = r.uniform(400e-9, 700e-9)
    k0 = 2 * np.pi / lam
    th = math.radians(r.uniform(0, 35))
    s.set_bloch_wavevector(k0 * math.sin(th), 0.0)
    a, b = s.alpha_beta()
    P, Q = lay.wave_eqn_coeff(np.diag(a / k0), np.diag(b / k0), 2 * np.pi, 2 * N + 1, 2 * N + 1)
    return P @ Q


@pytest.mark.parametrize("N", [1, 3, 7, 9])
def test_zgeev_rcwa_matrices(gpu_ctx, N):
    # (N = 9: n = 722, Hessenberg panels taller than 512 rows)
    A = np.stack([rcwa_pq(N, s) for s in range(3)])
    lam, W, resid, info = g.zgeev(gpu_ctx, A)
    assert not info.any()
    for b in range(3):
        ref = np.linalg.eigvals(A[b])
        assert match_eigenvalues(lam[b], ref) < 1e-10 * np.abs(ref).max()
        assert resid[b] < 1e-11


def test_zgeev_degenerate_block(gpu_ctx):
    # exactly degenerate pairs (a homogeneous layer's PQ is diag(kz^2) x 2)
    d = np.repeat(crand(20), 2)
    Q, _ = np.linalg.qr(crand(40, 40))
    A = (Q @ np.diag(d) @ np.linalg.inv(Q))[None]
    lam, W, resid, info = g.zgeev(gpu_ctx, A)
    assert not info.any() and resid[0] < 1e-11
    assert np.linalg.cond(W[0]) < 1e6


def test_zgeev_residual_status(gpu_ctx):
    """k_eig_residual sets ORCWA_STATUS_EIG_RESIDUAL exactly for the matrices
    whose relative residual ||AW - W diag(lam)||max / ||A||max exceeds the
    acceptance tolerance (rcwa/MKLEigenSolver.hpp:66-79): with the tolerance
    set between the residuals of a batch, the larger ones are flagged and
    nothing else changes; a negative tolerance flags every matrix."""
    from openrcwa_amd import _lib
    A = crand(6, 40, 40)
    _, W0, r0, info0 = g.zgeev(gpu_ctx, A)
    assert not info0.any() and r0.max() < 1e-12
    tol = float(np.median(r0))
    try:
        g.set_eig_tolerance(gpu_ctx, tol)
        _, W1, r1, info1 = g.zgeev(gpu_ctx, A)
        assert np.abs(r1 - r0).max() < 1e-15 and np.abs(W1 - W0).max() < 1e-13
        want = np.where(r1 > tol, _lib.STATUS_EIG_RESIDUAL, 0)
        assert np.array_equal(info1, want) and 0 < (want != 0).sum() < 6
        g.set_eig_tolerance(gpu_ctx, -1.0)
        _, _, _, info2 = g.zgeev(gpu_ctx, A)
        assert (info2 == _lib.STATUS_EIG_RESIDUAL).all()
    finally:
        g.set_eig_tolerance(gpu_ctx, 1e-8)


@pytest.mark.parametrize("nX,nY,xe,ye", [
    (5, 0, [0.0, 1.5e-7, 3e-7], [0.0, 3e-7]),
    (3, 2, [0.0, 1e-7, 2.2e-7, 4e-7], [0.0, 1.7e-7, 4e-7]),
    (7, 7, [0.0, 0.9e-7, 2e-7, 3.1e-7, 4e-7], [0.0, 1.2e-7, 2e-7, 2.9e-7, 4e-7]),
])
def test_convolution_matrices(gpu_ctx, nX, nY, xe, ye):
    eps = (rng.uniform(1, 12, (len(ye) - 1, len(xe) - 1)) +
           1j * rng.uniform(0, 0.2, (len(ye) - 1, len(xe) - 1)))
    got = g.convolution(gpu_ctx, nX, nY, xe, ye, eps)
    ref = o.convolution_matrices(xe, ye, eps, 2 * nX + 1, 2 * nY + 1)
    for a, b in zip(got, ref):
        assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()


def test_zgeev_largest_rcwa_matrix(gpu_ctx):
    # n = 1922 (31 x 31 orders): eigenvalues against LAPACK and the residual
    A = rcwa_pq(15, 0)[None]
    lam, W, resid, info = g.zgeev(gpu_ctx, A)
    assert not info.any()
    ref = np.linalg.eigvals(A[0])
    assert match_eigenvalues(lam[0], ref) < 1e-10 * np.abs(ref).max()
    assert resid[0] < 1e-11
