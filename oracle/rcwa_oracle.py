"""FP64 CPU oracle for the RCWA layer-stack solve.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``openrcwa_amd``) imports
this module; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker / the
timed CPU baseline.  It is a clean-room numpy restatement of the reference's
hot path (Sirokujira/OpenRCWA, ``rcwa/``), in the reference's own operation
order:

  a1  K matrices ............ rcwa/RCWASolver.cpp:150-185, 242-266
  a2  1D Fourier series ..... rcwa/FourierSolver1D.cpp:11-41, rcwa/MathFunction.h:6-9
  a3  Toeplitz / block ...... rcwa/ToeplitzMatrixXcs.cpp:8-44,
                              rcwa/BlockToeplitzMatrixXcs.cpp:5-65
  a4  Li factorisation ...... rcwa/FourierSolver2D.cpp:12-351
  a5  P, Q .................. rcwa/Layer.cpp:75-253
  a6  wave equation ......... rcwa/WaveEquationSolver.cpp:8-37, rcwa/Layer.cpp:255-273
  a7  eigen backend ......... rcwa/MKLEigenSolver.hpp:40-179 (LAPACK zgeev +
                              residual acceptance 1e-8 + perturbed fallback)
  a8  interfaces ............ rcwa/RCWASolver.cpp:495-614 (COD solves)
  a9  RTCM cascade .......... rcwa/RCWASolver.cpp:378-492
  a10 slab S-matrix+Redheffer rcwa/RCWAScatterMatrix.cpp:26-85,
                              core/RCWAIntegrator.cpp:57-67
  a11 plane wave ............ rcwa/RCWASolver.cpp:1778-1789, rcwa/Layer.cpp:453-480
  a12 per-order R/T ......... rcwa/RCWASolver.cpp:1792-1908
  a13 .ofd bridge glue ...... sol/rcwa_bridge.cpp:52-141, 179-366
  integ.  z-varying slabs ... core/RCWAIntegrator.cpp:20-86, core/VarLayerSampler.cpp:52-68
  fields  mode amplitudes ... rcwa/RCWASolver.cpp:1505-1645 (evaluateIntermediateField)
          field/device images rcwa/RCWASolver.cpp:707-1011, 1322-1487

Third-party arithmetic: the reference calls LAPACKE ``zgeev`` (unpinned
version) and Eigen3 (unpinned) ``BDCSVD``, ``CompleteOrthogonalDecomposition``
and ``PartialPivLU``.  Here ``numpy.linalg.eig`` is LAPACK ``zgeev`` (OpenBLAS
0.3.29 as shipped with numpy 2.2), COD solves are LAPACK ``zgelsy`` (the
complete-orthogonal-factorisation least-squares driver) and BDCSVD solves are
LAPACK ``zgelsd`` (divide-and-conquer SVD least squares), via scipy.

Pinning: every function is checked in ``tests/test_oracle.py`` against the
reference's published values (AGENTS.md:94-95, 114-120), closed forms
(Fresnel, Brewster, Airy film, uniform-layer eigenvalues, exact Fourier
integrals by quadrature) and the invariances the reference's CI asserts.  Two
items stay "parity unpinned": the basis-dependent intermediates (W, V,
gamma order, interface matrices, Rud, Tdd) and the Eigen-fallback
perturbation path -- only basis-invariant outputs (per-order R/T) are compared.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.linalg as sla

PI = 3.1415926535897932        # rcwa/defns.h:25
C0 = 2.99792458e8              # sol/rcwa_bridge.cpp:44 (speed of light, m/s)

CONTINUOUS_XY = 0              # rcwa/FourierSolver2D.h:7-12
DISCONTINUOUS_X = 1
DISCONTINUOUS_Y = 2


# --------------------------------------------------------------------------
# a2: exact Fourier coefficients of a piecewise-constant function
# --------------------------------------------------------------------------
def sinc(x):
    """sin(pi x)/(pi x) with sinc(0)=1 (rcwa/MathFunction.h:6-9)."""
    x = np.asarray(x, dtype=float)
    out = np.ones_like(x)
    nz = x != 0.0
    out[nz] = np.sin(PI * x[nz]) / (PI * x[nz])
    return out


def fourier_1d(coords, f, d, M):
    """Coefficients F_m, m = -(M-1)/2..(M-1)/2, of f on cells [x_{j-1}, x_j]
    with period d (rcwa/FourierSolver1D.cpp:11-41).  Phases use absolute
    coordinates: F_m = sum_j f_j dx_j/d exp(-i pi m (x_j+x_{j-1})/d) sinc(m dx_j/d).
    """
    x = np.asarray(coords, dtype=float)
    f = np.asarray(f, dtype=complex)
    if f.size != x.size - 1:
        raise ValueError("The function configuration does not match the coordinates")
    m = np.arange(M) - (M - 1) // 2
    x0, x1 = x[:-1], x[1:]
    dx = x1 - x0
    mm = m[:, None].astype(float)
    ph = np.exp(-1j * mm * PI * (x1 + x0)[None, :] / d)
    term = f[None, :] / d * dx[None, :] * ph * sinc(mm * dx[None, :] / d)
    # m == 0 row: f/d*dx without the phase factor (exactly 1 there anyway)
    term[m == 0, :] = (f / d * dx)[None, :]
    return term.sum(axis=1)


# --------------------------------------------------------------------------
# a3: Toeplitz and block-Toeplitz expansion
# --------------------------------------------------------------------------
def toeplitz_dense(vec):
    """T[r,c] = F[(r-c) + (n-1)], n = (len+1)/2 (rcwa/ToeplitzMatrixXcs.cpp:8-37)."""
    vec = np.asarray(vec, dtype=complex)
    n = (vec.size + 1) // 2
    r = np.arange(n)
    return vec[(r[:, None] - r[None, :]) + n - 1]


def toeplitz_inverse(vec):
    """Dense inverse of the Toeplitz matrix (rcwa/ToeplitzMatrixXcs.cpp:41-44)."""
    return np.linalg.inv(toeplitz_dense(vec))


def block_toeplitz_dense(mat):
    """B[(m,n),(j,l)] = f2d[m-j+nx-1, n-l+ny-1], row = m*ny+n
    (rcwa/BlockToeplitzMatrixXcs.cpp:5-52)."""
    mat = np.asarray(mat, dtype=complex)
    nx = (mat.shape[0] + 1) // 2
    ny = (mat.shape[1] + 1) // 2
    m = np.repeat(np.arange(nx), ny)
    n = np.tile(np.arange(ny), nx)
    return mat[(m[:, None] - m[None, :]) + nx - 1, (n[:, None] - n[None, :]) + ny - 1]


# --------------------------------------------------------------------------
# a4: Li factorisation rules for the three permittivity convolution matrices
# --------------------------------------------------------------------------
def _check_sorted(c):
    c = np.asarray(c, dtype=float)
    if np.any(np.diff(c) < 0):
        # FourierSolver1D::reorder/refill (rcwa/FourierSolver1D.cpp:145-193)
        # only matter for unsorted coordinates, which no caller on this path builds.
        raise ValueError("coordinates must be sorted")
    return c


def fourier_2d(coord_x, coord_y, f, nx, ny, opt):
    """FourierSolver2D::solve (rcwa/FourierSolver2D.cpp:12-351).

    f is (nCy-1) x (nCx-1), row = y cell.  nx, ny = number of harmonics per
    axis (2N+1).  Returns the dense (nx*ny) x (nx*ny) convolution matrix.
    """
    cx = _check_sorted(coord_x)
    cy = _check_sorted(coord_y)
    f = np.asarray(f, dtype=complex)
    dx = cx[-1] - cx[0]
    dy = cy[-1] - cy[0]
    h = nx * ny
    Ff = np.zeros((h, h), dtype=complex)
    ii = np.repeat(np.arange(nx), ny)       # x-harmonic index of row
    jj = np.tile(np.arange(ny), nx)         # y-harmonic index of row
    if opt == DISCONTINUOUS_X:              # :43-144
        for i in range(cy.size - 1):
            inv_fi = 1.0 / f[i, :]
            tffi = toeplitz_inverse(fourier_1d(cx, inv_fi, dx, 2 * nx - 1))
            fy1 = fourier_1d([cy[i], cy[i + 1]], [1.0], dy, 2 * ny - 1)
            Ff += fy1[(jj[:, None] - jj[None, :]) + ny - 1] * tffi[ii[:, None], ii[None, :]]
    elif opt == DISCONTINUOUS_Y:            # :145-249
        for i in range(cx.size - 1):
            inv_fi = 1.0 / f[:, i]
            tffi = toeplitz_inverse(fourier_1d(cy, inv_fi, dy, 2 * ny - 1))
            fx1 = fourier_1d([cx[i], cx[i + 1]], [1.0], dx, 2 * nx - 1)
            Ff += fx1[(ii[:, None] - ii[None, :]) + nx - 1] * tffi[jj[:, None], jj[None, :]]
    elif opt == CONTINUOUS_XY:              # :250-349
        f2 = np.zeros((2 * nx - 1, 2 * ny - 1), dtype=complex)
        for i in range(cy.size - 1):
            ffi = fourier_1d(cx, f[i, :], dx, 2 * nx - 1)
            fy1 = fourier_1d([cy[i], cy[i + 1]], [1.0], dy, 2 * ny - 1)
            f2 += np.outer(ffi, fy1)
        Ff = block_toeplitz_dense(f2)
    else:
        raise ValueError("unknown FourierSolverOption")
    return Ff


def convolution_matrices(coord_x, coord_y, eps, nx, ny):
    """(fe11, fe22, fe33) as built at the top of Layer::waveEqnCoeff
    (rcwa/Layer.cpp:99-116)."""
    return (fourier_2d(coord_x, coord_y, eps, nx, ny, DISCONTINUOUS_X),
            fourier_2d(coord_x, coord_y, eps, nx, ny, DISCONTINUOUS_Y),
            fourier_2d(coord_x, coord_y, eps, nx, ny, CONTINUOUS_XY))


# --------------------------------------------------------------------------
# a6/a7: branch-cut square root and the eigen backend
# --------------------------------------------------------------------------
def gamma_sqrt(g):
    """GammaSqrtOp (rcwa/WaveEquationSolver.cpp:8-22): principal sqrt, negate if
    Im<0, then negate again if |Im|<|Re| and Re<0."""
    gv = np.sqrt(np.asarray(g, dtype=complex))
    gv = np.where(gv.imag < 0, -gv, gv)
    flip = (np.abs(gv.imag) < np.abs(gv.real)) & (gv.real < 0)
    return np.where(flip, -gv, gv)


def rel_residual(A, V, d):
    """||A V - V diag(d)||_max / ||A||_max (rcwa/MKLEigenSolver.hpp:66-73)."""
    nA = np.abs(A).max()
    if not nA > 0:
        return 0.0
    return float(np.abs(A @ V - V * d[None, :]).max() / nA)


class EigenStats:
    """Counts of eigen solves and fallbacks (the reference's sticky flag is
    process-global, rcwa/MKLEigenSolver.hpp:112; here it is per oracle)."""

    def __init__(self):
        self.solves = 0
        self.fallbacks = 0
        self.max_residual = 0.0


# residual acceptance of an eigen solve (rcwa/MKLEigenSolver.hpp:75-79); tests
# lower it to force the recovery branch on ordinary matrices
EIG_TOL = 1e-8


def eig_backend(A, stats: EigenStats | None = None):
    """MKLEigenSolver<MatrixXcs>::compute (rcwa/MKLEigenSolver.hpp:103-179):
    zgeev('N','V'); accept if the relative residual <= EIG_TOL (1e-8),
    otherwise solve A + diag(1e-11 (i+1) max|A|) (the Eigen fallback's
    perturbation, :81-93) and report its residual against the unperturbed A
    (:170-176)."""
    d, V = np.linalg.eig(A)
    res = rel_residual(A, V, d)
    if stats is not None:
        stats.solves += 1
    if not (res <= EIG_TOL):
        scale = np.abs(A).max()
        Ap = A + np.diag(scale * 1e-11 * np.arange(1, A.shape[0] + 1))
        d, V = np.linalg.eig(Ap)
        res = rel_residual(A, V, d)
        if stats is not None:
            stats.fallbacks += 1
    if stats is not None:
        stats.max_residual = max(stats.max_residual, res)
    return d, V


def cod_solve(A, B):
    """CompleteOrthogonalDecomposition<MatrixXcs>(A).solve(B)
    (rcwa/RCWASolver.cpp:537-543): minimum-norm least squares by complete
    orthogonal factorisation (LAPACK zgelsy)."""
    return sla.lstsq(A, B, lapack_driver="gelsy", check_finite=False)[0]


def svd_solve(A, B):
    """BDCSVD<MatrixXcs>(A).solve(B) (rcwa/Layer.cpp:122-124, 477-479):
    minimum-norm least squares by divide-and-conquer SVD (LAPACK zgelsd)."""
    return sla.lstsq(A, B, lapack_driver="gelsd", check_finite=False)[0]


# --------------------------------------------------------------------------
# a5/a6: Layer
# --------------------------------------------------------------------------
class Layer:
    """rcwa/Layer.{h,cpp}: eps grid (nCy-1) x (nCx-1), row = y cell."""

    def __init__(self, coord_x, coord_y, eps, mu=None):
        self.coord_x = np.asarray(coord_x, dtype=float)
        self.coord_y = np.asarray(coord_y, dtype=float)
        self.eps_grid = np.atleast_2d(np.asarray(eps, dtype=complex))
        if self.eps_grid.shape != (self.coord_y.size - 1, self.coord_x.size - 1):
            raise ValueError("The configuration of permittivity does not match coordinates")
        if mu is None:
            self.mu_grid = np.ones_like(self.eps_grid)
        else:
            self.mu_grid = np.atleast_2d(np.asarray(mu, dtype=complex))
        # rcwa/Layer.cpp:66-72: magnetic iff some |mu-1| > 1e-12
        self.is_magnetic = bool(np.any(np.abs(self.mu_grid - 1.0) > 1e-12))
        self.W = self.V = self.gamma = None
        self.lam = 0.0

    # geometry accessors (rcwa/Layer.h:56-67)
    @property
    def Lx(self):
        return self.coord_x[-1] - self.coord_x[0]

    @property
    def Ly(self):
        return self.coord_y[-1] - self.coord_y[0]

    def edges(self):
        cx, cy = self.coord_x, self.coord_y
        return (cy[0], cy[1], cy[-2], cy[-1], cx[0], cx[1], cx[-2], cx[-1])

    def eps_at(self, x, y):
        """Layer::eps(x, y) cell lookup (rcwa/Layer.cpp:383-428): x wrapped
        into the period, y outside the cell range gives 0."""
        return self._cell_value(self.eps_grid, x, y)

    def mu_at(self, x, y):
        return self._cell_value(self.mu_grid, x, y)

    def _cell_value(self, grid, x, y):
        cx, cy = self.coord_x, self.coord_y
        while x < cx[0]:
            x += self.Lx
        while x > cx[-1]:
            x -= self.Lx
        if y < cy[0] or y > cy[-1]:
            return 0.0 + 0.0j
        ix = next(i for i in range(cx.size - 1) if x <= cx[i + 1])
        iy = next(j for j in range(cy.size - 1) if y <= cy[j + 1])
        return grid[iy, ix]

    def wave_eqn_coeff(self, Kx, Ky, lam, nx, ny):
        """P and Q (rcwa/Layer.cpp:75-253); Kx, Ky dense diagonal h x h."""
        fe11, fe22, fe33 = convolution_matrices(self.coord_x, self.coord_y,
                                                self.eps_grid, nx, ny)
        F1h = svd_solve(fe33, Kx)
        F2h = svd_solve(fe33, Ky)
        k0 = 2 * PI / lam
        k02 = k0 * k0
        h = nx * ny
        I = np.eye(h)
        if self.is_magnetic:
            fmu11, fmu22, fmu33 = convolution_matrices(self.coord_x, self.coord_y,
                                                       self.mu_grid, nx, ny)
            G1h = svd_solve(fmu33, Kx)
            G2h = svd_solve(fmu33, Ky)
            P = np.block([[Kx @ F2h, -Kx @ F1h + k02 * fmu22],
                          [Ky @ F2h - k02 * fmu11, -Ky @ F1h]])
            Q = np.block([[-Kx @ G2h, Kx @ G1h - k02 * fe22],
                          [-Ky @ G2h + k02 * fe11, Ky @ G1h]])
        else:
            P = np.block([[Kx @ F2h, -Kx @ F1h + k02 * I],
                          [Ky @ F2h - k02 * I, -Ky @ F1h]])
            Q = np.block([[-Kx @ Ky, Kx @ Kx - k02 * fe22],
                          [-Ky @ Ky + k02 * fe11, Ky @ Kx]])
        return P, Q

    def solve(self, Kx, Ky, lam, nx, ny, stats=None):
        """Layer::solve + WaveEquationSolver::solve (rcwa/Layer.cpp:255-273,
        rcwa/WaveEquationSolver.cpp:24-37)."""
        k0 = 2 * PI / lam
        P, Q = self.wave_eqn_coeff(Kx, Ky, lam, nx, ny)
        PQ = P @ Q
        d, W = eig_backend(PQ, stats)
        keff = gamma_sqrt(d)
        self.W = W
        self.V = Q @ W @ np.diag(1.0 / keff)
        self.gamma = keff / k0
        self.lam = lam
        return self

    def generate_plane_wave(self, delta, px, py):
        """Layer::generatePlaneWave (rcwa/Layer.cpp:453-480): c = W^+ s."""
        s = np.concatenate([px * delta, py * delta])
        return svd_solve(self.W, s)


# --------------------------------------------------------------------------
# a1, a8, a9, a11, a12: the stack solver
# --------------------------------------------------------------------------
class RCWASolver:
    """rcwa/RCWASolver.{h,cpp}, periodic (PML disabled) path."""

    def __init__(self, nX, nY):
        self.nx = 2 * nX + 1          # rcwa/RCWASolver.h:67-78
        self.ny = 2 * nY + 1
        self.layers: list[Layer] = []
        self.enable_pml = True
        self.kx0 = 0.0
        self.ky0 = 0.0
        self.stats = EigenStats()
        self.Rud = self.Tdd = None
        self.Kx = self.Ky = None

    def add_layer(self, layer: Layer):
        """RCWASolver::addLayer (rcwa/RCWASolver.cpp:108-137): a layer whose
        period or edge cells differ from the first is reported and dropped."""
        if self.layers:
            a = self.layers[0]
            if (a.Lx != layer.Lx or a.Ly != layer.Ly or a.edges() != layer.edges()):
                print("The new layer does not match layers before!")
                return False
        self.layers.append(layer)
        return True

    def disable_pml(self):
        self.enable_pml = False

    def set_bloch_wavevector(self, kx0, ky0):
        self.kx0, self.ky0 = kx0, ky0

    def alpha_beta(self):
        """RCWASolver::evaluateAlphaBeta (rcwa/RCWASolver.cpp:242-266)."""
        L0 = self.layers[0]
        Nx, Ny = (self.nx - 1) // 2, (self.ny - 1) // 2
        m = np.repeat(np.arange(self.nx) - Nx, self.ny).astype(float)
        n = np.tile(np.arange(self.ny) - Ny, self.nx).astype(float)
        alpha = (2 * PI / L0.Lx) * m + self.kx0
        beta = (2 * PI / L0.Ly) * n + self.ky0
        return alpha, beta

    def solve(self, lam, stack, thickness):
        """RCWASolver::solve (rcwa/RCWASolver.cpp:139-148)."""
        if self.enable_pml:
            raise NotImplementedError("PML path is out of scope (disabled on the plain solve)")
        alpha, beta = self.alpha_beta()              # :150-185
        self.Kx = alpha.astype(complex)
        self.Ky = beta.astype(complex)
        Kx, Ky = np.diag(self.Kx), np.diag(self.Ky)
        for L in self.layers:                         # :194-210
            L.solve(Kx, Ky, lam, self.nx, self.ny, self.stats)
        E, H, imap = self.evaluate_interfaces(stack)
        self.Tdd, self.Rud = self.evaluate_scatter_matrix(stack, thickness, E, H, imap)

    def evaluate_interfaces(self, stack):
        """rcwa/RCWASolver.cpp:495-569 (translation == nullptr branch)."""
        pairs = sorted({(stack[i], stack[i + 1]) for i in range(len(stack) - 1)})
        nD = 2 * self.nx * self.ny
        E, H = [], []
        for (i1, i2) in pairs:
            if i1 == i2:
                E.append(np.eye(nD, dtype=complex))
                H.append(np.eye(nD, dtype=complex))
            else:
                E.append(cod_solve(self.layers[i1].W, self.layers[i2].W))
                H.append(cod_solve(self.layers[i1].V, self.layers[i2].V))
        imap = [pairs.index((stack[i], stack[i + 1])) for i in range(len(stack) - 1)]
        return E, H, imap

    def evaluate_scatter_matrix(self, stack, thickness, E, H, imap):
        """RTCM recursion from the bottom (rcwa/RCWASolver.cpp:378-492)."""
        nD = 2 * self.nx * self.ny
        nL = len(stack)
        Tdd = np.eye(nD, dtype=complex)
        Rud = np.zeros((nD, nD), dtype=complex)
        I = np.eye(nD)
        # kept for the field reconstruction (the reference's taus_,
        # interfacesE_/H_ and interfaceMap_)
        self.taus, self.iface_E, self.iface_H, self.imap = [], E, H, imap
        for i in range(nL - 1):
            idx = nL - i - 1
            L = self.layers[stack[idx]]
            if i == 0:
                phi = np.ones(nD, dtype=complex)
            else:
                phi = np.exp(1j * L.gamma * thickness[idx])
            wR = phi[:, None] * Rud * phi[None, :]
            wT = Tdd * phi[None, :]
            k = imap[idx - 1]
            F = E[k] + E[k] @ wR
            G = H[k] - H[k] @ wR
            tau = sla.lu_solve(sla.lu_factor(F + G, check_finite=False), I,
                               check_finite=False)
            self.taus.append(tau)
            Rud = -2.0 * G @ tau + I
            Tdd = 2.0 * wT @ tau
        return Tdd, Rud

    def generate_horizontal_plane_wave(self, px, py, layer_type):
        """rcwa/RCWASolver.cpp:1778-1789: delta at k = ny(nx-1)/2 + (ny-1)/2."""
        h = self.nx * self.ny
        delta = np.zeros(h, dtype=complex)
        delta[self.ny * (self.nx - 1) // 2 + (self.ny - 1) // 2] = 1.0
        return self.layers[layer_type].generate_plane_wave(delta, px, py)

    def scatter_plane_wave(self, c_inc, ref, trn, k0):
        """Per-order efficiencies (rcwa/RCWASolver.cpp:1792-1908).  Returns
        (REF, TRN) or (None, None) on the reference's silent early returns."""
        h = self.nx * self.ny
        rxy = self.layers[ref].W @ (self.Rud @ c_inc)
        txy = self.layers[trn].W @ (self.Tdd @ c_inc)
        rx, ry = rxy[:h], rxy[h:]
        tx, ty = txy[:h], txy[h:]
        Lr, Lt = self.layers[ref], self.layers[trn]
        eps_ref, eps_trn = Lr.eps_at(0.0, 0.0), Lt.eps_at(0.0, 0.0)
        mu_ref, mu_trn = Lr.mu_at(0.0, 0.0), Lt.mu_at(0.0, 0.0)
        kref2 = k0 * k0 * eps_ref * mu_ref
        ktrn2 = k0 * k0 * eps_trn * mu_trn
        kzr = -np.sqrt(kref2 - self.Kx * self.Kx - self.Ky * self.Ky)
        kzt = np.sqrt(ktrn2 - self.Kx * self.Kx - self.Ky * self.Ky)
        if np.any(kzr == 0) or np.any(kzt == 0):
            return None, None
        rz = (self.Kx * rx + self.Ky * ry) / kzr
        tz = (self.Kx * tx + self.Ky * ty) / kzt
        r = np.abs(rx) ** 2 + np.abs(ry) ** 2 + np.abs(rz) ** 2
        t = np.abs(tx) ** 2 + np.abs(ty) ** 2 + np.abs(tz) ** 2
        inc = self.ny * (self.nx - 1) // 2 + (self.ny - 1) // 2
        kzinc = abs(kzr[inc].real)
        if kzinc < 1e-12 * k0:
            return None, None
        mu_ratio = abs(mu_ref / mu_trn)
        return (np.abs(kzr.real) * r / kzinc, mu_ratio * np.abs(kzt.real) * t / kzinc)


# --------------------------------------------------------------------------
# field reconstruction: mode amplitudes inside the stack and field images
# --------------------------------------------------------------------------
EX, EY, EZ, HX, HY, HZ = range(6)            # rcwa/defns.h:27-35
SLICE_XY, SLICE_XZ, SLICE_YZ = range(3)      # rcwa/defns.h:37-42
MODULATION, REALPART, IMAGPART = range(3)    # rcwa/defns.h:44-49


def intermediate_field(solver, c_inc, stack, thickness):
    """RCWASolver::evaluateIntermediateField, inputCoeffs form
    (rcwa/RCWASolver.cpp:1505-1645).  Returns (c_m, c_p) for the interior
    layers j = 1..L-2 (list index j-1): c_p the downward mode amplitudes at
    the layer's top, c_m the upward ones at its bottom.  Interface
    S-matrices from t11 = (E+H)/2, t12 = (E-H)/2 (:1574-1586), the RTCM
    reflection matrices rebuilt from taus_ (:1593-1598), the input entries
    below 1e-14 dropped as the driver does (rcwa/RCWADriver.cpp:344-346)."""
    nD = 2 * solver.nx * solver.ny
    nL = len(stack)
    if nL < 3:
        raise ValueError("Intermediate field requires at least one finite-thickness "
                         "interior layer")
    I = np.eye(nD)
    S11, S12, S21, S22, phases, REFs = [], [], [], [], [], []
    Rud = np.zeros((nD, nD), dtype=complex)
    REFs.append(Rud)
    for_inv_s22 = None
    for i in range(nL - 1):
        index = nL - i - 1
        L = solver.layers[stack[index]]
        phi = (np.ones(nD, dtype=complex) if i == 0
               else np.exp(1j * L.gamma * thickness[index]))
        phases.append(phi)
        k = solver.imap[index - 1]
        E, H = solver.iface_E[k], solver.iface_H[k]
        t11 = 0.5 * (E + H)
        t12 = 0.5 * (E - H)
        lu = sla.lu_factor(t11, check_finite=False)
        inv_t22 = sla.lu_solve(lu, I, check_finite=False)
        inv_t22_t12 = sla.lu_solve(lu, t12, check_finite=False)
        S11.append(t11 - t12 @ inv_t22_t12)
        S12.append(t12 @ inv_t22)
        S21.append(-inv_t22_t12)
        S22.append(inv_t22)
        if i == 0:
            for_inv_s22 = t11
        wR = phi[:, None] * Rud * phi[None, :]
        G = H - H @ wR
        Rud = -2.0 * G @ solver.taus[i] + I
        REFs.append(Rud)
    c_in = np.asarray(c_inc, dtype=complex)
    inp = np.where(np.abs(c_in) > 1e-14, c_in, 0.0)
    R = solver.Rud @ inp
    T = solver.Tdd @ inp
    nIn = nL - 2
    c_m, c_p, d = [None] * nIn, [None] * nIn, [None] * nIn
    c_m[nIn - 1] = S12[0] @ (for_inv_s22 @ T)
    temp = cod_solve(S11[nL - 2], R - S12[nL - 2] @ inp)
    c_p[0] = S21[nL - 2] @ temp + S22[nL - 2] @ inp
    d[nIn - 1] = phases[nL - 2] * c_p[0]
    for i in range(nIn - 2, -1, -1):
        c_m[nIn - i - 2] = REFs[i + 2] @ d[i + 1]
        tmp = cod_solve(S11[i + 1], c_m[nIn - i - 2] - S12[i + 1] @ d[i + 1])
        c_p[nIn - i - 1] = S21[i + 1] @ tmp + S22[i + 1] @ d[i + 1]
        d[i] = phases[i + 1] * c_p[nIn - i - 1]
    return c_m, c_p


def find_interior_layer(stack, thickness, z, closed=True):
    """Layer search of saveFieldImage (closed interval, rcwa/RCWASolver.cpp:
    788-802) or saveDeviceImage (half-open, :1396-1408): z from the top of
    the first interior layer, downwards.  Returns (j, z0, z1) or None."""
    z0 = 0.0
    for j in range(1, len(stack) - 1):
        z1 = z0 + thickness[j]
        if (z0 <= z <= z1) if closed else (z0 <= z < z1):
            return j, z0, z1
        z0 = z1
    return None


def layer_harmonics(solver, c_m, c_p, stack, j, z_p, z_m, comp):
    """Harmonics of one field component in interior layer j at z_p below its
    top (z_m = z_p - thickness <= 0), the SET_HARMONICS_2D lambda of
    rcwa/RCWASolver.cpp:781-895: u = c_p exp(i gamma z_p),
    d = c_m exp(-i gamma z_m); [Ex; Ey] = W (u + d), [Hx; Hy] = V (u - d);
    Hz = (i/k0^2)(Kx Ey - Ky Ex); Ez = i [[eps]]^-1 (Kx Hy - Ky Hx) with the
    ContinuousXY convolution matrix (rcwa/Layer.cpp:275-283; its
    colPivHouseholderQr solve is an exact solve here).  Returns the h-vector,
    index k = m*ny + n."""
    L = solver.layers[stack[j]]
    h = solver.nx * solver.ny
    u = c_p[j - 1] * np.exp(1j * z_p * L.gamma)
    d = c_m[j - 1] * np.exp(-1j * z_m * L.gamma)
    if comp in (EX, EY):
        v = L.W @ (u + d)
        return v[:h] if comp == EX else v[h:]
    if comp in (HX, HY):
        v = L.V @ (u - d)
        return v[:h] if comp == HX else v[h:]
    Kx, Ky = solver.Kx, solver.Ky
    k0 = 2 * PI / L.lam
    if comp == HZ:
        E = L.W @ (u + d)
        return 1j / (k0 * k0) * (Kx * E[h:] - Ky * E[:h])
    if comp == EZ:
        Hv = L.V @ (u - d)
        rhs = Kx * Hv[h:] - Ky * Hv[:h]
        fe33 = fourier_2d(L.coord_x, L.coord_y, L.eps_grid, solver.nx, solver.ny,
                          CONTINUOUS_XY)
        return 1j * np.linalg.solve(fe33, rhs)
    raise ValueError("Evaluation for this component is not implemented!")


def _image_axes(solver, stack, thickness, slice_type, nres1):
    """Sampling of saveFieldImage / saveDeviceImage (rcwa/RCWASolver.cpp:
    730-779, 1330-1380): cell-centred samples, nres2 = int(nres1 L2 / L1) >= 1."""
    L0 = solver.layers[0]
    if slice_type == SLICE_XZ:
        s1, l1 = L0.coord_x[0], L0.Lx
        s2, l2 = 0.0, float(sum(thickness[1:len(stack) - 1]))
    elif slice_type == SLICE_YZ:
        s1, l1 = L0.coord_y[0], L0.Ly
        s2, l2 = 0.0, float(sum(thickness[1:len(stack) - 1]))
    elif slice_type == SLICE_XY:
        s1, l1 = L0.coord_x[0], L0.Lx
        s2, l2 = L0.coord_y[0], L0.Ly
    else:
        raise ValueError("unknown slice")
    nres2 = max(1, int(nres1 * l2 / l1))
    st1, st2 = l1 / nres1, l2 / nres2
    return (s1 + st1 * (np.arange(nres1) + 0.5), s2 + st2 * (np.arange(nres2) + 0.5))


def field_image(solver, c_inc, stack, thickness, slice_type, coord, comp, nres1=500):
    """RCWASolver::saveFieldImage, inputCoeffs form (rcwa/RCWASolver.cpp:
    707-1011): the complex field on the slice (rows = second axis), before
    writeFieldCSV takes |f|, Re f or Im f.  f(r) = sum_mn f_mn
    exp(i (kx0 + 2 pi m / Lx) x + i (ky0 + 2 pi n / Ly) y); XZ/YZ rows run
    down the interior layers, XY is the plane z = coord."""
    c_m, c_p = intermediate_field(solver, c_inc, stack, thickness)
    a1, a2 = _image_axes(solver, stack, thickness, slice_type, nres1)
    nx, ny = solver.nx, solver.ny
    Lx, Ly = solver.layers[0].Lx, solver.layers[0].Ly
    mx = np.arange(nx) - (nx - 1) // 2
    my = np.arange(ny) - (ny - 1) // 2
    kxm = 2 * PI * mx / Lx + solver.kx0
    kyn = 2 * PI * my / Ly + solver.ky0

    def harm2d(z):
        hit = find_interior_layer(stack, thickness, z)
        if hit is None:
            raise ValueError("slice outside the interior layers")
        j, z0, z1 = hit
        return layer_harmonics(solver, c_m, c_p, stack, j, z - z0, z - z1,
                               comp).reshape(nx, ny)

    out = np.zeros((a2.size, a1.size), dtype=complex)
    if slice_type in (SLICE_XZ, SLICE_YZ):
        for i, z in enumerate(a2):
            H2 = harm2d(z)
            if slice_type == SLICE_XZ:
                one = H2 @ np.exp(1j * kyn * coord)
                out[i] = np.exp(1j * np.outer(a1, kxm)) @ one
            else:
                one = H2.T @ np.exp(1j * kxm * coord)
                out[i] = np.exp(1j * np.outer(a1, kyn)) @ one
    else:
        H2 = harm2d(coord)
        wx = np.exp(1j * np.outer(a1, kxm))          # [x, m]
        wy = np.exp(1j * np.outer(a2, kyn))          # [y, n]
        out = wy @ H2.T @ wx.T
    return out


def device_image(solver, stack, thickness, slice_type, coord, nres1=2000):
    """RCWASolver::saveDeviceImage (rcwa/RCWASolver.cpp:1322-1487): the layer's
    permittivity at the sample points (Layer::eps cell lookup), layer search
    on the half-open interval.  Returns the complex grid (rows = second axis)."""
    a1, a2 = _image_axes(solver, stack, thickness, slice_type, nres1)
    out = np.zeros((a2.size, a1.size), dtype=complex)

    def layer_at(z):
        hit = find_interior_layer(stack, thickness, z, closed=False)
        if hit is None:
            raise ValueError("Device image requires at least one finite-thickness "
                             "interior layer!")
        return solver.layers[stack[hit[0]]]

    if slice_type in (SLICE_XZ, SLICE_YZ):
        for i, z in enumerate(a2):
            L = layer_at(z)
            for j, v in enumerate(a1):
                out[i, j] = L.eps_at(v, coord) if slice_type == SLICE_XZ else L.eps_at(coord, v)
    else:
        L = layer_at(coord)
        for i, y in enumerate(a2):
            for j, x in enumerate(a1):
                out[i, j] = L.eps_at(x, y)
    return out


# --------------------------------------------------------------------------
# a10: slab S-matrix in a reference basis and the Redheffer star product
# --------------------------------------------------------------------------
def slab_scatter_matrix(W, V, gamma, W0, V0, wz, lam):
    """dtmm::RCWAScatterMatrix::compute (rcwa/RCWAScatterMatrix.cpp:26-85):
    returns (Rud, Tuu) of a slab of thickness wz embedded in the basis (W0, V0)."""
    F0 = np.linalg.solve(W, W0)
    G0 = np.linalg.solve(V, V0)
    A = F0 + G0
    B = F0 - G0
    X = np.exp(1j * gamma * wz * lam / (2 * PI))
    Ai = np.linalg.inv(A)
    AiXB = Ai @ (X[:, None] * B)     # A^-1 X B
    AiXA = Ai @ (X[:, None] * A)     # A^-1 X A
    AiB = Ai @ B
    n = W.shape[0]
    Minv = np.linalg.inv(np.eye(n) - AiXB @ AiXB)
    Rud = Minv @ (AiXB @ AiXA - AiB)
    Tuu = Minv @ (AiXA - AiXB @ AiB)
    return Rud, Tuu


def redheffer_star_general(top, bottom):
    """Redheffer star product top * bottom of S-matrices (rud, tuu, rdu, tdd),
    in the form of core/RCWAIntegrator.cpp:57-67 (there the new slab is `top`
    and the accumulated stack `bottom`)."""
    rud, tuu, rdu, tdd = top
    Rud, Tuu, Rdu, Tdd = bottom
    n = Rud.shape[0]
    I = np.eye(n)
    C1 = np.linalg.inv(I - rdu @ Rud)
    C2 = I + Rud @ C1 @ rdu
    return (rud + tuu @ Rud @ C1 @ tdd, tuu @ C2 @ Tuu, Rdu + Tdd @ rdu @ C2 @ Tuu, Tdd @ C1 @ tdd)


def redheffer_star(state, Rud, Tuu):
    """The accumulated (rud, tuu, rdu, tdd) above a symmetric slab
    (Rdu = Rud, Tdd = Tuu): redheffer_star_general(state, slab)."""
    return redheffer_star_general(state, (Rud, Tuu, Rud, Tuu))


# --------------------------------------------------------------------------
# z-varying structures: the midpoint-staircase integrator
# --------------------------------------------------------------------------
def wave_equation(P, Q, stats=None):
    """WaveEquationSolver::solve (rcwa/WaveEquationSolver.cpp:24-37): modes of
    P Q with keff = GammaSqrtOp(eig) in the units of P and Q (no 1/k0)."""
    d, W = eig_backend(P @ Q, stats)
    keff = gamma_sqrt(d)
    return W, Q @ W @ np.diag(1.0 / keff), keff


def sampler_pq(layer, lam, nX, nY, kx0=0.0, ky0=0.0):
    """VarLayerSampler::sample (core/VarLayerSampler.cpp:52-68) on the
    periodic branch of LayerSampler::evaluateKMatrices
    (core/LayerSampler.cpp:114-144): P and Q of the layer at this z with
    K = diag(2 pi m / Lx + kx0), diag(2 pi n / Ly + ky0)."""
    nx, ny = 2 * nX + 1, 2 * nY + 1
    m = np.repeat(np.arange(nx) - nX, ny).astype(float)
    n = np.tile(np.arange(ny) - nY, nx).astype(float)
    Kx = np.diag((2 * PI / layer.Lx) * m + kx0).astype(complex)
    Ky = np.diag((2 * PI / layer.Ly) * n + ky0).astype(complex)
    return layer.wave_eqn_coeff(Kx, Ky, lam, nx, ny)


def integrator_sample_points(z0, z1, N):
    """The z of every sample RCWAIntegrator::compute takes, in the order of
    core/RCWAIntegrator.cpp:22-75: z1 (the top basis), the N midpoints
    z0 + (i + 1/2) dz, then z0 (the bottom basis)."""
    dz = (z1 - z0) / float(N)
    return [z1] + [z0 + (i + 0.5) * dz for i in range(N)] + [z0]


def rcwa_integrator(sample, lam, nX, nY, z0, z1, N, kx0=0.0, ky0=0.0):
    """RCWAIntegrator::compute(N) (core/RCWAIntegrator.cpp:20-86): the
    structure between z0 and z1, `sample(z)` -> Layer, as N midpoint slabs
    in the basis of the modes at z1, each put on top of the accumulated
    S-matrix by the Redheffer star (:57-67), then the bottom side taken into
    the basis of the modes at z0 (:74-85).  Returns dict(Rud, Tuu, Rdu, Tdd,
    W0, V0, gam0, W1, V1, gam1) with gam the raw keff of WaveEquationSolver."""
    P1, Q1 = sampler_pq(sample(z1), lam, nX, nY, kx0, ky0)
    W1, V1, g1 = wave_equation(P1, Q1)
    n = W1.shape[0]
    I = np.eye(n)
    Tdd, Tuu = I.astype(complex), I.astype(complex)
    Rud, Rdu = np.zeros((n, n), complex), np.zeros((n, n), complex)
    dz = (z1 - z0) / float(N)
    for i in range(N):
        z = z0 + (i + 0.5) * dz
        P, Q = sampler_pq(sample(z), lam, nX, nY, kx0, ky0)
        W, V, g = wave_equation(P, Q)
        rud, tuu = slab_scatter_matrix(W, V, g, W1, V1, dz, lam)
        # the accessors Rdu() / Tdd() return Rud_ / Tuu_ (rcwa/RCWAScatterMatrix.h:44-47)
        Rud, Tuu, Rdu, Tdd = redheffer_star_general((rud, tuu, rud, tuu), (Rud, Tuu, Rdu, Tdd))
    P0, Q0 = sampler_pq(sample(z0), lam, nX, nY, kx0, ky0)
    W0, V0, g0 = wave_equation(P0, Q0)
    F = np.linalg.solve(W1, W0)
    G = np.linalg.solve(V1, V0)
    A, B = 0.5 * (F + G), 0.5 * (F - G)
    lu = sla.lu_factor(A - Rdu @ B)
    Rdu = sla.lu_solve(lu, Rdu @ A - B)
    Tdd = sla.lu_solve(lu, Tdd)
    Rud = Tuu @ B @ Tdd + Rud
    Tuu = Tuu @ (B @ Rdu + A)
    return dict(Rud=Rud, Tuu=Tuu, Rdu=Rdu, Tdd=Tdd, W0=W0, V0=V0, gam0=g0, W1=W1, V1=V1,
                gam1=g1)


def _perturbation(Pr, Qr, W, V, gam, z0, z1, k0, P0, Pm, P1, Q0, Qm, Q1):
    """RedhefferIntegrator::evaluatePerturbationAndDifference
    (core/RedhefferIntegrator.cpp:317-399): first-order S-matrix of [z0, z1]
    around the reference modes (W, V, gam) of (Pr, Qr), the z-integral by
    Simpson's rule over the samples at z0, the middle and z1.  Returns
    (tdd, tuu, rud, rdu, diffError)."""
    luW, luV = sla.lu_factor(W), sla.lu_factor(V)

    def dAB(P, Q):
        a = sla.lu_solve(luW, (P - Pr) @ V)
        b = sla.lu_solve(luV, (Q - Qr) @ W)
        return a + b, a - b
    dA0, dB0 = dAB(P0, Q0)
    dA1, dB1 = dAB(Pm, Qm)
    dA2, dB2 = dAB(P1, Q1)
    ph = np.exp(1j * 0.5 * (z1 - z0) / k0 * gam)
    pf = np.exp(1j * (z1 - z0) / k0 * gam)
    c = 1j * (z1 - z0) / (6 * 2 * k0)
    tuu = c * (pf[:, None] * dA0 + 4.0 * ph[:, None] * dA1 * ph[None, :] + dA2 * pf[None, :])
    rud = -c * (pf[:, None] * dB0 * pf[None, :] + 4.0 * ph[:, None] * dB1 * ph[None, :] + dB2)
    rdu = -c * (dB0 + 4.0 * ph[:, None] * dB1 * ph[None, :] + pf[:, None] * dB2 * pf[None, :])
    tdd = c * (dA0 * pf[None, :] + 4.0 * ph[:, None] * dA1 * ph[None, :] + pf[:, None] * dA2)
    err = max(np.abs(tuu).max(), np.abs(rud).max(), np.abs(rdu).max(), np.abs(tdd).max())
    tuu = tuu + np.diag(pf)
    tdd = tdd + np.diag(pf)
    return tdd, tuu, rud, rdu, err


def redheffer_integrator(sample, lam, nX, nY, z0, z1, pe, kx0=0.0, ky0=0.0, max_segments=None):
    """RedhefferIntegrator::compute (core/RedhefferIntegrator.cpp:173-314, the
    active triple-division branch): first-order perturbation around the
    segment's midpoint modes; a segment whose perturbation terms exceed pe is
    cut in three (the middle third keeps the parent's reference); accepted
    segments, from z0 upwards, are taken from the previous segment's basis
    into their own (:242-262) and put on top by the Redheffer star
    (:264-277); finally the top side goes into the basis at z1 (:198-213).
    Returns the integrator's dict plus N (segments) and nEigVal."""
    k0 = 2 * PI / lam

    def pq(z):
        return sampler_pq(sample(z), lam, nX, nY, kx0, ky0)
    P0, Q0 = pq(z0)
    W0, V0, g0 = wave_equation(P0, Q0)
    P1, Q1 = pq(z1)
    W1, V1, g1 = wave_equation(P1, Q1)
    n = W0.shape[0]
    I = np.eye(n)
    st = dict(Tdd=I.astype(complex), Tuu=I.astype(complex), Rud=np.zeros((n, n), complex),
              Rdu=np.zeros((n, n), complex), leftW=W0, leftV=V0, N=0, nEig=2)
    zm = 0.5 * (z0 + z1)
    Pm, Qm = pq(zm)
    Wm, Vm, gm = wave_equation(Pm, Qm)

    def accept(W, V, t):
        tdd, tuu, rud, rdu = t
        F = np.linalg.solve(W, st["leftW"])
        G = np.linalg.solve(V, st["leftV"])
        A, B = 0.5 * (F + G), 0.5 * (F - G)
        lu = sla.lu_factor(A - rdu @ B)
        rdu = sla.lu_solve(lu, rdu @ A - B)
        tdd = sla.lu_solve(lu, tdd)
        rud = tuu @ B @ tdd + rud
        tuu = tuu @ A + tuu @ B @ rdu
        st["leftW"], st["leftV"] = W, V
        st["Rud"], st["Tuu"], st["Rdu"], st["Tdd"] = redheffer_star_general(
            (rud, tuu, rdu, tdd), (st["Rud"], st["Tuu"], st["Rdu"], st["Tdd"]))

    def rec(Pr, Qr, W, V, gam, a, b, Pa, Pmid, Pb, Qa, Qmid, Qb):
        tdd, tuu, rud, rdu, err = _perturbation(Pr, Qr, W, V, gam, a, b, k0,
                                                Pa, Pmid, Pb, Qa, Qmid, Qb)
        if err <= pe:
            st["N"] += 1
            if max_segments is not None and st["N"] > max_segments:
                raise RuntimeError("too many segments")
            accept(W, V, (tdd, tuu, rud, rdu))
            return
        m0 = 2.0 / 3.0 * a + 1.0 / 3.0 * b
        m1 = 1.0 / 3.0 * a + 2.0 / 3.0 * b
        Pt, Qt = pq(0.5 * (a + m0))
        Pk, Qk = pq(m0)
        Wt, Vt, gt = wave_equation(Pt, Qt)
        rec(Pt, Qt, Wt, Vt, gt, a, m0, Pa, Pt, Pk, Qa, Qt, Qk)
        Pt, Qt = pq(m1)
        rec(Pr, Qr, W, V, gam, m0, m1, Pk, Pmid, Pt, Qk, Qmid, Qt)
        Pk, Qk = pq(0.5 * (m1 + b))
        Wt, Vt, gt = wave_equation(Pk, Qk)
        rec(Pk, Qk, Wt, Vt, gt, m1, b, Pt, Pk, Pb, Qt, Qk, Qb)

    rec(Pm, Qm, Wm, Vm, gm, z0, z1, P0, Pm, P1, Q0, Qm, Q1)
    # the top side into the basis at z1
    F = np.linalg.solve(st["leftW"], W1)
    G = np.linalg.solve(st["leftV"], V1)
    A, B = 0.5 * (F + G), 0.5 * (F - G)
    lu = sla.lu_factor(A - st["Rud"] @ B)
    Tuu = sla.lu_solve(lu, st["Tuu"])
    Rud = sla.lu_solve(lu, st["Rud"] @ A - B)
    Rdu = st["Rdu"] + st["Tdd"] @ B @ Tuu
    Tdd = st["Tdd"] @ (B @ Rud + A)
    return dict(Rud=Rud, Tuu=Tuu, Rdu=Rdu, Tdd=Tdd, W0=W0, V0=V0, gam0=g0, W1=W1, V1=V1,
                gam1=g1, N=st["N"], nEigVal=st["nEig"])


def integrator_harmonic_operators(r):
    """The integrator's S-matrix in harmonic (E-field) space, independent of
    the mode bases: T_down = W0 Tdd W1^-1, R_top = W1 Rud W1^-1,
    T_up = W1 Tuu W0^-1, R_bottom = W0 Rdu W0^-1."""
    W0, W1 = r["W0"], r["W1"]
    return (W0 @ r["Tdd"] @ np.linalg.inv(W1), W1 @ r["Rud"] @ np.linalg.inv(W1),
            W1 @ r["Tuu"] @ np.linalg.inv(W0), W0 @ r["Rdu"] @ np.linalg.inv(W0))


# --------------------------------------------------------------------------
# a13: .ofd bridge glue
# --------------------------------------------------------------------------
def common_edges(period, fills):
    """commonEdges (sol/rcwa_bridge.cpp:52-67)."""
    e = [0.0] + [f * period for f in fills if f < 1.0] + [period]
    e.sort()
    u = []
    for v in e:
        if not u or (v - u[-1]) > 1e-12 * period:
            u.append(v)
    return u


def make_ofd_layer(fill, fill_y, e1, e2, xe, ye):
    """makeLayer (sol/rcwa_bridge.cpp:72-89): eps1 on [0, fill Lx] x [0, fill_y Ly]."""
    px, py = xe[-1], ye[-1]
    eps = np.empty((len(ye) - 1, len(xe) - 1), dtype=complex)
    for j in range(len(ye) - 1):
        yc = 0.5 * (ye[j] + ye[j + 1])
        for i in range(len(xe) - 1):
            xc = 0.5 * (xe[i] + xe[i + 1])
            eps[j, i] = e1 if (xc < fill * px and yc < fill_y * py) else e2
    return Layer(xe, ye, eps)


def lorentz_eps(static, poles, omega):
    """layerEps (sol/rcwa_bridge.cpp:105-123).  poles = [(einf, ae, be, ce), ...]
    for one (layer, side); einf is the last row's value; ae == 0 rows are not poles."""
    active = [p for p in poles if p[1] != 0.0]
    if not active:
        return complex(static)
    e = complex(poles[-1][0], 0.0)
    for (_, ae, be, ce) in active:
        e += complex(ae * ae, 0.0) / complex(ce * ce - omega * omega, -be * omega)
    return e


def polarisation_basis(theta_deg, phi_deg):
    """TE/TM transverse amplitudes (sol/rcwa_bridge.cpp:187-196).
    Returns ((teX, teY), (tmX, tmY))."""
    th = theta_deg * math.pi / 180.0
    ph = phi_deg * math.pi / 180.0
    if abs(math.sin(th)) < 1e-9:
        return (0.0, 1.0), (1.0, 0.0)
    return ((-math.sin(ph), math.cos(ph)),
            (math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph)))


def ofd_solve_frequency(nX, period, layers, freq, theta=0.0, phi=0.0,
                        nY=0, period_y=None, poles=None):
    """One iteration of the frequency loop of rcwa_run
    (sol/rcwa_bridge.cpp:244-334), for both polarisations.

    layers: list of dicts with keys eps1, eps2, eps1i, eps2i, fill, fill_y,
    thickness (rcwalayer_t, include/orcwa_rcwa.h:27-35).
    poles: dict {(layer, side): [(einf, ae, be, ce), ...]}.
    Returns dict with R, T (index 0 = TE, 1 = TM), per-order REF/TRN (masked
    as in the CSV), raw REF/TRN, propagation masks and the eigen stats.
    """
    if period_y is None or nY == 0:
        period_y = period          # input_data.c:580-583: 1D uses Ly = Lx
    poles = poles or {}
    lam = C0 / freq
    k0 = 2.0 * math.pi / lam
    omega = 2.0 * math.pi * freq
    xe = common_edges(period, [l["fill"] for l in layers])
    ye = common_edges(period_y, [l.get("fill_y", 1.0) for l in layers])
    nL = len(layers)

    def eps_of(n, side):
        l = layers[n]
        st = complex(l["eps1"], l.get("eps1i", 0.0)) if side == 1 else \
            complex(l["eps2"], l.get("eps2i", 0.0))
        return lorentz_eps(st, poles.get((n, side), []), omega)

    eps_inc = eps_of(0, 1)
    n_inc = max(np.sqrt(complex(eps_inc)).real, 1e-6)
    solver = RCWASolver(nX, nY)
    solver.disable_pml()
    for n in range(nL):
        solver.add_layer(make_ofd_layer(layers[n]["fill"], layers[n].get("fill_y", 1.0),
                                        eps_of(n, 1), eps_of(n, 2), xe, ye))
    th = theta * math.pi / 180.0
    ph = phi * math.pi / 180.0
    kx0 = k0 * n_inc * math.sin(th) * math.cos(ph)
    ky0 = k0 * n_inc * math.sin(th) * math.sin(ph)
    solver.set_bloch_wavevector(kx0, ky0)
    stack = list(range(nL))
    thick = [0.0 if n in (0, nL - 1) else layers[n]["thickness"] for n in range(nL)]
    solver.solve(lam, stack, thick)

    nOX, nOY = 2 * nX + 1, 2 * nY + 1
    kr2 = k0 * k0 * eps_inc.real
    kt2 = k0 * k0 * eps_of(nL - 1, 1).real
    k = np.arange(nOX * nOY)
    m = k // nOY - nX
    nn = k % nOY - nY
    kx = kx0 + 2.0 * math.pi * m / period
    ky = ky0 + 2.0 * math.pi * nn / period_y
    kt = kx * kx + ky * ky
    propR = kt < kr2
    propT = kt < kt2
    te, tm = polarisation_basis(theta, phi)
    R, T, REF, TRN, rawR, rawT = [], [], [], [], [], []
    for (px, py) in (te, tm):
        c = solver.generate_horizontal_plane_wave(px, py, 0)
        ref, trn = solver.scatter_plane_wave(c, 0, nL - 1, k0)
        if ref is None:
            R.append(0.0); T.append(0.0)
            REF.append(np.zeros(nOX * nOY)); TRN.append(np.zeros(nOX * nOY))
            rawR.append(None); rawT.append(None)
            continue
        R.append(float(ref.sum()))
        T.append(float(trn.sum()))
        REF.append(np.where(propR, ref, 0.0))
        TRN.append(np.where(propT, trn, 0.0))
        rawR.append(ref); rawT.append(trn)
    return dict(freq=freq, lam=lam, R=R, T=T, REF=REF, TRN=TRN, rawREF=rawR,
                rawTRN=rawT, propR=propR, propT=propT, m=m, n=nn,
                stats=solver.stats, kx0=kx0, ky0=ky0)


def solve_stack(nX, nY, period_x, period_y, coord_x, coord_y, eps_layers,
                thickness, lam, theta, phi, pols=None, mu_layers=None):
    """Generic one-instance solve used by the parity tests: layer i has eps grid
    eps_layers[i] on the common (coord_x, coord_y) edges, stack = 0..L-1, the
    Bloch vector from the superstrate's cell-(0,0) permittivity as in
    sol/rcwa_bridge.cpp:256-282 (real part of sqrt(eps), clamped at 1e-6).
    pols: list of (px, py) (default TE then TM of sol/rcwa_bridge.cpp:187-196).
    Returns (REF, TRN) lists (one entry per polarisation, unmasked)."""
    k0 = 2 * PI / lam
    solver = RCWASolver(nX, nY)
    solver.disable_pml()
    for i, e in enumerate(eps_layers):
        mu = None if mu_layers is None else mu_layers[i]
        solver.add_layer(Layer(coord_x, coord_y, e, mu))
    eps_inc = complex(np.atleast_2d(eps_layers[0])[0, 0])
    n_inc = max(np.sqrt(eps_inc).real, 1e-6)
    th = theta * math.pi / 180.0
    ph = phi * math.pi / 180.0
    solver.set_bloch_wavevector(k0 * n_inc * math.sin(th) * math.cos(ph),
                                k0 * n_inc * math.sin(th) * math.sin(ph))
    nL = len(eps_layers)
    thick = [0.0 if i in (0, nL - 1) else thickness[i] for i in range(nL)]
    solver.solve(lam, list(range(nL)), thick)
    if pols is None:
        te, tm = polarisation_basis(theta, phi)
        pols = [te, tm]
    REF, TRN = [], []
    for (px, py) in pols:
        c = solver.generate_horizontal_plane_wave(px, py, 0)
        r, t = solver.scatter_plane_wave(c, 0, nL - 1, k0)
        REF.append(r)
        TRN.append(t)
    return REF, TRN, solver
