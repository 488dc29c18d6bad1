// RCWA-specific kernels of the layer-stack solve on gfx950.
//
//   k_conv_strip / k_conv_expand : Li-rule convolution matrices
//       fe11 (DiscontinuousX), fe22 (DiscontinuousY), fe33 (ContinuousXY)
//       rcwa/FourierSolver2D.cpp:12-351, rcwa/FourierSolver1D.cpp:11-41,
//       rcwa/ToeplitzMatrixXcs.cpp:8-44, rcwa/BlockToeplitzMatrixXcs.cpp:5-65
//   k_build_pq   : P*Q of rcwa/Layer.cpp:150-220 + rcwa/WaveEquationSolver.cpp:30
//   k_build_v    : V = Q W diag(1/keff), rcwa/WaveEquationSolver.cpp:33-36
//   k_uniform    : closed-form modes of a homogeneous layer
//   k_phase_scale, k_bdinv_apply, k_sum_diff_t, k_cascade_tail,
//   k_cascade_chain : pieces of the RTCM cascade and interfaces
//       rcwa/RCWASolver.cpp:378-614
//   k_rt         : plane-wave response and per-order efficiencies
//       rcwa/RCWASolver.cpp:1778-1908
//
// Units: every wave-vector is divided by k0 = 2 pi / lambda, so P*Q is
// dimensionless; eigenvectors, V, interface matrices and R/T are unchanged
// by that scaling, and the phase uses keff * (k0 d).
#include <algorithm>

#include "rcwa_kernels.h"

namespace orcwa {

static constexpr double kPi = 3.1415926535897932;  // rcwa/defns.h:25

// exact Fourier coefficient F_m of f on cells [x_{j-1}, x_j] with period d
// (rcwa/FourierSolver1D.cpp:11-41; sinc of rcwa/MathFunction.h:6-9)
__device__ inline cd fourier_coeff(int m, const double* x, const cd* f, int ncell, double d,
                                   int fstride) {
  cd acc = mk(0.0, 0.0);
  for (int j = 1; j <= ncell; ++j) {
    const double x1 = x[j], x0 = x[j - 1];
    const cd fj = f[(j - 1) * fstride];
    if (m == 0) {
      acc += ((x1 - x0) / d) * fj;
    } else {
      const double arg = m * (x1 - x0) / d;
      const double sc = (arg == 0.0) ? 1.0 : sin(kPi * arg) / (kPi * arg);
      double sn, cs;
      sincos(-(double)m * kPi * (x1 + x0) / d, &sn, &cs);
      acc += (sc * (x1 - x0) / d) * (fj * mk(cs, sn));
    }
  }
  return acc;
}

// One workgroup per strip:
//   type 0 (y-strip i, x-direction series): tX = inv(Toeplitz(F_x[1/eps_row])),
//            fy1 = F_y[indicator of strip], ffi = F_x[eps_row]
//   type 1 (x-strip i, y-direction series): tY = inv(Toeplitz(F_y[1/eps_col])),
//            fx1 = F_x[indicator of strip]
__global__ __launch_bounds__(256) void k_conv_strip(const ConvGeom g, const StripTask* tasks,
                                                    const cd* grids, cd* strips) {
  extern __shared__ cd sm[];
  const StripTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x;
  const int nxh = g.nx, nyh = g.ny;
  const int ncx = g.ncx, ncy = g.ncy;
  const cd* eps = grids + (long long)t.grid * g.ncell;
  cd* out = strips + t.out_off;
  const int mdim = (t.type == 0) ? nxh : nyh;
  const int M = 2 * mdim - 1;
  const double* cx = g.cx + (long long)t.grid * g.sx;   // this grid's cell edges
  const double* cy = g.cy + (long long)t.grid * g.sy;
  cd* sF = sm;                 // [M] coefficients of 1/eps
  cd* sT = sm + M;             // [mdim*mdim] Toeplitz -> inverse
  cd* sf = sT + mdim * mdim;   // [max(ncx,ncy)] 1/eps values
  __shared__ int s_piv;
  __shared__ int perm[64];
  // 1/eps along the strip
  if (t.type == 0) {
    for (int j = tid; j < ncx - 1; j += 256) sf[j] = cinv(eps[t.index * (ncx - 1) + j]);
  } else {
    for (int j = tid; j < ncy - 1; j += 256) sf[j] = cinv(eps[j * (ncx - 1) + t.index]);
  }
  __syncthreads();
  for (int i = tid; i < M; i += 256) {
    const int m = i - (M - 1) / 2;
    sF[i] = (t.type == 0) ? fourier_coeff(m, cx, sf, ncx - 1, g.Lx, 1)
                          : fourier_coeff(m, cy, sf, ncy - 1, g.Ly, 1);
  }
  __syncthreads();
  // T[r,c] = F[(r-c) + mdim-1], column-major in LDS
  for (int e = tid; e < mdim * mdim; e += 256) {
    const int r = e % mdim, c = e / mdim;
    sT[e] = sF[(r - c) + mdim - 1];
  }
  __syncthreads();
  // in-place Gauss-Jordan inverse with partial pivoting (rows), columns
  // unscrambled at the end
  for (int k = 0; k < mdim; ++k) {
    if (tid == 0) {
      int p = k;
      double best = cabs(sT[k + k * mdim]);
      for (int i = k + 1; i < mdim; ++i) {
        double v = cabs(sT[i + k * mdim]);
        if (v > best) { best = v; p = i; }
      }
      s_piv = p;
      perm[k] = p;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != k)
      for (int c = tid; c < mdim; c += 256) {
        cd a = sT[k + c * mdim];
        sT[k + c * mdim] = sT[p + c * mdim];
        sT[p + c * mdim] = a;
      }
    __syncthreads();
    const cd piv = sT[k + k * mdim];
    const cd ip = cinv(piv);
    __syncthreads();
    for (int c = tid; c < mdim; c += 256) {
      if (c == k) sT[k + c * mdim] = ip;
      else sT[k + c * mdim] = sT[k + c * mdim] * ip;
    }
    __syncthreads();
    for (int e = tid; e < mdim * mdim; e += 256) {
      const int i = e % mdim, c = e / mdim;
      if (i == k) continue;
      const cd f = sT[i + k * mdim];
      if (c == k) continue;
      sT[e] -= f * sT[k + c * mdim];
    }
    __syncthreads();
    for (int i = tid; i < mdim; i += 256)
      if (i != k) sT[i + k * mdim] = -(sT[i + k * mdim] * ip);
    __syncthreads();
  }
  if (tid == 0) {
    for (int k = mdim - 1; k >= 0; --k) {
      const int p = perm[k];
      if (p != k)
        for (int r = 0; r < mdim; ++r) {
          cd a = sT[r + k * mdim];
          sT[r + k * mdim] = sT[r + p * mdim];
          sT[r + p * mdim] = a;
        }
    }
  }
  __syncthreads();
  for (int e = tid; e < mdim * mdim; e += 256) out[e] = sT[e];
  // indicator series of the strip along the other axis, and (type 0) the
  // series of eps along x
  cd one = mk(1.0, 0.0);
  if (t.type == 0) {
    const int My = 2 * nyh - 1, Mx = 2 * nxh - 1;
    double yy[2] = {cy[t.index], cy[t.index + 1]};
    for (int i = tid; i < My; i += 256)
      out[mdim * mdim + i] = fourier_coeff(i - (My - 1) / 2, yy, &one, 1, g.Ly, 0);
    for (int i = tid; i < Mx; i += 256)
      out[mdim * mdim + My + i] =
          fourier_coeff(i - (Mx - 1) / 2, cx, eps + t.index * (ncx - 1), ncx - 1, g.Lx, 1);
  } else {
    const int Mx = 2 * nxh - 1;
    double xx[2] = {cx[t.index], cx[t.index + 1]};
    for (int i = tid; i < Mx; i += 256)
      out[mdim * mdim + i] = fourier_coeff(i - (Mx - 1) / 2, xx, &one, 1, g.Lx, 0);
  }
}

// fe11, fe22, fe33 (h x h, column-major) of every grid.  Row r = ii*ny+jj,
// column c = kk*ny+ll.
__global__ void k_conv_expand(const ConvGeom g, const cd* strips, const long long* strip_base,
                              cd* const* E11, cd* const* E22, cd* const* E33) {
  const int gi = blockIdx.y;
  const int h = g.nx * g.ny;
  const long long hh = (long long)h * h;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= hh) return;
  const int r = (int)(e % h), c = (int)(e / h);
  const int ii = r / g.ny, jj = r % g.ny, kk = c / g.ny, ll = c % g.ny;
  const int nx = g.nx, ny = g.ny;
  const int My = 2 * ny - 1, Mx = 2 * nx - 1;
  const long long sx = (long long)nx * nx + My + Mx;  // stride of a type-0 strip record
  const long long sy = (long long)ny * ny + Mx;       // stride of a type-1 strip record
  const cd* base0 = strips + strip_base[2 * gi];
  const cd* base1 = strips + strip_base[2 * gi + 1];
  cd f11 = mk(0.0, 0.0), f22 = mk(0.0, 0.0), f33 = mk(0.0, 0.0);
  for (int i = 0; i < g.ncy - 1; ++i) {
    const cd* rec = base0 + i * sx;
    const cd fy = rec[nx * nx + (jj - ll) + ny - 1];
    f11 += fy * rec[ii + kk * nx];
    f33 += rec[nx * nx + My + (ii - kk) + nx - 1] * fy;
  }
  for (int i = 0; i < g.ncx - 1; ++i) {
    const cd* rec = base1 + i * sy;
    f22 += rec[ny * ny + (ii - kk) + nx - 1] * rec[jj + ll * ny];
  }
  E11[gi][e] = f11;
  E22[gi][e] = f22;
  E33[gi][e] = f33;
}

// out[r, c] = d_r * in[r, c] with d_r = 2 pi (order index) / period along
// the chosen axis (Dm for axis 0, Dn for axis 1)
__global__ void k_scale_rows_order(const ConvGeom g, int axis, const cd* const* in,
                                   cd* const* out) {
  const int h = g.nx * g.ny;
  const long long hh = (long long)h * h;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= hh) return;
  const int r = (int)(e % h);
  const double d = (axis == 0) ? 2.0 * kPi * (r / g.ny - (g.nx - 1) / 2) / g.Lx
                               : 2.0 * kPi * (r % g.ny - (g.ny - 1) / 2) / g.Ly;
  out[blockIdx.y][e] = d * in[blockIdx.y][e];
}

__device__ inline void k_vectors(const ConvGeom& g, const ProbParam& p, int r, double& kx,
                                 double& ky) {
  // alpha_k = 2 pi m / Lx + kx0, beta_k = 2 pi n / Ly + ky0 (rcwa/RCWASolver.cpp:242-266),
  // divided by k0
  const int m = r / g.ny - (g.nx - 1) / 2;
  const int n = r % g.ny - (g.ny - 1) / 2;
  kx = (2.0 * kPi * m / g.Lx + p.kx0) * p.inv_k0;
  ky = (2.0 * kPi * n / g.Ly + p.ky0) * p.inv_k0;
}

// P*Q for a patterned non-magnetic layer, in k0 units, from the per-grid
// products S1 = M E11, S2 = M Dm E11, S3 = M E22, S4 = M Dn E22 (M = fe33^-1):
//   A1 = M Kx E11 = (kx0/k0) S1 + S2/k0,  A2 = M Ky E22 = (ky0/k0) S3 + S4/k0
//   PQ = [[E11 - Ky^2 - Kx A1, KxKy - Kx A2], [KxKy - Ky A1, E22 - Kx^2 - Ky A2]]
// which equals P*Q of rcwa/Layer.cpp:150-220 after cancelling the diagonal
// terms (Kx, Ky diagonal).
__global__ void k_build_pq(const ConvGeom g, const ProbParam* params, const GridMats* gm,
                           cd* const* PQ, cd* const* PQcopy) {
  const int pi = blockIdx.y;
  const ProbParam p = params[pi];
  const GridMats G = gm[p.grid];
  const int h = g.nx * g.ny, n = 2 * h;
  const long long hh = (long long)h * h;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= hh) return;
  const int r = (int)(e % h), c = (int)(e / h);
  double kxr, kyr;
  k_vectors(g, p, r, kxr, kyr);
  const double ax = p.kx0 * p.inv_k0, ay = p.ky0 * p.inv_k0;
  const cd a1 = ax * G.S1[e] + p.inv_k0 * G.S2[e];
  const cd a2 = ay * G.S3[e] + p.inv_k0 * G.S4[e];
  cd v11 = G.E11[e] - kxr * a1;
  cd v21 = -(kyr * a1);
  cd v12 = -(kxr * a2);
  cd v22 = G.E22[e] - kyr * a2;
  if (r == c) {
    v11.re -= kyr * kyr;
    v21.re += kxr * kyr;
    v12.re += kxr * kyr;
    v22.re -= kxr * kxr;
  }
  cd* P = PQ[pi];
  cd* P2 = PQcopy[pi];
  const long long i11 = r + (long long)c * n, i21 = (h + r) + (long long)c * n;
  const long long i12 = r + (long long)(h + c) * n, i22 = (h + r) + (long long)(h + c) * n;
  P[i11] = v11; P[i21] = v21; P[i12] = v12; P[i22] = v22;
  P2[i11] = v11; P2[i21] = v21; P2[i12] = v12; P2[i22] = v22;
}

// keff = GammaSqrtOp(lambda) (rcwa/WaveEquationSolver.cpp:32), then
// V[:, c] = (Q W)[:, c] / keff_c with
//   Q = [[-KxKy, Kx^2 - E22], [E11 - Ky^2, KyKx]]  (rcwa/Layer.cpp:214-220, k0 = 1)
// QW1 holds E22 W_bottom (rows 0..h-1) and E11 W_top (rows h..n-1).
__global__ void k_build_v(const ConvGeom g, const ProbParam* params, const cd* const* W,
                          const cd* const* QW1, const cd* lam, cd* keff_all, cd* const* V,
                          int* info, const int* inst_of) {
  const int pi = blockIdx.y;
  const ProbParam p = params[pi];
  const int h = g.nx * g.ny, n = 2 * h;
  const long long nn = (long long)n * n;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const cd kf = gamma_sqrt(lam[(long long)pi * n + c]);
  if (r == 0) {
    keff_all[(long long)pi * n + c] = kf;
    if (kf.re == 0.0 && kf.im == 0.0) atomicOr(&info[inst_of[pi]], ORCWA_STATUS_SINGULAR);
  }
  const int rr = (r < h) ? r : r - h;
  double kx, ky;
  k_vectors(g, p, rr, kx, ky);
  const cd* w = W[pi] + (long long)c * n;
  const cd* q = QW1[pi] + (long long)c * n;
  cd v;
  if (r < h) v = (-(kx * ky)) * w[rr] + (kx * kx) * w[h + rr] - q[rr];
  else v = q[h + rr] - (ky * ky) * w[rr] + (ky * kx) * w[h + rr];
  V[pi][e] = (kf.re == 0.0 && kf.im == 0.0) ? v : cdiv(v, kf);
}

// P and Q of a magnetic patterned layer (rcwa/Layer.cpp:129-145, 150-164,
// 206-213) in k0 units, with F1h = fe33^-1 Kx, F2h = fe33^-1 Ky,
// G1h = fmu33^-1 Kx, G2h = fmu33^-1 Ky (Kx, Ky diagonal):
//   P = [[Kx F2h, fmu22 - Kx F1h], [Ky F2h - fmu11, -Ky F1h]]
//   Q = [[-Kx G2h, Kx G1h - fe22], [fe11 - Ky G2h, Ky G1h]]
__global__ void k_build_pq_mag(const ConvGeom g, const ProbParam* params, const MagMats* mm,
                               cd* const* Pp, cd* const* Qp) {
  const int pi = blockIdx.y;
  const ProbParam p = params[pi];
  const MagMats M = mm[pi];
  const int h = g.nx * g.ny, n = 2 * h;
  const long long hh = (long long)h * h;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= hh) return;
  const int r = (int)(e % h), c = (int)(e / h);
  double kxr, kyr, kxc, kyc;
  k_vectors(g, p, r, kxr, kyr);
  k_vectors(g, p, c, kxc, kyc);
  const cd ei = M.Ei[e], mi = M.Mi[e];
  cd* P = Pp[pi];
  cd* Q = Qp[pi];
  const long long i11 = r + (long long)c * n, i21 = (h + r) + (long long)c * n;
  const long long i12 = r + (long long)(h + c) * n, i22 = (h + r) + (long long)(h + c) * n;
  P[i11] = (kxr * kyc) * ei;
  P[i12] = M.M22[e] - (kxr * kxc) * ei;
  P[i21] = (kyr * kyc) * ei - M.M11[e];
  P[i22] = -((kyr * kxc) * ei);
  Q[i11] = -((kxr * kyc) * mi);
  Q[i12] = (kxr * kxc) * mi - M.E22[e];
  Q[i21] = M.E11[e] - (kyr * kyc) * mi;
  Q[i22] = (kyr * kxc) * mi;
}

// keff = GammaSqrtOp(lambda), V[:, c] = (Q W)[:, c] / keff_c
// (rcwa/WaveEquationSolver.cpp:32-36) from the product Q W of a magnetic layer
__global__ void k_build_v_mag(int n, const cd* const* QW, const cd* lam, cd* keff_all,
                              cd* const* V, int* info, const int* inst_of) {
  const int pi = blockIdx.y;
  const long long nn = (long long)n * n;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= nn) return;
  const int r = (int)(e % n), c = (int)(e / n);
  const cd kf = gamma_sqrt(lam[(long long)pi * n + c]);
  if (r == 0) {
    keff_all[(long long)pi * n + c] = kf;
    if (kf.re == 0.0 && kf.im == 0.0) atomicOr(&info[inst_of[pi]], ORCWA_STATUS_SINGULAR);
  }
  const cd v = QW[pi][e];
  V[pi][e] = (kf.re == 0.0 && kf.im == 0.0) ? v : cdiv(v, kf);
}

__global__ void k_zero_mats(long long count, cd* const* A) {
  cd* a = A[blockIdx.y];
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < count;
       e += (long long)gridDim.x * blockDim.x)
    a[e] = mk(0.0, 0.0);
}

// Homogeneous layer (eps, mu): W = I, keff = GammaSqrtOp(eps mu - kx^2 - ky^2)
// for both field components of every order (P Q = (eps mu - kx^2 - ky^2) I),
// V = Q diag(1/keff) with Q = [[-kx ky/mu, kx^2/mu - eps], [eps - ky^2/mu,
// kx ky/mu]] per order, 2x2-block diagonal.  Writes keff (n), the non-zero entries of the
// dense V (zeroed beforehand by k_zero_mats) and the inverse 2x2 blocks
// (4h: b00, b01, b10, b11 of V^{-1}).
__global__ void k_uniform(const ConvGeom g, const UniParam* params, cd* keff_all,
                          cd* const* Vdense, cd* binv_all, int* info) {
  const int ui = blockIdx.y;
  const UniParam u = params[ui];
  const int h = g.nx * g.ny, n = 2 * h;
  const long long nn = (long long)n * n;
  cd* V = Vdense[ui];
  (void)nn;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= h) return;
  ProbParam p;
  p.kx0 = u.kx0; p.ky0 = u.ky0; p.inv_k0 = u.inv_k0; p.grid = 0;
  double kx, ky;
  k_vectors(g, p, k, kx, ky);
  const cd kz2 = u.eps * u.mu - mk(kx * kx + ky * ky, 0.0);
  const cd kf = gamma_sqrt(kz2);
  keff_all[(long long)ui * n + k] = kf;
  keff_all[(long long)ui * n + h + k] = kf;
  if (kf.re == 0.0 && kf.im == 0.0) {
    atomicOr(&info[u.inst], ORCWA_STATUS_SINGULAR);
    return;
  }
  const cd ik = cinv(kf);
  const cd imu = cinv(u.mu);
  const cd v00 = (-(kx * ky)) * imu * ik;
  const cd v01 = ((kx * kx) * imu - u.eps) * ik;
  const cd v10 = (u.eps - (ky * ky) * imu) * ik;
  const cd v11 = (ky * kx) * imu * ik;
  V[k + (long long)k * n] = v00;
  V[k + (long long)(h + k) * n] = v01;
  V[(h + k) + (long long)k * n] = v10;
  V[(h + k) + (long long)(h + k) * n] = v11;
  const cd det = v00 * v11 - v01 * v10;
  cd* bi = binv_all + (long long)ui * 4 * h;
  if (det.re == 0.0 && det.im == 0.0) {
    atomicOr(&info[u.inst], ORCWA_STATUS_SINGULAR);
    bi[k] = bi[h + k] = bi[2 * h + k] = bi[3 * h + k] = mk(0.0, 0.0);
    return;
  }
  const cd idet = cinv(det);
  bi[k] = v11 * idet;
  bi[h + k] = -(v01 * idet);
  bi[2 * h + k] = -(v10 * idet);
  bi[3 * h + k] = v00 * idet;
}

// H = V_a^{-1} V_b with V_a 2x2-block diagonal (inverse blocks binv)
__global__ void k_bdinv_apply(int h, const cd* const* binv, const cd* const* Vb, cd* const* Hout) {
  const int b = blockIdx.y;
  const int n = 2 * h;
  const long long cnt = (long long)h * n;
  const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (e >= cnt) return;
  const int k = (int)(e % h), c = (int)(e / h);
  const cd* bi = binv[b];
  const cd* v = Vb[b] + (long long)c * n;
  const cd x0 = v[k], x1 = v[h + k];
  cd* o = Hout[b] + (long long)c * n;
  o[k] = bi[k] * x0 + bi[h + k] * x1;
  o[h + k] = bi[2 * h + k] * x0 + bi[3 * h + k] * x1;
}

// R <- Phi R Phi, T <- T Phi with phi = exp(i keff (k0 d))
// (rcwa/RCWASolver.cpp:415-434)
// A mode that grows by more than this factor across one layer (Im keff < 0
// under the GammaSqrtOp branch of rcwa/WaveEquationSolver.cpp:13-21) makes
// the cascade amplify rounding by about its square: ORCWA_STATUS_ILLCOND.
__device__ constexpr double kGrowthLog = 9.2103403719761836;   // ln(1e4)

__global__ void k_phase_scale(int n, const cd* const* keff, const double* k0d, cd* const* R,
                              cd* const* T, int* info) {
  const int b = blockIdx.y;
  const long long nn = (long long)n * n;
  const double kd = k0d[b];
  const cd* kf = keff[b];
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nn;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    const cd ar = kd * kf[r], ac = kd * kf[c];
    if (c == 0 && info && -ar.im > kGrowthLog) atomicOr(&info[b], ORCWA_STATUS_ILLCOND);
    const cd pr = cexp(mk(-ar.im, ar.re));
    const cd pc = cexp(mk(-ac.im, ac.re));
    R[b][e] = pr * R[b][e] * pc;
    if (T) T[b][e] = T[b][e] * pc;
  }
}

// ---------------------------------------------------------------------------
// RTCM cascade pieces (rcwa/RCWASolver.cpp:378-492), rewritten so that no
// explicit inverse is formed:
//   with A = E + H, B = E - H and R~ = phi R phi,
//   F + G = A + B R~,  F - G = B + A R~,
//   R_new = I - 2 G tau = (F - G)(F + G)^-1,      tau = (F + G)^-1,
//   T_new = 2 (T phi) tau.
// R_new is obtained from the transposed system (F+G)^T R_new^T = (F-G)^T, so
// the cascade keeps R^T; its LU factors M^T = P L U (M = F + G) are kept per
// step, and T enters only through two columns, carried right to left:
//   Tdd e_c = 2 tau_0 (phi_1 2 tau_1 (... (phi_last 2 tau_last e_c))).
// ---------------------------------------------------------------------------
// A = E + D H, B = E - D H with D = diag(1/keff_a) where the instance's
// upper layer is patterned (H holds W_a^-1 P_a V_b there) and D = I otherwise;
// A and B are written in place and also transposed, A^T into At and B^T into
// Bt (when Bt is given): the cascade's next operands, in one pass over E and H
// (32 x 32 tiles through LDS).
__global__ void k_sum_diff_t(int n, cd* const* Ep, cd* const* Hp, const cd* const* keff_a,
                             cd* const* Atp, cd* const* Btp) {
  __shared__ cd ta[32][33], tb[32][33];
  const int z = blockIdx.z;
  cd* E = Ep[z];
  cd* H = Hp[z];
  const cd* kf = keff_a[z];
  cd* At = Atp[z];
  cd* Bt = Btp ? Btp[z] : nullptr;
  const int r0 = blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 256 threads: 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = r0 + tx, c = c0 + k;
    if (r < n && c < n) {
      const long long e = r + (long long)c * n;
      const cd a = E[e];
      cd h = H[e];
      if (kf) h = cdiv(h, kf[r]);
      const cd sa = a + h, sb = a - h;
      E[e] = sa;
      H[e] = sb;
      ta[k][tx] = sa;
      tb[k][tx] = sb;
    }
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = c0 + tx, c = r0 + k;   // At(r, c) = A(c, r)
    if (r < n && c < n) {
      At[r + (long long)c * n] = ta[tx][k];
      if (Bt) Bt[r + (long long)c * n] = tb[tx][k];
    }
  }
}

// Interface H = V_a^-1 V_b for a patterned upper layer a without inverting V_a:
// P_a Q_a W_a = W_a K_a^2 and V_a = Q_a W_a K_a^-1 give V_a^-1 = K_a^-1 W_a^-1 P_a,
// and with P = [[Kx M Ky, I - Kx M Kx], [Ky M Ky - I, -Ky M Kx]] (M = fe33^-1,
// k0 units, rcwa/Layer.cpp:150-169):
//   P V = [Kx U + V2; Ky U - V1],  U = M (Ky V1 - Kx V2),  V = [V1; V2].
// k_pv_in writes Ky V1 - Kx V2 (h x n, ld h); k_pv_out forms P V from U.
__global__ void k_pv_in(const ConvGeom g, const ProbParam* params, const cd* const* Vb,
                        cd* const* out) {
  const int b = blockIdx.y;
  const ProbParam p = params[b];
  const int h = g.nx * g.ny, n = 2 * h;
  const cd* V = Vb[b];
  cd* O = out[b];
  const long long cnt = (long long)h * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < cnt;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e % h);
    const long long col = e / h;
    double kx, ky;
    k_vectors(g, p, r, kx, ky);
    O[e] = ky * V[r + col * n] - kx * V[h + r + col * n];
  }
}

__global__ void k_pv_out(const ConvGeom g, const ProbParam* params, const cd* const* Up,
                         long long offU, const cd* const* Vb, cd* const* Hp) {
  const int b = blockIdx.y;
  const ProbParam p = params[b];
  const int h = g.nx * g.ny, n = 2 * h;
  const cd* U = Up[b] + offU;
  const cd* V = Vb[b];
  cd* H = Hp[b];
  const long long cnt = (long long)n * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < cnt;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n);
    const long long col = e / n;
    const int rr = r < h ? r : r - h;
    double kx, ky;
    k_vectors(g, p, rr, kx, ky);
    const cd u = U[rr + col * h];
    H[e] = (r < h) ? kx * u + V[h + rr + col * n] : ky * u - V[rr + col * n];
  }
}

// x <- M^-1 x for the two vectors x0, x1 (LDS), given zgetrf of M^T:
// M^T = P L U  =>  M = U^T L^T P^T  =>  x = P L^-T U^-T x.
// Solved by the whole 256-thread workgroup, in blocks of TS rows:
// per block the four waves first subtract the contributions of the entries
// already solved (coalesced column reads of LU, one wave per row), then wave
// 0 finishes the TS x TS triangle from LDS with one lane per row, each solved
// entry broadcast by a lane shuffle.  2 n / TS barriers instead of a global
// load latency per row.  Every thread of the workgroup must call it.
constexpr int TS = 32;
__device__ void lu_t_solve2(int n, const cd* __restrict__ LU,
                            const int* __restrict__ piv, cd* x0, cd* x1,
                            cd (*blk)[TS + 1]) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nblk = (n + TS - 1) / TS;
  // U^T y = x, forward: y_i = (x_i - sum_{j<i} U(j,i) y_j) / U(i,i)
  for (int bk = 0; bk < nblk; ++bk) {
    const int r0 = bk * TS, nb = min(TS, n - r0);
    for (int e = tid; e < nb * nb; e += 256) {
      const int a = e % nb, c = e / nb;
      blk[a][c] = LU[(r0 + a) + (long long)(r0 + c) * n];
    }
    for (int r = r0 + wave; r < r0 + nb; r += 4) {
      const cd* col = LU + (long long)r * n;
      double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
      for (int j = lane; j < r0; j += 64) {
        const cd u = col[j], y0 = x0[j], y1 = x1[j];
        a0r += u.re * y0.re - u.im * y0.im;
        a0i += u.re * y0.im + u.im * y0.re;
        a1r += u.re * y1.re - u.im * y1.im;
        a1i += u.re * y1.im + u.im * y1.re;
      }
      a0r = wave_sum(a0r);
      a0i = wave_sum(a0i);
      a1r = wave_sum(a1r);
      a1i = wave_sum(a1i);
      if (lane == 0) {
        x0[r] -= mk(a0r, a0i);
        x1[r] -= mk(a1r, a1i);
      }
    }
    __syncthreads();
    if (wave == 0) {
      cd v0 = mk(0.0, 0.0), v1 = mk(0.0, 0.0);
      if (lane < nb) {
        v0 = x0[r0 + lane];
        v1 = x1[r0 + lane];
      }
      for (int k = 0; k < nb; ++k) {
        if (lane == k) {
          const cd d = cinv(blk[k][k]);
          v0 = v0 * d;
          v1 = v1 * d;
        }
        const cd y0 = mk(__shfl(v0.re, k, 64), __shfl(v0.im, k, 64));
        const cd y1 = mk(__shfl(v1.re, k, 64), __shfl(v1.im, k, 64));
        if (lane > k && lane < nb) {
          const cd u = blk[k][lane];   // U(r0 + k, r0 + lane)
          v0 -= u * y0;
          v1 -= u * y1;
        }
      }
      if (lane < nb) {
        x0[r0 + lane] = v0;
        x1[r0 + lane] = v1;
      }
    }
    __syncthreads();
  }
  // L^T z = y, backward, unit diagonal: z_i = y_i - sum_{j>i} L(j,i) z_j
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int r0 = bk * TS, nb = min(TS, n - r0), r1 = r0 + nb;
    for (int e = tid; e < nb * nb; e += 256) {
      const int a = e % nb, c = e / nb;
      blk[a][c] = LU[(r0 + a) + (long long)(r0 + c) * n];
    }
    for (int r = r0 + wave; r < r1; r += 4) {
      const cd* col = LU + (long long)r * n;
      double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
      for (int j = r1 + lane; j < n; j += 64) {
        const cd l = col[j], z0 = x0[j], z1 = x1[j];
        a0r += l.re * z0.re - l.im * z0.im;
        a0i += l.re * z0.im + l.im * z0.re;
        a1r += l.re * z1.re - l.im * z1.im;
        a1i += l.re * z1.im + l.im * z1.re;
      }
      a0r = wave_sum(a0r);
      a0i = wave_sum(a0i);
      a1r = wave_sum(a1r);
      a1i = wave_sum(a1i);
      if (lane == 0) {
        x0[r] -= mk(a0r, a0i);
        x1[r] -= mk(a1r, a1i);
      }
    }
    __syncthreads();
    if (wave == 0) {
      cd v0 = mk(0.0, 0.0), v1 = mk(0.0, 0.0);
      if (lane < nb) {
        v0 = x0[r0 + lane];
        v1 = x1[r0 + lane];
      }
      for (int k = nb - 1; k >= 0; --k) {
        const cd z0 = mk(__shfl(v0.re, k, 64), __shfl(v0.im, k, 64));
        const cd z1 = mk(__shfl(v1.re, k, 64), __shfl(v1.im, k, 64));
        if (lane < k) {
          const cd l = blk[k][lane];   // L(r0 + k, r0 + lane)
          v0 -= l * z0;
          v1 -= l * z1;
        }
      }
      if (lane < nb) {
        x0[r0 + lane] = v0;
        x1[r0 + lane] = v1;
      }
    }
    __syncthreads();
  }
  // x = P z: the interchanges of zgetrf in reverse order
  if (tid == 0) {
    for (int k = n - 1; k >= 0; --k) {
      const int p = piv[k];
      if (p != k) {
        cd t = x0[k]; x0[k] = x0[p]; x0[p] = t;
        t = x1[k]; x1[k] = x1[p]; x1[p] = t;
      }
    }
  }
  __syncthreads();
}

// Last cascade step (the interface at the incident half-space): for
// c in {i0, h+i0}: x = (F+G)^-1 e_c, Rud e_c = (F-G) x = B x + A (R~ x),
// 2 x starts the Tdd column chain.  Rt = R~^T of the previous steps, or null
// for a two-layer stack.  One 256-thread workgroup per instance, LDS 4n cd.
__global__ __launch_bounds__(256) void k_cascade_tail(int n, int i0, const cd* const* LUt,
                                                      const int* piv_all, const cd* const* Ap,
                                                      const cd* const* Bp, const cd* const* Rtp,
                                                      cd* rt_all) {
  extern __shared__ cd sm[];
  cd* x0 = sm;
  cd* x1 = sm + n;
  cd* u0 = sm + 2 * n;
  cd* u1 = sm + 3 * n;
  const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int h = n / 2;
  for (int i = tid; i < n; i += 256) {
    x0[i] = mk(i == i0 ? 1.0 : 0.0, 0.0);
    x1[i] = mk(i == h + i0 ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  // (the triangle blocks of the solve live where u0, u1 go afterwards)
  lu_t_solve2(n, LUt[b], piv_all + (long long)b * n, x0, x1,
              reinterpret_cast<cd(*)[TS + 1]>(u0));
  // u = R~ x, R~(i, j) = Rt(j, i): column i of Rt
  const cd* Rt = Rtp ? Rtp[b] : nullptr;
  for (int i = wave; i < n; i += 4) {
    double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
    if (Rt) {
      const cd* col = Rt + (long long)i * n;
      for (int j = lane; j < n; j += 64) {
        const cd r = col[j], y0 = x0[j], y1 = x1[j];
        a0r += r.re * y0.re - r.im * y0.im;
        a0i += r.re * y0.im + r.im * y0.re;
        a1r += r.re * y1.re - r.im * y1.im;
        a1i += r.re * y1.im + r.im * y1.re;
      }
      a0r = wave_sum(a0r);
      a0i = wave_sum(a0i);
      a1r = wave_sum(a1r);
      a1i = wave_sum(a1i);
    }
    if (lane == 0) {
      u0[i] = mk(a0r, a0i);
      u1[i] = mk(a1r, a1i);
    }
  }
  __syncthreads();
  const cd* A = Ap[b];
  const cd* B = Bp[b];
  cd* out = rt_all + (long long)b * 4 * n;
  for (int i = tid; i < n; i += 256) {
    cd r0 = mk(0.0, 0.0), r1 = mk(0.0, 0.0);
    for (int j = 0; j < n; ++j) {
      const cd bij = B[i + (long long)j * n], aij = A[i + (long long)j * n];
      r0 += bij * x0[j] + aij * u0[j];
      r1 += bij * x1[j] + aij * u1[j];
    }
    out[i] = r0;
    out[n + i] = r1;
    out[2 * n + i] = 2.0 * x0[i];
    out[3 * n + i] = 2.0 * x1[i];
  }
}

// One link of the Tdd column chain: t <- 2 (F+G)_s^-1 (phi_{s+1} t), with
// phi = exp(i keff (k0 d)) of the layer propagated at step s+1.
__global__ __launch_bounds__(256) void k_cascade_chain(int n, const cd* const* LUt,
                                                       const int* piv_all, const cd* const* keff,
                                                       const double* k0d, cd* rt_all) {
  extern __shared__ cd sm[];
  __shared__ cd blk[TS][TS + 1];
  cd* x0 = sm;
  cd* x1 = sm + n;
  const int b = blockIdx.x, tid = threadIdx.x;
  cd* t = rt_all + (long long)b * 4 * n + 2 * n;
  const cd* kf = keff[b];
  const double kd = k0d[b];
  for (int i = tid; i < n; i += 256) {
    const cd a = kd * kf[i];
    const cd ph = cexp(mk(-a.im, a.re));
    x0[i] = t[i] * ph;
    x1[i] = t[n + i] * ph;
  }
  __syncthreads();
  lu_t_solve2(n, LUt[b], piv_all + (long long)b * n, x0, x1, blk);
  for (int i = tid; i < n; i += 256) {
    t[i] = 2.0 * x0[i];
    t[n + i] = 2.0 * x1[i];
  }
}

// Plane-wave response and efficiencies (rcwa/RCWASolver.cpp:1778-1908) for a
// stack whose first and last layers are homogeneous (W = I there): the
// incident coefficient vector is s = [px delta; py delta], r = Rud s,
// t = Tdd s, so only columns i0 and h+i0 of Rud and Tdd enter; they arrive as
// [4][n] per instance (Rud col i0, Rud col h+i0, Tdd col i0, Tdd col h+i0).
// One workgroup per instance.
__global__ __launch_bounds__(256) void k_rt(const ConvGeom g, const RtParam* params,
                                            const cd* rt_all, double* ref_out,
                                            double* trn_out, cd* amp_out, int* info,
                                            int max_pol) {
  const int b = blockIdx.x;
  const RtParam p = params[b];
  const int h = g.nx * g.ny, n = 2 * h;
  const int i0 = g.ny * (g.nx - 1) / 2 + (g.ny - 1) / 2;
  // columns i0 and h+i0 of Rud and Tdd
  const cd* R0 = rt_all + (long long)b * 4 * n;
  const cd* R1 = R0 + n;
  const cd* T0 = R0 + 2 * n;
  const cd* T1 = R0 + 3 * n;
  cd* amp = amp_out + (long long)b * 8 * h;
  __shared__ int s_graze;
  if (threadIdx.x == 0) s_graze = 0;
  __syncthreads();
  ProbParam pp;
  pp.kx0 = p.kx0; pp.ky0 = p.ky0; pp.inv_k0 = p.inv_k0; pp.grid = 0;
  double kzinc = 0.0;
  {
    double kx, ky;
    k_vectors(g, pp, i0, kx, ky);
    const cd kzi = -csqrt_principal(p.eps_ref * p.mu_ref - mk(kx * kx + ky * ky, 0.0));
    kzinc = fabs(kzi.re);
  }
  for (int k = threadIdx.x; k < h; k += blockDim.x) {
    // basis responses: column i0 (x-polarised) and h+i0 (y-polarised)
    amp[0 * h + k] = R0[k];
    amp[1 * h + k] = R0[h + k];
    amp[2 * h + k] = T0[k];
    amp[3 * h + k] = T0[h + k];
    amp[4 * h + k] = R1[k];
    amp[5 * h + k] = R1[h + k];
    amp[6 * h + k] = T1[k];
    amp[7 * h + k] = T1[h + k];
    double kx, ky;
    k_vectors(g, pp, k, kx, ky);
    const cd kzr = -csqrt_principal(p.eps_ref * p.mu_ref - mk(kx * kx + ky * ky, 0.0));
    const cd kzt = csqrt_principal(p.eps_trn * p.mu_trn - mk(kx * kx + ky * ky, 0.0));
    if ((kzr.re == 0.0 && kzr.im == 0.0) || (kzt.re == 0.0 && kzt.im == 0.0)) s_graze = 1;
  }
  __syncthreads();
  const bool graze = s_graze || (kzinc < 1e-12);
  // |mu_ref / mu_trn| (rcwa/RCWASolver.cpp:1903-1907)
  const double mu_ratio = sqrt(norm2(p.mu_ref) / norm2(p.mu_trn));
  if (threadIdx.x == 0 && graze) info[b] |= ORCWA_STATUS_GRAZING;
  for (int pol = 0; pol < max_pol; ++pol) {
    double* ro = ref_out + ((long long)b * max_pol + pol) * h;
    double* to = trn_out + ((long long)b * max_pol + pol) * h;
    if (pol >= p.npol) continue;
    const cd px = p.pol[pol][0], py = p.pol[pol][1];
    for (int k = threadIdx.x; k < h; k += blockDim.x) {
      if (graze) {
        ro[k] = 0.0;
        to[k] = 0.0;
        continue;
      }
      const cd rx = px * R0[k] + py * R1[k];
      const cd ry = px * R0[h + k] + py * R1[h + k];
      const cd tx = px * T0[k] + py * T1[k];
      const cd ty = px * T0[h + k] + py * T1[h + k];
      double kx, ky;
      k_vectors(g, pp, k, kx, ky);
      const cd kzr = -csqrt_principal(p.eps_ref * p.mu_ref - mk(kx * kx + ky * ky, 0.0));
      const cd kzt = csqrt_principal(p.eps_trn * p.mu_trn - mk(kx * kx + ky * ky, 0.0));
      const cd rz = cdiv(kx * rx + ky * ry, kzr);
      const cd tz = cdiv(kx * tx + ky * ty, kzt);
      const double rr = norm2(rx) + norm2(ry) + norm2(rz);
      const double tt = norm2(tx) + norm2(ty) + norm2(tz);
      ro[k] = fabs(kzr.re) * rr / kzinc;
      to[k] = mu_ratio * fabs(kzt.re) * tt / kzinc;
    }
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(long long count, int per = 256) {
  long long b = (count + per - 1) / per;
  return (unsigned)(b < 1 ? 1 : b);
}

void launch_conv_strip(hipStream_t s, const ConvGeom& g, const StripTask* tasks, int ntasks,
                       const cd* grids, cd* strips, size_t lds_bytes) {
  if (ntasks <= 0) return;
  hipLaunchKernelGGL(k_conv_strip, dim3(ntasks), dim3(256), lds_bytes, s, g, tasks, grids, strips);
}

void launch_conv_expand(hipStream_t s, const ConvGeom& g, const cd* strips,
                        const long long* strip_base, cd* const* E11, cd* const* E22,
                        cd* const* E33, int ngrid) {
  if (ngrid <= 0) return;
  const long long hh = (long long)g.nx * g.ny * g.nx * g.ny;
  hipLaunchKernelGGL(k_conv_expand, dim3(blocks_for(hh), ngrid), dim3(256), 0, s, g, strips,
                     strip_base, E11, E22, E33);
}

void launch_scale_rows_order(hipStream_t s, const ConvGeom& g, int axis, const cd* const* in,
                             cd* const* out, int ngrid) {
  if (ngrid <= 0) return;
  const long long hh = (long long)g.nx * g.ny * g.nx * g.ny;
  hipLaunchKernelGGL(k_scale_rows_order, dim3(blocks_for(hh), ngrid), dim3(256), 0, s, g, axis,
                     in, out);
}

void launch_build_pq(hipStream_t s, const ConvGeom& g, const ProbParam* params, const GridMats* gm,
                     cd* const* PQ, cd* const* PQcopy, int nprob) {
  if (nprob <= 0) return;
  const long long hh = (long long)g.nx * g.ny * g.nx * g.ny;
  hipLaunchKernelGGL(k_build_pq, dim3(blocks_for(hh), nprob), dim3(256), 0, s, g, params, gm, PQ,
                     PQcopy);
}

void launch_build_v(hipStream_t s, const ConvGeom& g, const ProbParam* params, const cd* const* W,
                    const cd* const* QW1, const cd* lam, cd* keff, cd* const* V, int* info,
                    const int* inst_of, int nprob) {
  if (nprob <= 0) return;
  const long long n = 2LL * g.nx * g.ny;
  hipLaunchKernelGGL(k_build_v, dim3(blocks_for(n * n), nprob), dim3(256), 0, s, g, params, W,
                     QW1, lam, keff, V, info, inst_of);
}

void launch_build_pq_mag(hipStream_t s, const ConvGeom& g, const ProbParam* params,
                         const MagMats* mm, cd* const* P, cd* const* Q, int nprob) {
  if (nprob <= 0) return;
  const long long hh = (long long)g.nx * g.ny * g.nx * g.ny;
  hipLaunchKernelGGL(k_build_pq_mag, dim3(blocks_for(hh), nprob), dim3(256), 0, s, g, params, mm,
                     P, Q);
}

void launch_build_v_mag(hipStream_t s, int n, const cd* const* QW, const cd* lam, cd* keff,
                        cd* const* V, int* info, const int* inst_of, int nprob) {
  if (nprob <= 0) return;
  const long long nn = (long long)n * n;
  hipLaunchKernelGGL(k_build_v_mag, dim3(blocks_for(nn), nprob), dim3(256), 0, s, n, QW, lam,
                     keff, V, info, inst_of);
}

void launch_zero_mats(hipStream_t s, long long count, cd* const* A, int batch) {
  if (batch <= 0) return;
  long long b = blocks_for(count);
  hipLaunchKernelGGL(k_zero_mats, dim3((unsigned)(b > 2048 ? 2048 : b), batch), dim3(256), 0, s,
                     count, A);
}

void launch_uniform(hipStream_t s, const ConvGeom& g, const UniParam* params, cd* keff,
                    cd* const* V, cd* binv, int* info, int nuni) {
  if (nuni <= 0) return;
  const long long h = (long long)g.nx * g.ny;
  launch_zero_mats(s, 4 * h * h, V, nuni);
  hipLaunchKernelGGL(k_uniform, dim3(blocks_for(h), nuni), dim3(256), 0, s, g, params, keff, V,
                     binv, info);
}

void launch_bdinv_apply(hipStream_t s, int h, const cd* const* binv, const cd* const* Vb,
                        cd* const* H, int batch) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(k_bdinv_apply, dim3(blocks_for(2LL * h * h), batch), dim3(256), 0, s, h,
                     binv, Vb, H);
}

void launch_phase_scale(hipStream_t s, int n, const cd* const* keff, const double* k0d,
                        cd* const* R, cd* const* T, int* info, int batch) {
  if (batch <= 0) return;
  long long b = blocks_for((long long)n * n);
  hipLaunchKernelGGL(k_phase_scale, dim3((unsigned)(b > 2048 ? 2048 : b), batch), dim3(256), 0, s,
                     n, keff, k0d, R, T, info);
}

void launch_rt(hipStream_t s, const ConvGeom& g, const RtParam* params, const cd* rt,
               double* ref, double* trn, cd* amp, int* info, int max_pol, int batch) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(k_rt, dim3(batch), dim3(256), 0, s, g, params, rt, ref, trn, amp, info,
                     max_pol);
}

void launch_sum_diff_t(hipStream_t s, int n, cd* const* E, cd* const* H,
                       const cd* const* keff_a, cd* const* At, cd* const* Bt, int batch) {
  if (batch <= 0) return;
  const unsigned t = (unsigned)((n + 31) / 32);
  hipLaunchKernelGGL(k_sum_diff_t, dim3(t, t, batch), dim3(256), 0, s, n, E, H, keff_a, At, Bt);
}

void launch_pv_in(hipStream_t s, const ConvGeom& g, const ProbParam* params, const cd* const* Vb,
                  cd* const* out, int batch) {
  if (batch <= 0) return;
  const int h = g.nx * g.ny;
  long long b = blocks_for(2LL * h * h);
  hipLaunchKernelGGL(k_pv_in, dim3((unsigned)(b > 2048 ? 2048 : b), batch), dim3(256), 0, s, g,
                     params, Vb, out);
}

void launch_pv_out(hipStream_t s, const ConvGeom& g, const ProbParam* params, const cd* const* U,
                   long long offU, const cd* const* Vb, cd* const* H, int batch) {
  if (batch <= 0) return;
  const int h = g.nx * g.ny;
  long long b = blocks_for(4LL * h * h);
  hipLaunchKernelGGL(k_pv_out, dim3((unsigned)(b > 2048 ? 2048 : b), batch), dim3(256), 0, s, g,
                     params, U, offU, Vb, H);
}

void launch_cascade_tail(hipStream_t s, int n, int i0, const cd* const* LUt, const int* piv,
                         const cd* const* A, const cd* const* B, const cd* const* Rt, cd* rt,
                         int batch) {
  if (batch <= 0) return;
  // x0, x1, u0, u1; u0 onwards also holds the TS x (TS + 1) blocks of the solve
  const size_t lds = (size_t)std::max(4 * n, 2 * n + TS * (TS + 1)) * sizeof(cd);
  hipLaunchKernelGGL(k_cascade_tail, dim3(batch), dim3(256), lds, s, n, i0, LUt, piv, A, B, Rt,
                     rt);
}

void launch_cascade_chain(hipStream_t s, int n, const cd* const* LUt, const int* piv,
                          const cd* const* keff, const double* k0d, cd* rt, int batch) {
  if (batch <= 0) return;
  hipLaunchKernelGGL(k_cascade_chain, dim3(batch), dim3(256), 2 * n * sizeof(cd), s, n, LUt, piv,
                     keff, k0d, rt);
}

}  // namespace orcwa
