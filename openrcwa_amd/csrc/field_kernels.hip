// Field reconstruction inside the stack (RCWASolver::saveFieldImage with
// evaluateIntermediateField, rcwa/RCWASolver.cpp:707-1011, 1505-1645).
//
// Mode amplitudes per interior layer j: c_p[j] (downward, at the layer top)
// comes from the Tdd column chain of the cascade; c_m[j] (upward, at the
// layer bottom) from the interface relation
//   c_m[j] = (B c_p[j+1] + A phi_{j+1} c_m[j+1]) / 2,   A = E + H, B = E - H,
// walked from the exit side (c_m = 0 in the exit half-space), where
// phi = exp(i keff k0 d) only ever damps.  Fields at depth z_p below the top
// (z_m = z_p - d above the bottom):
//   [Ex; Ey] = W (u + v), [Hx; Hy] = V (u - v),
//   u = c_p exp(i keff k0 z_p), v = c_m exp(-i keff k0 z_m),
//   Hz = (i / k0^2)(Kx Ey - Ky Ex), Ez = i [[eps]]^-1 (Kx Hy - Ky Hx)
// (rcwa/RCWASolver.cpp:849-891), then the inverse Fourier synthesis
//   f(x, y) = sum_mn f_mn exp(i (kx0 + 2 pi m / Lx) x + i (ky0 + 2 pi n / Ly) y)
// on the slice (:898-1008).
#include <algorithm>

#include "rcwa_kernels.h"

namespace orcwa {

static constexpr double kPi = 3.1415926535897932;  // rcwa/defns.h:25
static constexpr int kTile = 1024;                  // vector entries staged in LDS at a time

// cp[l] = px t_x[l] + py t_y[l] for layers 1..L-1 (t: [L][2][n])
__global__ void k_field_combine(int n, int L, const cd* tv, cd px, cd py, cd* cp) {
  const long long cnt = (long long)L * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < cnt;
       e += (long long)gridDim.x * blockDim.x) {
    const long long l = e / n, k = e % n;
    cp[e] = (l == 0) ? mk(0.0, 0.0)
                     : px * tv[(l * 2 + 0) * n + k] + py * tv[(l * 2 + 1) * n + k];
  }
}

// cm = (B cp_next + A (phi cm_next)) / 2 with phi = exp(i keff_next k0d_next);
// thread i owns row i, the vectors are staged through LDS tile by tile
__global__ __launch_bounds__(256) void k_field_up(int n, const cd* A, const cd* B, const cd* cp1,
                                                  const cd* cm1, const cd* keff1, double k0d1,
                                                  cd* cm) {
  __shared__ cd s_x[256], s_y[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  cd acc = mk(0.0, 0.0);
  for (int j0 = 0; j0 < n; j0 += 256) {
    const int j = j0 + threadIdx.x;
    if (j < n) {
      const cd a = k0d1 * keff1[j];
      s_x[threadIdx.x] = cp1[j];
      s_y[threadIdx.x] = cm1[j] * cexp(mk(-a.im, a.re));
    }
    __syncthreads();
    const int jn = min(256, n - j0);
    if (i < n)
      for (int jj = 0; jj < jn; ++jj) {
        const long long col = (long long)(j0 + jj) * n;
        acc += B[i + col] * s_x[jj] + A[i + col] * s_y[jj];
      }
    __syncthreads();
  }
  if (i < n) cm[i] = 0.5 * acc;
}

// X[r] = M_l (u +- v) for image row r in layer l = rows[r].layer (all n rows
// of M: W for E-type components, V for H-type)
__global__ __launch_bounds__(256) void k_field_modes(int n, const FieldLayer* layers,
                                                     const FieldRow* rows, const cd* cp,
                                                     const cd* cm, int sign, cd* X) {
  __shared__ cd s_a[kTile];
  const int r = blockIdx.x;
  const FieldRow fr = rows[r];
  const FieldLayer fl = layers[fr.layer];
  const cd* up = cp + (long long)fr.layer * n;
  const cd* dn = cm + (long long)fr.layer * n;
  cd acc[8];
  for (int q = 0; q < 8; ++q) acc[q] = mk(0.0, 0.0);
  for (int k0 = 0; k0 < n; k0 += kTile) {
    const int kn = min(kTile, n - k0);
    for (int k = threadIdx.x; k < kn; k += 256) {
      const cd kf = fl.keff[k0 + k];
      const cd a = fr.k0zp * kf, b = fr.k0zm * kf;
      const cd u = up[k0 + k] * cexp(mk(-a.im, a.re));
      const cd v = dn[k0 + k] * cexp(mk(b.im, -b.re));
      s_a[k] = sign > 0 ? u + v : u - v;
    }
    __syncthreads();
    for (int q = 0; q < 8; ++q) {
      const int i = threadIdx.x + 256 * q;
      if (i < n)
        for (int k = 0; k < kn; ++k) acc[q] += fl.M[i + (long long)(k0 + k) * n] * s_a[k];
    }
    __syncthreads();
  }
  for (int q = 0; q < 8; ++q) {
    const int i = threadIdx.x + 256 * q;
    if (i < n) X[(long long)r * n + i] = acc[q];
  }
}

// Hm[r] = harmonics (h) of the requested component from X[r]
__global__ __launch_bounds__(256) void k_field_component(FieldGeom g, const FieldLayer* layers,
                                                         const FieldRow* rows, const cd* X,
                                                         cd* Hm) {
  __shared__ cd s_r[kTile];
  const int r = blockIdx.x;
  const int h = g.nx * g.ny, n = 2 * h;
  const cd* x = X + (long long)r * n;
  cd* out = Hm + (long long)r * h;
  const FieldLayer fl = layers[rows[r].layer];
  auto kx = [&](int k) { return 2.0 * kPi * (k / g.ny - (g.nx - 1) / 2) / g.Lx + g.kx0; };
  auto ky = [&](int k) { return 2.0 * kPi * (k % g.ny - (g.ny - 1) / 2) / g.Ly + g.ky0; };
  const int c = g.component;
  if (c == 0 || c == 1 || c == 3 || c == 4) {
    const int off = (c == 1 || c == 4) ? h : 0;
    for (int k = threadIdx.x; k < h; k += 256) out[k] = x[off + k];
    return;
  }
  if (c == 5) {
    // Hz = (i / k0^2)(Kx Ey - Ky Ex)
    const double s = 1.0 / (g.k0 * g.k0);
    for (int k = threadIdx.x; k < h; k += 256) {
      const cd t = kx(k) * x[h + k] - ky(k) * x[k];
      out[k] = mk(-s * t.im, s * t.re);
    }
    return;
  }
  // Ez = i [[eps]]^-1 (Kx Hy - Ky Hx)
  if (!fl.m33inv) {
    const cd ie = cinv(fl.eps);
    for (int k = threadIdx.x; k < h; k += 256) {
      const cd t = (kx(k) * x[h + k] - ky(k) * x[k]) * ie;
      out[k] = mk(-t.im, t.re);
    }
    return;
  }
  cd acc[16];
  for (int q = 0; q < 16; ++q) acc[q] = mk(0.0, 0.0);
  for (int j0 = 0; j0 < h; j0 += kTile) {
    const int jn = min(kTile, h - j0);
    for (int j = threadIdx.x; j < jn; j += 256)
      s_r[j] = kx(j0 + j) * x[h + j0 + j] - ky(j0 + j) * x[j0 + j];
    __syncthreads();
    for (int q = 0; q < 16; ++q) {
      const int k = threadIdx.x + 256 * q;
      if (k < h)
        for (int j = 0; j < jn; ++j) acc[q] += fl.m33inv[k + (long long)(j0 + j) * h] * s_r[j];
    }
    __syncthreads();
  }
  for (int q = 0; q < 16; ++q) {
    const int k = threadIdx.x + 256 * q;
    if (k < h) out[k] = mk(-acc[q].im, acc[q].re);
  }
}

// img[r][c]: XZ sums over n at y = coord, then over m at x_c; YZ over m at
// x = coord, then n at y_c; XY over n at y_r, then m at x_c (the nesting of
// rcwa/RCWASolver.cpp:916-1005)
__global__ __launch_bounds__(256) void k_field_synth(FieldGeom g, const cd* Hm, int ncols,
                                                     cd* img) {
  __shared__ cd s_one[64];
  const int r = blockIdx.y;
  const int nx = g.nx, ny = g.ny;
  const int Mx = (nx - 1) / 2, My = (ny - 1) / 2;
  const cd* H = Hm + (long long)(g.slice == 0 ? 0 : r) * nx * ny;
  const bool outer_x = (g.slice != 2);           // XZ and XY: columns run along x
  const int nouter = outer_x ? nx : ny;
  if (threadIdx.x < nouter) {
    const int o = threadIdx.x;
    cd acc = mk(0.0, 0.0);
    if (outer_x) {
      const double y = (g.slice == 0) ? g.start2 + g.step2 * (r + 0.5) : g.coord;
      for (int nn = 0; nn < ny; ++nn) {
        double sn, cs;
        sincos((2.0 * kPi * (nn - My) / g.Ly + g.ky0) * y, &sn, &cs);
        acc += H[o * ny + nn] * mk(cs, sn);
      }
    } else {
      for (int m = 0; m < nx; ++m) {
        double sn, cs;
        sincos((2.0 * kPi * (m - Mx) / g.Lx + g.kx0) * g.coord, &sn, &cs);
        acc += H[m * ny + o] * mk(cs, sn);
      }
    }
    s_one[o] = acc;
  }
  __syncthreads();
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncols) return;
  const double t = g.start1 + g.step1 * (c + 0.5);
  const double per = outer_x ? g.Lx : g.Ly, k0o = outer_x ? g.kx0 : g.ky0;
  const int Mo = outer_x ? Mx : My;
  cd acc = mk(0.0, 0.0);
  for (int o = 0; o < nouter; ++o) {
    double sn, cs;
    sincos((2.0 * kPi * (o - Mo) / per + k0o) * t, &sn, &cs);
    acc += mk(cs, sn) * s_one[o];
  }
  img[(long long)r * ncols + c] = acc;
}

void launch_field_combine(hipStream_t s, int n, int L, const cd* tv, cd px, cd py, cd* cp) {
  const long long cnt = (long long)L * n;
  hipLaunchKernelGGL(k_field_combine, dim3((unsigned)std::min<long long>(4096, (cnt + 255) / 256)),
                     dim3(256), 0, s, n, L, tv, px, py, cp);
}

void launch_field_up(hipStream_t s, int n, const cd* A, const cd* B, const cd* cp1, const cd* cm1,
                     const cd* keff1, double k0d1, cd* cm) {
  hipLaunchKernelGGL(k_field_up, dim3((n + 255) / 256), dim3(256), 0, s, n, A, B, cp1, cm1, keff1,
                     k0d1, cm);
}

void launch_field_modes(hipStream_t s, int n, const FieldLayer* layers, const FieldRow* rows,
                        int nrows, const cd* cp, const cd* cm, int sign, cd* X) {
  hipLaunchKernelGGL(k_field_modes, dim3(nrows), dim3(256), 0, s, n, layers, rows, cp, cm, sign,
                     X);
}

void launch_field_component(hipStream_t s, const FieldGeom& g, const FieldLayer* layers,
                            const FieldRow* rows, int nrows, const cd* X, cd* Hm) {
  hipLaunchKernelGGL(k_field_component, dim3(nrows), dim3(256), 0, s, g, layers, rows, X, Hm);
}

void launch_field_synth(hipStream_t s, const FieldGeom& g, const cd* Hm, int nrows, int ncols,
                        cd* img) {
  hipLaunchKernelGGL(k_field_synth, dim3((ncols + 255) / 256, nrows), dim3(256), 0, s, g, Hm,
                     ncols, img);
}

}  // namespace orcwa
