// Batched FP64 complex non-Hermitian eigensolver for gfx950.
//
// Replaces MKLEigenSolver<MatrixXcs>::compute (rcwa/MKLEigenSolver.hpp:103-179,
// LAPACK zgeev('N','V')) for the per-layer wave-equation matrix P*Q
// (rcwa/WaveEquationSolver.cpp:24-37).  Same mathematics as zgeev:
//   1. unitary Householder reduction to upper Hessenberg form, H = Q^H A Q
//   2. explicit Q
//   3. single-shift complex QR iteration with Wilkinson shifts, deflation by
//      the Ahues-Tisseur small-subdiagonal test, full Schur form T and Schur
//      vectors Z = Q * (rotations)
//   4. eigenvectors of the triangular T by back substitution (with the
//      smin = max(ulp |t_jj|, smlnum) pivot guard), W = Z X, unit 2-norm columns
//   5. the reference's acceptance test ||A W - W diag(l)||max / ||A||max <= 1e-8
// One workgroup owns one matrix; matrices live in HBM (column-major).
#include <cfloat>

#include "engine.h"

namespace orcwa {

static constexpr double kUlp = DBL_EPSILON;
static constexpr double kSafmin = DBL_MIN;

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// block reduction of a double (256 threads); result valid in all threads
__device__ inline double block_sum(double v, double* sred) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sred[wave] = v;
  __syncthreads();
  double r = sred[0] + sred[1] + sred[2] + sred[3];
  __syncthreads();
  return r;
}

// ---------------------------------------------------------------------------
// 1. Hessenberg reduction (zgehd2 order), fused two-sided rank-2 update.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_hess(int n, cd* const* Ap, cd* tau_all) {
  extern __shared__ cd sm[];
  cd* sv = sm;          // v, rows k+1..n-1
  cd* sy = sm + n;      // y = A v   -> p
  cd* sw = sm + 2 * n;  // w = v^H A -> q
  __shared__ double sred[4];
  __shared__ cd s_tau, s_scale;

  const int b = blockIdx.x;
  cd* __restrict__ A = Ap[b];
  cd* tau = tau_all + (long long)b * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

  for (int k = 0; k + 2 < n; ++k) {
    const long long ck = (long long)k * n;
    double part = 0.0;
    for (int i = k + 2 + tid; i < n; i += 256) part += norm2(A[i + ck]);
    const double xn2 = block_sum(part, sred);
    if (tid == 0) {
      cd alpha = A[k + 1 + ck];
      if (xn2 == 0.0 && alpha.im == 0.0) {
        s_tau = mk(0.0, 0.0);
        s_scale = mk(0.0, 0.0);
      } else {
        double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
        double beta = (alpha.re >= 0.0) ? -nrm : nrm;
        s_tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
        s_scale = cinv(alpha - mk(beta, 0.0));
        A[k + 1 + ck] = mk(beta, 0.0);
      }
      tau[k] = s_tau;
    }
    __syncthreads();
    const cd t = s_tau, sc = s_scale;
    if (t.re == 0.0 && t.im == 0.0) {
      __syncthreads();
      continue;
    }
    for (int i = k + 2 + tid; i < n; i += 256) {
      cd v = A[i + ck] * sc;
      A[i + ck] = v;
      sv[i] = v;
    }
    if (tid == 0) sv[k + 1] = mk(1.0, 0.0);
    __syncthreads();
    // y_i = sum_{j>k} A[i,j] v_j, all rows (thread per row, coalesced)
    for (int i = tid; i < n; i += 256) {
      cd acc = mk(0.0, 0.0);
      for (int j = k + 1; j < n; ++j) acc += A[i + (long long)j * n] * sv[j];
      sy[i] = acc;
    }
    // w_j = sum_{i>k} conj(v_i) A[i,j] (wave per column)
    for (int j = k + 1 + wave; j < n; j += 4) {
      const cd* col = A + (long long)j * n;
      double ar = 0.0, ai = 0.0;
      for (int i = k + 1 + lane; i < n; i += 64) {
        cd z = cmul(sv[i], col[i]);
        ar += z.re;
        ai += z.im;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) sw[j] = mk(ar, ai);
    }
    __syncthreads();
    // s = w . v
    double sr = 0.0, si = 0.0;
    for (int j = k + 1 + tid; j < n; j += 256) {
      cd z = sw[j] * sv[j];
      sr += z.re;
      si += z.im;
    }
    sr = block_sum(sr, sred);
    si = block_sum(si, sred);
    const cd sv_s = mk(sr, si);
    const double t2 = norm2(t);
    for (int i = tid; i < n; i += 256) {
      cd p = t * sy[i];
      if (i > k) p -= t2 * (sv_s * sv[i]);
      sy[i] = p;
    }
    for (int j = k + 1 + tid; j < n; j += 256) sw[j] = conj(t) * sw[j];
    __syncthreads();
    // A[i,j] -= p_i conj(v_j) + v_i q_j   (rows 0..n-1, columns k+1..n-1)
    for (int j = k + 1 + wave; j < n; j += 4) {
      cd* col = A + (long long)j * n;
      const cd vj = conj(sv[j]);
      const cd qj = sw[j];
      for (int i = lane; i < n; i += 64) {
        cd upd = sy[i] * vj;
        if (i > k) upd += sv[i] * qj;
        col[i] -= upd;
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (n >= 2) tau[n - 2] = mk(0.0, 0.0);
    tau[n - 1] = mk(0.0, 0.0);
  }
}

// ---------------------------------------------------------------------------
// 2. Z = Q = H_0 H_1 ... H_{n-3} (backward accumulation, zung2r order).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_formq(int n, const cd* const* Ap, const cd* tau_all,
                                               cd* const* Zp) {
  extern __shared__ cd sm[];
  cd* sv = sm;
  cd* sw = sm + n;
  const int b = blockIdx.x;
  const cd* __restrict__ A = Ap[b];
  cd* __restrict__ Z = Zp[b];
  const cd* tau = tau_all + (long long)b * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long nn = (long long)n * n;
  for (long long e = tid; e < nn; e += 256) Z[e] = (e % n == e / n) ? mk(1.0, 0.0) : mk(0.0, 0.0);
  __syncthreads();
  for (int k = n - 3; k >= 0; --k) {
    const cd t = tau[k];
    if (t.re == 0.0 && t.im == 0.0) continue;
    const long long ck = (long long)k * n;
    for (int i = k + 2 + tid; i < n; i += 256) sv[i] = A[i + ck];
    if (tid == 0) sv[k + 1] = mk(1.0, 0.0);
    __syncthreads();
    for (int j = k + 1 + wave; j < n; j += 4) {
      const cd* col = Z + (long long)j * n;
      double ar = 0.0, ai = 0.0;
      for (int i = k + 1 + lane; i < n; i += 64) {
        cd z = cmul(sv[i], col[i]);
        ar += z.re;
        ai += z.im;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) sw[j] = t * mk(ar, ai);
    }
    __syncthreads();
    for (int j = k + 1 + wave; j < n; j += 4) {
      cd* col = Z + (long long)j * n;
      const cd wj = sw[j];
      for (int i = k + 1 + lane; i < n; i += 64) col[i] -= sv[i] * wj;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 3. single-shift complex QR (zlahqr mathematics; wantt, wantz), one
//    workgroup per matrix.  Each 2x2 unitary G = [[c, -conj(s)], [s, conj(c)]]
//    with (c, s) = x/||x|| maps x to (||x||, 0) under G^H.
// ---------------------------------------------------------------------------
__device__ inline bool small_subdiag(const cd* H, int n, int k, int lo, int hi, double smlnum) {
  const cd h = H[k + (long long)(k - 1) * n];
  const double ah = cabs1(h);
  if (ah <= smlnum) return true;
  const cd hkk = H[k + (long long)k * n];
  const cd hk1 = H[(k - 1) + (long long)(k - 1) * n];
  double tst = cabs1(hk1) + cabs1(hkk);
  if (tst == 0.0) {
    if (k - 2 >= lo) tst += cabs1(H[(k - 1) + (long long)(k - 2) * n]);
    if (k + 1 <= hi) tst += cabs1(H[(k + 1) + (long long)k * n]);
  }
  if (ah <= kUlp * tst) {
    const double hup = cabs1(H[(k - 1) + (long long)k * n]);
    const double ab = fmax(ah, hup), ba = fmin(ah, hup);
    const double d2 = cabs1(hk1 - hkk);
    const double aa = fmax(cabs1(hkk), d2), bb = fmin(cabs1(hkk), d2);
    const double s = aa + ab;
    if (ba * (ab / s) <= fmax(smlnum, kUlp * (bb * (aa / s)))) return true;
  }
  return false;
}

__device__ inline cd wilkinson_shift(const cd* H, int n, int hi) {
  cd t = H[hi + (long long)hi * n];
  cd u = csqrt_principal(H[(hi - 1) + (long long)hi * n]) *
         csqrt_principal(H[hi + (long long)(hi - 1) * n]);
  double s = cabs1(u);
  if (s != 0.0) {
    cd x = 0.5 * (H[(hi - 1) + (long long)(hi - 1) * n] - t);
    double sx = cabs1(x);
    s = fmax(s, sx);
    cd xs = (1.0 / s) * x, us = (1.0 / s) * u;
    cd y = s * csqrt_principal(xs * xs + us * us);
    if (sx > 0.0) {
      cd xx = (1.0 / sx) * x;
      if (xx.re * y.re + xx.im * y.im < 0.0) y = -y;
    }
    t = t - u * cdiv(u, x + y);
  }
  return t;
}

__global__ __launch_bounds__(256) void k_qr_simple(int n, cd* const* Hp, cd* const* Zp,
                                                   int* info, int maxit) {
  __shared__ int s_l;
  __shared__ cd s_c, s_s;
  __shared__ int s_red[256];
  const int b = blockIdx.x;
  cd* __restrict__ H = Hp[b];
  cd* __restrict__ Z = Zp[b];
  const int tid = threadIdx.x;
  const double smlnum = kSafmin * ((double)n / kUlp);

  int hi = n - 1, its = 0, total = 0;
  while (hi > 0) {
    // largest k in [1, hi] whose subdiagonal is negligible (0 if none)
    int best = 0;
    for (int k = 1 + tid; k <= hi; k += 256)
      if (small_subdiag(H, n, k, 0, hi, smlnum)) best = max(best, k);
    s_red[tid] = best;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (tid < st) s_red[tid] = max(s_red[tid], s_red[tid + st]);
      __syncthreads();
    }
    const int l = s_red[0];
    __syncthreads();
    if (l > 0 && tid == 0) H[l + (long long)(l - 1) * n] = mk(0.0, 0.0);
    if (l == hi) {
      --hi;
      its = 0;
      __syncthreads();
      continue;
    }
    ++its;
    ++total;
    if (total > maxit) {
      if (tid == 0) info[b] |= ORCWA_STATUS_NOCONV;
      break;
    }
    cd shift;
    if (its % 30 == 10) {
      shift = mk(0.75 * cabs1(H[(l + 1) + (long long)l * n]), 0.0) + H[l + (long long)l * n];
    } else if (its % 30 == 20) {
      shift = mk(0.75 * cabs1(H[hi + (long long)(hi - 1) * n]), 0.0) + H[hi + (long long)hi * n];
    } else {
      shift = wilkinson_shift(H, n, hi);
    }
    __syncthreads();
    for (int k = l; k < hi; ++k) {
      if (tid == 0) {
        cd x0, x1;
        if (k == l) {
          x0 = H[l + (long long)l * n] - shift;
          x1 = H[(l + 1) + (long long)l * n];
        } else {
          x0 = H[k + (long long)(k - 1) * n];
          x1 = H[(k + 1) + (long long)(k - 1) * n];
        }
        const double nr = hypot(cabs(x0), cabs(x1));
        if (nr == 0.0) {
          s_c = mk(1.0, 0.0);
          s_s = mk(0.0, 0.0);
        } else {
          s_c = (1.0 / nr) * x0;
          s_s = (1.0 / nr) * x1;
          if (k > l) {
            H[k + (long long)(k - 1) * n] = mk(nr, 0.0);
            H[(k + 1) + (long long)(k - 1) * n] = mk(0.0, 0.0);
          }
        }
      }
      __syncthreads();
      const cd c = s_c, s = s_s;
      // rows k, k+1 from the left (columns k..n-1; column k-1 set above)
      for (int j = k + tid; j < n; j += 256) {
        const long long cj = (long long)j * n;
        const cd a = H[k + cj], bb = H[k + 1 + cj];
        H[k + cj] = cmul(c, a) + cmul(s, bb);
        H[k + 1 + cj] = c * bb - s * a;
      }
      __syncthreads();
      // columns k, k+1 from the right (rows 0..min(k+2,hi)), and Z
      const int iend = min(k + 2, hi);
      const long long ck = (long long)k * n, ck1 = (long long)(k + 1) * n;
      for (int i = tid; i <= iend; i += 256) {
        const cd p = H[i + ck], q = H[i + ck1];
        H[i + ck] = p * c + q * s;
        H[i + ck1] = q * conj(c) - p * conj(s);
      }
      for (int i = tid; i < n; i += 256) {
        const cd p = Z[i + ck], q = Z[i + ck1];
        Z[i + ck] = p * c + q * s;
        Z[i + ck1] = q * conj(c) - p * conj(s);
      }
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------------------
// 4. eigenvectors of the upper triangular T (ztrevc mathematics, right
//    vectors): X upper triangular with unit diagonal.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trevc(int n, const cd* const* Tp, cd* const* Xp) {
  const int b = blockIdx.x;
  const cd* __restrict__ T = Tp[b];
  cd* __restrict__ X = Xp[b];
  const int tid = threadIdx.x;
  const long long nn = (long long)n * n;
  const double smlnum = kSafmin * ((double)n / kUlp);
  for (long long e = tid; e < nn; e += 256) {
    const long long i = e % n, j = e / n;
    X[e] = (i < j) ? -T[e] : ((i == j) ? mk(1.0, 0.0) : mk(0.0, 0.0));
  }
  __syncthreads();
  for (int k = n - 2; k >= 0; --k) {
    const cd tkk = T[k + (long long)k * n];
    for (int j = k + 1 + tid; j < n; j += 256) {
      const cd tjj = T[j + (long long)j * n];
      const double smin = fmax(kUlp * cabs1(tjj), smlnum);
      cd d = tkk - tjj;
      if (cabs1(d) < smin) d = mk(smin, 0.0);
      X[k + (long long)j * n] = cdiv(X[k + (long long)j * n], d);
    }
    __syncthreads();
    if (k > 0) {
      const long long cnt = (long long)k * (n - k - 1);
      for (long long e = tid; e < cnt; e += 256) {
        const int i = (int)(e % k);
        const int j = k + 1 + (int)(e / k);
        X[i + (long long)j * n] -= T[i + (long long)k * n] * X[k + (long long)j * n];
      }
    }
    __syncthreads();
  }
}

// lam[j] = T[j,j]; unit 2-norm columns of W; relative residual.
__global__ __launch_bounds__(256) void k_eig_finish(int n, const cd* const* Tp, cd* const* Wp,
                                                    cd* lam) {
  const int b = blockIdx.x;
  const cd* T = Tp[b];
  cd* W = Wp[b];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = tid; j < n; j += 256) lam[(long long)b * n + j] = T[j + (long long)j * n];
  for (int j = wave; j < n; j += 4) {
    cd* col = W + (long long)j * n;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += norm2(col[i]);
    s = wave_sum(s);
    const double inv = (s > 0.0) ? 1.0 / sqrt(s) : 1.0;
    for (int i = lane; i < n; i += 64) col[i] = inv * col[i];
  }
}

// R = A W already in Rp; residual = max|R - W diag(lam)| / max|A|
__global__ __launch_bounds__(256) void k_eig_residual(int n, const cd* const* Ap,
                                                      const cd* const* Wp, const cd* const* Rp,
                                                      const cd* lam, double* resid, int* info) {
  __shared__ double sred[4];
  const int b = blockIdx.x;
  const cd* A = Ap[b];
  const cd* W = Wp[b];
  const cd* R = Rp[b];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long nn = (long long)n * n;
  double ma = 0.0, mr = 0.0;
  for (long long e = tid; e < nn; e += 256) {
    ma = fmax(ma, cabs(A[e]));
    const cd l = lam[(long long)b * n + e / n];
    mr = fmax(mr, cabs(R[e] - W[e] * l));
  }
  ma = wave_max(ma);
  mr = wave_max(mr);
  __shared__ double sa[4];
  if (lane == 0) {
    sa[wave] = ma;
    sred[wave] = mr;
  }
  __syncthreads();
  if (tid == 0) {
    const double a = fmax(fmax(sa[0], sa[1]), fmax(sa[2], sa[3]));
    const double r = fmax(fmax(sred[0], sred[1]), fmax(sred[2], sred[3]));
    const double rel = (a > 0.0) ? r / a : 0.0;
    resid[b] = rel;
    if (!(rel <= 1e-8)) info[b] |= ORCWA_STATUS_EIG_RESIDUAL;
  }
}

__global__ void k_zero_below_subdiag(int n, cd* const* Ap) {
  cd* Am = Ap[blockIdx.y];
  const long long nn = (long long)n * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nn;
       e += (long long)gridDim.x * blockDim.x) {
    const long long i = e % n, j = e / n;
    if (i > j + 1) Am[e] = mk(0.0, 0.0);
  }
}

void zgeev_batched(hipStream_t s, int n, cd* const* A, const cd* const* Aorig, cd* const* W,
                   cd* lam, const EigWorkspace& ws, int* info, double* resid,
                   cd* const* scratch, int batch, hipEvent_t* ev) {
  if (batch <= 0 || n <= 0) return;
  const cd* const* Ac = (const cd* const*)A;
  if (n >= 3) {
    hipLaunchKernelGGL(k_hess, dim3(batch), dim3(256), 3 * n * sizeof(cd), s, n, A, ws.tau);
    hipLaunchKernelGGL(k_formq, dim3(batch), dim3(256), 2 * n * sizeof(cd), s, n, Ac,
                       (const cd*)ws.tau, ws.Z);
  } else {
    set_identity_batched(s, n, ws.Z, batch);
  }
  if (ev) (void)hipEventRecord(ev[0], s);
  // zero everything below the first subdiagonal (the Householder vectors)
  hipLaunchKernelGGL(k_zero_below_subdiag, dim3(64, batch), dim3(256), 0, s, n, A);
  hipLaunchKernelGGL(k_qr_simple, dim3(batch), dim3(256), 0, s, n, A, ws.Z, info, 60 * n);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_trevc, dim3(batch), dim3(256), 0, s, n, Ac, ws.X);
  zgemm_batched(s, n, n, n, mk(1.0, 0.0), (const cd* const*)ws.Z, 0, 1, n, 0,
                (const cd* const*)ws.X, 0, 1, n, 0, mk(0.0, 0.0), W, 0, 1, n, batch);
  hipLaunchKernelGGL(k_eig_finish, dim3(batch), dim3(256), 0, s, n, Ac, W, lam);
  if (Aorig) {
    zgemm_batched(s, n, n, n, mk(1.0, 0.0), Aorig, 0, 1, n, 0, (const cd* const*)W, 0, 1, n, 0,
                  mk(0.0, 0.0), scratch, 0, 1, n, batch);
    hipLaunchKernelGGL(k_eig_residual, dim3(batch), dim3(256), 0, s, n, Aorig,
                       (const cd* const*)W, (const cd* const*)scratch, (const cd*)lam, resid,
                       info);
  }
  if (ev) (void)hipEventRecord(ev[2], s);
}

}  // namespace orcwa
