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
#include <algorithm>
#include <cfloat>

#include "engine.h"

namespace orcwa {

typedef double v4d __attribute__((ext_vector_type(4)));

static constexpr double kUlp = DBL_EPSILON;
static constexpr double kSafmin = DBL_MIN;


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
// 1b. blocked Hessenberg reduction (zgehrd / zlahr2 order) for large n.
//     Per block of HB columns [i0, i0+nb): k_hess_panel reduces the panel
//     columns one by one, keeping the rest of the matrix at its block-start
//     value and representing the block's transformation Q_b = I - V T V^H by
//     V (explicit, n x HB), T (HB x HB) and Y = A V T (rows i0+1..n-1); the
//     trailing matrix is then updated with batched MFMA GEMMs:
//        Y(0:i0+1) = A(0:i0+1, i0+1:n) V T
//        A(0:i0+1, i0+1:n)   -= Y(0:i0+1) V^H             (right, top rows)
//        A(i0+1:n, i0+nb:n)  -= Y(i0+1:n) V(i0+nb:n)^H    (right)
//        A(i0+1:n, i0+nb:n)  -= V T^H (V^H A(i0+1:n, i0+nb:n))   (left)
//     so each trailing element is read once per column (the matvec A v) and
//     twice per block, instead of twice and written once per column.
//     Workspace per matrix (X): V | Y | Ytmp (n x HB each) | T (HB x HB);
//     (Z): W1 | W2 (HB x n each).
// ---------------------------------------------------------------------------
constexpr int HB = 32;

__global__ __launch_bounds__(256) void k_hess_panel(int n, int i0, int nb, cd* const* Ap,
                                                    cd* tau_all, cd* const* Xp) {
  extern __shared__ cd sm[];
  cd* sb = sm;          // current column, rows i0+1..n-1
  cd* sv = sm + n;      // current reflector
  // T of the panel's block reflector, upper triangular, packed by columns:
  // T(r, c) at c (c + 1) / 2 + r, r <= c
  __shared__ cd Ts[HB * (HB + 1) / 2];
  __shared__ cd s_w[HB], s_t[HB], s_row[HB];
  __shared__ double sred[4];
  __shared__ cd s_tau, s_scale;

  const int mat = blockIdx.x;
  cd* __restrict__ A = Ap[mat];
  cd* __restrict__ V = Xp[mat];
  cd* __restrict__ Y = V + (long long)n * HB;
  cd* __restrict__ Tg = V + 3LL * n * HB;
  cd* tau = tau_all + (long long)mat * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int r0 = i0 + 1;

  for (int j = 0; j < nb; ++j) {
    const int c = i0 + j;
    const long long cc = (long long)c * n;
    for (int r = r0 + tid; r < n; r += 256) sb[r] = A[r + cc];
    // the reflector below reads sb with another row-to-thread mapping
    __syncthreads();
    if (j > 0) {
      // b -= Y(:, 0:j) conj(V(c, 0:j))
      if (tid < j) s_row[tid] = conj(V[c + (long long)tid * n]);
      __syncthreads();
      for (int r = r0 + tid; r < n; r += 256) {
        cd acc = mk(0.0, 0.0);
        for (int p = 0; p < j; ++p) acc += Y[r + (long long)p * n] * s_row[p];
        sb[r] -= acc;
      }
      __syncthreads();
      // b <- (I - V T^H V^H) b
      for (int p = wave; p < j; p += 4) {
        const cd* vp = V + (long long)p * n;
        double ar = 0.0, ai = 0.0;
        for (int r = r0 + lane; r < n; r += 64) {
          const cd z = cmul(vp[r], sb[r]);
          ar += z.re;
          ai += z.im;
        }
        ar = wave_sum(ar);
        ai = wave_sum(ai);
        if (lane == 0) s_w[p] = mk(ar, ai);
      }
      __syncthreads();
      if (tid < j) {
        cd acc = mk(0.0, 0.0);
        for (int q = 0; q <= tid; ++q) acc += conj(Ts[tid * (tid + 1) / 2 + q]) * s_w[q];
        s_t[tid] = acc;
      }
      __syncthreads();
      for (int r = r0 + tid; r < n; r += 256) {
        cd acc = mk(0.0, 0.0);
        for (int p = 0; p < j; ++p) acc += V[r + (long long)p * n] * s_t[p];
        sb[r] -= acc;
      }
      __syncthreads();
    }
    // reflector of b(c+1 .. n-1) (zlarfg: H^H x = beta e1)
    double part = 0.0;
    for (int r = c + 2 + tid; r < n; r += 256) part += norm2(sb[r]);
    const double xn2 = block_sum(part, sred);
    if (tid == 0) {
      const cd alpha = sb[c + 1];
      cd beta = alpha;
      if (xn2 == 0.0 && alpha.im == 0.0) {
        s_tau = mk(0.0, 0.0);
        s_scale = mk(0.0, 0.0);
      } else {
        const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
        const double bt = (alpha.re >= 0.0) ? -nrm : nrm;
        s_tau = mk((bt - alpha.re) / bt, -alpha.im / bt);
        s_scale = cinv(alpha - mk(bt, 0.0));
        beta = mk(bt, 0.0);
      }
      tau[c] = s_tau;
      sb[c + 1] = beta;
    }
    __syncthreads();
    const cd t = s_tau, sc = s_scale;
    cd* Vj = V + (long long)j * n;
    for (int r = tid; r < n; r += 256) {
      cd v = mk(0.0, 0.0);
      if (r == c + 1) v = mk(1.0, 0.0);
      else if (r > c + 1) v = sb[r] * sc;
      sv[r] = v;
      Vj[r] = v;
      // column c of the reduced matrix: updated entries above the
      // subdiagonal, beta on it, the reflector below it
      if (r >= r0) A[r + cc] = (r > c + 1) ? v : sb[r];
    }
    __syncthreads();
    // t = V(:, 0:j)^H v
    for (int p = wave; p < j; p += 4) {
      const cd* vp = V + (long long)p * n;
      double ar = 0.0, ai = 0.0;
      for (int r = c + 1 + lane; r < n; r += 64) {
        const cd z = cmul(vp[r], sv[r]);
        ar += z.re;
        ai += z.im;
      }
      ar = wave_sum(ar);
      ai = wave_sum(ai);
      if (lane == 0) s_t[p] = mk(ar, ai);
    }
    __syncthreads();
    // y = A(r0:n, c+1:n) v(c+1:n) over the block-start columns, then
    // Y(:, j) = tau (y - Y(:, 0:j) t): two rows per thread, 512 rows a pass
    for (int base = r0; base < n; base += 512) {
      const int ra = base + tid, rb = base + tid + 256;
      if (ra >= n) break;
      const bool hb = rb < n;
      cd ya = mk(0.0, 0.0), yb = mk(0.0, 0.0);
#pragma unroll 4
      for (int q = c + 1; q < n; ++q) {
        const cd vq = sv[q];
        const cd* col = A + (long long)q * n;
        ya += col[ra] * vq;
        if (hb) yb += col[rb] * vq;
      }
      cd acc = ya;
      for (int p = 0; p < j; ++p) acc -= Y[ra + (long long)p * n] * s_t[p];
      Y[ra + (long long)j * n] = t * acc;
      if (hb) {
        acc = yb;
        for (int p = 0; p < j; ++p) acc -= Y[rb + (long long)p * n] * s_t[p];
        Y[rb + (long long)j * n] = t * acc;
      }
    }
    // T(0:j, j) = -tau T(0:j, 0:j) t,  T(j, j) = tau
    if (tid < j) {
      cd acc = mk(0.0, 0.0);
      for (int q = tid; q < j; ++q) acc += Ts[q * (q + 1) / 2 + tid] * s_t[q];
      Ts[j * (j + 1) / 2 + tid] = -(t * acc);
    }
    if (tid == j) Ts[j * (j + 1) / 2 + j] = t;
    __syncthreads();
  }
  for (int e = tid; e < HB * HB; e += 256) {
    const int r = e % HB, q = e / HB;
    Tg[e] = (r <= q && q < nb) ? Ts[q * (q + 1) / 2 + r] : mk(0.0, 0.0);
  }
  // columns outside the panel of V stay zero for the GEMMs
  for (long long e = (long long)nb * n + tid; e < (long long)HB * n; e += 256) V[e] = mk(0.0, 0.0);
  if (tid == 0 && i0 + nb == n - 2) {
    tau[n - 2] = mk(0.0, 0.0);
    tau[n - 1] = mk(0.0, 0.0);
  }
}

static void hess_blocked(hipStream_t s, int n, cd* const* A, cd* tau, cd* const* X,
                         cd* const* Z, int batch) {
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  const cd* const* Ac = (const cd* const*)A;
  const cd* const* Xc = (const cd* const*)X;
  const cd* const* Zc = (const cd* const*)Z;
  const long long oV = 0, oY = (long long)n * HB, oYt = 2LL * n * HB, oT = 3LL * n * HB;
  const long long oW1 = 0, oW2 = (long long)HB * n;
  for (int i0 = 0; i0 + 2 < n; i0 += HB) {
    const int nb = std::min(HB, n - 2 - i0);
    const int top = i0 + 1, m = n - i0 - 1, nt = n - i0 - nb;
    hipLaunchKernelGGL(k_hess_panel, dim3(batch), dim3(256), 2 * n * sizeof(cd), s, n, i0, nb,
                       A, tau, X);
    // Ytmp(0:top) = A(0:top, i0+1:n) V(i0+1:n, :);  Y(0:top) = Ytmp T
    zgemm_batched(s, top, nb, m, one, Ac, (long long)top * n, 1, n, 0, Xc, oV + top, 1, n, 0,
                  zero, X, oYt, 1, n, batch);
    zgemm_batched(s, top, nb, nb, one, Xc, oYt, 1, n, 0, Xc, oT, 1, HB, 0, zero, X, oY, 1, n,
                  batch);
    // A(0:top, i0+1:n) -= Y(0:top) V(i0+1:n)^H
    zgemm_batched(s, top, m, nb, mone, Xc, oY, 1, n, 0, Xc, oV + top, n, 1, 1, one, A,
                  (long long)top * n, 1, n, batch);
    if (nt > 0) {
      // A(top:n, i0+nb:n) -= Y(top:n) V(i0+nb:n)^H
      zgemm_batched(s, m, nt, nb, mone, Xc, oY + top, 1, n, 0, Xc, oV + i0 + nb, n, 1, 1, one, A,
                    (long long)top + (long long)(i0 + nb) * n, 1, n, batch);
      // W1 = V(top:n)^H A(top:n, i0+nb:n);  W2 = T^H W1;  A -= V W2
      zgemm_batched(s, nb, nt, m, one, Xc, oV + top, n, 1, 1, Ac,
                    (long long)top + (long long)(i0 + nb) * n, 1, n, 0, zero, Z, oW1, 1, HB,
                    batch);
      zgemm_batched(s, nb, nt, nb, one, Xc, oT, HB, 1, 1, Zc, oW1, 1, HB, 0, zero, Z, oW2, 1, HB,
                    batch);
      zgemm_batched(s, m, nt, nb, mone, Xc, oV + top, 1, n, 0, Zc, oW2, 1, HB, 0, one, A,
                    (long long)top + (long long)(i0 + nb) * n, 1, n, batch);
    }
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
// 2b. blocked Q: Q = H_0 ... H_{n-3} = Q_0 Q_1 ... with Q_b = I - V_b T_b V_b^H
//     (compact WY, zlarft forward columnwise), accumulated backward with three
//     MFMA GEMMs per block.  V_b is extracted from the reflectors stored below
//     the subdiagonal by k_hess.
// ---------------------------------------------------------------------------
constexpr int QB = 32;

// Vall[:, j] = reflector j as a dense column (zeros above row j+1, 1 at j+1)
__global__ void k_extract_v(int n, const cd* const* Ap, cd* const* Vp) {
  const cd* A = Ap[blockIdx.y];
  cd* V = Vp[blockIdx.y];
  const long long cnt = (long long)n * (n - 2);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < cnt;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    cd v = mk(0.0, 0.0);
    if (i == j + 1) v = mk(1.0, 0.0);
    else if (i > j + 1) v = A[i + (long long)j * n];
    V[e] = v;
  }
}

// T factor of reflector block b (columns b*QB .. b*QB+nbk-1 of Vall):
//   T[j,j] = tau_j,  T[0:j, j] = -tau_j T[0:j, 0:j] (V^H v_j)[0:j]
// one workgroup per (block, matrix); T stored QB x QB (ld QB) per block.
__global__ __launch_bounds__(256) void k_larft(int n, const cd* const* Vp, const cd* tau_all,
                                               cd* const* Tp) {
  __shared__ cd G[QB][QB + 1];
  __shared__ cd Tl[QB][QB + 1];
  const int blk = blockIdx.x, b = blockIdx.y;
  const int nref = n - 2;
  const int j0 = blk * QB;
  const int nbk = min(QB, nref - j0);
  const cd* V = Vp[b];
  const cd* tau = tau_all + (long long)b * n;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // G = V_b^H V_b (upper part needed), rows from j0+1 (above: zeros)
  for (int pr = wave; pr < nbk * nbk; pr += 4) {
    const int r = pr / nbk, c = pr % nbk;
    if (r > c) continue;
    const cd* vr = V + (long long)(j0 + r) * n;
    const cd* vc = V + (long long)(j0 + c) * n;
    double ar = 0.0, ai = 0.0;
    for (int i = j0 + r + 1 + lane; i < n; i += 64) {
      const cd z = cmul(vr[i], vc[i]);
      ar += z.re;
      ai += z.im;
    }
    ar = wave_sum(ar);
    ai = wave_sum(ai);
    if (lane == 0) G[r][c] = mk(ar, ai);
  }
  __syncthreads();
  if (tid < 64) {
    for (int j = 0; j < nbk; ++j) {
      const cd tj = tau[j0 + j];
      // lane i < j: T[i, j] = -tau_j sum_{k=i}^{j-1} T[i,k] G[k,j]
      cd acc = mk(0.0, 0.0);
      if (lane < j)
        for (int k = lane; k < j; ++k) acc += Tl[lane][k] * G[k][j];
      if (lane < j) Tl[lane][j] = -(tj * acc);
      if (lane == j) Tl[j][j] = tj;
      if (lane > j && lane < QB) Tl[lane][j] = mk(0.0, 0.0);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  __syncthreads();
  cd* T = Tp[b] + (long long)blk * QB * QB;
  for (int e = tid; e < QB * QB; e += 256) {
    const int r = e % QB, c = e / QB;
    T[e] = (r < nbk && c < nbk) ? Tl[r][c] : mk(0.0, 0.0);
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

// ---------------------------------------------------------------------------
// 3b. windowed small-bulge multishift QR, in three variants of window /
// shifts / workgroups per CU (qr_ms.inc); the plan's autotuner picks one per
// matrix size, 48 / 12 / 2 by default.
// ---------------------------------------------------------------------------
namespace qr_w48 {
#define ORCWA_QR_W 48
#define ORCWA_QR_NS 12
#define ORCWA_QR_WG_PER_CU 2
#include "qr_ms.inc"
#undef ORCWA_QR_W
#undef ORCWA_QR_NS
#undef ORCWA_QR_WG_PER_CU
}  // namespace qr_w48
namespace qr_w32 {
#define ORCWA_QR_W 32
#define ORCWA_QR_NS 8
#define ORCWA_QR_WG_PER_CU 3
#include "qr_ms.inc"
#undef ORCWA_QR_W
#undef ORCWA_QR_NS
#undef ORCWA_QR_WG_PER_CU
}  // namespace qr_w32
namespace qr_w64 {
#define ORCWA_QR_W 64
#define ORCWA_QR_NS 16
#define ORCWA_QR_WG_PER_CU 1
#include "qr_ms.inc"
#undef ORCWA_QR_W
#undef ORCWA_QR_NS
#undef ORCWA_QR_WG_PER_CU
}  // namespace qr_w64

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

// Blocked form of the same recurrence, by row blocks [i0, i1) of 32 from the
// bottom: with X = I initially,
//   X(i0:i1, i1:n) = -T(i0:i1, i1:n) X(i1:n, i1:n)          (one batched GEMM)
// then, per column j >= i0, the rows k = min(j, i1)-1 .. i0 of the block:
//   x(k) = (x(k) - sum_{m=k+1}^{i1-1} T(k, m) x(m)) / (T(k,k) - T(j,j))
// (same smin guard), with the 32 x 32 diagonal block of T in LDS and the
// block of column j in registers.
constexpr int TB = 32;

__global__ __launch_bounds__(256) void k_trevc_block(int n, int i0, int i1, const cd* const* Tp,
                                                     cd* const* Xp) {
  __shared__ cd Tb[TB][TB + 1];
  const int b = blockIdx.y;
  const cd* __restrict__ T = Tp[b];
  cd* __restrict__ X = Xp[b];
  const int nb = i1 - i0;
  for (int e = threadIdx.x; e < TB * TB; e += 256) {
    const int r = e % TB, c = e / TB;
    Tb[r][c] = (r < nb && c < nb) ? T[(i0 + r) + (long long)(i0 + c) * n] : mk(0.0, 0.0);
  }
  __syncthreads();
  const int j = i0 + blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const double smlnum = kSafmin * ((double)n / kUlp);
  const cd lam = T[j + (long long)j * n];
  const double smin = fmax(kUlp * cabs1(lam), smlnum);
  const int kend = min(j, i1) - i0;       // rows [0, kend) of the block are solved
  cd* xc = X + (long long)j * n + i0;
  cd x[TB];
#pragma unroll
  for (int r = 0; r < TB; ++r) x[r] = (r < nb) ? xc[r] : mk(0.0, 0.0);
#pragma unroll
  for (int kk = TB - 1; kk >= 0; --kk) {
    if (kk < kend) {
      cd acc = x[kk];
#pragma unroll
      for (int m = kk + 1; m < TB; ++m) acc -= Tb[kk][m] * x[m];
      cd d = Tb[kk][kk] - lam;
      if (cabs1(d) < smin) d = mk(smin, 0.0);
      x[kk] = cdiv(acc, d);
    }
  }
#pragma unroll
  for (int r = 0; r < TB; ++r)
    if (r < kend) xc[r] = x[r];
}

static void trevc_blocked(hipStream_t s, int n, const cd* const* T, cd* const* X, int batch) {
  set_identity_batched(s, n, X, batch);
  const int nblk = (n + TB - 1) / TB;
  for (int bk = nblk - 1; bk >= 0; --bk) {
    const int i0 = bk * TB, i1 = std::min(n, i0 + TB);
    // (X(i1:n, i1:n) is upper triangular: the identity's zeros below the
    // diagonal stay untouched)
    if (i1 < n)
      zgemm_batched_upper_b(s, i1 - i0, n - i1, n - i1, mk(-1.0, 0.0), T,
                            (long long)i0 + (long long)i1 * n, 1, n, 0, (const cd* const*)X,
                            (long long)i1 + (long long)i1 * n, 1, n, 0, mk(0.0, 0.0), X,
                            (long long)i0 + (long long)i1 * n, 1, n, batch);
    hipLaunchKernelGGL(k_trevc_block, dim3((n - i0 + 255) / 256, batch), dim3(256), 0, s, n, i0,
                       i1, T, X);
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

// R = A W already in Rp; residual = max|R - W diag(lam)| / max|A|, accepted
// when <= tol (1e-8 in the reference, rcwa/MKLEigenSolver.hpp:75-79)
__global__ __launch_bounds__(256) void k_eig_residual(int n, const cd* const* Ap,
                                                      const cd* const* Wp, const cd* const* Rp,
                                                      const cd* lam, double* resid, int* info,
                                                      double tol) {
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
    if (!(rel <= tol)) info[b] |= ORCWA_STATUS_EIG_RESIDUAL;
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


// The perturbed matrix of the reference's recovery solve
// (rcwa/MKLEigenSolver.hpp:81-93): A = Aorig + diag(1e-11 (i+1) max|Aorig|),
// one workgroup per matrix.
__global__ __launch_bounds__(256) void k_perturb_diag(int n, const cd* const* Aop,
                                                      cd* const* Ap) {
  __shared__ double sm[4];
  const cd* Ao = Aop[blockIdx.x];
  cd* A = Ap[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const long long nn = (long long)n * n;
  double m = 0.0;
  for (long long e = tid; e < nn; e += 256) m = fmax(m, cabs(Ao[e]));
  m = wave_max(m);
  if (lane == 0) sm[wave] = m;
  __syncthreads();
  const double scale = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
  for (long long e = tid; e < nn; e += 256) {
    cd v = Ao[e];
    const long long i = e % n, j = e / n;
    if (i == j) v.re += scale * 1e-11 * (double)(i + 1);
    A[e] = v;
  }
}

void eig_perturb_batched(hipStream_t s, int n, const cd* const* Aorig, cd* const* A, int batch) {
  if (batch <= 0 || n <= 0) return;
  hipLaunchKernelGGL(k_perturb_diag, dim3(batch), dim3(256), 0, s, n, Aorig, A);
}

// results of the recovery solves back into the slots of their problems:
// problem idx[k] gets lam_k, status info_k | EIG_PERTURBED (replacing the
// failed solve's bits) and residual res_k
__global__ __launch_bounds__(256) void k_eig_recover_scatter(int n, const int* idx,
                                                             const cd* lam_k, const int* info_k,
                                                             const double* res_k, cd* lam,
                                                             int* info, double* resid) {
  const int k = blockIdx.x, q = idx[k];
  for (int j = threadIdx.x; j < n; j += 256)
    lam[(long long)q * n + j] = lam_k[(long long)k * n + j];
  if (threadIdx.x == 0) {
    info[q] = info_k[k] | ORCWA_STATUS_EIG_PERTURBED;
    resid[q] = res_k[k];
  }
}

void eig_recover_scatter(hipStream_t s, int n, int m, const int* idx, const cd* lam_k,
                         const int* info_k, const double* res_k, cd* lam, int* info,
                         double* resid) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_eig_recover_scatter, dim3(m), dim3(256), 0, s, n, idx, lam_k, info_k,
                     res_k, lam, info, resid);
}

void zgeev_batched(hipStream_t s, int n, cd* const* A, const cd* const* Aorig, cd* const* W,
                   cd* lam, const EigWorkspace& ws, int* info, double* resid,
                   cd* const* scratch, int batch, hipEvent_t* ev, double tol, int qr_variant) {
  if (batch <= 0 || n <= 0) return;
  const cd* const* Ac = (const cd* const*)A;
  if (n >= 3) {
    if (n >= 128)
      hess_blocked(s, n, A, ws.tau, ws.X, ws.Z, batch);
    else
      hipLaunchKernelGGL(k_hess, dim3(batch), dim3(256), 3 * n * sizeof(cd), s, n, A, ws.tau);
    // blocked explicit Q into Z; X holds the dense reflectors, W the T factors
    // and the two GEMM intermediates until the eigenvectors are formed (W must
    // hold nblk QB^2 + 2 QB n entries: n >= 104 or so, else the unblocked loop)
    const int nref = n - 2, nblk = (nref + QB - 1) / QB;
    if ((long long)nblk * QB * QB + 2LL * QB * n > (long long)n * n) {
      hipLaunchKernelGGL(k_formq, dim3(batch), dim3(256), 2 * n * sizeof(cd), s, n, Ac,
                         (const cd*)ws.tau, ws.Z);
    } else {
      hipLaunchKernelGGL(k_extract_v, dim3(64, batch), dim3(256), 0, s, n, Ac, ws.X);
      hipLaunchKernelGGL(k_larft, dim3(nblk, batch), dim3(256), 0, s, n, (const cd* const*)ws.X,
                         (const cd*)ws.tau, W);
      set_identity_batched(s, n, ws.Z, batch);
      const long long offW1 = (long long)nblk * QB * QB, offW2 = offW1 + (long long)QB * n;
      const cd* const* Xc = (const cd* const*)ws.X;
      const cd* const* Wc = (const cd* const*)W;
      for (int blk = nblk - 1; blk >= 0; --blk) {
        const int j0 = blk * QB, nbk = min(QB, nref - j0), m = n - j0 - 1;
        const long long offV = (long long)(j0 + 1) + (long long)j0 * n;
        const long long offZ = (long long)(j0 + 1) + (long long)(j0 + 1) * n;
        // W1 = V^H Zt   (nbk x m)
        zgemm_batched(s, nbk, m, m, mk(1.0, 0.0), Xc, offV, n, 1, 1, (const cd* const*)ws.Z, offZ, 1,
                      n, 0, mk(0.0, 0.0), W, offW1, 1, QB, batch);
        // W2 = T W1
        zgemm_batched(s, nbk, m, nbk, mk(1.0, 0.0), Wc, (long long)blk * QB * QB, 1, QB, 0, Wc,
                      offW1, 1, QB, 0, mk(0.0, 0.0), W, offW2, 1, QB, batch);
        // Zt -= V W2
        zgemm_batched(s, m, m, nbk, mk(-1.0, 0.0), Xc, offV, 1, n, 0, Wc, offW2, 1, QB, 0,
                      mk(1.0, 0.0), ws.Z, offZ, 1, n, batch);
      }
    }
  } else {
    set_identity_batched(s, n, ws.Z, batch);
  }
  if (ev) (void)hipEventRecord(ev[0], s);
  // zero everything below the first subdiagonal (the Householder vectors)
  hipLaunchKernelGGL(k_zero_below_subdiag, dim3(64, batch), dim3(256), 0, s, n, A);
  switch (qr_variant) {
    case QR_W32:
      hipLaunchKernelGGL(qr_w32::k_qr_ms, dim3(batch), dim3(256), 0, s, n, A, ws.Z, info,
                         40 + 3 * n);
      break;
    case QR_W64:
      hipLaunchKernelGGL(qr_w64::k_qr_ms, dim3(batch), dim3(256), 0, s, n, A, ws.Z, info,
                         40 + 3 * n);
      break;
    default:
      hipLaunchKernelGGL(qr_w48::k_qr_ms, dim3(batch), dim3(256), 0, s, n, A, ws.Z, info,
                         40 + 3 * n);
  }
  if (ev) (void)hipEventRecord(ev[1], s);
  if (n >= 64)
    trevc_blocked(s, n, Ac, ws.X, batch);
  else
    hipLaunchKernelGGL(k_trevc, dim3(batch), dim3(256), 0, s, n, Ac, ws.X);
  // W = Z X, X upper triangular (the eigenvectors of T)
  zgemm_batched_upper_b(s, n, n, n, mk(1.0, 0.0), (const cd* const*)ws.Z, 0, 1, n, 0,
                        (const cd* const*)ws.X, 0, 1, n, 0, mk(0.0, 0.0), W, 0, 1, n, batch);
  hipLaunchKernelGGL(k_eig_finish, dim3(batch), dim3(256), 0, s, n, Ac, W, lam);
  if (Aorig) {
    zgemm_batched(s, n, n, n, mk(1.0, 0.0), Aorig, 0, 1, n, 0, (const cd* const*)W, 0, 1, n, 0,
                  mk(0.0, 0.0), scratch, 0, 1, n, batch);
    hipLaunchKernelGGL(k_eig_residual, dim3(batch), dim3(256), 0, s, n, Aorig,
                       (const cd* const*)W, (const cd* const*)scratch, (const cd*)lam, resid,
                       info, tol);
  }
  if (ev) (void)hipEventRecord(ev[2], s);
}

}  // namespace orcwa
