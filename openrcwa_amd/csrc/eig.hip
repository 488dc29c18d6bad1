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
  __shared__ cd Ts[HB][HB + 1];
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
        for (int q = 0; q <= tid; ++q) acc += conj(Ts[q][tid]) * s_w[q];
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
      for (int q = tid; q < j; ++q) acc += Ts[tid][q] * s_t[q];
      Ts[tid][j] = -(t * acc);
    }
    if (tid == j) Ts[j][j] = t;
    __syncthreads();
  }
  for (int e = tid; e < HB * HB; e += 256) {
    const int r = e % HB, q = e / HB;
    Tg[e] = (r <= q && q < nb) ? Ts[r][q] : mk(0.0, 0.0);
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
// 3b. windowed small-bulge multishift QR (one workgroup per matrix).
//
// A sweep chases QNS single-shift bulges (2x2 unitaries G, as above) as a
// train through the active block [l, hi].  The chase runs inside a QW x QW
// window held in LDS; the window's rotations are accumulated into a QW x QW
// unitary U (LDS) and applied afterwards, with FP64 MFMA, to the rows right of
// the window, the columns above it and the Schur vectors Z.  A bulge may
// process pair p only when the bulge ahead of it has already processed
// p + 2 or later; with that spacing the interleaved train is the same
// similarity as QNS consecutive single-shift sweeps.  Blocks of at most QW
// rows are finished entirely in LDS.  Shifts are the eigenvalues of the
// trailing QNS x QNS submatrix.
// ---------------------------------------------------------------------------

// A 48-wide window with 12 shifts keeps the LDS of one workgroup under 80 KB,
// so two matrices run on each CU and one's latency-bound chase and window
// Schur overlap the other's MFMA updates (1.28x over a 64-wide window with 16
// shifts at one workgroup per CU on n = 450, though each matrix takes more
// windows and sweeps).
#ifndef ORCWA_QR_W
#define ORCWA_QR_W 48
#endif
#ifndef ORCWA_QR_NS
#define ORCWA_QR_NS 12
#endif
constexpr int QW = ORCWA_QR_W;   // window (a multiple of 16)
constexpr int QLD = QW + 1;   // padded LDS leading dimension (bank spread)
constexpr int QT = QW / 16;   // 16-wide blocks of the window
constexpr int QNS = ORCWA_QR_NS;
static_assert(QW % 16 == 0 && QW <= 64 && QNS <= 64, "QR window");
constexpr int QSCHED = QNS * QW;
#ifndef ORCWA_AED_NW
#define ORCWA_AED_NW 24
#endif
constexpr int AED_NW = ORCWA_AED_NW;   // early-deflation window

struct QrShared {
  cd Hw[QLD * QW];      // window / small block (column-major, ld QLD)
  cd U[QLD * QW];       // accumulated unitary; trailing block for the shifts
  short sched_b[QSCHED];
  short sched_p[QSCHED];
  short pass_start[QSCHED + 1];
  cd rot_c[QNS], rot_s[QNS];
  int rot_r[QNS];
  cd shifts[QNS];
  int bp[QNS];         // next pair of each bulge (-1: not introduced)
  int bdone[QNS];
  cd c, s, sh;
  cd hv[QW];           // Householder vector (early deflation)
  cd spike[QW];        // early-deflation spike, reduced to its first entry
  int nsched, npass, l, ok, ired[4];
};

__device__ inline void lds_rot_left(cd* B, int r, int j0, int j1, cd c, cd s) {
  for (int j = j0 + (int)threadIdx.x; j < j1; j += 256) {
    const cd a = B[r + j * QLD], b = B[r + 1 + j * QLD];
    B[r + j * QLD] = cmul(c, a) + cmul(s, b);
    B[r + 1 + j * QLD] = c * b - s * a;
  }
}
__device__ inline void lds_rot_right(cd* B, int col, int i0, int i1, cd c, cd s) {
  for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
    const cd p = B[i + col * QLD], q = B[i + (col + 1) * QLD];
    B[i + col * QLD] = p * c + q * s;
    B[i + (col + 1) * QLD] = q * conj(c) - p * conj(s);
  }
}

__device__ inline bool small_subdiag_lds(const cd* B, int k, int lo, int hi, double smlnum) {
  const cd h = B[k + (k - 1) * QLD];
  const double ah = cabs1(h);
  if (ah <= smlnum) return true;
  const cd hkk = B[k + k * QLD], hk1 = B[(k - 1) + (k - 1) * QLD];
  double tst = cabs1(hk1) + cabs1(hkk);
  if (tst == 0.0) {
    if (k - 2 >= lo) tst += cabs1(B[(k - 1) + (k - 2) * QLD]);
    if (k + 1 <= hi) tst += cabs1(B[(k + 1) + k * QLD]);
  }
  if (ah <= kUlp * tst) {
    const double hup = cabs1(B[(k - 1) + k * QLD]);
    const double ab = fmax(ah, hup), ba = fmin(ah, hup);
    const double d2 = cabs1(hk1 - hkk);
    const double aa = fmax(cabs1(hkk), d2), bb = fmin(cabs1(hkk), d2);
    const double ss = aa + ab;
    if (ba * (ab / ss) <= fmax(smlnum, kUlp * (bb * (aa / ss)))) return true;
  }
  return false;
}

__device__ inline cd wilkinson_lds(const cd* B, int hi) {
  cd t = B[hi + hi * QLD];
  cd u = csqrt_principal(B[(hi - 1) + hi * QLD]) * csqrt_principal(B[hi + (hi - 1) * QLD]);
  double s = cabs1(u);
  if (s != 0.0) {
    cd x = 0.5 * (B[(hi - 1) + (hi - 1) * QLD] - t);
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

// Single-shift QR of the m x m (m <= 64) Hessenberg block B (LDS) to upper
// triangular form, accumulating the rotations into U when wantu.  Run by
// wave 0 only: every lane computes the (uniform) deflation point, shift and
// rotation redundantly, lane j owns column j / row j of the updates, and LDS
// ordering inside the wave is enforced with wavefront-scope fences.

__device__ __forceinline__ bool wave_small_qr(cd* B, cd* U, int m, bool wantu, int maxit) {
  const int lane = threadIdx.x & 63;
  const double smlnum = kSafmin * ((double)m / kUlp);
  int hi = m - 1, its = 0, tot = 0;
  while (hi > 0) {
    const bool flag = (lane >= 1 && lane <= hi) ? small_subdiag_lds(B, lane, 0, hi, smlnum) : false;
    const unsigned long long mask = __ballot(flag);
    const int l = mask ? 63 - __clzll(mask) : 0;
    wave_fence();
    if (l > 0 && lane == 0) B[l + (l - 1) * QLD] = mk(0.0, 0.0);
    wave_fence();
    if (l == hi) { --hi; its = 0; continue; }
    ++its;
    if (++tot > maxit) return false;
    cd sh;
    if (its % 30 == 10) sh = mk(0.75 * cabs1(B[(l + 1) + l * QLD]), 0.0) + B[l + l * QLD];
    else if (its % 30 == 20) sh = mk(0.75 * cabs1(B[hi + (hi - 1) * QLD]), 0.0) + B[hi + hi * QLD];
    else sh = wilkinson_lds(B, hi);
    for (int k = l; k < hi; ++k) {
      cd x0, x1;
      if (k == l) { x0 = B[l + l * QLD] - sh; x1 = B[(l + 1) + l * QLD]; }
      else { x0 = B[k + (k - 1) * QLD]; x1 = B[(k + 1) + (k - 1) * QLD]; }
      const double nr = sqrt(norm2(x0) + norm2(x1));
      cd c = mk(1.0, 0.0), sn = mk(0.0, 0.0);
      if (nr != 0.0) {
        const double inv = 1.0 / nr;
        c = inv * x0;
        sn = inv * x1;
      }
      wave_fence();
      if (nr != 0.0 && k > l && lane == 0) {
        B[k + (k - 1) * QLD] = mk(nr, 0.0);
        B[(k + 1) + (k - 1) * QLD] = mk(0.0, 0.0);
      }
      if (lane >= k && lane < m) {
        const cd a = B[k + lane * QLD], b = B[k + 1 + lane * QLD];
        B[k + lane * QLD] = cmul(c, a) + cmul(sn, b);
        B[k + 1 + lane * QLD] = c * b - sn * a;
      }
      if (wantu && lane < m) {
        const cd p = U[lane + k * QLD], q = U[lane + (k + 1) * QLD];
        U[lane + k * QLD] = p * c + q * sn;
        U[lane + (k + 1) * QLD] = q * conj(c) - p * conj(sn);
      }
      wave_fence();
      if (lane <= min(k + 2, hi)) {
        const cd p = B[lane + k * QLD], q = B[lane + (k + 1) * QLD];
        B[lane + k * QLD] = p * c + q * sn;
        B[lane + (k + 1) * QLD] = q * conj(c) - p * conj(sn);
      }
      wave_fence();
    }
  }
  return true;
}

// block-wide wrapper: wave 0 works, the result (ok) is shared
__device__ __forceinline__ bool lds_small_qr(QrShared& S, cd* B, int m, bool wantu, int maxit) {
  __syncthreads();
  if (threadIdx.x < 64) {
    const bool ok = wave_small_qr(B, S.U, m, wantu, maxit);
    if (threadIdx.x == 0) S.ok = ok ? 1 : 0;
  }
  __syncthreads();
  return S.ok != 0;
}

// X <- U^H X for rows [a, a+wa) of global X (ld), columns [c0, c1).  Each
// wave owns 16-column blocks; the block's operand is loaded in two halves of
// 32 rows whose loads are issued together, so two memory latencies are paid
// per block.  Complex products in three real MFMAs (Gauss): with
// conj(U) = a + ib and X = x + iy,
//   P = a (x + y),  Q = (a + b) y,  R = (b - a) x,  re = P - Q,  im = P + R.
__device__ __forceinline__ void ms_apply_uh_rows(const QrShared& S, cd* X, long long ld, int a, int wa, int c0,
                                 int c1) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int HALF = QW / 8;
  for (int j0 = c0 + 16 * wave; j0 < c1; j0 += 64) {
    const int j = j0 + (lane & 15);
    v4d accP[QT], accQ[QT], accR[QT];
#pragma unroll
    for (int t = 0; t < QT; ++t) {
      accP[t] = (v4d){0, 0, 0, 0};
      accQ[t] = (v4d){0, 0, 0, 0};
      accR[t] = (v4d){0, 0, 0, 0};
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      cd cur[HALF];
#pragma unroll
      for (int q = 0; q < HALF; ++q) {
        const int k = 4 * (HALF * h + q) + (lane >> 4);
        cur[q] = (k < wa && j < c1) ? X[(a + k) + (long long)j * ld] : mk(0.0, 0.0);
      }
#pragma unroll
      for (int q = 0; q < HALF; ++q) {
        const int k = 4 * (HALF * h + q) + (lane >> 4);
        const double x = cur[q].re, y = cur[q].im, xy = x + y;
#pragma unroll
        for (int t = 0; t < QT; ++t) {
          const cd u = S.U[k + (16 * t + (lane & 15)) * QLD];   // A[i][k] = conj(U[k][i])
          const double ua = u.re, ub = -u.im;
          accP[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua, xy, accP[t], 0, 0, 0);
          accQ[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua + ub, y, accQ[t], 0, 0, 0);
          accR[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ub - ua, x, accR[t], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < QT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = 16 * t + (lane >> 4) + 4 * r;
        if (i < wa && j < c1)
          X[(a + i) + (long long)j * ld] = mk(accP[t][r] - accQ[t][r], accP[t][r] + accR[t][r]);
      }
  }
}

// Y <- Y U for rows [r0, r1) of global Y (ld), columns [a, a+wa); same
// blocking over 16-row blocks and the same three-product form with
// Y = a + ib, U = x + iy.
__device__ __forceinline__ void ms_apply_u_cols(const QrShared& S, cd* Y, long long ld, int r0, int r1, int a,
                                int wa) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int HALF = QW / 8;
  for (int i0 = r0 + 16 * wave; i0 < r1; i0 += 64) {
    const int i = i0 + (lane & 15);
    v4d accP[QT], accQ[QT], accR[QT];
#pragma unroll
    for (int t = 0; t < QT; ++t) {
      accP[t] = (v4d){0, 0, 0, 0};
      accQ[t] = (v4d){0, 0, 0, 0};
      accR[t] = (v4d){0, 0, 0, 0};
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      cd cur[HALF];
#pragma unroll
      for (int q = 0; q < HALF; ++q) {
        const int k = 4 * (HALF * h + q) + (lane >> 4);
        cur[q] = (k < wa && i < r1) ? Y[i + (long long)(a + k) * ld] : mk(0.0, 0.0);
      }
#pragma unroll
      for (int q = 0; q < HALF; ++q) {
        const int k = 4 * (HALF * h + q) + (lane >> 4);
        const double ya = cur[q].re, yb = cur[q].im, yab = ya + yb, yba = yb - ya;
#pragma unroll
        for (int t = 0; t < QT; ++t) {
          const cd u = S.U[k + (16 * t + (lane & 15)) * QLD];   // B[k][j] = U[k][j]
          accP[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ya, u.re + u.im, accP[t], 0, 0, 0);
          accQ[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(yab, u.im, accQ[t], 0, 0, 0);
          accR[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(yba, u.re, accR[t], 0, 0, 0);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < QT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ii = i0 + (lane >> 4) + 4 * r;
        const int jj = 16 * t + (lane & 15);
        if (ii < r1 && jj < wa)
          Y[ii + (long long)(a + jj) * ld] = mk(accP[t][r] - accQ[t][r], accP[t][r] + accR[t][r]);
      }
  }
}

// load H[a:a+wa, a:a+wa] into S.Hw and reset S.U to the identity (zero padded)
__device__ __forceinline__ void ms_load_window(QrShared& S, const cd* H, int n, int a, int wa) {
  // QW*QW entries over 256 threads, loads issued in two batches
  constexpr int PER = (QW * QW + 511) / 512;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    cd v[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = threadIdx.x + 256 * (PER * h + t);
      const int i = e % QW, j = e / QW;
      v[t] = (e < QW * QW && i < wa && j < wa) ? H[(a + i) + (long long)(a + j) * n]
                                              : mk(0.0, 0.0);
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int e = threadIdx.x + 256 * (PER * h + t);
      const int i = e % QW, j = e / QW;
      if (e < QW * QW) {
        S.Hw[i + j * QLD] = v[t];
        S.U[i + j * QLD] = (i == j && i < wa) ? mk(1.0, 0.0) : mk(0.0, 0.0);
      }
    }
  }
}

__device__ __forceinline__ void ms_store_window(const QrShared& S, cd* H, int n, int a, int wa) {
  for (int e = threadIdx.x; e < QW * QW; e += 256) {
    const int i = e % QW, j = e / QW;
    if (i < wa && j < wa) H[(a + i) + (long long)(a + j) * n] = S.Hw[i + j * QLD];
  }
}

// the window's deferred updates: rows right, columns above, Schur vectors
__device__ __forceinline__ void ms_apply_window(const QrShared& S, cd* H, cd* Z, int n, int a, int wa) {
  if (a + wa < n) ms_apply_uh_rows(S, H, n, a, wa, a + wa, n);
  if (a > 0) ms_apply_u_cols(S, H, n, 0, a, a, wa);
  ms_apply_u_cols(S, Z, n, 0, n, a, wa);
}

// Householder reflector of x (zlarfg): H^H x = beta e1, H = I - tau v v^H, v0 = 1.
// Wave 0, lanes own the entries x[i] = B[(r0 + i) + col * QLD], i < len; the
// reflector is written to hv and the column to (beta, 0, ...).
__device__ inline cd wave_house_col(cd* B, int r0, int col, int len, cd* hv) {
  const int lane = threadIdx.x & 63;
  const cd alpha = B[r0 + col * QLD];
  double p = (lane >= 1 && lane < len) ? norm2(B[(r0 + lane) + col * QLD]) : 0.0;
  const double xn2 = wave_sum(p);
  cd tau = mk(0.0, 0.0), scale = mk(0.0, 0.0), beta = alpha;
  if (!(xn2 == 0.0 && alpha.im == 0.0)) {
    const double nrm = sqrt(norm2(alpha) + xn2);
    const double bt = (alpha.re >= 0.0) ? -nrm : nrm;
    tau = mk((bt - alpha.re) / bt, -alpha.im / bt);
    scale = cinv(alpha - mk(bt, 0.0));
    beta = mk(bt, 0.0);
  }
  if (lane < len) hv[lane] = (lane == 0) ? mk(1.0, 0.0) : B[(r0 + lane) + col * QLD] * scale;
  wave_fence();
  if (lane < len) B[(r0 + lane) + col * QLD] = (lane == 0) ? beta : mk(0.0, 0.0);
  wave_fence();
  return tau;
}

// T(r0:r0+len, c0:nw) <- (I - tau v v^H)^H T(...) and T(0:nw, r0:r0+len), V(...) <- (...)(I - tau v v^H)
__device__ inline void wave_house_apply(cd* T, cd* V, int nw, int r0, int len, int c0, cd tau,
                                        const cd* hv) {
  const int lane = threadIdx.x & 63;
  if (lane >= c0 && lane < nw) {
    cd d = mk(0.0, 0.0);
    for (int i = 0; i < len; ++i) d += cmul(hv[i], T[(r0 + i) + lane * QLD]);
    const cd f = conj(tau) * d;
    for (int i = 0; i < len; ++i) T[(r0 + i) + lane * QLD] -= hv[i] * f;
  }
  wave_fence();
  if (lane < nw) {
    cd d = mk(0.0, 0.0), e = mk(0.0, 0.0);
    for (int j = 0; j < len; ++j) {
      d += T[lane + (r0 + j) * QLD] * hv[j];
      e += V[lane + (r0 + j) * QLD] * hv[j];
    }
    const cd fd = tau * d, fe = tau * e;
    for (int j = 0; j < len; ++j) {
      T[lane + (r0 + j) * QLD] -= fd * conj(hv[j]);
      V[lane + (r0 + j) * QLD] -= fe * conj(hv[j]);
    }
  }
  wave_fence();
}

// Aggressive early deflation (zlaqr3 mathematics, without its reordering) on
// the trailing nw x nw window [kw, hi] of the active block [l, hi], kw > l:
// the window's Schur form T = V^H H_w V in LDS; its spike s conj(V(0, :))
// (s = H(kw, kw-1)) is tested from the bottom up to the first undeflatable
// eigenvalue; the undeflated part is returned to Hessenberg form and V is
// applied to the rest of H and to Z.
// Returns the number of deflated eigenvalues, or -1 when the window's QR did
// not converge (H untouched).  S.ok = 1 and S.shifts = the bottom QNS
// undeflated eigenvalues when there are at least QNS of them.
__device__ int aed_window(QrShared& S, cd* H, cd* Z, int n, int hi, int nw, double smlnum) {
  const int kw = hi - nw + 1;
  const int tid = threadIdx.x, lane = tid & 63;
  ms_load_window(S, H, n, kw, nw);
  if (!lds_small_qr(S, S.Hw, nw, true, 30 * nw + 100)) return -1;
  const cd s = H[kw + (long long)(kw - 1) * n];
  cd* T = S.Hw;
  cd* V = S.U;
  if (tid < 64) {
    // trailing eigenvalues whose spike entry is negligible deflate; the first
    // undeflatable one from the bottom ends the test (no reordering)
    const double as = cabs1(s);
    int ns = nw;
    while (ns > 0) {
      double foo = cabs1(T[(ns - 1) + (ns - 1) * QLD]);
      if (foo == 0.0) foo = as;
      if (as * cabs1(V[(ns - 1) * QLD]) <= fmax(smlnum, kUlp * foo)) --ns;
      else break;
    }
    if (ns >= QNS) {
      if (lane < QNS) S.shifts[lane] = T[(ns - QNS + lane) + (ns - QNS + lane) * QLD];
    }
    wave_fence();
    const cd sp = (ns == 0) ? mk(0.0, 0.0) : s;
    if (ns < nw || (sp.re == 0.0 && sp.im == 0.0)) {
      if (ns > 1 && !(sp.re == 0.0 && sp.im == 0.0)) {
        // reflect the spike s conj(V(0, 0:ns)) onto its first entry, then
        // return T(0:ns, 0:ns) to Hessenberg form
        if (lane < ns) S.spike[lane] = sp * conj(V[lane * QLD]);
        wave_fence();
        const cd tau = wave_house_col(S.spike, 0, 0, ns, S.hv);
        wave_house_apply(T, V, nw, 0, ns, 0, tau, S.hv);
        for (int k = 0; k + 2 < ns; ++k) {
          const cd t2 = wave_house_col(T, k + 1, k, ns - k - 1, S.hv);
          wave_house_apply(T, V, nw, k + 1, ns - k - 1, k + 1, t2, S.hv);
        }
      }
      if (lane == 0) S.c = sp * conj(V[0]);   // new spike H(kw, kw-1)
      if (lane == 0) S.nsched = ns;
    } else if (lane == 0) {
      S.nsched = -nw;    // nothing deflated: H unchanged
    }
    if (lane == 0) S.ok = (ns >= QNS) ? 1 : 0;
  }
  __syncthreads();
  const int code = S.nsched;
  if (code < 0) return 0;
  ms_store_window(S, H, n, kw, nw);
  if (tid == 0) H[kw + (long long)(kw - 1) * n] = S.c;
  ms_apply_window(S, H, Z, n, kw, nw);
  __syncthreads();
  return nw - code;
}

#ifndef ORCWA_QR_WG_PER_CU
#define ORCWA_QR_WG_PER_CU 2
#endif
static_assert(sizeof(QrShared) * ORCWA_QR_WG_PER_CU <= 160 * 1024, "QR LDS per CU");
__global__ __launch_bounds__(256, ORCWA_QR_WG_PER_CU) void k_qr_ms(int n, cd* const* Hp, cd* const* Zp, int* info,
                                               int maxsweep) {
  __shared__ QrShared S;
  const int b = blockIdx.x;
  cd* __restrict__ H = Hp[b];
  cd* __restrict__ Z = Zp[b];
  const int tid = threadIdx.x;
  const double smlnum = kSafmin * ((double)n / kUlp);
  int hi = n - 1, sweeps = 0, noprog = 0;
  while (hi >= 0) {
    int best = 0;
    for (int k = 1 + tid; k <= hi; k += 256)
      if (small_subdiag(H, n, k, 0, hi, smlnum)) best = max(best, k);
    for (int off = 32; off > 0; off >>= 1) best = max(best, __shfl_xor(best, off, 64));
    if ((tid & 63) == 0) S.ired[tid >> 6] = best;
    __syncthreads();
    const int l = max(max(S.ired[0], S.ired[1]), max(S.ired[2], S.ired[3]));
    __syncthreads();
    if (l > 0 && tid == 0) H[l + (long long)(l - 1) * n] = mk(0.0, 0.0);
    __syncthreads();
    if (l == hi) { --hi; noprog = 0; continue; }
    const int m = hi - l + 1;
    if (m <= QW) {
      ms_load_window(S, H, n, l, m);
      __syncthreads();
      const bool ok = lds_small_qr(S, S.Hw, m, true, 30 * m + 100);
      ms_store_window(S, H, n, l, m);
      __syncthreads();
      ms_apply_window(S, H, Z, n, l, m);
      __syncthreads();
      if (!ok) { if (tid == 0) info[b] |= ORCWA_STATUS_NOCONV; return; }
      hi = l - 1;
      noprog = 0;
      continue;
    }
    // aggressive early deflation of the trailing window; a sweep follows
    // unless it deflated more than 14 % of the window
    bool aed_shifts = false;
    {
      const int nd = aed_window(S, H, Z, n, hi, AED_NW, smlnum);
      if (nd > 0) {
        hi -= nd;
        noprog = 0;
        if (100 * nd > 14 * AED_NW || hi - l + 1 <= QW) continue;
      }
      aed_shifts = (nd >= 0) && (S.ok != 0);
    }
    // shifts: the bottom undeflated eigenvalues of the early-deflation
    // window, else the eigenvalues of the trailing QNS x QNS block (in S.U)
    constexpr int ns = QNS;
    if (!aed_shifts) {
      for (int e = tid; e < QW * QW; e += 256) {
        const int i = e % QW, j = e / QW;
        cd v = mk(0.0, 0.0);
        if (i < ns && j < ns && i <= j + 1)
          v = H[(hi - ns + 1 + i) + (long long)(hi - ns + 1 + j) * n];
        S.U[i + j * QLD] = v;
      }
      __syncthreads();
      lds_small_qr(S, S.U, ns, false, 30 * ns + 100);
    }
    ++noprog;
    if (tid < ns) {
      cd sh = aed_shifts ? S.shifts[tid] : S.U[tid + tid * QLD];
      if (noprog % 6 == 0) {
        const double pert = 0.75 * cabs1(H[hi + (long long)(hi - 1) * n]);
        double sn, cs;
        sincos(2.0 * 3.14159265358979 * tid / ns, &sn, &cs);
        sh = sh + mk(pert * cs, pert * sn);
      }
      S.shifts[tid] = sh;
      S.bp[tid] = -1;
      S.bdone[tid] = 0;
    }
    __syncthreads();
    if (++sweeps > maxsweep) { if (tid == 0) info[b] |= ORCWA_STATUS_NOCONV; return; }
    int a = l;
    while (true) {
      const int bend = min(a + QW, hi + 1);
      const int wa = bend - a;
      ms_load_window(S, H, n, a, wa);
      if (tid < 64) {
        // Schedule of (bulge, pair) moves inside this window, grouped in
        // passes (the moves of one pass touch disjoint row and column pairs).
        // Lane i owns bulge i.  Within a pass bulge i moves iff it can move
        // on its own (A) or it may follow (B) and bulge i-1 moved: a boolean
        // carry chain solved on ballot masks.  A follower's spacing to its
        // predecessor is >= 2 pairs, so "predecessor moved" always suffices.
        const int lane = tid;
        const int lim = (bend - 1 < hi) ? bend - 3 : hi - 1;
        int bpv = (lane < QNS) ? S.bp[lane] : 0;
        int bdv = (lane < QNS) ? S.bdone[lane] : 1;
        int cnt = 0, np = 0;
        while (cnt + QNS <= QSCHED) {
          const int pbp = __shfl_up(bpv, 1, 64), pbd = __shfl_up(bdv, 1, 64);
          const bool first = (lane == 0);
          const bool intro = (bpv < 0);
          const bool own = !bdv && (intro ? (a == l) : (bpv <= lim && bpv >= a + 1));
          bool A, B;
          if (intro) {
            A = first || pbd || pbp >= l + 3;
            B = !first && pbp >= 0 && pbp + 1 >= l + 3;
          } else {
            A = first || pbd || bpv <= pbp - 3;
            B = true;
          }
          const unsigned long long mA = __ballot(own && A), mB = __ballot(own && B);
          unsigned long long mv = mA;
#pragma unroll
          for (int it = 0; it < QNS; ++it) mv = mA | (mB & (mv << 1));
          if (mv == 0ull) break;
          if ((mv >> lane) & 1ull) {
            const int pos = cnt + __popcll(mv & ((1ull << lane) - 1ull));
            const int pp = intro ? l : bpv;
            S.sched_b[pos] = (short)(intro ? -(lane + 1) : lane + 1);
            S.sched_p[pos] = (short)pp;
            bpv = pp + 1;
            if (bpv >= hi) bdv = 1;
          }
          if (lane == 0) S.pass_start[np] = (short)cnt;
          cnt += __popcll(mv);
          ++np;
        }
        int cand = (lane < QNS && !bdv) ? (bpv < 0 ? l + 1 : bpv) : (1 << 30);
        for (int off = 32; off > 0; off >>= 1) cand = min(cand, __shfl_xor(cand, off, 64));
        if (lane < QNS) { S.bp[lane] = bpv; S.bdone[lane] = bdv; }
        if (lane == 0) {
          S.pass_start[np] = (short)cnt;
          S.nsched = cnt;
          S.npass = np;
          S.l = cand;
        }
      }
      __syncthreads();
      const int cnt = S.nsched, npass = S.npass;
      const int rlim = min(hi - a + 1, wa);
      for (int ps = 0; ps < npass; ++ps) {
        const int e0 = S.pass_start[ps], ne = S.pass_start[ps + 1] - e0;
        if (tid < ne) {
          const int e = e0 + tid;
          const int r = S.sched_p[e] - a;
          const int code = S.sched_b[e];
          cd x0, x1;
          if (code < 0) { x0 = S.Hw[r + r * QLD] - S.shifts[-code - 1]; x1 = S.Hw[(r + 1) + r * QLD]; }
          else { x0 = S.Hw[r + (r - 1) * QLD]; x1 = S.Hw[(r + 1) + (r - 1) * QLD]; }
          const double nr = sqrt(norm2(x0) + norm2(x1));
          cd c = mk(1.0, 0.0), sn = mk(0.0, 0.0);
          if (nr != 0.0) {
            const double inv = 1.0 / nr;
            c = inv * x0;
            sn = inv * x1;
            if (code > 0) { S.Hw[r + (r - 1) * QLD] = mk(nr, 0.0); S.Hw[(r + 1) + (r - 1) * QLD] = mk(0.0, 0.0); }
          }
          S.rot_c[tid] = c;
          S.rot_s[tid] = sn;
          S.rot_r[tid] = r;
        }
        __syncthreads();
        // all left rotations (disjoint row pairs) and the U columns
        for (int idx = tid; idx < ne * QW; idx += 256) {
          const int q = idx / QW, j = idx % QW;
          const int r = S.rot_r[q];
          const cd c = S.rot_c[q], sn = S.rot_s[q];
          if (j >= r && j < wa) {
            const cd x = S.Hw[r + j * QLD], y = S.Hw[r + 1 + j * QLD];
            S.Hw[r + j * QLD] = cmul(c, x) + cmul(sn, y);
            S.Hw[r + 1 + j * QLD] = c * y - sn * x;
          }
          if (j < wa) {
            const cd pu = S.U[j + r * QLD], qu = S.U[j + (r + 1) * QLD];
            S.U[j + r * QLD] = pu * c + qu * sn;
            S.U[j + (r + 1) * QLD] = qu * conj(c) - pu * conj(sn);
          }
        }
        __syncthreads();
        // all right rotations (disjoint column pairs)
        for (int idx = tid; idx < ne * QW; idx += 256) {
          const int q = idx / QW, i = idx % QW;
          const int r = S.rot_r[q];
          if (i < min(r + 3, rlim)) {
            const cd c = S.rot_c[q], sn = S.rot_s[q];
            const cd pv = S.Hw[i + r * QLD], qv = S.Hw[i + (r + 1) * QLD];
            S.Hw[i + r * QLD] = pv * c + qv * sn;
            S.Hw[i + (r + 1) * QLD] = qv * conj(c) - pv * conj(sn);
          }
        }
        __syncthreads();
      }
      ms_store_window(S, H, n, a, wa);
      ms_apply_window(S, H, Z, n, a, wa);
      __syncthreads();
      const int next = S.l;
      __syncthreads();
      if (next == (1 << 30)) break;
      if (cnt == 0) { if (tid == 0) info[b] |= ORCWA_STATUS_NOCONV; return; }
      a = next - 1;
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
    if (i1 < n)
      zgemm_batched(s, i1 - i0, n - i1, n - i1, mk(-1.0, 0.0), T,
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
                   cd* const* scratch, int batch, hipEvent_t* ev, double tol) {
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
  hipLaunchKernelGGL(k_qr_ms, dim3(batch), dim3(256), 0, s, n, A, ws.Z, info, 40 + 3 * n);
  if (ev) (void)hipEventRecord(ev[1], s);
  if (n >= 64)
    trevc_blocked(s, n, Ac, ws.X, batch);
  else
    hipLaunchKernelGGL(k_trevc, dim3(batch), dim3(256), 0, s, n, Ac, ws.X);
  zgemm_batched(s, n, n, n, mk(1.0, 0.0), (const cd* const*)ws.Z, 0, 1, n, 0,
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
