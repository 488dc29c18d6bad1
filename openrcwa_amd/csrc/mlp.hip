// Learned surrogate: training and inference kernels of a small multilayer
// perceptron ensemble, and the C ABI of include/openrcwa_surrogate.h.
//
// Every product is one batched launch of k_gemm (FP32 MFMA 16x16x4, FP32
// accumulation; the members of the ensemble are the batch) with the bias,
// GELU and GELU-derivative work fused into its epilogue:
//   forward   Z_l = A_{l-1} W_l + b_l,  A_l = GELU(Z_l)      (Z kept for training)
//   output    Y = A_{D-1} W_D + b_D
//   loss      G_D = 2 (Y - T) / (B n_out)
//   backward  dW_l = A_{l-1}^T G_l,  db_l = colsum(G_l),
//             G_{l-1} = (G_l W_l^T) * GELU'(Z_{l-1})
//   update    Adam with a cosine learning-rate schedule
// One training step (gather of the minibatch, forward, loss, backward, Adam,
// step counter) reads its step number from device memory, so it is captured
// once in a hipGraph and replayed.  Inference can use BF16 products
// (MFMA 16x16x32 bf16, FP32 accumulation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "context.h"
#include "../../include/openrcwa_surrogate.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;    // FP32 tile
constexpr int BKH = 32;                     // BF16 tile depth
constexpr int LDA = BM + 4, LDB = BN + 4;   // padded LDS rows (floats)
constexpr int LDH = BKH + 8;                // padded LDS rows of the BF16 tiles (bf16)
constexpr int kMaxRowsPredict = 32768;      // rows per inference pass
constexpr int kMaxGradRows = 65536;

enum Epi { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_BWD = 3 };

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  // d/dz z Phi(z) = Phi(z) + z phi(z)
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * expf(-0.5f * z * z);
}

struct GemmArgs {
  int M, N, K;
  const float* A;              // A(m, k) = A[z*sabat + m*sam + k*sak]
  long long sam, sak, sabat;
  const float* B;              // B(k, n) = B[z*sbbat + k*sbk + n*sbn]
  long long sbk, sbn, sbbat;
  float* C;                    // C[z*scbat + m*ldc + n]
  long long ldc, scbat;
  const float* bias;           // bias[z*sbias + n]
  long long sbias;
  float* Z;                    // EPI_BIAS_GELU: pre-activation, layout of C
  const float* Zp;             // EPI_GELU_BWD: pre-activation of the layer below, layout of C
};

// C = op(A) op(B) with a fused epilogue.  Workgroup of 4 waves on a 64 x 64
// tile, each wave 32 x 32 (2 x 2 MFMA blocks of 16 x 16), K in steps of 16
// through LDS (k-major rows padded by 4), the next step's operands loaded into
// registers while the current one is multiplied.  AK: A is contiguous along k
// (else along m); BN_: B is contiguous along n (else along k).  The
// contiguous extent of each operand is a multiple of 4 and 16-byte aligned
// (checked by the launcher), so operands move as float4.
// MFMA 16x16x4 f32 lane layout: A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k, D[i][j] in lane j + 16 (i / 4), element i % 4.
template <bool AK, bool BNC, int EPI>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g) {
  __shared__ float As[BK][LDA];
  __shared__ float Bs[BK][LDB];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int z = blockIdx.z;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const float* A = g.A + z * g.sabat;
  const float* B = g.B + z * g.sbbat;
  // this thread's load coordinates inside the tile
  const int am = AK ? (t >> 2) : ((t & 15) * 4), ak = AK ? ((t & 3) * 4) : (t >> 4);
  const int bk = BNC ? (t >> 4) : ((t & 3) * 4), bn = BNC ? ((t & 15) * 4) : (t >> 2);
  auto loadA = [&](int k0) -> f4 {
    f4 v = {0.f, 0.f, 0.f, 0.f};
    const int m = m0 + am, k = k0 + ak;
    if (AK) {
      if (m < g.M && k < g.K) v = *(const f4*)(A + m * g.sam + k);
    } else {
      if (m < g.M && k < g.K) v = *(const f4*)(A + k * g.sak + m);
    }
    return v;
  };
  auto loadB = [&](int k0) -> f4 {
    f4 v = {0.f, 0.f, 0.f, 0.f};
    const int k = k0 + bk, n = n0 + bn;
    if (BNC) {
      if (k < g.K && n < g.N) v = *(const f4*)(B + k * g.sbk + n);
    } else {
      if (k < g.K && n < g.N) v = *(const f4*)(B + n * g.sbn + k);
    }
    return v;
  };
  auto storeA = [&](f4 v) {
    if (AK) {
      As[ak + 0][am] = v[0]; As[ak + 1][am] = v[1]; As[ak + 2][am] = v[2]; As[ak + 3][am] = v[3];
    } else {
      *(f4*)&As[ak][am] = v;
    }
  };
  auto storeB = [&](f4 v) {
    if (BNC) {
      *(f4*)&Bs[bk][bn] = v;
    } else {
      Bs[bk + 0][bn] = v[0]; Bs[bk + 1][bn] = v[1]; Bs[bk + 2][bn] = v[2]; Bs[bk + 3][bn] = v[3];
    }
  };

  f4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;

  f4 ra = loadA(0), rb = loadB(0);
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    storeA(ra);
    storeB(rb);
    __syncthreads();
    if (k0 + BK < g.K) {
      ra = loadA(k0 + BK);
      rb = loadB(k0 + BK);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      float av[2], bv[2];
      for (int a = 0; a < 2; ++a) av[a] = As[kk + lk][wm + a * 16 + li];
      for (int b = 0; b < 2; ++b) bv[b] = Bs[kk + lk][wn + b * 16 + li];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }

  float* C = g.C + z * g.scbat;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + a * 16 + lk * 4 + r, n = n0 + wn + b * 16 + li;
        if (m >= g.M || n >= g.N) continue;
        const long long o = (long long)m * g.ldc + n;
        float v = acc[a][b][r];
        if (EPI == EPI_BIAS) {
          v += g.bias[z * g.sbias + n];
        } else if (EPI == EPI_BIAS_GELU) {
          v += g.bias[z * g.sbias + n];
          if (g.Z) g.Z[z * g.scbat + o] = v;
          v = gelu(v);
        } else if (EPI == EPI_GELU_BWD) {
          v *= gelu_grad(g.Zp[z * g.scbat + o]);
        }
        C[o] = v;
      }
}

// Inference product in BF16: C = A B (+ bias, GELU), A contiguous along k
// (activations, rows of K floats), B contiguous along n (weights).  Operands
// are rounded to BF16 when staged into LDS, k-fastest (rows padded by 8), and
// multiplied with MFMA 16x16x32 bf16: A[i][k] in lane i + 16 (k / 8), element
// k % 8; B[k][j] in lane j + 16 (k / 8), element k % 8; D as the FP32 form.
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) __bf16 Ah[BM][LDH];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[BN][LDH];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int z = blockIdx.z;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const float* A = g.A + z * g.sabat;
  const float* B = g.B + z * g.sbbat;
  f4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // The next K step's operands are loaded into registers while the current
  // one is multiplied.  A: two float4 per thread along k (rows of K floats);
  // B: four rows k..k+3 of one column n per thread and q, so each load of a
  // wave reads 64 consecutive floats of a row and each thread stores its four
  // values as one 8-byte run of the k-fastest LDS row.
  f4 ra[2], rb[2];
  auto load = [&](int k0) {
    for (int q = 0; q < 2; ++q) {
      const int idx = t + q * 256, mm = idx >> 3, kq = (idx & 7) * 4;
      ra[q] = (f4){0.f, 0.f, 0.f, 0.f};
      if (m0 + mm < g.M && k0 + kq < g.K) ra[q] = *(const f4*)(A + (m0 + mm) * g.sam + k0 + kq);
    }
    for (int q = 0; q < 2; ++q) {
      const int idx = t + q * 256, nn = idx & 63, kq = (idx >> 6) * 4;
      for (int e = 0; e < 4; ++e)
        rb[q][e] = (k0 + kq + e < g.K && n0 + nn < g.N)
                       ? B[(long long)(k0 + kq + e) * g.sbk + n0 + nn] : 0.f;
    }
  };
  auto stage = [&]() {
    for (int q = 0; q < 2; ++q) {
      const int idx = t + q * 256, mm = idx >> 3, kq = (idx & 7) * 4;
      bf4 h;
      for (int e = 0; e < 4; ++e) h[e] = (__bf16)ra[q][e];
      *(bf4*)&Ah[mm][kq] = h;
    }
    for (int q = 0; q < 2; ++q) {
      const int idx = t + q * 256, nn = idx & 63, kq = (idx >> 6) * 4;
      bf4 h;
      for (int e = 0; e < 4; ++e) h[e] = (__bf16)rb[q][e];
      *(bf4*)&Bh[nn][kq] = h;
    }
  };
  load(0);
  for (int k0 = 0; k0 < g.K; k0 += BKH) {
    stage();
    __syncthreads();
    if (k0 + BKH < g.K) load(k0 + BKH);
    bf8 av[2], bv[2];
    for (int a = 0; a < 2; ++a) av[a] = *(const bf8*)&Ah[wm + a * 16 + li][lk * 8];
    for (int b = 0; b < 2; ++b) bv[b] = *(const bf8*)&Bh[wn + b * 16 + li][lk * 8];
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  float* C = g.C + z * g.scbat;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + a * 16 + lk * 4 + r, n = n0 + wn + b * 16 + li;
        if (m >= g.M || n >= g.N) continue;
        float v = acc[a][b][r] + g.bias[z * g.sbias + n];
        if (EPI == EPI_BIAS_GELU) v = gelu(v);
        C[(long long)m * g.ldc + n] = v;
      }
}

// Minibatch of every member: rows perm[member][step * bs + r] of the training
// set, inputs padded to kin_pad columns and targets to kout_pad with zeros.
// The step comes from device memory (graph replays); the permutation buffer
// alternates by epoch parity.
__global__ void k_gather(const float* X, const float* T, const int* perm0, const int* perm1,
                         long long count, const long long* step, int spe, int bs, int nin,
                         int nin_pad, int nout, int nout_pad, float* Xb, float* Tb) {
  const int r = blockIdx.x, m = blockIdx.y;
  const long long st = *step;
  const int* perm = ((st / spe) & 1) ? perm1 : perm0;
  const long long src = perm[(long long)m * count + (st % spe) * bs + r];
  float* xo = Xb + ((long long)m * bs + r) * nin_pad;
  float* to = Tb + ((long long)m * bs + r) * nout_pad;
  for (int c = threadIdx.x; c < nin_pad; c += blockDim.x) xo[c] = c < nin ? X[src * nin + c] : 0.f;
  for (int c = threadIdx.x; c < nout_pad; c += blockDim.x)
    to[c] = c < nout ? T[src * nout + c] : 0.f;
}

// Rows of x [rows][nin] into [rows][nin_pad] (zero columns beyond nin).
__global__ void k_pad_rows(const float* x, long long rows, int nin, int nin_pad, float* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * nin_pad) return;
  const long long r = i / nin_pad;
  const int c = (int)(i % nin_pad);
  out[i] = c < nin ? x[r * nin + c] : 0.f;
}

// MSE loss and its gradient: G = 2 (Y - T) / (bs nout) on the nout real
// columns (0 on the padding); loss[slot * members + m] += sum (Y - T)^2 / (bs nout)
// where slot = step / spe (the epoch) when step is given, else 0.
__global__ void k_mse(const float* Y, const float* T, int bs, int nout, int nout_pad, float* G,
                      float* loss, const long long* step, int spe, int members) {
  const int m = blockIdx.y;
  const float sc = 1.0f / ((float)bs * (float)nout);
  float part = 0.f;
  const long long base = (long long)m * bs * nout_pad;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < (long long)bs * nout_pad;
       i += (long long)gridDim.x * blockDim.x) {
    const int c = (int)(i % nout_pad);
    float gv = 0.f;
    if (c < nout) {
      const float d = Y[base + i] - T[base + i];
      gv = 2.0f * sc * d;
      part += sc * d * d;
    }
    G[base + i] = gv;
  }
  __shared__ float red[256];
  red[threadIdx.x] = part;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const long long slot = step ? (*step / spe) : 0;
    atomicAdd(&loss[slot * members + m], red[0]);
  }
}

// db[m][n] = sum over rows of G[m][r][n].  A block takes 16 columns of one
// member (a wave reads four 64-byte row segments per load), so the launch
// has N/16 x members blocks; its 16 row groups keep four independent
// partial sums each (loads in flight instead of one dependent add chain),
// combined in a fixed order: the result does not depend on timing.
constexpr int CS_COLS = 16;
__global__ __launch_bounds__(256) void k_colsum(const float* G, int rows, int N,
                                                long long sgbat, float* db, long long sdbat) {
  const int m = blockIdx.y, cl = threadIdx.x % CS_COLS, rg = threadIdx.x / CS_COLS;
  const int c = blockIdx.x * CS_COLS + cl;
  constexpr int RG = 256 / CS_COLS;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  if (c < N) {
    const float* g = G + m * sgbat + c;
    int r = rg;
    for (; r + 3 * RG < rows; r += 4 * RG) {
      s0 += g[(long long)r * N];
      s1 += g[(long long)(r + RG) * N];
      s2 += g[(long long)(r + 2 * RG) * N];
      s3 += g[(long long)(r + 3 * RG) * N];
    }
    for (; r < rows; r += RG) s0 += g[(long long)r * N];
  }
  __shared__ float red[RG][CS_COLS + 1];
  red[rg][cl] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (rg == 0 && c < N) {
    float t = 0.f;
    for (int k = 0; k < RG; ++k) t += red[k][cl];
    db[m * sdbat + c] = t;
  }
}

struct AdamArgs {
  float lr_max, lr_min, b1, b2, eps;
  long long total_steps;
};

// Adam on every parameter of every member; the step number t = *step + 1
// comes from device memory, the learning rate follows the cosine schedule
// lr_min + (lr_max - lr_min) (1 + cos(pi (t - 1) / total)) / 2.
__global__ void k_adam(float* P, const float* G, float* Mo, float* V, long long n,
                       const long long* step, AdamArgs a) {
  const long long t = *step + 1;
  const float frac = (float)(t - 1) / (float)(a.total_steps > 1 ? a.total_steps : 1);
  const float lr = a.lr_min + 0.5f * (a.lr_max - a.lr_min) * (1.0f + cosf(3.14159265358979f * frac));
  const float bc1 = 1.0f - powf(a.b1, (float)t), bc2 = 1.0f - powf(a.b2, (float)t);
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float g = G[i];
    const float m = a.b1 * Mo[i] + (1.0f - a.b1) * g;
    const float v = a.b2 * V[i] + (1.0f - a.b2) * g * g;
    Mo[i] = m;
    V[i] = v;
    P[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + a.eps);
  }
}

__global__ void k_step_inc(long long* step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

__global__ void k_fill(float* p, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_ll(long long* p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0;
}

// Ensemble mean over members and err = max over outputs c of
// scale[c] (max - min over members).
__global__ void k_ensemble(const float* Y, int members, long long rows, int nout, int nout_pad,
                           const float* scale, float* mean, float* err) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  float e = 0.f;
  for (int c = 0; c < nout; ++c) {
    float s = 0.f, lo = 3.4e38f, hi = -3.4e38f;
    for (int m = 0; m < members; ++m) {
      const float v = Y[((long long)m * rows + r) * nout_pad + c];
      s += v;
      lo = fminf(lo, v);
      hi = fmaxf(hi, v);
    }
    mean[r * nout + c] = s / (float)members;
    e = fmaxf(e, scale[c] * (hi - lo));
  }
  err[r] = e;
}

#define MCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      fprintf(stderr, "[openrcwa] HIP error %s at %s:%d: %s\n", hipGetErrorName(_e), __FILE__, \
              __LINE__, #expr);                                                              \
      return ORCWA_ERR_HIP;                                                                  \
    }                                                                                        \
  } while (0)

int fail(const char* msg) {
  fprintf(stderr, "[openrcwa] %s\n", msg);
  return ORCWA_ERR_ARG;
}

long long round_up(long long x, long long m) { return (x + m - 1) / m * m; }

}  // namespace

struct orcwa_mlp {
  orcwa_gpu_context* ctx = nullptr;
  int nin = 0, nin_pad = 0, width = 0, depth = 0, nout = 0, nout_pad = 0, members = 0;
  int precision = ORCWA_MLP_FP32;
  // internal (padded) layout of one member: W_0 [nin_pad][width] (rows >= nin
  // zero), b_0, ..., W_depth [width][nout_pad] (columns >= nout zero), b_depth
  std::vector<long long> offW, offB;
  long long P = 0;                 // internal parameters per member
  long long P_ext = 0;             // the header's layout
  float *d_param = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr;
  // training set
  float *d_X = nullptr, *d_T = nullptr;
  long long count = 0;
  // per-step buffers sized for `bs_cap` rows per member
  int bs_cap = 0;
  float *d_Xb = nullptr, *d_Tb = nullptr, *d_Y = nullptr, *d_G = nullptr;
  std::vector<float*> d_Z, d_A;   // [depth] each [members][bs][width]
  float *d_H0 = nullptr, *d_H1 = nullptr;   // backward ping-pong [members][bs][width]
  float* d_loss = nullptr;
  long long loss_cap = 0;
  long long* d_step = nullptr;
  // inference buffers
  float *d_in = nullptr, *d_inp = nullptr, *d_P0 = nullptr, *d_P1 = nullptr, *d_Yp = nullptr;
  float *d_mean = nullptr, *d_err = nullptr;
  float* d_escale = nullptr;       // [n_out] scale of each output's spread in err
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = -1.0;
};

namespace {

int in_dim(const orcwa_mlp* n, int l) { return l == 0 ? n->nin_pad : n->width; }
int out_dim(const orcwa_mlp* n, int l) { return l == n->depth ? n->nout_pad : n->width; }

template <bool AK, bool BNC, int EPI>
void launch_gemm(hipStream_t s, const GemmArgs& g, int batch) {
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, batch);
  hipLaunchKernelGGL((k_gemm<AK, BNC, EPI>), grid, dim3(256), 0, s, g);
}

template <int EPI>
void launch_gemm_bf16(hipStream_t s, const GemmArgs& g, int batch) {
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, batch);
  hipLaunchKernelGGL((k_gemm_bf16<EPI>), grid, dim3(256), 0, s, g);
}

GemmArgs gemm_args() {
  GemmArgs g;
  std::memset(&g, 0, sizeof(g));
  return g;
}

// forward through every layer for `rows` rows per member.  x: [rows][nin_pad]
// shared by the members (xbat = 0) or per member (xbat = rows * nin_pad).
// train: pre-activations into d_Z, activations into d_A; else the ping-pong
// buffers P0/P1 ([members][rows][width] each) hold the activations.  Output
// into Y [members][rows][nout_pad].
void forward(orcwa_mlp* n, hipStream_t s, const float* x, long long xbat, int rows, bool train,
             float* Y, bool bf16, float* P0 = nullptr, float* P1 = nullptr) {
  const int M = n->members;
  const float* in = x;
  long long inbat = xbat;
  for (int l = 0; l <= n->depth; ++l) {
    GemmArgs g = gemm_args();
    g.M = rows;
    g.N = out_dim(n, l);
    g.K = in_dim(n, l);
    g.A = in;
    g.sam = g.K;
    g.sak = 1;
    g.sabat = inbat;
    g.B = n->d_param + n->offW[l];
    g.sbk = g.N;
    g.sbn = 1;
    g.sbbat = n->P;
    g.bias = n->d_param + n->offB[l];
    g.sbias = n->P;
    g.ldc = g.N;
    g.scbat = (long long)rows * g.N;
    if (l == n->depth) {
      g.C = Y;
      if (bf16) launch_gemm_bf16<EPI_BIAS>(s, g, M);
      else launch_gemm<true, true, EPI_BIAS>(s, g, M);
    } else {
      float* out = train ? n->d_A[l] : ((l & 1) ? P1 : P0);
      g.C = out;
      g.Z = train ? n->d_Z[l] : nullptr;
      if (bf16) launch_gemm_bf16<EPI_BIAS_GELU>(s, g, M);
      else launch_gemm<true, true, EPI_BIAS_GELU>(s, g, M);
      in = out;
      inbat = (long long)rows * n->width;
    }
  }
}

// backward from G (the output gradient, [members][rows][nout_pad]) into d_grad.
// x: the inputs of layer 0 ([members][rows][nin_pad]).
void backward(orcwa_mlp* n, hipStream_t s, const float* x, int rows) {
  const int M = n->members;
  const float* G = n->d_G;
  for (int l = n->depth; l >= 0; --l) {
    const int K = in_dim(n, l), N = out_dim(n, l);
    const float* Ain = l == 0 ? x : n->d_A[l - 1];
    // dW_l = Ain^T G: C [K][N], reduction over rows
    GemmArgs g = gemm_args();
    g.M = K;
    g.N = N;
    g.K = rows;
    g.A = Ain;
    g.sam = 1;
    g.sak = K;
    g.sabat = (long long)rows * K;
    g.B = G;
    g.sbk = N;
    g.sbn = 1;
    g.sbbat = (long long)rows * N;
    g.C = n->d_grad + n->offW[l];
    g.ldc = N;
    g.scbat = n->P;
    launch_gemm<false, true, EPI_STORE>(s, g, M);
    hipLaunchKernelGGL(k_colsum, dim3((N + CS_COLS - 1) / CS_COLS, M), dim3(256), 0, s, G, rows, N,
                       (long long)rows * N, n->d_grad + n->offB[l], n->P);
    if (l > 0) {
      // G_{l-1} = (G W_l^T) * GELU'(Z_{l-1}): C [rows][K]
      float* Gn = (l & 1) ? n->d_H1 : n->d_H0;
      GemmArgs h = gemm_args();
      h.M = rows;
      h.N = K;
      h.K = N;
      h.A = G;
      h.sam = N;
      h.sak = 1;
      h.sabat = (long long)rows * N;
      h.B = n->d_param + n->offW[l];    // B(k = j, n = i) = W_l[i][j]
      h.sbk = 1;
      h.sbn = N;
      h.sbbat = n->P;
      h.C = Gn;
      h.ldc = K;
      h.scbat = (long long)rows * K;
      h.Zp = n->d_Z[l - 1];
      launch_gemm<true, false, EPI_GELU_BWD>(s, h, M);
      G = Gn;
    }
  }
}

void free_step_buffers(orcwa_mlp* n) {
  for (float* p : {n->d_Xb, n->d_Tb, n->d_Y, n->d_G, n->d_H0, n->d_H1})
    if (p) (void)hipFree(p);
  for (float* p : n->d_Z) (void)hipFree(p);
  for (float* p : n->d_A) (void)hipFree(p);
  n->d_Z.clear();
  n->d_A.clear();
  n->d_Xb = n->d_Tb = n->d_Y = n->d_G = n->d_H0 = n->d_H1 = nullptr;
  n->bs_cap = 0;
}

int ensure_step_buffers(orcwa_mlp* n, int bs) {
  if (bs <= n->bs_cap) return ORCWA_OK;
  (void)hipStreamSynchronize(n->ctx->stream);
  free_step_buffers(n);
  const long long M = n->members, H = n->width;
  auto al = [&](float** p, long long cnt) { return hipMalloc(p, std::max(1LL, cnt) * sizeof(float)); };
  if (al(&n->d_Xb, M * bs * n->nin_pad) || al(&n->d_Tb, M * bs * n->nout_pad) ||
      al(&n->d_Y, M * bs * n->nout_pad) || al(&n->d_G, M * bs * n->nout_pad) ||
      al(&n->d_H0, M * bs * H) || al(&n->d_H1, M * bs * H)) {
    free_step_buffers(n);
    return ORCWA_ERR_NOMEM;
  }
  for (int l = 0; l < n->depth; ++l) {
    float *z = nullptr, *a = nullptr;
    if (al(&z, M * bs * H) || al(&a, M * bs * H)) {
      if (z) (void)hipFree(z);
      free_step_buffers(n);
      return ORCWA_ERR_NOMEM;
    }
    n->d_Z.push_back(z);
    n->d_A.push_back(a);
  }
  n->bs_cap = bs;
  return ORCWA_OK;
}

int ensure_loss(orcwa_mlp* n, long long slots) {
  if (slots * n->members <= n->loss_cap) return ORCWA_OK;
  (void)hipStreamSynchronize(n->ctx->stream);
  if (n->d_loss) (void)hipFree(n->d_loss);
  n->d_loss = nullptr;
  n->loss_cap = 0;
  if (hipMalloc(&n->d_loss, slots * n->members * sizeof(float)) != hipSuccess)
    return ORCWA_ERR_NOMEM;
  n->loss_cap = slots * n->members;
  return ORCWA_OK;
}

void fill(hipStream_t s, float* p, long long cnt, float v) {
  const int blocks = (int)std::min<long long>(4096, (cnt + 255) / 256);
  hipLaunchKernelGGL(k_fill, dim3(std::max(1, blocks)), dim3(256), 0, s, p, cnt, v);
}

// external (unpadded) <-> internal (padded) parameter layout of one member
void to_internal(const orcwa_mlp* n, const float* ext, float* in) {
  std::fill(in, in + n->P, 0.f);
  const float* e = ext;
  for (int l = 0; l <= n->depth; ++l) {
    const int K = l == 0 ? n->nin : n->width, N = l == n->depth ? n->nout : n->width;
    const int Np = out_dim(n, l);
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < N; ++j) in[n->offW[l] + (long long)i * Np + j] = *e++;
    for (int j = 0; j < N; ++j) in[n->offB[l] + j] = *e++;
  }
}

void to_external(const orcwa_mlp* n, const float* in, float* ext) {
  float* e = ext;
  for (int l = 0; l <= n->depth; ++l) {
    const int K = l == 0 ? n->nin : n->width, N = l == n->depth ? n->nout : n->width;
    const int Np = out_dim(n, l);
    for (int i = 0; i < K; ++i)
      for (int j = 0; j < N; ++j) *e++ = in[n->offW[l] + (long long)i * Np + j];
    for (int j = 0; j < N; ++j) *e++ = in[n->offB[l] + j];
  }
}

int ensure_infer_buffers(orcwa_mlp* n) {
  if (n->d_P0) return ORCWA_OK;
  const long long M = n->members, R = kMaxRowsPredict;
  auto al = [&](float** p, long long cnt) { return hipMalloc(p, cnt * sizeof(float)); };
  if (al(&n->d_in, R * n->nin) || al(&n->d_inp, R * n->nin_pad) ||
      al(&n->d_P0, M * R * n->width) || al(&n->d_P1, M * R * n->width) ||
      al(&n->d_Yp, M * R * n->nout_pad) || al(&n->d_mean, R * n->nout) || al(&n->d_err, R))
    return ORCWA_ERR_NOMEM;
  return ORCWA_OK;
}

// one inference pass over `rows` <= kMaxRowsPredict rows already padded in d_inp
void infer_pass(orcwa_mlp* n, hipStream_t s, int rows) {
  forward(n, s, n->d_inp, 0, rows, false, n->d_Yp, n->precision == ORCWA_MLP_BF16, n->d_P0,
          n->d_P1);
}

}  // namespace

extern "C" {

int orcwa_mlp_create(orcwa_gpu_context* ctx, int n_in, int width, int depth, int n_out,
                     int members, unsigned long long seed, orcwa_mlp** net) {
  if (!ctx || !net) return fail("orcwa_mlp_create: null argument");
  *net = nullptr;
  if (n_in < 1 || n_in > 64 || width < 32 || width > 1024 || width % 32 || depth < 1 ||
      depth > 8 || n_out < 1 || n_out > 64 || members < 1 || members > 16)
    return fail("orcwa_mlp_create: sizes out of range");
  MCHK(hipSetDevice(ctx->device));
  orcwa_mlp* n = new orcwa_mlp;
  n->ctx = ctx;
  n->nin = n_in;
  n->nin_pad = (int)round_up(n_in, 16);
  n->width = width;
  n->depth = depth;
  n->nout = n_out;
  n->nout_pad = (int)round_up(n_out, 32);
  n->members = members;
  long long off = 0;
  for (int l = 0; l <= depth; ++l) {
    n->offW.push_back(off);
    off += (long long)in_dim(n, l) * out_dim(n, l);
    n->offB.push_back(off);
    off += out_dim(n, l);
  }
  n->P = off;
  n->P_ext = 0;
  for (int l = 0; l <= depth; ++l) {
    const int K = l == 0 ? n_in : width, N = l == depth ? n_out : width;
    n->P_ext += (long long)K * N + N;
  }
  const long long tot = n->P * members;
  if (hipMalloc(&n->d_param, tot * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->d_grad, tot * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->d_m, tot * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->d_v, tot * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->d_step, sizeof(long long)) != hipSuccess ||
      hipMalloc(&n->d_escale, 64 * sizeof(float)) != hipSuccess ||
      hipEventCreate(&n->ev0) != hipSuccess || hipEventCreate(&n->ev1) != hipSuccess) {
    orcwa_mlp_destroy(n);
    return ORCWA_ERR_NOMEM;
  }
  // He initialisation, zero biases, from one seeded generator
  std::mt19937_64 rng(seed);
  std::vector<float> ext((size_t)n->P_ext), in((size_t)n->P);
  std::vector<float> all((size_t)tot);
  for (int m = 0; m < members; ++m) {
    float* e = ext.data();
    for (int l = 0; l <= depth; ++l) {
      const int K = l == 0 ? n_in : width, N = l == depth ? n_out : width;
      std::normal_distribution<float> nd(0.f, std::sqrt(2.0f / (float)K));
      for (long long i = 0; i < (long long)K * N; ++i) *e++ = nd(rng);
      for (int j = 0; j < N; ++j) *e++ = 0.f;
    }
    to_internal(n, ext.data(), in.data());
    std::copy(in.begin(), in.end(), all.begin() + (size_t)m * n->P);
  }
  hipStream_t s = ctx->stream;
  const std::vector<float> ones(64, 1.0f);
  if (hipMemcpyAsync(n->d_param, all.data(), tot * sizeof(float), hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipMemcpyAsync(n->d_escale, ones.data(), 64 * sizeof(float), hipMemcpyHostToDevice, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    orcwa_mlp_destroy(n);
    return ORCWA_ERR_HIP;
  }
  ctx->live_plans++;
  *net = n;
  return ORCWA_OK;
}

int orcwa_mlp_destroy(orcwa_mlp* n) {
  if (!n) return ORCWA_OK;
  (void)hipSetDevice(n->ctx->device);
  (void)hipStreamSynchronize(n->ctx->stream);
  free_step_buffers(n);
  for (float* p : {n->d_param, n->d_grad, n->d_m, n->d_v, n->d_X, n->d_T, n->d_loss, n->d_in,
                   n->d_inp, n->d_P0, n->d_P1, n->d_Yp, n->d_mean, n->d_err})
    if (p) (void)hipFree(p);
  if (n->d_step) (void)hipFree(n->d_step);
  if (n->d_escale) (void)hipFree(n->d_escale);
  if (n->ev0) (void)hipEventDestroy(n->ev0);
  if (n->ev1) (void)hipEventDestroy(n->ev1);
  if (n->d_param) n->ctx->live_plans--;   // counted once creation succeeded
  delete n;
  return ORCWA_OK;
}

long long orcwa_mlp_param_count(const orcwa_mlp* n) { return n ? n->P_ext : -1; }

int orcwa_mlp_get_params(orcwa_mlp* n, float* params) {
  if (!n || !params) return fail("orcwa_mlp_get_params: null argument");
  MCHK(hipSetDevice(n->ctx->device));
  std::vector<float> all((size_t)(n->P * n->members));
  MCHK(hipMemcpyAsync(all.data(), n->d_param, all.size() * sizeof(float), hipMemcpyDeviceToHost,
                      n->ctx->stream));
  MCHK(hipStreamSynchronize(n->ctx->stream));
  for (int m = 0; m < n->members; ++m)
    to_external(n, all.data() + (size_t)m * n->P, params + (size_t)m * n->P_ext);
  return ORCWA_OK;
}

int orcwa_mlp_set_params(orcwa_mlp* n, const float* params) {
  if (!n || !params) return fail("orcwa_mlp_set_params: null argument");
  MCHK(hipSetDevice(n->ctx->device));
  std::vector<float> all((size_t)(n->P * n->members));
  for (int m = 0; m < n->members; ++m)
    to_internal(n, params + (size_t)m * n->P_ext, all.data() + (size_t)m * n->P);
  MCHK(hipMemcpyAsync(n->d_param, all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice,
                      n->ctx->stream));
  MCHK(hipStreamSynchronize(n->ctx->stream));
  return ORCWA_OK;
}

int orcwa_mlp_set_data(orcwa_mlp* n, const float* x, const float* t, long long count) {
  if (!n || !x || !t || count < 1) return fail("orcwa_mlp_set_data: bad argument");
  if (count > (1LL << 30)) return fail("orcwa_mlp_set_data: at most 2^30 rows");
  MCHK(hipSetDevice(n->ctx->device));
  hipStream_t s = n->ctx->stream;
  MCHK(hipStreamSynchronize(s));
  if (n->d_X) (void)hipFree(n->d_X);
  if (n->d_T) (void)hipFree(n->d_T);
  n->d_X = n->d_T = nullptr;
  n->count = 0;
  if (hipMalloc(&n->d_X, count * n->nin * sizeof(float)) != hipSuccess ||
      hipMalloc(&n->d_T, count * n->nout * sizeof(float)) != hipSuccess)
    return ORCWA_ERR_NOMEM;
  MCHK(hipMemcpyAsync(n->d_X, x, count * n->nin * sizeof(float), hipMemcpyHostToDevice, s));
  MCHK(hipMemcpyAsync(n->d_T, t, count * n->nout * sizeof(float), hipMemcpyHostToDevice, s));
  MCHK(hipStreamSynchronize(s));
  n->count = count;
  return ORCWA_OK;
}

int orcwa_mlp_set_error_scale(orcwa_mlp* n, const float* scale) {
  if (!n || !scale) return fail("orcwa_mlp_set_error_scale: null argument");
  for (int c = 0; c < n->nout; ++c)
    if (!(scale[c] >= 0.0f)) return fail("orcwa_mlp_set_error_scale: scales must be >= 0");
  MCHK(hipSetDevice(n->ctx->device));
  MCHK(hipMemcpyAsync(n->d_escale, scale, n->nout * sizeof(float), hipMemcpyHostToDevice,
                      n->ctx->stream));
  MCHK(hipStreamSynchronize(n->ctx->stream));
  return ORCWA_OK;
}

int orcwa_mlp_set_precision(orcwa_mlp* n, int precision) {
  if (!n || (precision != ORCWA_MLP_FP32 && precision != ORCWA_MLP_BF16))
    return fail("orcwa_mlp_set_precision: bad argument");
  n->precision = precision;
  return ORCWA_OK;
}

int orcwa_mlp_train(orcwa_mlp* n, const orcwa_mlp_train_opts* o, float* loss) {
  if (!n || !o) return fail("orcwa_mlp_train: null argument");
  if (!n->d_X) return fail("orcwa_mlp_train: no training set (orcwa_mlp_set_data)");
  const int bs = o->batch;
  if (bs < 16 || bs > 65536 || bs > n->count || o->epochs < 1)
    return fail("orcwa_mlp_train: batch must be 16..65536 and at most the training rows");
  MCHK(hipSetDevice(n->ctx->device));
  hipStream_t s = n->ctx->stream;
  int rc = ensure_step_buffers(n, bs);
  if (rc) return rc;
  if ((rc = ensure_loss(n, o->epochs))) return rc;
  const int M = n->members;
  const long long cnt = n->count;
  const int spe = (int)(cnt / bs);
  const long long total = (long long)spe * o->epochs;
  AdamArgs aa;
  aa.lr_max = o->lr_max;
  aa.lr_min = o->lr_min;
  aa.b1 = o->beta1 > 0 ? o->beta1 : 0.9f;
  aa.b2 = o->beta2 > 0 ? o->beta2 : 0.999f;
  aa.eps = o->eps > 0 ? o->eps : 1e-8f;
  aa.total_steps = total;
  // device permutations (two, alternating by epoch) and their pinned staging
  int *d_perm[2] = {nullptr, nullptr}, *h_perm[2] = {nullptr, nullptr};
  hipEvent_t copied[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(s);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int i = 0; i < 2; ++i) {
      if (d_perm[i]) (void)hipFree(d_perm[i]);
      if (h_perm[i]) (void)hipHostFree(h_perm[i]);
      if (copied[i]) (void)hipEventDestroy(copied[i]);
    }
  };
  for (int i = 0; i < 2; ++i) {
    if (hipMalloc(&d_perm[i], (size_t)M * cnt * sizeof(int)) != hipSuccess ||
        hipHostMalloc(&h_perm[i], (size_t)M * cnt * sizeof(int), 0) != hipSuccess ||
        hipEventCreateWithFlags(&copied[i], hipEventDisableTiming) != hipSuccess) {
      cleanup();
      return ORCWA_ERR_NOMEM;
    }
  }
  const long long tot = n->P * M;
  fill(s, n->d_m, tot, 0.f);
  fill(s, n->d_v, tot, 0.f);
  fill(s, n->d_loss, (long long)o->epochs * M, 0.f);
  hipLaunchKernelGGL(k_fill_ll, dim3(1), dim3(64), 0, s, n->d_step, 1LL);

  auto step_kernels = [&]() {
    hipLaunchKernelGGL(k_gather, dim3(bs, M), dim3(32), 0, s, n->d_X, n->d_T, d_perm[0],
                       d_perm[1], cnt, n->d_step, spe, bs, n->nin, n->nin_pad, n->nout,
                       n->nout_pad, n->d_Xb, n->d_Tb);
    forward(n, s, n->d_Xb, (long long)bs * n->nin_pad, bs, true, n->d_Y, false);
    hipLaunchKernelGGL(k_mse, dim3(std::min(1024, (bs * n->nout_pad + 255) / 256), M), dim3(256),
                       0, s, n->d_Y, n->d_Tb, bs, n->nout, n->nout_pad, n->d_G, n->d_loss,
                       n->d_step, spe, M);
    backward(n, s, n->d_Xb, bs);
    hipLaunchKernelGGL(k_adam, dim3(2048), dim3(256), 0, s, n->d_param, n->d_grad, n->d_m,
                       n->d_v, tot, n->d_step, aa);
    hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, n->d_step);
  };

  if (o->use_graph) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      cleanup();
      return ORCWA_ERR_HIP;
    }
    step_kernels();
    if (hipStreamEndCapture(s, &graph) != hipSuccess ||
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGetLastError();
      cleanup();
      fprintf(stderr, "[openrcwa] orcwa_mlp_train: graph capture failed\n");
      return ORCWA_ERR_HIP;
    }
  }

  std::mt19937_64 rng(o->seed);
  std::vector<int> base((size_t)cnt);
  for (long long i = 0; i < cnt; ++i) base[(size_t)i] = (int)i;
  (void)hipEventRecord(n->ev0, s);
  for (int e = 0; e < o->epochs; ++e) {
    const int b = e & 1;
    // the staging buffer of epoch e - 2 must have been copied before it is refilled
    if (e >= 2 && hipEventSynchronize(copied[b]) != hipSuccess) {
      cleanup();
      return ORCWA_ERR_HIP;
    }
    for (int m = 0; m < M; ++m) {
      int* dst = h_perm[b] + (size_t)m * cnt;
      std::memcpy(dst, base.data(), (size_t)cnt * sizeof(int));
      std::shuffle(dst, dst + cnt, rng);
    }
    if (hipMemcpyAsync(d_perm[b], h_perm[b], (size_t)M * cnt * sizeof(int),
                       hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(copied[b], s) != hipSuccess) {
      cleanup();
      return ORCWA_ERR_HIP;
    }
    for (int st = 0; st < spe; ++st) {
      if (exec) {
        if (hipGraphLaunch(exec, s) != hipSuccess) {
          cleanup();
          return ORCWA_ERR_HIP;
        }
      } else {
        step_kernels();
      }
    }
  }
  (void)hipEventRecord(n->ev1, s);
  std::vector<float> lh((size_t)o->epochs * M);
  hipError_t err = hipMemcpyAsync(lh.data(), n->d_loss, lh.size() * sizeof(float),
                                  hipMemcpyDeviceToHost, s);
  if (err == hipSuccess) err = hipStreamSynchronize(s);
  if (err == hipSuccess) err = hipGetLastError();
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, n->ev0, n->ev1);
  n->last_ms = ms;
  cleanup();
  if (err != hipSuccess) {
    fprintf(stderr, "[openrcwa] orcwa_mlp_train: %s\n", hipGetErrorName(err));
    return ORCWA_ERR_HIP;
  }
  if (loss)
    for (int e = 0; e < o->epochs; ++e) {
      double acc = 0.0;
      for (int m = 0; m < M; ++m) acc += lh[(size_t)e * M + m];
      loss[e] = (float)(acc / spe / M);
    }
  return ORCWA_OK;
}

int orcwa_mlp_gradient(orcwa_mlp* n, const float* x, const float* t, int count, float* grad,
                       float* loss) {
  if (!n || !x || !t || !grad || count < 1 || count > kMaxGradRows)
    return fail("orcwa_mlp_gradient: bad argument");
  MCHK(hipSetDevice(n->ctx->device));
  hipStream_t s = n->ctx->stream;
  int rc = ensure_step_buffers(n, count);
  if (rc) return rc;
  if ((rc = ensure_loss(n, 1))) return rc;
  const int M = n->members;
  // every member sees the same rows
  std::vector<float> xb((size_t)M * count * n->nin_pad, 0.f), tb((size_t)M * count * n->nout_pad, 0.f);
  for (int m = 0; m < M; ++m)
    for (int r = 0; r < count; ++r) {
      for (int c = 0; c < n->nin; ++c)
        xb[((size_t)m * count + r) * n->nin_pad + c] = x[(size_t)r * n->nin + c];
      for (int c = 0; c < n->nout; ++c)
        tb[((size_t)m * count + r) * n->nout_pad + c] = t[(size_t)r * n->nout + c];
    }
  MCHK(hipMemcpyAsync(n->d_Xb, xb.data(), xb.size() * sizeof(float), hipMemcpyHostToDevice, s));
  MCHK(hipMemcpyAsync(n->d_Tb, tb.data(), tb.size() * sizeof(float), hipMemcpyHostToDevice, s));
  fill(s, n->d_loss, M, 0.f);
  forward(n, s, n->d_Xb, (long long)count * n->nin_pad, count, true, n->d_Y, false);
  hipLaunchKernelGGL(k_mse, dim3(std::min(1024, (count * n->nout_pad + 255) / 256), M), dim3(256),
                     0, s, n->d_Y, n->d_Tb, count, n->nout, n->nout_pad, n->d_G, n->d_loss,
                     (const long long*)nullptr, 1, M);
  backward(n, s, n->d_Xb, count);
  std::vector<float> g((size_t)(n->P * M));
  std::vector<float> lh((size_t)M);
  MCHK(hipMemcpyAsync(g.data(), n->d_grad, g.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  MCHK(hipMemcpyAsync(lh.data(), n->d_loss, M * sizeof(float), hipMemcpyDeviceToHost, s));
  MCHK(hipStreamSynchronize(s));
  MCHK(hipGetLastError());
  for (int m = 0; m < M; ++m) to_external(n, g.data() + (size_t)m * n->P, grad + (size_t)m * n->P_ext);
  if (loss)
    for (int m = 0; m < M; ++m) loss[m] = lh[(size_t)m];
  return ORCWA_OK;
}

int orcwa_mlp_forward(orcwa_mlp* n, const float* x, long long count, float* y) {
  if (!n || !x || !y || count < 1) return fail("orcwa_mlp_forward: bad argument");
  MCHK(hipSetDevice(n->ctx->device));
  hipStream_t s = n->ctx->stream;
  int rc = ensure_infer_buffers(n);
  if (rc) return rc;
  const int M = n->members;
  std::vector<float> yb;
  for (long long r0 = 0; r0 < count; r0 += kMaxRowsPredict) {
    const int rows = (int)std::min<long long>(kMaxRowsPredict, count - r0);
    MCHK(hipMemcpyAsync(n->d_in, x + r0 * n->nin, (size_t)rows * n->nin * sizeof(float),
                        hipMemcpyHostToDevice, s));
    const long long tot = (long long)rows * n->nin_pad;
    hipLaunchKernelGGL(k_pad_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, n->d_in,
                       (long long)rows, n->nin, n->nin_pad, n->d_inp);
    infer_pass(n, s, rows);
    yb.resize((size_t)M * rows * n->nout_pad);
    MCHK(hipMemcpyAsync(yb.data(), n->d_Yp, yb.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    MCHK(hipStreamSynchronize(s));
    for (int m = 0; m < M; ++m)
      for (int r = 0; r < rows; ++r)
        for (int c = 0; c < n->nout; ++c)
          y[((size_t)m * count + r0 + r) * n->nout + c] =
              yb[((size_t)m * rows + r) * n->nout_pad + c];
  }
  MCHK(hipGetLastError());
  return ORCWA_OK;
}

int orcwa_mlp_predict_device(orcwa_mlp* n, const float* x, long long count, float* mean,
                             float* err) {
  if (!n || !x || !mean || !err || count < 1) return fail("orcwa_mlp_predict_device: bad argument");
  MCHK(hipSetDevice(n->ctx->device));
  hipStream_t s = n->ctx->stream;
  int rc = ensure_infer_buffers(n);
  if (rc) return rc;
  for (long long r0 = 0; r0 < count; r0 += kMaxRowsPredict) {
    const int rows = (int)std::min<long long>(kMaxRowsPredict, count - r0);
    const long long tot = (long long)rows * n->nin_pad;
    hipLaunchKernelGGL(k_pad_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       x + r0 * n->nin, (long long)rows, n->nin, n->nin_pad, n->d_inp);
    infer_pass(n, s, rows);
    hipLaunchKernelGGL(k_ensemble, dim3((rows + 255) / 256), dim3(256), 0, s, n->d_Yp,
                       n->members, (long long)rows, n->nout, n->nout_pad, n->d_escale,
                       mean + r0 * n->nout, err + r0);
  }
  MCHK(hipGetLastError());
  return ORCWA_OK;
}

}  // extern "C"

// Inference scratch of orcwa_mlp_predict_device_noalloc: the padded inputs,
// the two activation buffers and the members' outputs of one pass, in a
// static device array (every GPU the library runs on holds its own copy until
// the process ends).  Host code never uses the array's name as an address:
// its device address comes from hipGetSymbolAddress, once per device.
static constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// Device address of `scratch` on `device` (current); 0 on success.
int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return fail("scratch: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      fprintf(stderr, "[openrcwa] hipGetSymbolAddress(scratch) failed: %s\n",
              hipGetErrorName(e));
      return ORCWA_ERR_HIP;
    }
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return ORCWA_OK;
}

}  // namespace

extern "C" {

int orcwa_mlp_predict_device_noalloc(orcwa_mlp* n, const float* x, long long count, float* mean,
                                     float* err) {
  if (!n || !x || !mean || !err || count < 1)
    return fail("orcwa_mlp_predict_device_noalloc: bad argument");
  MCHK(hipSetDevice(n->ctx->device));
  void* base = nullptr;
  int rc = scratch_address(n->ctx->device, &base);
  if (rc) return rc;
  // rows per pass: inputs [R][nin_pad], activations 2 x [members][R][width],
  // outputs [members][R][nout_pad], R a multiple of 64 (every part then stays
  // 256-byte aligned)
  const size_t per_row = sizeof(float) * ((size_t)n->nin_pad + 2 * (size_t)n->members * n->width +
                                          (size_t)n->members * n->nout_pad);
  const long long R = (long long)(kNoAllocScratchBytes / per_row) / 64 * 64;
  if (R < 64)
    return fail("orcwa_mlp_predict_device_noalloc: the network needs more than the 64 MiB "
                "inference scratch for 64 rows (fewer members or a narrower network)");
  float* inp = static_cast<float*>(base);
  float* P0 = inp + R * n->nin_pad;
  float* P1 = P0 + R * n->members * n->width;
  float* Yp = P1 + R * n->members * n->width;
  hipStream_t s = n->ctx->stream;
  for (long long r0 = 0; r0 < count; r0 += R) {
    const int rows = (int)std::min<long long>(R, count - r0);
    const long long tot = (long long)rows * n->nin_pad;
    hipLaunchKernelGGL(k_pad_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       x + r0 * n->nin, (long long)rows, n->nin, n->nin_pad, inp);
    forward(n, s, inp, 0, rows, false, Yp, n->precision == ORCWA_MLP_BF16, P0, P1);
    hipLaunchKernelGGL(k_ensemble, dim3((rows + 255) / 256), dim3(256), 0, s, Yp, n->members,
                       (long long)rows, n->nout, n->nout_pad, n->d_escale, mean + r0 * n->nout,
                       err + r0);
  }
  MCHK(hipGetLastError());
  return ORCWA_OK;
}

int orcwa_mlp_predict(orcwa_mlp* n, const float* x, long long count, float* mean, float* err) {
  if (!n || !x || !mean || !err || count < 1) return fail("orcwa_mlp_predict: bad argument");
  MCHK(hipSetDevice(n->ctx->device));
  hipStream_t s = n->ctx->stream;
  float *dx = nullptr, *dm = nullptr, *de = nullptr;
  auto cleanup = [&]() {
    if (dx) (void)hipFree(dx);
    if (dm) (void)hipFree(dm);
    if (de) (void)hipFree(de);
  };
  if (hipMalloc(&dx, count * n->nin * sizeof(float)) != hipSuccess ||
      hipMalloc(&dm, count * n->nout * sizeof(float)) != hipSuccess ||
      hipMalloc(&de, count * sizeof(float)) != hipSuccess) {
    cleanup();
    return ORCWA_ERR_NOMEM;
  }
  hipError_t e = hipMemcpyAsync(dx, x, count * n->nin * sizeof(float), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipEventRecord(n->ev0, s);
  int rc = e == hipSuccess ? orcwa_mlp_predict_device(n, dx, count, dm, de) : ORCWA_ERR_HIP;
  if (rc == ORCWA_OK) {
    e = hipEventRecord(n->ev1, s);
    if (e == hipSuccess)
      e = hipMemcpyAsync(mean, dm, count * n->nout * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(err, de, count * sizeof(float), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float ms = 0.f;
    if (e == hipSuccess && hipEventElapsedTime(&ms, n->ev0, n->ev1) == hipSuccess) n->last_ms = ms;
    if (e != hipSuccess) rc = ORCWA_ERR_HIP;
  }
  (void)hipStreamSynchronize(s);
  cleanup();
  return rc;
}

double orcwa_mlp_last_ms(const orcwa_mlp* n) { return n ? n->last_ms : -1.0; }

}  // extern "C"
