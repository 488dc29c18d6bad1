// Slab scattering matrices and the Redheffer star product on batches
// (row a10: rcwa/RCWAScatterMatrix.cpp:88-147, core/RCWAIntegrator.cpp:57-67),
// C ABI of include/openrcwa_gpu.h.  Host buffers in and out; every product is
// a batched MFMA GEMM and every inverse an LU solve.
#include <vector>

#include "context.h"
#include "engine.h"

using namespace orcwa;

namespace {

#define SCHK(expr)                                                                          \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      fprintf(stderr, "[openrcwa] HIP error %s at %s:%d\n", hipGetErrorName(_e), __FILE__,  \
              __LINE__);                                                                    \
      return ORCWA_ERR_HIP;                                                                 \
    }                                                                                       \
  } while (0)

constexpr double kPi = 3.1415926535897932;  // rcwa/defns.h:25

// `nslot` n x n matrices per instance, one allocation; tab(k) is the device
// table of the k-th slot of every instance
struct Slots {
  cd* data = nullptr;
  cd** tables = nullptr;
  int n = 0, nslot = 0, batch = 0;
  ~Slots() {
    if (data) (void)hipFree(data);
    if (tables) (void)hipFree(tables);
  }
  int init(hipStream_t s, int n_, int nslot_, int batch_) {
    n = n_;
    nslot = nslot_;
    batch = batch_;
    const size_t nn = (size_t)n * n;
    if (hipMalloc(&data, sizeof(cd) * nn * nslot * batch) != hipSuccess) return ORCWA_ERR_NOMEM;
    if (hipMalloc(&tables, sizeof(cd*) * nslot * batch) != hipSuccess) return ORCWA_ERR_NOMEM;
    std::vector<cd*> t((size_t)nslot * batch);
    for (int k = 0; k < nslot; ++k)
      for (int b = 0; b < batch; ++b) t[(size_t)k * batch + b] = data + ((size_t)b * nslot + k) * nn;
    SCHK(hipMemcpyAsync(tables, t.data(), sizeof(cd*) * t.size(), hipMemcpyHostToDevice, s));
    return ORCWA_OK;
  }
  cd* const* tab(int k) const { return tables + (size_t)k * batch; }
  const cd* const* ctab(int k) const { return (const cd* const*)(tables + (size_t)k * batch); }
  cd* slot(int b, int k) const { return data + ((size_t)b * nslot + k) * (size_t)n * n; }
};

// A = F0 + G0, B = F0 - G0, XA = diag(X) A, XB = diag(X) B with
// X_i = exp(i gamma_i wz lambda / 2 pi) (rcwa/RCWAScatterMatrix.cpp:111-113, 126-133)
__global__ void k_slab_ab(int n, const cd* const* F0p, const cd* const* G0p, const cd* gamma_all,
                          const double* lam, const double* wz, cd* const* Ap, cd* const* Bp,
                          cd* const* XAp, cd* const* XBp) {
  const int b = blockIdx.y;
  const cd* F0 = F0p[b];
  const cd* G0 = G0p[b];
  const cd* g = gamma_all + (long long)b * n;
  const double ss = lam[b] / (2.0 * kPi) * wz[b];
  const long long nn = (long long)n * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nn;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n);
    const cd a = F0[e] + G0[e], c = F0[e] - G0[e];
    const cd gi = g[r];
    const cd x = cexp(mk(-gi.im * ss, gi.re * ss));
    Ap[b][e] = a;
    Bp[b][e] = c;
    XAp[b][e] = x * a;
    XBp[b][e] = x * c;
  }
}

// device allocations released on every return path
struct DevAllocs {
  std::vector<void*> p;
  ~DevAllocs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int get(T** out, size_t count) {
    *out = nullptr;
    if (hipMalloc((void**)out, sizeof(T) * (count ? count : 1)) != hipSuccess) return ORCWA_ERR_NOMEM;
    p.push_back(*out);
    return ORCWA_OK;
  }
};

int upload(hipStream_t s, const Slots& S, int k, const double* host) {
  const size_t nn = (size_t)S.n * S.n;
  for (int b = 0; b < S.batch; ++b)
    SCHK(hipMemcpyAsync(S.slot(b, k), host + 2 * nn * b, sizeof(cd) * nn, hipMemcpyHostToDevice,
                        s));
  return ORCWA_OK;
}

int download(hipStream_t s, const Slots& S, int k, double* host, size_t stride_elems) {
  const size_t nn = (size_t)S.n * S.n;
  for (int b = 0; b < S.batch; ++b)
    SCHK(hipMemcpyAsync(host + 2 * stride_elems * b, S.slot(b, k), sizeof(cd) * nn,
                        hipMemcpyDeviceToHost, s));
  return ORCWA_OK;
}

}  // namespace

extern "C" {

int orcwa_gpu_slab_smatrix(orcwa_gpu_context* ctx, int n, int batch, const double* lambda,
                           const double* wz, const double* W, const double* V,
                           const double* gamma, const double* W0, const double* V0, double* Rud,
                           double* Tuu, int* info) {
  if (!ctx || n <= 0 || batch <= 0 || !lambda || !wz || !W || !V || !gamma || !W0 || !V0 ||
      !Rud || !Tuu) {
    fprintf(stderr, "[openrcwa] orcwa_gpu_slab_smatrix: invalid argument\n");
    return ORCWA_ERR_ARG;
  }
  SCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // slots: 0 W/LU, 1 V/LU, 2 F0, 3 G0, 4-6 [invAB | invAXA | invAXB], 7 A/LU,
  //        8-9 [Rud | Tuu], 10 I - invAXB^2 / LU
  Slots S;
  int rc = S.init(s, n, 11, batch);
  if (rc) return rc;
  DevAllocs mem;
  cd* d_gamma;
  double *d_lam, *d_wz;
  int *d_ipiv, *d_perm, *d_info;
  if ((rc = mem.get(&d_gamma, (size_t)n * batch)) || (rc = mem.get(&d_lam, batch)) ||
      (rc = mem.get(&d_wz, batch)) || (rc = mem.get(&d_ipiv, (size_t)n * batch)) ||
      (rc = mem.get(&d_perm, (size_t)n * batch)) || (rc = mem.get(&d_info, batch)))
    return rc;
  SCHK(hipMemsetAsync(d_info, 0, sizeof(int) * batch, s));
  SCHK(hipMemcpyAsync(d_gamma, gamma, sizeof(cd) * n * batch, hipMemcpyHostToDevice, s));
  SCHK(hipMemcpyAsync(d_lam, lambda, sizeof(double) * batch, hipMemcpyHostToDevice, s));
  SCHK(hipMemcpyAsync(d_wz, wz, sizeof(double) * batch, hipMemcpyHostToDevice, s));
  if ((rc = upload(s, S, 0, W)) || (rc = upload(s, S, 1, V)) || (rc = upload(s, S, 2, W0)) ||
      (rc = upload(s, S, 3, V0)))
    return rc;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  // F0 = W^-1 W0, G0 = V^-1 V0 (:120-129)
  zgetrf_batched(s, n, S.tab(0), d_ipiv, d_info, nullptr, batch);
  zgetrs_batched(s, n, n, S.ctab(0), d_ipiv, S.tab(2), batch, d_perm);
  zgetrf_batched(s, n, S.tab(1), d_ipiv, d_info, nullptr, batch);
  zgetrs_batched(s, n, n, S.ctab(1), d_ipiv, S.tab(3), batch, d_perm);
  // A, B, XA, XB, then [A^-1 B | A^-1 X A | A^-1 X B] with one LU of A (:131-137)
  const long long nn = (long long)n * n;
  unsigned gx = (unsigned)((nn + 255) / 256 > 2048 ? 2048 : (nn + 255) / 256);
  hipLaunchKernelGGL(k_slab_ab, dim3(gx, batch), dim3(256), 0, s, n, S.ctab(2), S.ctab(3),
                     (const cd*)d_gamma, (const double*)d_lam, (const double*)d_wz, S.tab(7),
                     S.tab(4), S.tab(5), S.tab(6));
  zgetrf_batched(s, n, S.tab(7), d_ipiv, d_info, nullptr, batch);
  zgetrs_batched(s, n, 3 * n, S.ctab(7), d_ipiv, S.tab(4), batch, d_perm);
  // I - invAXB^2; Rud rhs = invAXB invAXA - invAB; Tuu rhs = invAXA - invAXB invAB (:139-146)
  set_identity_batched(s, n, S.tab(10), batch);
  zgemm_batched(s, n, n, n, mone, S.ctab(6), 0, 1, n, 0, S.ctab(6), 0, 1, n, 0, one, S.tab(10), 0,
                1, n, batch);
  zgemm_batched(s, n, n, n, one, S.ctab(6), 0, 1, n, 0, S.ctab(5), 0, 1, n, 0, zero, S.tab(8), 0,
                1, n, batch);
  axpby_batched(s, n, one, S.ctab(8), mone, S.ctab(4), S.tab(8), batch);
  copy_batched(s, nn, S.ctab(5), S.tab(9), batch);
  zgemm_batched(s, n, n, n, mone, S.ctab(6), 0, 1, n, 0, S.ctab(4), 0, 1, n, 0, one, S.tab(9), 0,
                1, n, batch);
  zgetrf_batched(s, n, S.tab(10), d_ipiv, d_info, nullptr, batch);
  zgetrs_batched(s, n, 2 * n, S.ctab(10), d_ipiv, S.tab(8), batch, d_perm);
  SCHK(hipGetLastError());
  if ((rc = download(s, S, 8, Rud, (size_t)nn)) || (rc = download(s, S, 9, Tuu, (size_t)nn)))
    return rc;
  if (info) SCHK(hipMemcpyAsync(info, d_info, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
  SCHK(hipStreamSynchronize(s));
  return ORCWA_OK;
}

int orcwa_gpu_redheffer_star(orcwa_gpu_context* ctx, int n, int batch, const double* top,
                             const double* bottom, double* out, int* info) {
  if (!ctx || n <= 0 || batch <= 0 || !top || !bottom || !out) {
    fprintf(stderr, "[openrcwa] orcwa_gpu_redheffer_star: invalid argument\n");
    return ORCWA_ERR_ARG;
  }
  SCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  // slots 0-3 top (rud, tuu, rdu, tdd), 4-7 bottom (Rud, Tuu, Rdu, Tdd),
  // 8 M1 = I - rdu Rud / LU, 9-10 [X | Y] = M1^-1 [tdd | rdu Tuu],
  // 11 C2 Tuu = Tuu + Rud Y, 12 scratch, 13-16 result (rud, tuu, rdu, tdd)
  Slots S;
  int rc = S.init(s, n, 17, batch);
  if (rc) return rc;
  DevAllocs mem;
  int *d_ipiv, *d_perm, *d_info;
  if ((rc = mem.get(&d_ipiv, (size_t)n * batch)) || (rc = mem.get(&d_perm, (size_t)n * batch)) ||
      (rc = mem.get(&d_info, batch)))
    return rc;
  SCHK(hipMemsetAsync(d_info, 0, sizeof(int) * batch, s));
  const long long nn = (long long)n * n;
  for (int b = 0; b < batch; ++b) {
    SCHK(hipMemcpyAsync(S.slot(b, 0), top + 2 * 4 * nn * b, sizeof(cd) * 4 * nn,
                        hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(S.slot(b, 4), bottom + 2 * 4 * nn * b, sizeof(cd) * 4 * nn,
                        hipMemcpyHostToDevice, s));
  }
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  auto G = [&](int dst, int a, int bb, cd alpha, cd beta) {
    zgemm_batched(s, n, n, n, alpha, S.ctab(a), 0, 1, n, 0, S.ctab(bb), 0, 1, n, 0, beta,
                  S.tab(dst), 0, 1, n, batch);
  };
  // C1 = (I - rdu Rud)^-1 (core/RCWAIntegrator.cpp:57-60)
  set_identity_batched(s, n, S.tab(8), batch);
  G(8, 2, 4, mone, one);
  zgetrf_batched(s, n, S.tab(8), d_ipiv, d_info, nullptr, batch);
  copy_batched(s, nn, S.ctab(3), S.tab(9), batch);
  G(10, 2, 5, one, zero);                       // rdu Tuu
  zgetrs_batched(s, n, 2 * n, S.ctab(8), d_ipiv, S.tab(9), batch, d_perm);
  // C2 Tuu = Tuu + Rud C1 rdu Tuu (C2 = I + Rud C1 rdu, :61-62)
  copy_batched(s, nn, S.ctab(5), S.tab(11), batch);
  G(11, 4, 10, one, one);
  // Rdu' = Rdu + Tdd rdu C2 Tuu (:64)
  G(12, 2, 11, one, zero);
  copy_batched(s, nn, S.ctab(6), S.tab(15), batch);
  G(15, 7, 12, one, one);
  // Tdd' = Tdd C1 tdd (:65)
  G(16, 7, 9, one, zero);
  // Rud' = rud + tuu Rud C1 tdd (:66)
  G(12, 4, 9, one, zero);
  copy_batched(s, nn, S.ctab(0), S.tab(13), batch);
  G(13, 1, 12, one, one);
  // Tuu' = tuu C2 Tuu (:67)
  G(14, 1, 11, one, zero);
  SCHK(hipGetLastError());
  for (int b = 0; b < batch; ++b)
    SCHK(hipMemcpyAsync(out + 2 * 4 * nn * b, S.slot(b, 13), sizeof(cd) * 4 * nn,
                        hipMemcpyDeviceToHost, s));
  if (info) SCHK(hipMemcpyAsync(info, d_info, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
  SCHK(hipStreamSynchronize(s));
  return ORCWA_OK;
}

}  // extern "C"
