// Slab scattering matrices and the Redheffer star product on batches
// (row a10: rcwa/RCWAScatterMatrix.cpp:88-147, core/RCWAIntegrator.cpp:57-67),
// and the midpoint-staircase integrator of a z-varying structure
// (RCWAIntegrator::compute, core/RCWAIntegrator.cpp:20-86), C ABI of
// include/openrcwa_gpu.h.  Host buffers in and out; every product is a
// batched MFMA GEMM and every inverse an LU solve.
#include <cmath>
#include <cstring>
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
// X_i = exp(i gamma_i ss), ss = wz lambda / 2 pi for gamma in the units of
// WaveEquationSolver (rcwa/RCWAScatterMatrix.cpp:111-113, 126-133)
__global__ void k_slab_ab(int n, const cd* const* F0p, const cd* const* G0p,
                          const cd* const* gamma_p, const double* ss_all, cd* const* Ap,
                          cd* const* Bp, cd* const* XAp, cd* const* XBp) {
  const int b = blockIdx.y;
  const cd* F0 = F0p[b];
  const cd* G0 = G0p[b];
  const cd* g = gamma_p[b];
  const double ss = ss_all[b];
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

// Slab S-matrices of a batch: slots 0-3 hold W, V, W0, V0 on entry (0, 1, 7
// and 10 are factored in place); Rud and Tuu land in slots 8 and 9
// (rcwa/RCWAScatterMatrix.cpp:120-146).  Slots: 0 W/LU, 1 V/LU, 2 F0, 3 G0,
// 4-6 [invAB | invAXA | invAXB], 7 A/LU, 8-9 [Rud | Tuu], 10 I - invAXB^2 / LU.
void slab_core(hipStream_t s, int n, int batch, const Slots& S, const cd* const* gamma,
               const double* ss, int* ipiv, int* perm, int* info) {
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  // F0 = W^-1 W0, G0 = V^-1 V0 (:120-129)
  zgetrf_batched(s, n, S.tab(0), ipiv, info, nullptr, batch);
  zgetrs_batched(s, n, n, S.ctab(0), ipiv, S.tab(2), batch, perm);
  zgetrf_batched(s, n, S.tab(1), ipiv, info, nullptr, batch);
  zgetrs_batched(s, n, n, S.ctab(1), ipiv, S.tab(3), batch, perm);
  // A, B, XA, XB, then [A^-1 B | A^-1 X A | A^-1 X B] with one LU of A (:131-137)
  const long long nn = (long long)n * n;
  unsigned gx = (unsigned)((nn + 255) / 256 > 2048 ? 2048 : (nn + 255) / 256);
  hipLaunchKernelGGL(k_slab_ab, dim3(gx, batch), dim3(256), 0, s, n, S.ctab(2), S.ctab(3),
                     gamma, ss, S.tab(7), S.tab(4), S.tab(5), S.tab(6));
  zgetrf_batched(s, n, S.tab(7), ipiv, info, nullptr, batch);
  zgetrs_batched(s, n, 3 * n, S.ctab(7), ipiv, S.tab(4), batch, perm);
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
  zgetrf_batched(s, n, S.tab(10), ipiv, info, nullptr, batch);
  zgetrs_batched(s, n, 2 * n, S.ctab(10), ipiv, S.tab(8), batch, perm);
}

// Star products top * bottom of a batch: slots 0-3 top (rud, tuu, rdu, tdd),
// 4-7 bottom (Rud, Tuu, Rdu, Tdd); the result lands in slots 13-16.  Slots
// 8 M1 = I - rdu Rud / LU, 9-10 [X | Y] = M1^-1 [tdd | rdu Tuu],
// 11 C2 Tuu = Tuu + Rud Y, 12 scratch.
void star_core(hipStream_t s, int n, int batch, const Slots& S, int* ipiv, int* perm, int* info) {
  const long long nn = (long long)n * n;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  auto G = [&](int dst, int a, int bb, cd alpha, cd beta) {
    zgemm_batched(s, n, n, n, alpha, S.ctab(a), 0, 1, n, 0, S.ctab(bb), 0, 1, n, 0, beta,
                  S.tab(dst), 0, 1, n, batch);
  };
  // C1 = (I - rdu Rud)^-1 (core/RCWAIntegrator.cpp:57-60)
  set_identity_batched(s, n, S.tab(8), batch);
  G(8, 2, 4, mone, one);
  zgetrf_batched(s, n, S.tab(8), ipiv, info, nullptr, batch);
  copy_batched(s, nn, S.ctab(3), S.tab(9), batch);
  G(10, 2, 5, one, zero);                       // rdu Tuu
  zgetrs_batched(s, n, 2 * n, S.ctab(8), ipiv, S.tab(9), batch, perm);
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
  double* d_ss;
  int *d_ipiv, *d_perm, *d_info;
  if ((rc = mem.get(&d_gamma, (size_t)n * batch)) || (rc = mem.get(&d_ss, batch)) ||
      (rc = mem.get(&d_ipiv, (size_t)n * batch)) || (rc = mem.get(&d_perm, (size_t)n * batch)) ||
      (rc = mem.get(&d_info, batch)))
    return rc;
  SCHK(hipMemsetAsync(d_info, 0, sizeof(int) * batch, s));
  SCHK(hipMemcpyAsync(d_gamma, gamma, sizeof(cd) * n * batch, hipMemcpyHostToDevice, s));
  if ((rc = upload(s, S, 0, W)) || (rc = upload(s, S, 1, V)) || (rc = upload(s, S, 2, W0)) ||
      (rc = upload(s, S, 3, V0)))
    return rc;
  std::vector<const cd*> gtab(batch);
  std::vector<double> ss(batch);
  for (int b = 0; b < batch; ++b) {
    gtab[b] = d_gamma + (size_t)b * n;
    ss[b] = lambda[b] / (2.0 * kPi) * wz[b];
  }
  const cd** d_gtab;
  if ((rc = mem.get(&d_gtab, batch))) return rc;
  SCHK(hipMemcpyAsync(d_gtab, gtab.data(), sizeof(cd*) * batch, hipMemcpyHostToDevice, s));
  SCHK(hipMemcpyAsync(d_ss, ss.data(), sizeof(double) * batch, hipMemcpyHostToDevice, s));
  slab_core(s, n, batch, S, d_gtab, d_ss, d_ipiv, d_perm, d_info);
  const long long nn = (long long)n * n;
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
  star_core(s, n, batch, S, d_ipiv, d_perm, d_info);
  SCHK(hipGetLastError());
  for (int b = 0; b < batch; ++b)
    SCHK(hipMemcpyAsync(out + 2 * 4 * nn * b, S.slot(b, 13), sizeof(cd) * 4 * nn,
                        hipMemcpyDeviceToHost, s));
  if (info) SCHK(hipMemcpyAsync(info, d_info, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
  SCHK(hipStreamSynchronize(s));
  return ORCWA_OK;
}

// A z-varying structure by the midpoint staircase (RCWAIntegrator::compute,
// core/RCWAIntegrator.cpp:20-86): the modes of all N + 2 samples in one
// batched solve, the N slab S-matrices in the basis of the top sample in one
// batch, their star products as a tree (level by level, each level one
// batch; the product is associative, so this equals the reference's
// slab-by-slab loop of :48-68 up to rounding), then the bottom side taken into
// the basis of the z0 sample (:74-85).
int orcwa_gpu_rcwa_integrate(orcwa_gpu_context* ctx, const orcwa_zstack* zs,
                             orcwa_zresult* res) {
  if (!ctx || !zs || !res || zs->n_slices < 1 || !(zs->z1 > zs->z0) || !(zs->lambda > 0) ||
      !zs->coord_x || !zs->coord_y || !zs->eps || zs->n_cx < 2 || zs->n_cy < 2) {
    fprintf(stderr, "[openrcwa] orcwa_gpu_rcwa_integrate: invalid argument\n");
    return ORCWA_ERR_ARG;
  }
  SCHK(hipSetDevice(ctx->device));
  hipStream_t s = ctx->stream;
  const int N = zs->n_slices, L = N + 2;
  std::vector<double> thick(L, 0.0);
  orcwa_stack st;
  st.nx_harm = zs->nx_harm;
  st.ny_harm = zs->ny_harm;
  st.n_cx = zs->n_cx;
  st.n_cy = zs->n_cy;
  st.coord_x = zs->coord_x;
  st.coord_y = zs->coord_y;
  st.n_layers = L;
  st.thickness = thick.data();
  const double pol[4] = {1.0, 0.0, 0.0, 0.0};
  orcwa_instance inst;
  std::memset(&inst, 0, sizeof(inst));
  inst.lambda = zs->lambda;
  inst.kx0 = zs->kx0;
  inst.ky0 = zs->ky0;
  inst.eps = zs->eps;
  inst.n_pol = 1;
  inst.pol = pol;
  LayerModes M;
  int rc = layer_modes(ctx, &st, &inst, zs->coord_x, zs->coord_y, M);
  if (rc != ORCWA_OK) return rc;
  struct Release {
    LayerModes& m;
    ~Release() { layer_modes_release(m); }
  } release{M};
  const int n = M.n;
  const long long nn = (long long)n * n;
  const double k0 = 2.0 * kPi / zs->lambda, dz = (zs->z1 - zs->z0) / N;
  DevAllocs mem;
  int *d_ipiv, *d_perm, *d_info;
  if ((rc = mem.get(&d_ipiv, (size_t)N * n)) || (rc = mem.get(&d_perm, (size_t)N * n)) ||
      (rc = mem.get(&d_info, (size_t)N)))
    return rc;
  SCHK(hipMemsetAsync(d_info, 0, sizeof(int) * N, s));
  // device pointer tables for batched copies
  auto copies = [&](const std::vector<const cd*>& src, const std::vector<cd*>& dst) -> int {
    const cd** d_src;
    cd** d_dst;
    int r;
    if ((r = mem.get(&d_src, src.size())) || (r = mem.get(&d_dst, dst.size()))) return r;
    SCHK(hipMemcpyAsync(d_src, src.data(), sizeof(cd*) * src.size(), hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(d_dst, dst.data(), sizeof(cd*) * dst.size(), hipMemcpyHostToDevice, s));
    copy_batched(s, nn, d_src, d_dst, (int)src.size());
    return ORCWA_OK;
  };

  // 1. slab S-matrices of samples 1..N in the basis (W, V) of sample 0
  Slots S;
  if ((rc = S.init(s, n, 11, N))) return rc;
  {
    std::vector<const cd*> src;
    std::vector<cd*> dst;
    std::vector<const cd*> gtab(N);
    std::vector<double> ss(N, k0 * dz);       // keff in k0 units: phase keff k0 dz
    for (int b = 0; b < N; ++b) {
      const int l = 1 + b;
      src.insert(src.end(), {M.W[l], M.V[l], M.W[0], M.V[0]});
      dst.insert(dst.end(), {S.slot(b, 0), S.slot(b, 1), S.slot(b, 2), S.slot(b, 3)});
      gtab[b] = M.keff[l];
    }
    if ((rc = copies(src, dst))) return rc;
    const cd** d_gtab;
    double* d_ss;
    if ((rc = mem.get(&d_gtab, N)) || (rc = mem.get(&d_ss, N))) return rc;
    SCHK(hipMemcpyAsync(d_gtab, gtab.data(), sizeof(cd*) * N, hipMemcpyHostToDevice, s));
    SCHK(hipMemcpyAsync(d_ss, ss.data(), sizeof(double) * N, hipMemcpyHostToDevice, s));
    slab_core(s, n, N, S, d_gtab, d_ss, d_ipiv, d_perm, d_info);
  }
  // 2. star products, the higher slab on top: a tree over z-ordered items
  struct Item {
    const cd* m[4];  // rud, tuu, rdu, tdd
  };
  std::vector<Item> items(N);
  for (int b = 0; b < N; ++b)
    items[b] = {{S.slot(b, 8), S.slot(b, 9), S.slot(b, 8), S.slot(b, 9)}};
  std::vector<Slots*> levels;
  struct FreeLevels {
    std::vector<Slots*>& v;
    ~FreeLevels() {
      for (Slots* q : v) delete q;
    }
  } free_levels{levels};
  while (items.size() > 1) {
    const int m = (int)items.size(), pairs = m / 2;
    Slots* T = new Slots;
    levels.push_back(T);
    if ((rc = T->init(s, n, 17, pairs))) return rc;
    std::vector<const cd*> src;
    std::vector<cd*> dst;
    for (int k = 0; k < pairs; ++k)
      for (int q = 0; q < 4; ++q) {
        src.push_back(items[2 * k + 1].m[q]);   // top: the higher slab
        dst.push_back(T->slot(k, q));
        src.push_back(items[2 * k].m[q]);       // bottom
        dst.push_back(T->slot(k, 4 + q));
      }
    if ((rc = copies(src, dst))) return rc;
    star_core(s, n, pairs, *T, d_ipiv, d_perm, d_info);
    std::vector<Item> next;
    for (int k = 0; k < pairs; ++k)
      next.push_back({{T->slot(k, 13), T->slot(k, 14), T->slot(k, 15), T->slot(k, 16)}});
    if (m & 1) next.push_back(items[m - 1]);
    items.swap(next);
  }
  // 3. the bottom side into the basis of the z0 sample (:74-85).  Slots:
  // 0 W1/LU, 1 V1/LU, 2 W1^-1 W0, 3 V1^-1 V0, 4 A, 5 B, 6 Rud, 7 Tuu, 8 Rdu,
  // 10 A - Rdu B / LU, 11-12 [Rdu A - B | Tdd] -> [Rdu' | Tdd'], 13 B Tdd',
  // 14 A + B Rdu', 15 Tuu'
  Slots F;
  if ((rc = F.init(s, n, 16, 1))) return rc;
  if ((rc = copies({M.W[0], M.V[0], M.W[L - 1], M.V[L - 1], items[0].m[0], items[0].m[1],
                    items[0].m[2], items[0].m[3]},
                   {F.slot(0, 0), F.slot(0, 1), F.slot(0, 2), F.slot(0, 3), F.slot(0, 6),
                    F.slot(0, 7), F.slot(0, 8), F.slot(0, 12)})))
    return rc;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0), half = mk(0.5, 0.0),
           mhalf = mk(-0.5, 0.0);
  auto G = [&](int dst, int a, int b, cd alpha, cd beta) {
    zgemm_batched(s, n, n, n, alpha, F.ctab(a), 0, 1, n, 0, F.ctab(b), 0, 1, n, 0, beta,
                  F.tab(dst), 0, 1, n, 1);
  };
  zgetrf_batched(s, n, F.tab(0), d_ipiv, d_info, nullptr, 1);
  zgetrs_batched(s, n, n, F.ctab(0), d_ipiv, F.tab(2), 1, d_perm);
  zgetrf_batched(s, n, F.tab(1), d_ipiv, d_info, nullptr, 1);
  zgetrs_batched(s, n, n, F.ctab(1), d_ipiv, F.tab(3), 1, d_perm);
  axpby_batched(s, n, half, F.ctab(2), half, F.ctab(3), F.tab(4), 1);
  axpby_batched(s, n, half, F.ctab(2), mhalf, F.ctab(3), F.tab(5), 1);
  copy_batched(s, nn, F.ctab(4), F.tab(10), 1);
  G(10, 8, 5, mone, one);                        // A - Rdu B
  G(11, 8, 4, one, zero);                        // Rdu A
  axpby_batched(s, n, one, F.ctab(11), mone, F.ctab(5), F.tab(11), 1);
  zgetrf_batched(s, n, F.tab(10), d_ipiv, d_info, nullptr, 1);
  zgetrs_batched(s, n, 2 * n, F.ctab(10), d_ipiv, F.tab(11), 1, d_perm);
  G(13, 5, 12, one, zero);                       // B Tdd'
  G(6, 7, 13, one, one);                         // Rud' = Rud + Tuu B Tdd'
  copy_batched(s, nn, F.ctab(4), F.tab(14), 1);
  G(14, 5, 11, one, one);                        // A + B Rdu'
  G(15, 7, 14, one, zero);                       // Tuu' = Tuu (B Rdu' + A)
  SCHK(hipGetLastError());
  // 4. results
  auto get = [&](double* host, const cd* dev, size_t count) -> int {
    if (host) SCHK(hipMemcpyAsync(host, dev, sizeof(cd) * count, hipMemcpyDeviceToHost, s));
    return ORCWA_OK;
  };
  if ((rc = get(res->Rud, F.slot(0, 6), nn)) || (rc = get(res->Tuu, F.slot(0, 15), nn)) ||
      (rc = get(res->Rdu, F.slot(0, 11), nn)) || (rc = get(res->Tdd, F.slot(0, 12), nn)) ||
      (rc = get(res->W0, M.W[L - 1], nn)) || (rc = get(res->V0, M.V[L - 1], nn)) ||
      (rc = get(res->W1, M.W[0], nn)) || (rc = get(res->V1, M.V[0], nn)) ||
      (rc = get(res->gam0, M.keff[L - 1], n)) || (rc = get(res->gam1, M.keff[0], n)))
    return rc;
  std::vector<int> info(N);
  SCHK(hipMemcpyAsync(info.data(), d_info, sizeof(int) * N, hipMemcpyDeviceToHost, s));
  SCHK(hipStreamSynchronize(s));
  // keff in k0 units -> the units of WaveEquationSolver (P and Q carry k0^2)
  for (double* g : {res->gam0, res->gam1})
    if (g)
      for (int i = 0; i < 2 * n; ++i) g[i] *= k0 * k0;
  res->status = M.status;
  for (int v : info) res->status |= v;
  return ORCWA_OK;
}

}  // extern "C"
