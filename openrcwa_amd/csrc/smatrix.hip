// Slab scattering matrices and the Redheffer star product on batches
// (row a10: rcwa/RCWAScatterMatrix.cpp:88-147, core/RCWAIntegrator.cpp:57-67),
// and the midpoint-staircase integrator of a z-varying structure
// (RCWAIntegrator::compute, core/RCWAIntegrator.cpp:20-86), C ABI of
// include/openrcwa_gpu.h.  Host buffers in and out; every product is a
// batched MFMA GEMM and every inverse an LU solve.
#include <algorithm>
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
// The nine products run at `prec` (GemmPrec: FP64, or FP32 / split BF16 on
// MFMA with FP64 accumulation of each K step); the LU and solve stay FP64.
void star_core(hipStream_t s, int n, int batch, const Slots& S, int* ipiv, int* perm, int* info,
               int prec) {
  const long long nn = (long long)n * n;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0);
  auto G = [&](int dst, int a, int bb, cd alpha, cd beta) {
    zgemm_batched_prec(prec, s, n, n, n, alpha, S.ctab(a), 0, 1, n, 0, S.ctab(bb), 0, 1, n, 0,
                       beta, S.tab(dst), 0, 1, n, batch);
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

// Device work of the z integrators: pointer tables, batched copies, the
// tree of star products; every allocation is released with the object.
struct ZWork {
  hipStream_t s;
  int n;
  long long nn;
  DevAllocs mem;
  std::vector<Slots*> slots;
  int* ipiv = nullptr;
  int* perm = nullptr;
  int* info = nullptr;
  int cap = 0;            // batch capacity of ipiv / perm / info
  int prec = 0;           // GemmPrec of the star products
  ~ZWork() {
    for (Slots* q : slots) delete q;
  }
  int init(hipStream_t s_, int n_, int cap_) {
    s = s_;
    n = n_;
    nn = (long long)n * n;
    cap = cap_ > 1 ? cap_ : 1;
    int rc;
    if ((rc = mem.get(&ipiv, (size_t)cap * n)) || (rc = mem.get(&perm, (size_t)cap * n)) ||
        (rc = mem.get(&info, (size_t)cap)))
      return rc;
    SCHK(hipMemsetAsync(info, 0, sizeof(int) * cap, s));
    return ORCWA_OK;
  }
  Slots* make(int nslot, int batch, int& rc) {
    Slots* q = new Slots;
    slots.push_back(q);
    rc = q->init(s, n, nslot, batch);
    return q;
  }
  template <class T>
  int table(const std::vector<T>& host, T** dev) {
    int rc;
    if ((rc = mem.get(dev, host.size()))) return rc;
    SCHK(hipMemcpyAsync(*dev, host.data(), sizeof(T) * host.size(), hipMemcpyHostToDevice, s));
    return ORCWA_OK;
  }
  int copies(const std::vector<const cd*>& src, const std::vector<cd*>& dst) {
    const cd** d_src;
    cd** d_dst;
    int rc;
    if ((rc = table(src, &d_src)) || (rc = table(dst, &d_dst))) return rc;
    copy_batched(s, nn, d_src, d_dst, (int)src.size());
    return ORCWA_OK;
  }
  // Star products of z-ordered S-matrices (rud, tuu, rdu, tdd), the higher
  // one on top, as a tree of batched levels (associative: the reference's
  // slab-by-slab loop up to rounding).  Returns the product in items[0].
  struct Item {
    const cd* m[4];
  };
  int star_tree(std::vector<Item>& items) {
    int rc = ORCWA_OK;
    while (items.size() > 1) {
      const int m = (int)items.size(), pairs = m / 2;
      Slots* T = make(17, pairs, rc);
      if (rc) return rc;
      std::vector<const cd*> src;
      std::vector<cd*> dst;
      for (int k = 0; k < pairs; ++k)
        for (int q = 0; q < 4; ++q) {
          src.push_back(items[2 * k + 1].m[q]);   // top: the higher one
          dst.push_back(T->slot(k, q));
          src.push_back(items[2 * k].m[q]);       // bottom
          dst.push_back(T->slot(k, 4 + q));
        }
      if ((rc = copies(src, dst))) return rc;
      star_core(s, n, pairs, *T, ipiv, perm, info, prec);
      std::vector<Item> next;
      for (int k = 0; k < pairs; ++k)
        next.push_back({{T->slot(k, 13), T->slot(k, 14), T->slot(k, 15), T->slot(k, 16)}});
      if (m & 1) next.push_back(items[m - 1]);
      items.swap(next);
    }
    return rc;
  }
  // batched GEMM over pointer tables: C = alpha A B + beta C
  void gemm(const cd* const* A, const cd* const* B, cd* const* C, cd alpha, cd beta, int batch) {
    zgemm_batched(s, n, n, n, alpha, A, 0, 1, n, 0, B, 0, 1, n, 0, beta, C, 0, 1, n, batch);
  }
  int status() {
    std::vector<int> h(cap);
    SCHK(hipMemcpyAsync(h.data(), info, sizeof(int) * cap, hipMemcpyDeviceToHost, s));
    SCHK(hipStreamSynchronize(s));
    int st = 0;
    for (int v : h) st |= v;
    return st;
  }
};

// the first-order perturbation terms of a batch of segments
// (core/RedhefferIntegrator.cpp:317-399) from X_i = W^-1 (P_i - Pr) V (slots
// 2-4) and Y_i = V^-1 (Q_i - Qr) W (slots 5-7), i = start, middle, end:
// dA_i = X_i + Y_i, dB_i = X_i - Y_i, Simpson's rule with ph = exp(i keff k0 dz / 2),
// pf = exp(i keff k0 dz), c = i k0 dz / 12 (keff, P, Q in k0 units).  Writes
// tdd, tuu, rud, rdu to slots 10-13 and the largest perturbation entry of
// each segment to err (as the bits of a non-negative double).
__global__ void k_redheffer_simpson(int n, cd* const* slots_p, const cd* const* keff_p,
                                    const double* k0dz_p, unsigned long long* err) {
  const int b = blockIdx.y;
  cd* base = slots_p[b];
  const long long nn = (long long)n * n;
  const cd* kf = keff_p[b];
  const double sdz = k0dz_p[b];
  const cd c = mk(0.0, sdz / 12.0);
  double emax = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < nn;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), col = (int)(e / n);
    const cd kr = kf[r], kc = kf[col];
    const cd phr = cexp(mk(-0.5 * sdz * kr.im, 0.5 * sdz * kr.re));
    const cd phc = cexp(mk(-0.5 * sdz * kc.im, 0.5 * sdz * kc.re));
    const cd pfr = cexp(mk(-sdz * kr.im, sdz * kr.re));
    const cd pfc = cexp(mk(-sdz * kc.im, sdz * kc.re));
    cd dA[3], dB[3];
    for (int i = 0; i < 3; ++i) {
      const cd x = base[(2 + i) * nn + e], y = base[(5 + i) * nn + e];
      dA[i] = x + y;
      dB[i] = x - y;
    }
    const cd mid_a = 4.0 * (phr * dA[1] * phc), mid_b = 4.0 * (phr * dB[1] * phc);
    cd tuu = c * (pfr * dA[0] + mid_a + dA[2] * pfc);
    cd rud = -(c * (pfr * dB[0] * pfc + mid_b + dB[2]));
    cd rdu = -(c * (dB[0] + mid_b + pfr * dB[2] * pfc));
    cd tdd = c * (dA[0] * pfc + mid_a + pfr * dA[2]);
    emax = fmax(emax, fmax(fmax(cabs(tuu), cabs(rud)), fmax(cabs(rdu), cabs(tdd))));
    if (r == col) {
      tuu += pfr;
      tdd += pfr;
    }
    base[10 * nn + e] = tdd;
    base[11 * nn + e] = tuu;
    base[12 * nn + e] = rud;
    base[13 * nn + e] = rdu;
  }
  emax = wave_max(emax);
  if ((threadIdx.x & 63) == 0) atomicMax(err + b, (unsigned long long)__double_as_longlong(emax));
}

// Modes (and with want_pq P and Q) of L cross-sections, each with its own
// edges, as one modes-only solve.
int sample_modes(orcwa_gpu_context* ctx, int nx_harm, int ny_harm, int ncx, int ncy, int L,
                 double lambda, double kx0, double ky0, const double* cx, const double* cy,
                 const double* eps, bool want_pq, LayerModes& out) {
  std::vector<double> thick(L, 0.0);
  orcwa_stack st;
  st.nx_harm = nx_harm;
  st.ny_harm = ny_harm;
  st.n_cx = ncx;
  st.n_cy = ncy;
  st.coord_x = cx;
  st.coord_y = cy;
  st.n_layers = L;
  st.thickness = thick.data();
  const double pol[4] = {1.0, 0.0, 0.0, 0.0};
  orcwa_instance inst;
  std::memset(&inst, 0, sizeof(inst));
  inst.lambda = lambda;
  inst.kx0 = kx0;
  inst.ky0 = ky0;
  inst.eps = eps;
  inst.n_pol = 1;
  inst.pol = pol;
  return layer_modes(ctx, &st, &inst, cx, cy, out, want_pq);
}

// copy an integrator's S-matrix and bases to the caller (NULL skips one);
// keff (k0 units) -> the units of WaveEquationSolver (P and Q carry k0^2)
int fetch_zresult(hipStream_t s, int n, double k0, const cd* Rud, const cd* Tuu, const cd* Rdu,
                  const cd* Tdd, const cd* W0, const cd* V0, const cd* g0, const cd* W1,
                  const cd* V1, const cd* g1, orcwa_zresult* res) {
  const size_t nn = (size_t)n * n;
  auto get = [&](double* host, const cd* dev, size_t count) -> int {
    if (host) SCHK(hipMemcpyAsync(host, dev, sizeof(cd) * count, hipMemcpyDeviceToHost, s));
    return ORCWA_OK;
  };
  int rc;
  if ((rc = get(res->Rud, Rud, nn)) || (rc = get(res->Tuu, Tuu, nn)) ||
      (rc = get(res->Rdu, Rdu, nn)) || (rc = get(res->Tdd, Tdd, nn)) || (rc = get(res->W0, W0, nn)) ||
      (rc = get(res->V0, V0, nn)) || (rc = get(res->W1, W1, nn)) || (rc = get(res->V1, V1, nn)) ||
      (rc = get(res->gam0, g0, n)) || (rc = get(res->gam1, g1, n)))
    return rc;
  SCHK(hipStreamSynchronize(s));
  for (double* g : {res->gam0, res->gam1})
    if (g)
      for (int i = 0; i < 2 * n; ++i) g[i] *= k0 * k0;
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
  ctx_trim(ctx);   // buffers cached by earlier plans go back to the device
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
  ctx_trim(ctx);   // buffers cached by earlier plans go back to the device
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
  star_core(s, n, batch, S, d_ipiv, d_perm, d_info, ctx->gemm_prec);
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
// batch, their star products as a tree of batched levels, then the bottom
// side taken into the basis of the z0 sample (:74-85).
int orcwa_gpu_rcwa_integrate(orcwa_gpu_context* ctx, const orcwa_zstack* zs,
                             orcwa_zresult* res) {
  if (!ctx || !zs || !res || zs->n_slices < 1 || !(zs->z1 > zs->z0) || !(zs->lambda > 0) ||
      !zs->coord_x || !zs->coord_y || !zs->eps || zs->n_cx < 2 || zs->n_cy < 2) {
    fprintf(stderr, "[openrcwa] orcwa_gpu_rcwa_integrate: invalid argument\n");
    return ORCWA_ERR_ARG;
  }
  SCHK(hipSetDevice(ctx->device));
  ctx_trim(ctx);   // buffers cached by earlier plans go back to the device
  hipStream_t s = ctx->stream;
  const int N = zs->n_slices, L = N + 2;
  LayerModes M;
  int rc = sample_modes(ctx, zs->nx_harm, zs->ny_harm, zs->n_cx, zs->n_cy, L, zs->lambda,
                        zs->kx0, zs->ky0, zs->coord_x, zs->coord_y, zs->eps, false, M);
  if (rc != ORCWA_OK) return rc;
  struct Release {
    LayerModes& m;
    ~Release() { layer_modes_release(m); }
  } release{M};
  const int n = M.n;
  const double k0 = 2.0 * kPi / zs->lambda, dz = (zs->z1 - zs->z0) / N;
  ZWork w;
  if ((rc = w.init(s, n, N))) return rc;
  w.prec = ctx->gemm_prec;
  // 1. slab S-matrices of samples 1..N in the basis (W, V) of sample 0
  Slots* S = w.make(11, N, rc);
  if (rc) return rc;
  {
    std::vector<const cd*> src, gtab(N);
    std::vector<cd*> dst;
    std::vector<double> ss(N, k0 * dz);       // keff in k0 units: phase keff k0 dz
    for (int b = 0; b < N; ++b) {
      const int l = 1 + b;
      src.insert(src.end(), {M.W[l], M.V[l], M.W[0], M.V[0]});
      dst.insert(dst.end(), {S->slot(b, 0), S->slot(b, 1), S->slot(b, 2), S->slot(b, 3)});
      gtab[b] = M.keff[l];
    }
    const cd** d_gtab;
    double* d_ss;
    if ((rc = w.copies(src, dst)) || (rc = w.table(gtab, &d_gtab)) || (rc = w.table(ss, &d_ss)))
      return rc;
    slab_core(s, n, N, *S, d_gtab, d_ss, w.ipiv, w.perm, w.info);
  }
  // 2. star products, the higher slab on top
  std::vector<ZWork::Item> items(N);
  for (int b = 0; b < N; ++b)
    items[b] = {{S->slot(b, 8), S->slot(b, 9), S->slot(b, 8), S->slot(b, 9)}};
  if ((rc = w.star_tree(items))) return rc;
  // 3. the bottom side into the basis of the z0 sample (:74-85).  Slots:
  // 0 W1/LU, 1 V1/LU, 2 W1^-1 W0, 3 V1^-1 V0, 4 A, 5 B, 6 Rud, 7 Tuu, 8 Rdu,
  // 10 A - Rdu B / LU, 11-12 [Rdu A - B | Tdd] -> [Rdu' | Tdd'], 13 B Tdd',
  // 14 A + B Rdu', 15 Tuu'
  Slots* F = w.make(16, 1, rc);
  if (rc) return rc;
  if ((rc = w.copies({M.W[0], M.V[0], M.W[L - 1], M.V[L - 1], items[0].m[0], items[0].m[1],
                      items[0].m[2], items[0].m[3]},
                     {F->slot(0, 0), F->slot(0, 1), F->slot(0, 2), F->slot(0, 3), F->slot(0, 6),
                      F->slot(0, 7), F->slot(0, 8), F->slot(0, 12)})))
    return rc;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0), half = mk(0.5, 0.0),
           mhalf = mk(-0.5, 0.0);
  auto G = [&](int dst, int a, int b, cd alpha, cd beta) {
    w.gemm(F->ctab(a), F->ctab(b), F->tab(dst), alpha, beta, 1);
  };
  zgetrf_batched(s, n, F->tab(0), w.ipiv, w.info, nullptr, 1);
  zgetrs_batched(s, n, n, F->ctab(0), w.ipiv, F->tab(2), 1, w.perm);
  zgetrf_batched(s, n, F->tab(1), w.ipiv, w.info, nullptr, 1);
  zgetrs_batched(s, n, n, F->ctab(1), w.ipiv, F->tab(3), 1, w.perm);
  axpby_batched(s, n, half, F->ctab(2), half, F->ctab(3), F->tab(4), 1);
  axpby_batched(s, n, half, F->ctab(2), mhalf, F->ctab(3), F->tab(5), 1);
  copy_batched(s, w.nn, F->ctab(4), F->tab(10), 1);
  G(10, 8, 5, mone, one);                        // A - Rdu B
  G(11, 8, 4, one, zero);                        // Rdu A
  axpby_batched(s, n, one, F->ctab(11), mone, F->ctab(5), F->tab(11), 1);
  zgetrf_batched(s, n, F->tab(10), w.ipiv, w.info, nullptr, 1);
  zgetrs_batched(s, n, 2 * n, F->ctab(10), w.ipiv, F->tab(11), 1, w.perm);
  G(13, 5, 12, one, zero);                       // B Tdd'
  G(6, 7, 13, one, one);                         // Rud' = Rud + Tuu B Tdd'
  copy_batched(s, w.nn, F->ctab(4), F->tab(14), 1);
  G(14, 5, 11, one, one);                        // A + B Rdu'
  G(15, 7, 14, one, zero);                       // Tuu' = Tuu (B Rdu' + A)
  SCHK(hipGetLastError());
  rc = fetch_zresult(s, n, k0, F->slot(0, 6), F->slot(0, 15), F->slot(0, 11), F->slot(0, 12),
                     M.W[L - 1], M.V[L - 1], M.keff[L - 1], M.W[0], M.V[0], M.keff[0], res);
  if (rc) return rc;
  res->status = M.status | w.status();
  return ORCWA_OK;
}

// The adaptive first-order perturbation integrator (RedhefferIntegrator::
// compute, core/RedhefferIntegrator.cpp:173-314, triple division): the
// reference's depth-first recursion taken level by level.  Each level
// evaluates the perturbation of all its open segments in one batch; a
// segment whose terms exceed pe is cut in three (the middle third keeps its
// parent's reference modes), and the new samples of the level are solved in
// one batch.  Whether a segment is cut depends only on the segment, so the
// accepted segments are those of the recursion; their projections onto the
// previous segment's basis run as one batch, their star products as a tree.
int orcwa_gpu_redheffer_integrate(orcwa_gpu_context* ctx, const orcwa_zsampler* zs, double pe,
                                  int max_segments, orcwa_zresult* res, int* n_segments) {
  if (!ctx || !zs || !res || !zs->sample || !(zs->z1 > zs->z0) || !(zs->lambda > 0) ||
      zs->n_cx < 2 || zs->n_cy < 2 || !(pe > 0)) {
    fprintf(stderr, "[openrcwa] orcwa_gpu_redheffer_integrate: invalid argument\n");
    return ORCWA_ERR_ARG;
  }
  SCHK(hipSetDevice(ctx->device));
  ctx_trim(ctx);   // buffers cached by earlier plans go back to the device
  hipStream_t s = ctx->stream;
  const int ncx = zs->n_cx, ncy = zs->n_cy, ncell = (ncx - 1) * (ncy - 1);
  const double k0 = 2.0 * kPi / zs->lambda;
  // samples: each batch of new z is one P/Q solve
  std::vector<LayerModes> sets;
  struct ReleaseSets {
    std::vector<LayerModes>& v;
    ~ReleaseSets() {
      for (LayerModes& m : v) layer_modes_release(m);
    }
  } release_sets{sets};
  struct SampleRef {
    int set, layer;
  };
  std::vector<SampleRef> samples;
  auto add_samples = [&](const std::vector<double>& zz, std::vector<int>& ids) -> int {
    const int L = (int)zz.size() < 2 ? 2 : (int)zz.size();
    std::vector<double> cx((size_t)L * ncx), cy((size_t)L * ncy), eps((size_t)L * 2 * ncell);
    for (int l = 0; l < L; ++l) {
      const int src = l < (int)zz.size() ? l : 0;
      if (zs->sample(zs->user, zz[src], cx.data() + (size_t)l * ncx, cy.data() + (size_t)l * ncy,
                     eps.data() + (size_t)l * 2 * ncell) != 0) {
        fprintf(stderr, "[openrcwa] orcwa_gpu_redheffer_integrate: the sampler failed\n");
        return ORCWA_ERR_ARG;
      }
    }
    sets.emplace_back();
    int rc = sample_modes(ctx, zs->nx_harm, zs->ny_harm, ncx, ncy, L, zs->lambda, zs->kx0,
                          zs->ky0, cx.data(), cy.data(), eps.data(), true, sets.back());
    if (rc) {
      sets.pop_back();
      return rc;
    }
    ids.clear();
    for (int l = 0; l < (int)zz.size(); ++l) {
      ids.push_back((int)samples.size());
      samples.push_back({(int)sets.size() - 1, l});
    }
    return ORCWA_OK;
  };
  auto Wp = [&](int i) { return sets[samples[i].set].W[samples[i].layer]; };
  auto Vp = [&](int i) { return sets[samples[i].set].V[samples[i].layer]; };
  auto Kp = [&](int i) { return sets[samples[i].set].keff[samples[i].layer]; };
  auto Pp = [&](int i) { return sets[samples[i].set].P[samples[i].layer]; };
  auto Qp = [&](int i) { return sets[samples[i].set].Q[samples[i].layer]; };

  const double z0 = zs->z0, z1 = zs->z1;
  std::vector<int> ids;
  int rc = add_samples({z0, z1, 0.5 * (z0 + z1)}, ids);
  if (rc) return rc;
  const int s0 = ids[0], s1 = ids[1], sm = ids[2];
  const int n = sets[0].n;
  ZWork w;
  if ((rc = w.init(s, n, 1))) return rc;
  w.prec = ctx->gemm_prec;
  struct Seg {
    double a, b;
    int ref, sa, sm, sb;
  };
  struct Leaf {
    Seg g;
    const cd* t[4];  // tdd, tuu, rud, rdu
  };
  std::vector<Seg> open{{z0, z1, sm, s0, sm, s1}};
  std::vector<Leaf> leaves;
  int level_status = 0;
  const cd one = mk(1.0, 0.0), mone = mk(-1.0, 0.0), zero = mk(0.0, 0.0), half = mk(0.5, 0.0),
           mhalf = mk(-0.5, 0.0);
  for (int level = 0; !open.empty(); ++level) {
    if (level > 40) {
      fprintf(stderr, "[openrcwa] orcwa_gpu_redheffer_integrate: no convergence in 40 levels\n");
      return ORCWA_ERR_UNSUPPORTED;
    }
    const int K = (int)open.size();
    ZWork lw;                 // level-local tables and pivots
    if ((rc = lw.init(s, n, K))) return rc;
    lw.prec = ctx->gemm_prec;
    // slots: 0 W/LU, 1 V/LU, 2-4 X_i, 5-7 Y_i, 8 P_i - Pr, 9 Q_i - Qr, 10-13 terms
    Slots* T = w.make(14, K, rc);
    if (rc) return rc;
    std::vector<const cd*> src, Wr(K), Vr(K), kf(K);
    std::vector<cd*> dst;
    std::vector<double> k0dz(K);
    for (int k = 0; k < K; ++k) {
      const Seg& g = open[k];
      src.insert(src.end(), {Wp(g.ref), Vp(g.ref)});
      dst.insert(dst.end(), {T->slot(k, 0), T->slot(k, 1)});
      Wr[k] = Wp(g.ref);
      Vr[k] = Vp(g.ref);
      kf[k] = Kp(g.ref);
      k0dz[k] = k0 * (g.b - g.a);
    }
    if ((rc = lw.copies(src, dst))) return rc;
    const cd **dW, **dV, **dK;
    double* dk0dz;
    if ((rc = lw.table(Wr, &dW)) || (rc = lw.table(Vr, &dV)) || (rc = lw.table(kf, &dK)) ||
        (rc = lw.table(k0dz, &dk0dz)))
      return rc;
    for (int i = 0; i < 3; ++i) {
      std::vector<const cd*> Pi(K), Pr(K), Qi(K), Qr(K);
      for (int k = 0; k < K; ++k) {
        const Seg& g = open[k];
        const int si = i == 0 ? g.sa : (i == 1 ? g.sm : g.sb);
        Pi[k] = Pp(si);
        Qi[k] = Qp(si);
        Pr[k] = Pp(g.ref);
        Qr[k] = Qp(g.ref);
      }
      const cd **dPi, **dPr, **dQi, **dQr;
      if ((rc = lw.table(Pi, &dPi)) || (rc = lw.table(Pr, &dPr)) || (rc = lw.table(Qi, &dQi)) ||
          (rc = lw.table(Qr, &dQr)))
        return rc;
      axpby_batched(s, n, one, dPi, mone, dPr, T->tab(8), K);
      axpby_batched(s, n, one, dQi, mone, dQr, T->tab(9), K);
      lw.gemm(T->ctab(8), dV, T->tab(2 + i), one, zero, K);     // (P_i - Pr) V
      lw.gemm(T->ctab(9), dW, T->tab(5 + i), one, zero, K);     // (Q_i - Qr) W
    }
    zgetrf_batched(s, n, T->tab(0), lw.ipiv, lw.info, nullptr, K);
    zgetrs_batched(s, n, 3 * n, T->ctab(0), lw.ipiv, T->tab(2), K, lw.perm);
    zgetrf_batched(s, n, T->tab(1), lw.ipiv, lw.info, nullptr, K);
    zgetrs_batched(s, n, 3 * n, T->ctab(1), lw.ipiv, T->tab(5), K, lw.perm);
    unsigned long long* derr;
    if ((rc = lw.mem.get(&derr, K))) return rc;
    SCHK(hipMemsetAsync(derr, 0, sizeof(unsigned long long) * K, s));
    cd** dbase;
    std::vector<cd*> bases(K);
    for (int k = 0; k < K; ++k) bases[k] = T->slot(k, 0);
    if ((rc = lw.table(bases, &dbase))) return rc;
    const long long nn = (long long)n * n;
    const unsigned gx = (unsigned)std::min<long long>(1024, (nn + 255) / 256);
    hipLaunchKernelGGL(k_redheffer_simpson, dim3(gx, K), dim3(256), 0, s, n, dbase, dK, dk0dz,
                       derr);
    SCHK(hipGetLastError());
    std::vector<unsigned long long> herr(K);
    SCHK(hipMemcpyAsync(herr.data(), derr, sizeof(unsigned long long) * K,
                        hipMemcpyDeviceToHost, s));
    level_status |= lw.status();
    // accept or cut in three (:216-314)
    std::vector<Seg> cut;
    for (int k = 0; k < K; ++k) {
      double e;
      std::memcpy(&e, &herr[k], sizeof(double));
      if (e <= pe) {
        leaves.push_back({open[k], {T->slot(k, 10), T->slot(k, 11), T->slot(k, 12),
                                    T->slot(k, 13)}});
        if (max_segments > 0 && (int)leaves.size() > max_segments) {
          fprintf(stderr, "[openrcwa] orcwa_gpu_redheffer_integrate: more than %d segments\n",
                  max_segments);
          return ORCWA_ERR_UNSUPPORTED;
        }
      } else {
        cut.push_back(open[k]);
      }
    }
    open.clear();
    if (cut.empty()) break;
    std::vector<double> zz;
    for (const Seg& g : cut) {
      const double m0 = 2.0 / 3.0 * g.a + 1.0 / 3.0 * g.b, m1 = 1.0 / 3.0 * g.a + 2.0 / 3.0 * g.b;
      zz.insert(zz.end(), {0.5 * (g.a + m0), m0, m1, 0.5 * (m1 + g.b)});
    }
    if ((rc = add_samples(zz, ids))) return rc;
    for (size_t c = 0; c < cut.size(); ++c) {
      const Seg& g = cut[c];
      const double m0 = 2.0 / 3.0 * g.a + 1.0 / 3.0 * g.b, m1 = 1.0 / 3.0 * g.a + 2.0 / 3.0 * g.b;
      const int t1 = ids[4 * c], k0s = ids[4 * c + 1], k1s = ids[4 * c + 2], t3 = ids[4 * c + 3];
      open.push_back({g.a, m0, t1, g.sa, t1, k0s});
      open.push_back({m0, m1, g.ref, k0s, g.sm, k1s});
      open.push_back({m1, g.b, t3, k1s, t3, g.sb});
    }
  }
  std::sort(leaves.begin(), leaves.end(),
            [](const Leaf& x, const Leaf& y) { return x.g.a < y.g.a; });
  const int M = (int)leaves.size();
  if (n_segments) *n_segments = M;
  ZWork pw;
  if ((rc = pw.init(s, n, M))) return rc;
  pw.prec = ctx->gemm_prec;
  // projections onto the previous segment's basis (:242-262).  Slots: 0 W/LU,
  // 1 V/LU, 2 W^-1 leftW, 3 V^-1 leftV, 4 A, 5 B, 6 A - rdu B / LU,
  // 7-8 [rdu A - B | tdd] -> [rdu' | tdd'], 9 B tdd', 10 rud', 11 A + B rdu',
  // 12 tuu'
  Slots* Pj = w.make(13, M, rc);
  if (rc) return rc;
  {
    std::vector<const cd*> src, rdu(M), tuu(M);
    std::vector<cd*> dst;
    for (int i = 0; i < M; ++i) {
      const int ref = leaves[i].g.ref;
      const int left = i == 0 ? s0 : leaves[i - 1].g.ref;
      src.insert(src.end(), {Wp(ref), Vp(ref), Wp(left), Vp(left), leaves[i].t[0],
                             leaves[i].t[2]});
      dst.insert(dst.end(), {Pj->slot(i, 0), Pj->slot(i, 1), Pj->slot(i, 2), Pj->slot(i, 3),
                             Pj->slot(i, 8), Pj->slot(i, 10)});
      rdu[i] = leaves[i].t[3];
      tuu[i] = leaves[i].t[1];
    }
    const cd **drdu, **dtuu;
    if ((rc = pw.copies(src, dst)) || (rc = pw.table(rdu, &drdu)) || (rc = pw.table(tuu, &dtuu)))
      return rc;
    zgetrf_batched(s, n, Pj->tab(0), pw.ipiv, pw.info, nullptr, M);
    zgetrs_batched(s, n, n, Pj->ctab(0), pw.ipiv, Pj->tab(2), M, pw.perm);
    zgetrf_batched(s, n, Pj->tab(1), pw.ipiv, pw.info, nullptr, M);
    zgetrs_batched(s, n, n, Pj->ctab(1), pw.ipiv, Pj->tab(3), M, pw.perm);
    axpby_batched(s, n, half, Pj->ctab(2), half, Pj->ctab(3), Pj->tab(4), M);
    axpby_batched(s, n, half, Pj->ctab(2), mhalf, Pj->ctab(3), Pj->tab(5), M);
    copy_batched(s, w.nn, Pj->ctab(4), Pj->tab(6), M);
    pw.gemm(drdu, Pj->ctab(5), Pj->tab(6), mone, one, M);        // A - rdu B
    pw.gemm(drdu, Pj->ctab(4), Pj->tab(7), one, zero, M);        // rdu A
    axpby_batched(s, n, one, Pj->ctab(7), mone, Pj->ctab(5), Pj->tab(7), M);
    zgetrf_batched(s, n, Pj->tab(6), pw.ipiv, pw.info, nullptr, M);
    zgetrs_batched(s, n, 2 * n, Pj->ctab(6), pw.ipiv, Pj->tab(7), M, pw.perm);
    pw.gemm(Pj->ctab(5), Pj->ctab(8), Pj->tab(9), one, zero, M);  // B tdd'
    pw.gemm(dtuu, Pj->ctab(9), Pj->tab(10), one, one, M);         // rud' = rud + tuu B tdd'
    copy_batched(s, w.nn, Pj->ctab(4), Pj->tab(11), M);
    pw.gemm(Pj->ctab(5), Pj->ctab(7), Pj->tab(11), one, one, M);  // A + B rdu'
    pw.gemm(dtuu, Pj->ctab(11), Pj->tab(12), one, zero, M);       // tuu' = tuu (A + B rdu')
  }
  std::vector<ZWork::Item> items(M);
  for (int i = 0; i < M; ++i)
    items[i] = {{Pj->slot(i, 10), Pj->slot(i, 12), Pj->slot(i, 7), Pj->slot(i, 8)}};
  if ((rc = pw.star_tree(items))) return rc;
  // the top side into the basis at z1 (:198-213).  Slots: 0 leftW/LU,
  // 1 leftV/LU, 2 leftW^-1 W1, 3 leftV^-1 V1, 4 A, 5 B, 6 A - Rud B / LU,
  // 7-8 [Tuu | Rud A - B] -> [Tuu' | Rud'], 9 B Tuu', 10 Rdu', 11 A + B Rud',
  // 12 Tdd', 13 Rud, 14 Tdd
  Slots* F = w.make(15, 1, rc);
  if (rc) return rc;
  const int last = leaves[M - 1].g.ref;
  if ((rc = pw.copies({Wp(last), Vp(last), Wp(s1), Vp(s1), items[0].m[1], items[0].m[2],
                       items[0].m[0], items[0].m[3]},
                      {F->slot(0, 0), F->slot(0, 1), F->slot(0, 2), F->slot(0, 3),
                       F->slot(0, 7), F->slot(0, 10), F->slot(0, 13), F->slot(0, 14)})))
    return rc;
  auto G = [&](int dst, int a, int b, cd alpha, cd beta) {
    pw.gemm(F->ctab(a), F->ctab(b), F->tab(dst), alpha, beta, 1);
  };
  zgetrf_batched(s, n, F->tab(0), pw.ipiv, pw.info, nullptr, 1);
  zgetrs_batched(s, n, n, F->ctab(0), pw.ipiv, F->tab(2), 1, pw.perm);
  zgetrf_batched(s, n, F->tab(1), pw.ipiv, pw.info, nullptr, 1);
  zgetrs_batched(s, n, n, F->ctab(1), pw.ipiv, F->tab(3), 1, pw.perm);
  axpby_batched(s, n, half, F->ctab(2), half, F->ctab(3), F->tab(4), 1);
  axpby_batched(s, n, half, F->ctab(2), mhalf, F->ctab(3), F->tab(5), 1);
  copy_batched(s, w.nn, F->ctab(4), F->tab(6), 1);
  G(6, 13, 5, mone, one);                        // A - Rud B
  G(8, 13, 4, one, zero);                        // Rud A
  axpby_batched(s, n, one, F->ctab(8), mone, F->ctab(5), F->tab(8), 1);
  zgetrf_batched(s, n, F->tab(6), pw.ipiv, pw.info, nullptr, 1);
  zgetrs_batched(s, n, 2 * n, F->ctab(6), pw.ipiv, F->tab(7), 1, pw.perm);
  G(9, 5, 7, one, zero);                         // B Tuu'
  G(10, 14, 9, one, one);                        // Rdu' = Rdu + Tdd B Tuu'
  copy_batched(s, w.nn, F->ctab(4), F->tab(11), 1);
  G(11, 5, 8, one, one);                         // A + B Rud'
  G(12, 14, 11, one, zero);                      // Tdd' = Tdd (B Rud' + A)
  SCHK(hipGetLastError());
  rc = fetch_zresult(s, n, k0, F->slot(0, 8), F->slot(0, 7), F->slot(0, 10), F->slot(0, 12),
                     Wp(s0), Vp(s0), Kp(s0), Wp(s1), Vp(s1), Kp(s1), res);
  if (rc) return rc;
  int st = level_status | pw.status();
  for (const LayerModes& m : sets) st |= m.status;
  res->status = st;
  return ORCWA_OK;
}

}  // extern "C"
