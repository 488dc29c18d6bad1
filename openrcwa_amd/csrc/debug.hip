// Component entry points of include/openrcwa_gpu_debug.h: one batched kernel
// family per call, on host buffers.
#include <cstring>
#include <vector>

#include "../../include/openrcwa_gpu_debug.h"
#include "context.h"
#include "rcwa_kernels.h"

using namespace orcwa;

namespace {

#define DCHK(expr)                                                                          \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      fprintf(stderr, "[openrcwa] HIP error %s at %s:%d\n", hipGetErrorName(_e), __FILE__,  \
              __LINE__);                                                                    \
      return ORCWA_ERR_HIP;                                                                 \
    }                                                                                       \
  } while (0)

// device buffer of `count` matrices of `elems` complex each + its pointer table
struct DBatch {
  cd* data = nullptr;
  cd** table = nullptr;
  ~DBatch() {
    if (data) (void)hipFree(data);
    if (table) (void)hipFree(table);
  }
  int init(hipStream_t s, long long elems, int count, const double* host) {
    if (hipMalloc(&data, sizeof(cd) * elems * count) != hipSuccess) return ORCWA_ERR_NOMEM;
    if (hipMalloc(&table, sizeof(cd*) * count) != hipSuccess) return ORCWA_ERR_NOMEM;
    std::vector<cd*> t(count);
    for (int i = 0; i < count; ++i) t[i] = data + (long long)i * elems;
    DCHK(hipMemcpyAsync(table, t.data(), sizeof(cd*) * count, hipMemcpyHostToDevice, s));
    if (host)
      DCHK(hipMemcpyAsync(data, host, sizeof(cd) * elems * count, hipMemcpyHostToDevice, s));
    else
      DCHK(hipMemsetAsync(data, 0, sizeof(cd) * elems * count, s));
    return ORCWA_OK;
  }
};

}  // namespace

extern "C" {

int orcwa_dbg_zgemm(orcwa_gpu_context* ctx, int M, int N, int K, const double* alpha,
                    const double* A, int ldA, int tA, int cA, const double* B, int ldB, int tB,
                    int cB, const double* beta, double* C, int ldC, int batch) {
  if (!ctx || M <= 0 || N <= 0 || K <= 0 || batch <= 0) return ORCWA_ERR_ARG;
  DCHK(hipSetDevice(ctx->device));
  orcwa::ctx_trim(ctx);   // cached plan buffers go back to the device
  hipStream_t s = ctx->stream;
  const long long eA = (long long)ldA * (tA ? M : K), eB = (long long)ldB * (tB ? K : N),
                  eC = (long long)ldC * N;
  DBatch a, b, c;
  int rc;
  if ((rc = a.init(s, eA, batch, A)) || (rc = b.init(s, eB, batch, B)) ||
      (rc = c.init(s, eC, batch, C)))
    return rc;
  // op(A)[i,k] = A[i*rs + k*cs]
  const long long rsA = tA ? ldA : 1, csA = tA ? 1 : ldA;
  const long long rsB = tB ? ldB : 1, csB = tB ? 1 : ldB;
  zgemm_batched_prec(ctx->gemm_prec, s, M, N, K, mk(alpha[0], alpha[1]),
                     (const cd* const*)a.table, 0, rsA, csA, cA, (const cd* const*)b.table, 0,
                     rsB, csB, cB, mk(beta[0], beta[1]), c.table, 0, 1, ldC, batch);
  DCHK(hipGetLastError());
  DCHK(hipMemcpyAsync(C, c.data, sizeof(cd) * eC * batch, hipMemcpyDeviceToHost, s));
  DCHK(hipStreamSynchronize(s));
  return ORCWA_OK;
}

int orcwa_dbg_zgesv(orcwa_gpu_context* ctx, int n, int nrhs, const double* A, const double* B,
                    double* X, int* info, int batch) {
  if (!ctx || n <= 0 || nrhs <= 0 || batch <= 0) return ORCWA_ERR_ARG;
  DCHK(hipSetDevice(ctx->device));
  orcwa::ctx_trim(ctx);   // cached plan buffers go back to the device
  hipStream_t s = ctx->stream;
  DBatch a, b;
  int rc;
  if ((rc = a.init(s, (long long)n * n, batch, A)) || (rc = b.init(s, (long long)n * nrhs, batch, B)))
    return rc;
  int *ipiv = nullptr, *dinfo = nullptr, *perm = nullptr;
  DCHK(hipMalloc(&ipiv, sizeof(int) * n * batch));
  DCHK(hipMalloc(&perm, sizeof(int) * n * batch));
  DCHK(hipMalloc(&dinfo, sizeof(int) * batch));
  DCHK(hipMemsetAsync(dinfo, 0, sizeof(int) * batch, s));
  zgetrf_batched(s, n, a.table, ipiv, dinfo, nullptr, batch);
  // odd batches exercise the swap-by-swap path, even ones the gather
  zgetrs_batched(s, n, nrhs, (const cd* const*)a.table, ipiv, b.table, batch,
                 (batch % 2 == 0) ? perm : nullptr);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(X, b.data, sizeof(cd) * n * nrhs * batch, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(info, dinfo, sizeof(int) * batch, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(ipiv);
  (void)hipFree(perm);
  (void)hipFree(dinfo);
  DCHK(e);
  return ORCWA_OK;
}

int orcwa_dbg_zgeev(orcwa_gpu_context* ctx, int n, const double* A, double* lam, double* W,
                    double* resid, int* info, int batch) {
  if (!ctx || n <= 0 || batch <= 0) return ORCWA_ERR_ARG;
  DCHK(hipSetDevice(ctx->device));
  orcwa::ctx_trim(ctx);   // cached plan buffers go back to the device
  hipStream_t s = ctx->stream;
  const long long nn = (long long)n * n;
  DBatch a, ao, w, z, x;
  int rc;
  if ((rc = a.init(s, nn, batch, A)) || (rc = ao.init(s, nn, batch, A)) ||
      (rc = w.init(s, nn, batch, nullptr)) || (rc = z.init(s, nn, batch, nullptr)) ||
      (rc = x.init(s, nn, batch, nullptr)))
    return rc;
  cd *dlam = nullptr, *dtau = nullptr;
  int* dinfo = nullptr;
  double* dres = nullptr;
  DCHK(hipMalloc(&dlam, sizeof(cd) * n * batch));
  DCHK(hipMalloc(&dtau, sizeof(cd) * n * batch));
  DCHK(hipMalloc(&dinfo, sizeof(int) * batch));
  DCHK(hipMalloc(&dres, sizeof(double) * batch));
  DCHK(hipMemsetAsync(dinfo, 0, sizeof(int) * batch, s));
  EigWorkspace ws;
  ws.Z = z.table;
  ws.X = x.table;
  ws.tau = dtau;
  zgeev_batched(s, n, a.table, (const cd* const*)ao.table, w.table, dlam, ws, dinfo, dres,
                x.table, batch, nullptr, ctx->eig_tol);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(lam, dlam, sizeof(cd) * n * batch, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(W, w.data, sizeof(cd) * nn * batch, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(resid, dres, sizeof(double) * batch, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(info, dinfo, sizeof(int) * batch, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dlam);
  (void)hipFree(dtau);
  (void)hipFree(dinfo);
  (void)hipFree(dres);
  DCHK(e);
  return ORCWA_OK;
}

int orcwa_dbg_set_eig_tolerance(orcwa_gpu_context* ctx, double tol) {
  if (!ctx || tol != tol) return ORCWA_ERR_ARG;
  ctx->eig_tol = tol;
  return ORCWA_OK;
}

int orcwa_dbg_convolution(orcwa_gpu_context* ctx, int nx_harm, int ny_harm, int n_cx, int n_cy,
                          const double* coord_x, const double* coord_y, const double* eps,
                          double* fe11, double* fe22, double* fe33) {
  if (!ctx || nx_harm < 0 || ny_harm < 0 || n_cx < 2 || n_cy < 2) return ORCWA_ERR_ARG;
  const int nx = 2 * nx_harm + 1, ny = 2 * ny_harm + 1;
  if (nx > 63 || ny > 63) return ORCWA_ERR_ARG;
  DCHK(hipSetDevice(ctx->device));
  orcwa::ctx_trim(ctx);   // cached plan buffers go back to the device
  hipStream_t s = ctx->stream;
  const int h = nx * ny, ncell = (n_cx - 1) * (n_cy - 1);
  const long long rec0 = (long long)nx * nx + (2 * ny - 1) + (2 * nx - 1);
  const long long rec1 = (long long)ny * ny + (2 * nx - 1);
  std::vector<StripTask> tasks;
  for (int i = 0; i < n_cy - 1; ++i) tasks.push_back({0, 0, i, 0, i * rec0});
  const long long b1 = (n_cy - 1) * rec0;
  for (int i = 0; i < n_cx - 1; ++i) tasks.push_back({0, 1, i, 0, b1 + i * rec1});
  long long sbase[2] = {0, b1};
  double *dcx = nullptr, *dcy = nullptr;
  StripTask* dtasks = nullptr;
  long long* dsb = nullptr;
  DBatch grid, strips, e11, e22, e33;
  int rc;
  if ((rc = grid.init(s, ncell, 1, eps)) || (rc = strips.init(s, b1 + (n_cx - 1) * rec1, 1, nullptr)) ||
      (rc = e11.init(s, (long long)h * h, 1, nullptr)) || (rc = e22.init(s, (long long)h * h, 1, nullptr)) ||
      (rc = e33.init(s, (long long)h * h, 1, nullptr)))
    return rc;
  DCHK(hipMalloc(&dcx, sizeof(double) * n_cx));
  DCHK(hipMalloc(&dcy, sizeof(double) * n_cy));
  DCHK(hipMalloc(&dtasks, sizeof(StripTask) * tasks.size()));
  DCHK(hipMalloc(&dsb, sizeof(sbase)));
  DCHK(hipMemcpyAsync(dcx, coord_x, sizeof(double) * n_cx, hipMemcpyHostToDevice, s));
  DCHK(hipMemcpyAsync(dcy, coord_y, sizeof(double) * n_cy, hipMemcpyHostToDevice, s));
  DCHK(hipMemcpyAsync(dtasks, tasks.data(), sizeof(StripTask) * tasks.size(), hipMemcpyHostToDevice, s));
  DCHK(hipMemcpyAsync(dsb, sbase, sizeof(sbase), hipMemcpyHostToDevice, s));
  ConvGeom g;
  g.nx = nx; g.ny = ny; g.ncx = n_cx; g.ncy = n_cy; g.ncell = ncell;
  g.Lx = coord_x[n_cx - 1] - coord_x[0];
  g.Ly = coord_y[n_cy - 1] - coord_y[0];
  g.cx = dcx; g.cy = dcy;
  g.sx = 0; g.sy = 0;
  const int mdim = nx > ny ? nx : ny;
  const size_t lds = (size_t)(2 * mdim - 1 + mdim * mdim + (n_cx > n_cy ? n_cx : n_cy)) * sizeof(cd);
  launch_conv_strip(s, g, dtasks, (int)tasks.size(), grid.data, strips.data, lds);
  launch_conv_expand(s, g, strips.data, dsb, e11.table, e22.table, e33.table, 1);
  hipError_t e = hipGetLastError();
  const size_t bytes = sizeof(cd) * (size_t)h * h;
  if (e == hipSuccess) e = hipMemcpyAsync(fe11, e11.data, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(fe22, e22.data, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(fe33, e33.data, bytes, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(dcx);
  (void)hipFree(dcy);
  (void)hipFree(dtasks);
  (void)hipFree(dsb);
  DCHK(e);
  return ORCWA_OK;
}

}  // extern "C"
