// Internal interfaces of the OpenRCWA MI355X engine (host-side launchers of
// the batched kernels).  Every launcher enqueues work on `s` and returns
// immediately.  Batched operands are device arrays of base pointers (one per
// batch entry) plus an element offset applied to every entry, so that
// sub-blocks of a batch of matrices can be addressed without new tables.
#pragma once

#include <vector>

#include "orcwa_common.h"
#include "../../include/openrcwa_gpu.h"

namespace orcwa {

// residual acceptance of an eigen solve (rcwa/MKLEigenSolver.hpp:75-79)
constexpr double kEigTol = 1e-8;

// Batched complex GEMM on FP64 MFMA:
//   C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b]
// op(X)[i,k] = X[off + i*rs + k*cs], conjugated when conj != 0.
// beta == 0 does not read C.
void zgemm_batched(hipStream_t s, int M, int N, int K, cd alpha,
                   const cd* const* A, long long offA, long long rsA, long long csA, int conjA,
                   const cd* const* B, long long offB, long long rsB, long long csB, int conjB,
                   cd beta, cd* const* C, long long offC, long long rsC, long long csC,
                   int batch);

// Precision of the products of zgemm_batched_prec (the cascade / star-product
// GEMMs); the values are those of orcwa_gemm_precision in openrcwa_gpu.h.
enum GemmPrec { GEMM_FP64 = 0, GEMM_FP32 = 1, GEMM_BF16X3 = 2, GEMM_BF16X6 = 3 };
// zgemm_batched with the products in FP64 (GEMM_FP64: the same kernels), FP32
// or split BF16 on MFMA, every K step's partial sums accumulated in FP64.
void zgemm_batched_prec(int prec, hipStream_t s, int M, int N, int K, cd alpha,
                        const cd* const* A, long long offA, long long rsA, long long csA,
                        int conjA, const cd* const* B, long long offB, long long rsB,
                        long long csB, int conjB, cd beta, cd* const* C, long long offC,
                        long long rsC, long long csC, int batch, int kup = -1);
// zgemm_batched for an upper-triangular op(B) (op(B)[k][j] == 0 for k > j,
// stored zeros included): each output tile's K range ends at its last column.
void zgemm_batched_upper_b(hipStream_t s, int M, int N, int K, cd alpha,
                           const cd* const* A, long long offA, long long rsA, long long csA,
                           int conjA, const cd* const* B, long long offB, long long rsB,
                           long long csB, int conjB, cd beta, cd* const* C, long long offC,
                           long long rsC, long long csC, int batch);

// Tile of the multiple-of-64 main block of the GEMMs the calling thread issues
// (the other tiles cover the edge strips); set per plan by the autotuner.
enum GemmTile { GEMM_TILE_64x64 = 0, GEMM_TILE_32x64 = 1, GEMM_TILE_64x32 = 2, GEMM_NTILES = 3 };
void set_gemm_tile(int tile);
int gemm_tile();

// Batched LU with partial pivoting of n x n column-major matrices (lda = n),
// in place.  ipiv[b*n + i] = global row swapped with row i.  info[b] gets
// ORCWA_STATUS_SINGULAR or-ed in when an exactly zero pivot is met
// (info[info_map[b]] when info_map is not null).
void zgetrf_batched(hipStream_t s, int n, cd* const* A, int* ipiv, int* info,
                    const int* info_map, int batch);

// Solve A X = B in place of B (n x nrhs, column-major, ldb = n) with the
// factors of zgetrf_batched.  perm_ws: optional batch x n ints of workspace
// (the row interchanges are then applied as one gather).
void zgetrs_batched(hipStream_t s, int n, int nrhs, const cd* const* LU, const int* ipiv,
                    cd* const* B, int batch, int* perm_ws = nullptr);

// Elementwise helpers over batches of n x n column-major matrices.
void set_identity_batched(hipStream_t s, int n, cd* const* A, int batch);
// C = a*A + b*B (+ d*I when addI)
void axpby_batched(hipStream_t s, int n, cd a, const cd* const* A, cd b, const cd* const* B,
                   cd* const* C, int batch);
void copy_batched(hipStream_t s, long long count, const cd* const* A, cd* const* B, int batch);

// Eigensolver for a batch of general complex n x n matrices (column-major):
// A[b] is overwritten (Schur form).  On exit W[b] holds unit-2-norm
// eigenvectors and lam[b*n + j] the eigenvalues.  work must hold 2 n*n per
// matrix (Z and X).  Aorig (may equal nullptr) is used for the residual test
// (accepted when ||Aorig W - W diag(lam)||max / ||Aorig||max <= tol, else
// ORCWA_STATUS_EIG_RESIDUAL); status bits are or-ed into info[b].  resid[b]
// receives the relative residual.
struct EigWorkspace {
  cd** Z;        // [batch] n*n
  cd** X;        // [batch] n*n
  cd* tau;       // [batch*n]
};
// Variants of the multishift QR (window / shifts / workgroups per CU):
// QR_W48 48 / 12 / 2 (default), QR_W32 32 / 8 / 3, QR_W64 64 / 16 / 1.
enum QrVariant { QR_W48 = 0, QR_W32 = 1, QR_W64 = 2, QR_NVARIANTS = 3 };
void zgeev_batched(hipStream_t s, int n, cd* const* A, const cd* const* Aorig, cd* const* W,
                   cd* lam, const EigWorkspace& ws, int* info, double* resid,
                   cd* const* scratch, int batch, hipEvent_t* phase_events,
                   double tol = kEigTol, int qr_variant = QR_W48);
// A[b] = Aorig[b] + diag(1e-11 (i+1) max|Aorig[b]|): the matrix of the
// reference's recovery solve after a failed residual test
// (rcwa/MKLEigenSolver.hpp:81-93).
void eig_perturb_batched(hipStream_t s, int n, const cd* const* Aorig, cd* const* A, int batch);
// For k < m: lam[idx[k]] (n values) = lam_k[k], info[idx[k]] = info_k[k] |
// ORCWA_STATUS_EIG_PERTURBED, resid[idx[k]] = res_k[k].
void eig_recover_scatter(hipStream_t s, int n, int m, const int* idx, const cd* lam_k,
                         const int* info_k, const double* res_k, cd* lam, int* info,
                         double* resid);

// Modes (W, V, keff in k0 units) of every layer of one instance, each layer
// with its own cell edges layer_cx [L][n_cx], layer_cy [L][n_cy] (same
// periods): a plan that stops after the layer modes.  The device pointers
// stay valid until layer_modes_release.  status: orcwa_status bits of the
// eigensolves.  With want_pq every layer (homogeneous ones too) is solved
// from its P and Q, which are returned as well (k0 units: P/k0^2, Q/k0^2).
struct LayerModes {
  orcwa_plan* plan = nullptr;
  int n = 0;
  int status = 0;
  std::vector<const cd*> W, V, keff, P, Q;
};
int layer_modes(orcwa_gpu_context* ctx, const orcwa_stack* st, const orcwa_instance* inst,
                const double* layer_cx, const double* layer_cy, LayerModes& out,
                bool want_pq = false);
void layer_modes_release(LayerModes& m);

}  // namespace orcwa
