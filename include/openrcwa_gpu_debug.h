/*
 * openrcwa_gpu_debug.h -- component entry points of the MI355X RCWA engine,
 * for tests that check one kernel family at a time against a CPU reference.
 *
 * These have no counterpart in the reference (whose components are C++
 * classes, not an ABI): each one runs a single batched kernel family of the
 * engine on host buffers (copied in and out) and returns an orcwa_error code.
 * Complex arrays are interleaved (re, im) doubles, matrices column-major.
 */
#ifndef OPENRCWA_GPU_DEBUG_H
#define OPENRCWA_GPU_DEBUG_H

#include "openrcwa_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* C[b] = alpha op(A[b]) op(B[b]) + beta C[b]; op = transpose (t) and/or
 * conjugate (c) flags per operand; A is M x K after op, B is K x N after op.
 * ldA/ldB/ldC are the column-major leading dimensions of the stored arrays.
 * The products run at the context's orcwa_gpu_set_gemm_precision setting. */
int orcwa_dbg_zgemm(orcwa_gpu_context *ctx, int M, int N, int K, const double *alpha,
                    const double *A, int ldA, int tA, int cA, const double *B, int ldB, int tB,
                    int cB, const double *beta, double *C, int ldC, int batch);

/* X = A^{-1} B for a batch of n x n A and n x nrhs B (LU with partial
 * pivoting); info[b] gets ORCWA_STATUS_SINGULAR for an exactly zero pivot. */
int orcwa_dbg_zgesv(orcwa_gpu_context *ctx, int n, int nrhs, const double *A, const double *B,
                    double *X, int *info, int batch);

/* Eigenvalues lam[b][n] and unit-norm right eigenvectors W[b] (n x n) of a
 * batch of general complex matrices; resid[b] = ||AW - W diag(lam)||max/||A||max;
 * info[b] status bits. */
int orcwa_dbg_zgeev(orcwa_gpu_context *ctx, int n, const double *A, double *lam, double *W,
                    double *resid, int *info, int batch);

/* Residual acceptance of the eigen solves of later plans and zgeev calls on
 * this context (default 1e-8, rcwa/MKLEigenSolver.hpp:75-79).  A negative
 * tolerance rejects every solve: it forces the EIG_RESIDUAL status and, in
 * plans, the perturbed recovery solve (ORCWA_STATUS_EIG_PERTURBED), so tests
 * can exercise those paths on ordinary inputs. */
int orcwa_dbg_set_eig_tolerance(orcwa_gpu_context *ctx, double tol);

/* fe11, fe22, fe33 ((2Nx+1)(2Ny+1) square, column-major) of one permittivity
 * grid eps[(n_cy-1)][(n_cx-1)] on the given edges (Li rules of
 * rcwa/FourierSolver2D.cpp:12-351). */
int orcwa_dbg_convolution(orcwa_gpu_context *ctx, int nx_harm, int ny_harm, int n_cx, int n_cy,
                          const double *coord_x, const double *coord_y, const double *eps,
                          double *fe11, double *fe22, double *fe33);

#ifdef __cplusplus
}
#endif

#endif /* OPENRCWA_GPU_DEBUG_H */
