/*
 * openrcwa_surrogate.h -- C ABI of the learned surrogate: a small multilayer
 * perceptron (an ensemble of M independently initialised members) trained and
 * evaluated on the GPU with the engine's own HIP kernels.
 *
 * The reference has no counterpart: it solves every design point with
 * RCWASolver.  The surrogate maps grating parameters to the per-order
 * efficiencies that rcwa/RCWASolver.cpp:1792-1908 (scatterPlaneWave) returns
 * for the driver loop of sol/rcwa_bridge.cpp:244-334; its training data comes
 * from orcwa_gpu_solve_batch / orcwa_plan_*, and the caller answers the
 * queries whose error estimate is too high with the plain solver.
 *
 * Network of one member: x (n_in) -> [Linear(width) -> GELU] x depth ->
 * Linear(n_out); GELU(z) = z (1 + erf(z / sqrt 2)) / 2.  FP32 parameters,
 * FP32 MFMA products with FP32 accumulation (BF16 products for inference
 * with orcwa_mlp_set_precision).  Loss of a member on a batch of B rows:
 * (1 / (B n_out)) sum (y - t)^2.  Inputs are expected scaled to about
 * [-1, 1] by the caller.
 *
 * Parameter layout of one member (FP32, contiguous, in this order):
 * W_0 [n_in][width], b_0 [width], W_1 [width][width], b_1, ...,
 * W_depth [width][n_out], b_depth [n_out]; row-major W_l[i][j] multiplies
 * input i into output j.  Members follow each other.
 *
 * Calls on one network must not overlap; they run on the stream of the
 * context the network was created on.
 */
#ifndef OPENRCWA_SURROGATE_H
#define OPENRCWA_SURROGATE_H

#include "openrcwa_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orcwa_mlp orcwa_mlp;

/* precision of the products of orcwa_mlp_forward / orcwa_mlp_predict
   (training always uses FP32) */
enum orcwa_mlp_precision { ORCWA_MLP_FP32 = 0, ORCWA_MLP_BF16 = 1 };

/* Create an ensemble of `members` networks (1..16), parameters drawn from
 * N(0, 2 / fan_in) (He) with zero biases from a seeded generator.
 * n_in 1..64, width 32..1024 (multiple of 32), depth 1..8, n_out 1..64. */
int orcwa_mlp_create(orcwa_gpu_context *ctx, int n_in, int width, int depth, int n_out,
                     int members, unsigned long long seed, orcwa_mlp **net);
int orcwa_mlp_destroy(orcwa_mlp *net);

/* Number of FP32 parameters of one member. */
long long orcwa_mlp_param_count(const orcwa_mlp *net);
/* Copy all members' parameters out ([members][param_count]) / in. */
int orcwa_mlp_get_params(orcwa_mlp *net, float *params);
int orcwa_mlp_set_params(orcwa_mlp *net, const float *params);

/* Training set: count rows of x [count][n_in] and targets t [count][n_out],
 * copied to device memory (replacing any earlier set). */
int orcwa_mlp_set_data(orcwa_mlp *net, const float *x, const float *t, long long count);

typedef struct {
  int epochs;                /* passes over the training set */
  int batch;                 /* rows per step (each member draws its own) */
  float lr_max, lr_min;      /* cosine schedule over all steps, from lr_max to lr_min */
  float beta1, beta2, eps;   /* Adam; 0 selects 0.9, 0.999, 1e-8 */
  unsigned long long seed;   /* shuffles: every member its own order per epoch */
  int use_graph;             /* 1: one training step is captured in a hipGraph once and
                                replayed; 0: launched kernel by kernel */
} orcwa_mlp_train_opts;

/* Adam on the MSE loss over the training set; every member trains on its own
 * shuffled order of the rows.  loss (optional, [epochs]) receives the mean
 * training loss of each epoch, averaged over members. */
int orcwa_mlp_train(orcwa_mlp *net, const orcwa_mlp_train_opts *opts, float *loss);

/* Loss and its gradient for one batch of rows (count <= 65536), every member:
 * grad [members][param_count], loss [members].  No parameter update. */
int orcwa_mlp_gradient(orcwa_mlp *net, const float *x, const float *t, int count, float *grad,
                       float *loss);

int orcwa_mlp_set_precision(orcwa_mlp *net, int precision);

/* Every member's output: y [members][count][n_out] (host arrays). */
int orcwa_mlp_forward(orcwa_mlp *net, const float *x, long long count, float *y);

/* Scale of each output's member spread in the error estimate (n_out values,
 * >= 0; all 1 after creation). */
int orcwa_mlp_set_error_scale(orcwa_mlp *net, const float *scale);

/* Ensemble prediction: mean [count][n_out] over members and the error
 * estimate err [count] = max over outputs c of scale[c] (max - min over
 * members). */
int orcwa_mlp_predict(orcwa_mlp *net, const float *x, long long count, float *mean, float *err);

/* Same on device pointers (x, mean, err in device memory), asynchronous on the
 * context stream. */
int orcwa_mlp_predict_device(orcwa_mlp *net, const float *x, long long count, float *mean,
                             float *err);

/* orcwa_mlp_predict_device without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops): the
 * padded inputs, activations and member outputs of each pass live in a static
 * 64 MiB device buffer that every GPU holds until the process ends.  Rows are
 * processed in passes of as many as fit, 4 (n_in + 2 members width + members
 * n_out) bytes per row with n_in and n_out padded (count is not limited; every
 * network orcwa_mlp_create accepts fits at least 448 rows, and a pass of fewer
 * than 64 rows would be refused with ORCWA_ERR_ARG).  Every call on one GPU
 * shares that GPU's buffer: calls that
 * overlap on one GPU race (from different threads or contexts), and the buffer
 * keeps the last call's contents.  Asynchronous on the context stream; the
 * network's parameters and error scales must already exist (orcwa_mlp_create).
 * No counterpart in the reference. */
int orcwa_mlp_predict_device_noalloc(orcwa_mlp *net, const float *x, long long count, float *mean,
                                     float *err);

/* GPU time in ms of the last orcwa_mlp_train / orcwa_mlp_predict (HIP events on
 * the context stream). */
double orcwa_mlp_last_ms(const orcwa_mlp *net);

#ifdef __cplusplus
}
#endif

#endif /* OPENRCWA_SURROGATE_H */
