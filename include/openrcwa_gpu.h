/*
 * openrcwa_gpu.h -- C ABI of the MI355X RCWA layer-stack engine.
 *
 * Two entry families:
 *
 *  1. rcwa_run(FILE*): the reference's C entry point for the .ofd RCWA path,
 *     same globals, same rcwa_efficiency.csv / rcwa_orders.csv bytes and log
 *     lines.  Replaces sol/rcwa_bridge.cpp:145 (declared in the reference at
 *     include/orcwa_rcwa.h:72-74, called from src/sol_Main.c:81-82).
 *
 *  2. A batched ABI behind the RCWASolver facade: one call solves many
 *     independent (wavelength, Bloch vector, permittivity) instances of one
 *     stack layout.  It replaces the per-wavelength sequence
 *       RCWASolver(nX,nY) / addLayer / disablePML / setBlochWavevector /
 *       solve / generateHorizontalPlaneWave / scatterPlaneWave
 *     (rcwa/RCWASolver.h:67-185, rcwa/RCWASolver.cpp:108-148, 1778-1908) that
 *     sol/rcwa_bridge.cpp:244-334 and rcwa/RCWADriver.cpp:282-368 run one
 *     wavelength at a time.
 *
 * Plain pointers and sizes only; no exceptions cross the ABI; every function
 * returns an orcwa_error code.  A context owns one HIP stream and its device
 * memory; calls on one context must not overlap, distinct contexts are
 * independent.
 */
#ifndef OPENRCWA_GPU_H
#define OPENRCWA_GPU_H

#include <stddef.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
enum orcwa_error {
  ORCWA_OK = 0,
  ORCWA_ERR_ARG = 1,          /* invalid argument (message on stderr) */
  ORCWA_ERR_HIP = 2,          /* HIP runtime error, or no usable GPU */
  ORCWA_ERR_NOMEM = 3,        /* device allocation failed */
  ORCWA_ERR_UNSUPPORTED = 4   /* input outside what the engine implements */
};

/* per-instance status bits (orcwa_result.status) */
enum orcwa_status {
  ORCWA_STATUS_OK = 0,
  ORCWA_STATUS_EIG_RESIDUAL = 1,  /* a layer's ||AW-W diag(l)||/||A|| exceeded 1e-8
                                     (acceptance test of rcwa/MKLEigenSolver.hpp:66-79) */
  ORCWA_STATUS_GRAZING = 2,       /* kz == 0 for some order, or kz_inc < 1e-12 k0: the
                                     reference returns empty REF/TRN
                                     (rcwa/RCWASolver.cpp:1843-1862, 1898-1901);
                                     ref/trn are left zero */
  ORCWA_STATUS_SINGULAR = 4,      /* exactly singular pivot in an interface/cascade solve */
  ORCWA_STATUS_NOCONV = 8,        /* the QR iteration hit its iteration limit */
  ORCWA_STATUS_EIG_PERTURBED = 16,/* a layer's eigen solve failed the residual test (or did
                                     not converge) and was redone, as the reference does, on
                                     A + diag(1e-11 (i+1) max|A|) (rcwa/MKLEigenSolver.hpp:
                                     81-93, 160-176); EIG_RESIDUAL stays set only if the
                                     redone solve also fails the test against A */
  ORCWA_STATUS_ILLCOND = 32       /* a layer carries a mode that grows by more than 1e4 across
                                     its thickness (Im keff < 0 under the GammaSqrtOp branch of
                                     rcwa/WaveEquationSolver.cpp:13-21): the reference's RTCM
                                     cascade (rcwa/RCWASolver.cpp:415-460) then amplifies
                                     rounding by about the square of that growth, and the
                                     result may differ from any other FP64 evaluation */
};

typedef struct orcwa_gpu_context orcwa_gpu_context;
typedef struct orcwa_plan orcwa_plan;

/* Stack layout shared by every instance of a batch.  Mirrors what
 * RCWASolver(nX, nY) + addLayer(Layer(coordX, coordY, eps)) fixes
 * (rcwa/RCWASolver.h:67-78, rcwa/Layer.h:29-37): all layers share the cell
 * edges, as RCWASolver::addLayer requires (rcwa/RCWASolver.cpp:125-133). */
typedef struct {
  int nx_harm;               /* Nx: x orders -Nx..Nx */
  int ny_harm;               /* Ny: y orders -Ny..Ny (0 = 1D grating) */
  int n_cx, n_cy;            /* number of cell edges along x, y (>= 2) */
  const double *coord_x;     /* [n_cx] sorted edges; Lx = last - first */
  const double *coord_y;     /* [n_cy] sorted edges; Ly = last - first */
  int n_layers;              /* L >= 2; layer 0 = incidence half-space, L-1 = exit */
  const double *thickness;   /* [L]; entries 0 and L-1 are not used */
} orcwa_stack;

/* One solve instance (one iteration of sol/rcwa_bridge.cpp:244-334). */
typedef struct {
  double lambda;             /* wavelength, in the length unit of the coordinates */
  double kx0, ky0;           /* Bloch offsets (RCWASolver::setBlochWavevector) */
  const double *eps;         /* [L][(n_cy-1)*(n_cx-1)] complex (re,im); cell (iy,ix)
                                at iy*(n_cx-1)+ix, row = y cell (rcwa/Layer.cpp:29-33).
                                Instances passing the same pointer share convolution
                                matrices. */
  int n_pol;                 /* 1..8 incident polarisations */
  const double *pol;         /* [n_pol][4] = Re px, Im px, Re py, Im py: the transverse
                                amplitudes of RCWASolver::generateHorizontalPlaneWave */
  const double *mu;          /* NULL (mu = 1 everywhere) or the relative permeability,
                                laid out as eps: Layer(coordX, coordY, eps, mu)
                                (rcwa/Layer.h:29-37).  Layers with mu != 1 take the
                                magnetic P and Q of rcwa/Layer.cpp:129-145, 150-164,
                                206-213; the half-spaces must stay homogeneous in eps
                                and mu. */
  /* NULL (the stack's edges) or this instance's n_cx / n_cy edges: same counts
     and periods as the stack, sorted, a repeated edge being an empty cell.  Lets
     stacks of one order count and layer count but different cell layouts share
     a batch (RCWASolver::addLayer takes each stack's own edges,
     rcwa/RCWASolver.cpp:108-137). */
  const double *coord_x;
  const double *coord_y;
  const double *thickness;   /* NULL (the stack's) or this instance's [L] thicknesses */
} orcwa_instance;

/* Output of one instance.  Orders k = (2Ny+1)(m+Nx) + (n+Ny), h = (2Nx+1)(2Ny+1)
 * (rcwa/RCWASolver.cpp:255-264). */
typedef struct {
  double *ref;               /* [n_pol][h] REF of scatterPlaneWave, or NULL */
  double *trn;               /* [n_pol][h] TRN, or NULL */
  double *amp;               /* optional [2][4][h] complex (re,im): for incident basis
                                s_x=(1,0), s_y=(0,1) the reflected (rx, ry) and
                                transmitted (tx, ty) harmonic amplitudes; or NULL */
  int status;                /* orcwa_status bits */
} orcwa_result;

/* Context: device selection, one stream, device memory pool.  The device
 * buffers of destroyed plans stay in the context's pool (its four largest) and
 * are reused by later plans, so plan_destroy never waits for the whole device
 * (hipFree does); orcwa_gpu_destroy frees them.  Calls on one context must not
 * overlap; separate contexts may be driven from separate threads.  Replaces the
 * per-solve construction of RCWASolver (rcwa/RCWASolver.h:67-80). */
int orcwa_gpu_create(orcwa_gpu_context **ctx, int device);
/* Destroy a context.  Every plan created on it must be destroyed first: with
 * live plans the call returns ORCWA_ERR_ARG and destroys nothing. */
int orcwa_gpu_destroy(orcwa_gpu_context *ctx);

/* Device-memory budget (bytes) of the plans later created on this context: a
 * plan splits its batch into chunks whose workspace fits it.  0 (the default)
 * means the ORCWA_WORKSPACE_GB environment value if set, else 70 % of the
 * device memory free at plan creation.  Lets several contexts driven from
 * separate threads share one device.  No reference counterpart (the
 * reference solves one instance at a time). */
int orcwa_gpu_set_workspace_limit(orcwa_gpu_context *ctx, size_t bytes);

/* Precision of the star-product GEMMs: the two products per step of the
 * RTCM cascade that form F + G = A + B R~ and F - G = B + A R~
 * (rcwa/RCWASolver.cpp:433-443) in plans created later on this context, and
 * the nine products of orcwa_gpu_redheffer_star and the integrators' star
 * products (core/RCWAIntegrator.cpp:57-67).  The eigensolves, the interface
 * (mode-matching) solves and the LU factorisations and solves of the cascade
 * stay FP64 in every mode.  Reduced modes multiply on FP32 or BF16 MFMA and
 * add each K step's partial sums in FP64:
 *   ORCWA_GEMM_FP32    operands rounded to FP32 (24 significant bits)
 *   ORCWA_GEMM_BF16X3  operands split into two BF16 slices (about 16 bits),
 *                      three slice products
 *   ORCWA_GEMM_BF16X6  three slices refined by the next slice's products (about
 *                      24 bits), six slice products
 * No reference counterpart (the reference multiplies in FP64). */
enum orcwa_gemm_precision {
  ORCWA_GEMM_FP64 = 0,
  ORCWA_GEMM_FP32 = 1,
  ORCWA_GEMM_BF16X3 = 2,
  ORCWA_GEMM_BF16X6 = 3
};
int orcwa_gpu_set_gemm_precision(orcwa_gpu_context *ctx, int precision);

/* One-shot batched solve: plan + execute + results + destroy.
 * Replaces the per-wavelength loop of sol/rcwa_bridge.cpp:244-334 and
 * rcwa/RCWADriver.cpp:282-368 (RCWASolver::solve + scatterPlaneWave). */
int orcwa_gpu_solve_batch(orcwa_gpu_context *ctx, const orcwa_stack *stack,
                          const orcwa_instance *inst, int count, orcwa_result *out);

/* Split form for callers that reuse one batch layout (benchmarks, sweeps):
 * plan_create copies all inputs to device memory and allocates; execute runs
 * the whole batch on the context stream (asynchronous); sync waits;
 * results copies the outputs back.  The plan keeps pointers to nothing the
 * caller owns. */
int orcwa_plan_create(orcwa_gpu_context *ctx, const orcwa_stack *stack,
                      const orcwa_instance *inst, int count, orcwa_plan **plan);
int orcwa_plan_execute(orcwa_plan *plan);
int orcwa_plan_sync(orcwa_plan *plan);
int orcwa_plan_results(orcwa_plan *plan, orcwa_result *out);
int orcwa_plan_destroy(orcwa_plan *plan);

/* A HIP graph of the executes of several plans of one context, captured once
 * and replayed: orcwa_graph_launch runs every plan's whole batch as one graph
 * launch on the context stream (plan i on fork i mod n_streams of the graph,
 * so plans of different order counts overlap on the device; each plan's
 * chunks stay in order), orcwa_graph_sync waits, and orcwa_plan_results reads
 * each plan's outputs as after an execute.  The captured form clears its
 * status words by kernel and does no host synchronisation, so a failed eigen
 * residual test is not recovered in the graph: orcwa_graph_sync then runs that
 * plan again with orcwa_plan_execute (recovery included) before it returns.
 * Plans used by a graph cannot be destroyed before the graph; a graph counts
 * as a live plan of its context.  No reference counterpart: it replaces the
 * per-wavelength loop of rcwa/RCWADriver.cpp:282 and sol/rcwa_bridge.cpp:244
 * for batches of several stack layouts. */
typedef struct orcwa_graph orcwa_graph;
int orcwa_graph_create(orcwa_gpu_context *ctx, orcwa_plan *const *plans, int nplans,
                       int n_streams, orcwa_graph **graph);
int orcwa_graph_launch(orcwa_graph *graph);
int orcwa_graph_sync(orcwa_graph *graph);
/* plans re-run eagerly by the last orcwa_graph_sync */
int orcwa_graph_reruns(const orcwa_graph *graph);
int orcwa_graph_destroy(orcwa_graph *graph);

/* Autotuning (no reference counterpart).  The multishift QR of the eigensolver
 * has three compiled variants (window / shifts per sweep / workgroups per CU:
 * 0 = 48 / 12 / 2, 1 = 32 / 8 / 3, 2 = 64 / 16 / 1) and the GEMMs three tiles
 * for their main block (0 = 64 x 64, 1 = 32 x 64, 2 = 64 x 32).
 * orcwa_plan_autotune executes the plan with each QR variant, then with each
 * tile under the fastest QR variant, keeps the fastest combination for the
 * plan and for later plans of the same layer-matrix size n on the context, and
 * appends "<arch> <n> <qr> <tile> <ms>" to the context's tuning file when it
 * is writable; the plan's results are then those of the chosen variant.
 * orcwa_gpu_set_tuning_file reads such a file (lines of other architectures
 * are skipped; later lines win) and, with append != 0, records later tuning
 * results in it.  Plans of sizes not in the table use variant 0 and tile 0. */
int orcwa_gpu_set_tuning_file(orcwa_gpu_context *ctx, const char *path, int append);
int orcwa_plan_autotune(orcwa_plan *plan, int *qr_variant, int *gemm_tile);

/* Elapsed GPU time (ms) of named phases of the last execute, measured with HIP
 * events on the plan's stream.  phase: 0 total, 1 convolution matrices,
 * 2 P/Q + eigensolve + V, 3 interfaces + cascade, 4 plane-wave R/T,
 * 5 eigensolver only.  Returns a negative value if unknown (also after a graph
 * launch, which records no events). */
double orcwa_plan_phase_ms(orcwa_plan *plan, int phase);

/* Slab scattering matrices in a reference basis, batched: for each instance b,
 * the S-matrix of a slab of thickness wz[b] with modes (W, V, gamma) embedded
 * in the basis (W0, V0): F0 = W^-1 W0, G0 = V^-1 V0, A = F0 + G0,
 * B = F0 - G0, X = exp(i gamma wz lambda / 2 pi),
 * Rud = (I - (A^-1 X B)^2)^-1 (A^-1 X B A^-1 X A - A^-1 B),
 * Tuu = (I - (A^-1 X B)^2)^-1 (A^-1 X A - A^-1 X B A^-1 B); the slab is
 * symmetric (Rdu = Rud, Tdd = Tuu).  Replaces
 * dtmm::RCWAScatterMatrix::compute_v2 (rcwa/RCWAScatterMatrix.cpp:88-147).
 * Matrices n x n column-major complex (re, im), batch-major; gamma [batch][n]
 * complex; info [batch] (optional) gets ORCWA_STATUS_SINGULAR bits. */
int orcwa_gpu_slab_smatrix(orcwa_gpu_context *ctx, int n, int batch, const double *lambda,
                           const double *wz, const double *W, const double *V,
                           const double *gamma, const double *W0, const double *V0,
                           double *Rud, double *Tuu, int *info);

/* Redheffer star product of S-matrices (rud, tuu, rdu, tdd), batched: `top`
 * above `bottom`, each [batch][4][n][n] complex in that order; out = top * bottom:
 *   C1 = (I - rdu_t Rud_b)^-1,  C2 = I + Rud_b C1 rdu_t,
 *   Rdu = Rdu_b + Tdd_b rdu_t C2 Tuu_b,  Tdd = Tdd_b C1 tdd_t,
 *   Rud = rud_t + tuu_t Rud_b C1 tdd_t,  Tuu = tuu_t C2 Tuu_b.
 * The slab-on-top update of RCWAIntegrator::compute
 * (core/RCWAIntegrator.cpp:57-67) is top = slab, bottom = accumulated stack. */
int orcwa_gpu_redheffer_star(orcwa_gpu_context *ctx, int n, int batch, const double *top,
                             const double *bottom, double *out, int *info);

/* A z-varying structure between z0 (bottom) and z1 (top), for the
 * midpoint-staircase integrator RCWAIntegrator::compute(N)
 * (core/RCWAIntegrator.cpp:20-86) with a sampler of cell edges and
 * permittivity per z (VarLayerSampler, core/VarLayerSampler.cpp:52-68; the
 * periodic K of core/LayerSampler.cpp:114-144).  The caller evaluates its
 * sampler at the N + 2 points the integrator takes, in this order: z1 (the top
 * basis), the slab midpoints z0 + (i + 1/2) dz for i = 0..N-1 (dz = (z1 - z0)
 * / N), then z0 (the bottom basis). */
typedef struct {
  int nx_harm, ny_harm;      /* orders -Nx..Nx, -Ny..Ny */
  int n_cx, n_cy;            /* edges per sample; every sample has the same periods */
  int n_slices;              /* N >= 1 */
  double lambda;
  double kx0, ky0;           /* Bloch offsets of K; the reference's sampler uses 0, 0 */
  double z0, z1;             /* z0 < z1 */
  const double *coord_x;     /* [N + 2][n_cx] sorted edges of each sample */
  const double *coord_y;     /* [N + 2][n_cy] */
  const double *eps;         /* [N + 2][(n_cy-1)*(n_cx-1)] complex, cell (iy, ix) at
                                iy*(n_cx-1)+ix as orcwa_instance.eps */
} orcwa_zstack;

/* The integrator's S-matrix (Rud, Tuu, Rdu, Tdd: the accessors of
 * core/RCWAIntegrator.h:29-32), its top side in the mode basis at z1 and its
 * bottom side in the basis at z0, and those bases (W0/V0/gam0 at z0, W1/V1/gam1
 * at z1, gam = keff of WaveEquationSolver).  Matrices [n][n] complex
 * column-major, n = 2 (2Nx+1)(2Ny+1); vectors [n] complex; NULL skips an output.
 * The mode bases are this engine's (homogeneous samples take W = I), so the
 * matrices match the reference's up to the choice of basis: W0 Tdd W1^-1,
 * W1 Rud W1^-1, W1 Tuu W0^-1 and W0 Rdu W0^-1 are basis-free. */
typedef struct {
  double *Rud, *Tuu, *Rdu, *Tdd;
  double *W0, *V0, *W1, *V1;
  double *gam0, *gam1;
  int status;                /* orcwa_status bits */
} orcwa_zresult;

int orcwa_gpu_rcwa_integrate(orcwa_gpu_context *ctx, const orcwa_zstack *zs,
                             orcwa_zresult *res);

/* A sampler of the cross-section at z (LayerSampler::sample,
 * core/LayerSampler.h:52-54, returning the geometry instead of P and Q):
 * writes n_cx and n_cy sorted edges (same periods at every z) and the
 * (n_cy-1)*(n_cx-1) complex permittivities; returns 0 on success. */
typedef int (*orcwa_sampler_fn)(void *user, double z, double *coord_x, double *coord_y,
                                double *eps);

typedef struct {
  int nx_harm, ny_harm;
  int n_cx, n_cy;
  double lambda;
  double kx0, ky0;           /* Bloch offsets of K (0, 0 in the reference's sampler) */
  double z0, z1;
  orcwa_sampler_fn sample;
  void *user;
} orcwa_zsampler;

/* The adaptive first-order perturbation integrator RedhefferIntegrator::
 * compute(p, maxN, pe) (core/RedhefferIntegrator.cpp:173-399, its active
 * triple-division form): segments whose first-order terms exceed pe are cut in
 * three.  Every segment is evaluated in the modes and P, Q of its midpoint
 * (homogeneous cross-sections included, as the reference does), so the sampler
 * is called at z0, z1, the midpoint and at the cuts.  max_segments > 0 stops
 * with ORCWA_ERR_UNSUPPORTED beyond that many accepted segments.  Results as
 * orcwa_gpu_rcwa_integrate; n_segments (optional) gets the reference's N(). */
int orcwa_gpu_redheffer_integrate(orcwa_gpu_context *ctx, const orcwa_zsampler *zs, double pe,
                                  int max_segments, orcwa_zresult *res, int *n_segments);

/* Field images inside the stack: the inputCoeffs form of
 * RCWASolver::saveFieldImage (rcwa/RCWASolver.cpp:707-1011, mode amplitudes of
 * evaluateIntermediateField :1505-1645) that the .orcwa driver calls for its
 * first wavelength (rcwa/RCWADriver.cpp:342-352).  Components and slices keep
 * the reference's numbering (rcwa/defns.h:27-42). */
enum orcwa_field_component {
  ORCWA_FIELD_EX = 0, ORCWA_FIELD_EY = 1, ORCWA_FIELD_EZ = 2,
  ORCWA_FIELD_HX = 3, ORCWA_FIELD_HY = 4, ORCWA_FIELD_HZ = 5
};
enum orcwa_slice { ORCWA_SLICE_XY = 0, ORCWA_SLICE_XZ = 1, ORCWA_SLICE_YZ = 2 };

typedef struct {
  int component;             /* orcwa_field_component */
  int slice;                 /* orcwa_slice */
  double coord;              /* the fixed coordinate: y of an XZ slice, x of a YZ slice,
                                z of an XY slice (measured down from the top of layer 1) */
  int nres1;                 /* samples along the image columns; 0 = 500 as the reference */
  int pol;                   /* which of the instance's polarisations excites the stack */
} orcwa_field_request;

/* Solve one instance and return its complex field on the slice: rows x cols
 * values (re, im), row-major, rows along z (XZ, YZ) or y (XY), samples at cell
 * centres start + step (i + 0.5) exactly as the reference.  Layer 0 and L-1
 * are the half-spaces: XZ/YZ rows span the interior layers 1..L-2.  Ex, Ey,
 * Hx, Hy are the modal fields ([Ex; Ey] = W (u + d), [Hx; Hy] = V (u - d));
 * Ez = i [[eps]]^-1 (Kx Hy - Ky Hx) and Hz = (i / k0^2)(Kx Ey - Ky Ex) as the
 * reference writes them (:849-891).  field == NULL returns only rows and cols.
 * status (optional) gets the instance's orcwa_status bits.  At most 1024
 * harmonics. */
int orcwa_gpu_field_image(orcwa_gpu_context *ctx, const orcwa_stack *stack,
                          const orcwa_instance *inst, const orcwa_field_request *req,
                          int *rows, int *cols, double *field, int *status);

/* Layout-compatible records of the .ofd RCWA input globals
 * (include/orcwa_rcwa.h:27-35 and :45-52 of the reference): same field order
 * and types, so rcwa_run reads the arrays the reference's input module fills. */
typedef struct {
  double eps1, eps2;         /* real parts inside / outside the eps1 rectangle */
  double eps1i, eps2i;       /* imaginary parts (loss > 0, exp(-i w t)) */
  double fill, fill_y;       /* eps1 rectangle [0, fill Lx] x [0, fill_y Ly] */
  double thickness;          /* [m]; ignored for the two half-spaces */
} orcwa_rcwalayer;

typedef struct {
  int layer;                 /* 0-based layer */
  int side;                  /* 1 = eps1, 2 = eps2 */
  double einf, ae, be, ce;   /* eps = einf + sum ae^2 / (ce^2 - w^2 - i be w) */
} orcwa_rcwapole;

/* The reference's .ofd entry point (sol/rcwa_bridge.cpp:145; declared at
 * include/orcwa_rcwa.h:72-74).  Reads the RCWA globals of
 * include/orcwa_rcwa.h:54-70 (defined by the caller's input module,
 * sol/input_data.c:13-24) and NFreq1/Freq1/Title (sol/rcwa_bridge.cpp:35-40),
 * writes rcwa_efficiency.csv and rcwa_orders.csv in the working directory and
 * log lines through monitor1().  Returns 0 on success, 1 on error. */
int rcwa_run(FILE *fp_log);

#ifdef __cplusplus
}
#endif

#endif /* OPENRCWA_GPU_H */
