// Parameter records and launchers of the RCWA-specific kernels.
#pragma once

#include "engine.h"

namespace orcwa {

// geometry shared by a batch (device pointers for the edges)
struct ConvGeom {
  int nx, ny;          // harmonics per axis, 2N+1
  int ncx, ncy;        // number of edges
  int ncell;           // (ncx-1)*(ncy-1)
  double Lx, Ly;       // periods
  const double* cx;    // device [ncx], or [grids][ncx] with sx = ncx
  const double* cy;    // device [ncy], or [grids][ncy] with sy = ncy
  int sx, sy;          // edge stride per grid: 0 when every grid shares the edges
};

// one strip of one grid for the Li-rule series
struct StripTask {
  int grid;            // unique permittivity grid
  int type;            // 0: y-strip (x series), 1: x-strip (y series)
  int index;           // strip index
  int pad;
  long long out_off;   // element offset of the output record
};

// per-grid convolution products (device pointers, h x h column-major)
struct GridMats {
  const cd* E11;
  const cd* E22;
  const cd* S1;        // M E11        (M = fe33^-1)
  const cd* S2;        // M Dm E11
  const cd* S3;        // M E22
  const cd* S4;        // M Dn E22
};

// one patterned-layer problem (instance, layer)
struct ProbParam {
  double kx0, ky0;     // Bloch offsets (1/length)
  double inv_k0;       // 1/k0 (length)
  int grid;            // grid index (permittivity)
  int mgrid;           // grid index of the permeability (magnetic problems), else -1
};

// convolution matrices of a magnetic problem: fe11, fe22, fe33^-1 of the
// permittivity grid and fmu11, fmu22, fmu33^-1 of the permeability grid
struct MagMats {
  const cd* E11;
  const cd* E22;
  const cd* Ei;
  const cd* M11;
  const cd* M22;
  const cd* Mi;
};

// one homogeneous-layer problem
struct UniParam {
  cd eps;
  cd mu;
  double kx0, ky0, inv_k0;
  int inst;
  int pad;
};

// plane-wave response of one instance
struct RtParam {
  cd eps_ref, eps_trn;
  cd mu_ref, mu_trn;
  double kx0, ky0, inv_k0;
  int npol;
  int pad;
  cd pol[8][2];        // (px, py)
};

// field reconstruction: one layer of the instance whose fields are imaged
struct FieldLayer {
  const cd* M;         // W (E-type components) or V (H-type), n x n column-major
  const cd* keff;      // [n] modal keff in k0 units
  const cd* m33inv;    // [[eps]]^-1 (h x h) of a patterned layer, else null
  cd eps;              // permittivity of a homogeneous layer
};

// one image row: its layer and depth below the layer top (z_p) and above the
// layer bottom (z_m <= 0), both times k0
struct FieldRow {
  int layer;
  int pad;
  double k0zp, k0zm;
};

struct FieldGeom {
  int nx, ny;
  int component;       // 0..5 = Ex, Ey, Ez, Hx, Hy, Hz (rcwa/defns.h:27-35)
  int slice;           // 0 XY, 1 XZ, 2 YZ (rcwa/defns.h:37-42)
  double Lx, Ly, kx0, ky0, k0;
  double start1, step1;  // column samples start1 + step1 (c + 0.5)
  double start2, step2;  // row samples of an XY slice
  double coord;          // the fixed coordinate of an XZ (y) or YZ (x) slice
};

// host launchers (all asynchronous on stream s)
void launch_conv_strip(hipStream_t s, const ConvGeom& g, const StripTask* tasks, int ntasks,
                       const cd* grids, cd* strips, size_t lds_bytes);
void launch_conv_expand(hipStream_t s, const ConvGeom& g, const cd* strips,
                        const long long* strip_base, cd* const* E11, cd* const* E22,
                        cd* const* E33, int ngrid);
void launch_scale_rows_order(hipStream_t s, const ConvGeom& g, int axis, const cd* const* in,
                             cd* const* out, int ngrid);
void launch_build_pq(hipStream_t s, const ConvGeom& g, const ProbParam* params, const GridMats* gm,
                     cd* const* PQ, cd* const* PQcopy, int nprob);
void launch_build_v(hipStream_t s, const ConvGeom& g, const ProbParam* params, const cd* const* W,
                    const cd* const* QW1, const cd* lam, cd* keff, cd* const* V, int* info,
                    const int* inst_of, int nprob);
void launch_build_pq_mag(hipStream_t s, const ConvGeom& g, const ProbParam* params,
                         const MagMats* mm, cd* const* P, cd* const* Q, int nprob);
void launch_build_v_mag(hipStream_t s, int n, const cd* const* QW, const cd* lam, cd* keff,
                        cd* const* V, int* info, const int* inst_of, int nprob);
void launch_zero_mats(hipStream_t s, long long count, cd* const* A, int batch);
void launch_uniform(hipStream_t s, const ConvGeom& g, const UniParam* params, cd* keff,
                    cd* const* V, cd* binv, int* info, int nuni);
void launch_bdinv_apply(hipStream_t s, int h, const cd* const* binv, const cd* const* Vb,
                        cd* const* H, int batch);
void launch_phase_scale(hipStream_t s, int n, const cd* const* keff, const double* k0d,
                        cd* const* R, cd* const* T, int* info, int batch);
void launch_rt(hipStream_t s, const ConvGeom& g, const RtParam* params, const cd* rt,
               double* ref, double* trn, cd* amp, int* info, int max_pol, int batch);
// E <- A = E + D H, H <- B = E - D H (D = diag(1/keff_a) where keff_a is given),
// At <- A^T and, when Bt is not null, Bt <- B^T
void launch_sum_diff_t(hipStream_t s, int n, cd* const* E, cd* const* H,
                       const cd* const* keff_a, cd* const* At, cd* const* Bt, int batch);
void launch_pv_in(hipStream_t s, const ConvGeom& g, const ProbParam* params, const cd* const* Vb,
                  cd* const* out, int batch);
void launch_pv_out(hipStream_t s, const ConvGeom& g, const ProbParam* params, const cd* const* U,
                   long long offU, const cd* const* Vb, cd* const* H, int batch);
void launch_cascade_tail(hipStream_t s, int n, int i0, const cd* const* LUt, const int* piv,
                         const cd* const* A, const cd* const* B, const cd* const* Rt, cd* rt,
                         int batch);
void launch_cascade_chain(hipStream_t s, int n, const cd* const* LUt, const int* piv,
                          const cd* const* keff, const double* k0d, cd* rt, int batch);
void launch_field_combine(hipStream_t s, int n, int L, const cd* tv, cd px, cd py, cd* cp);
void launch_field_up(hipStream_t s, int n, const cd* A, const cd* B, const cd* cp1, const cd* cm1,
                     const cd* keff1, double k0d1, cd* cm);
void launch_field_modes(hipStream_t s, int n, const FieldLayer* layers, const FieldRow* rows,
                        int nrows, const cd* cp, const cd* cm, int sign, cd* X);
void launch_field_component(hipStream_t s, const FieldGeom& g, const FieldLayer* layers,
                            const FieldRow* rows, int nrows, const cd* X, cd* Hm);
void launch_field_synth(hipStream_t s, const FieldGeom& g, const cd* Hm, int nrows, int ncols,
                        cd* img);

}  // namespace orcwa
