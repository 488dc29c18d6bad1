// CPU harness of rcwa_run (openrcwa_amd/csrc/ofd_bridge.cpp): defines the
// .ofd input globals, a stand-in for the batched engine's C ABI that records
// the batch it is given and returns made-up per-order efficiencies whose
// values encode (frequency, polarisation, order), and, unless built with
// -DNO_H5, stand-ins for the host program's HDF5 writers that record their
// arguments.  Everything recorded goes to harness_*.txt in the working
// directory; tests/test_ofd_bridge.py builds this with g++ and checks it.
//
//   harness <scenario>   0: 1D grating, 3 frequencies, one status bit set
//                        1: 2D grating with a Lorentz layer, oblique
//                        2: an invalid (zero) frequency in the list
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/openrcwa_gpu.h"

extern "C" {
int NRcwaPole = 0;
orcwa_rcwapole* RcwaPole = nullptr;
int NRcwaHarmonics = 0;
double RcwaPeriod = 0.0;
int NRcwaHarmonicsY = 0;
double RcwaPeriodY = 0.0;
int NRcwaLayer = 0;
orcwa_rcwalayer* RcwaLayer = nullptr;
double RcwaTheta = 0.0;
double RcwaPhi = 0.0;
int IRcwaIncidence = 0;
int NFreq1 = 0;
double* Freq1 = nullptr;
char Title[256] = "harness title";

void monitor1(FILE* fp, const char msg[]) {
  fprintf(fp, "%s\n", msg);
  fflush(fp);
}

struct orcwa_gpu_context {
  int dummy;
};

int orcwa_gpu_create(orcwa_gpu_context** ctx, int) {
  static orcwa_gpu_context c;
  *ctx = &c;
  return ORCWA_OK;
}
int orcwa_gpu_destroy(orcwa_gpu_context*) { return ORCWA_OK; }

// per-order efficiency of the stand-in engine
static double fake(int f, int p, int k, int trn) {
  return 1e-4 * (1 + k) + 1e-2 * p + 1e-1 * f + 5e-5 * trn;
}

int orcwa_gpu_solve_batch(orcwa_gpu_context*, const orcwa_stack* st, const orcwa_instance* inst,
                          int count, orcwa_result* out) {
  FILE* fp = fopen("harness_batch.txt", "w");
  const int h = (2 * st->nx_harm + 1) * (2 * st->ny_harm + 1);
  const int ncell = (st->n_cx - 1) * (st->n_cy - 1);
  fprintf(fp, "stack %d %d %d %d %d\n", st->nx_harm, st->ny_harm, st->n_cx, st->n_cy,
          st->n_layers);
  for (int i = 0; i < st->n_cx; ++i) fprintf(fp, "cx %.17g\n", st->coord_x[i]);
  for (int i = 0; i < st->n_cy; ++i) fprintf(fp, "cy %.17g\n", st->coord_y[i]);
  for (int l = 0; l < st->n_layers; ++l) fprintf(fp, "thick %.17g\n", st->thickness[l]);
  for (int f = 0; f < count; ++f) {
    const orcwa_instance& in = inst[f];
    fprintf(fp, "inst %d %.17g %.17g %.17g %d\n", f, in.lambda, in.kx0, in.ky0, in.n_pol);
    for (int q = 0; q < 4 * in.n_pol; ++q) fprintf(fp, "pol %.17g\n", in.pol[q]);
    for (int e = 0; e < st->n_layers * ncell; ++e)
      fprintf(fp, "eps %.17g %.17g\n", in.eps[2 * e], in.eps[2 * e + 1]);
    for (int p = 0; p < in.n_pol; ++p)
      for (int k = 0; k < h; ++k) {
        out[f].ref[p * h + k] = fake(f, p, k, 0);
        out[f].trn[p * h + k] = fake(f, p, k, 1);
      }
    out[f].status = (f == 1) ? (ORCWA_STATUS_ILLCOND | ORCWA_STATUS_EIG_PERTURBED) : 0;
  }
  fclose(fp);
  return ORCWA_OK;
}

#ifndef NO_H5
typedef struct {
  double frequency, lambda, R, T, A;
} spectrum_row;
typedef struct {
  const char* title;
  int nharmonics;
  double period;
  int nharmonics_y;
  double period_y;
  int nlayer;
  double theta, phi;
} h5meta;

int rcwa_write_hdf5(const char* filename, const h5meta* meta, const spectrum_row* rows, int npol,
                    int nfreq, const char* const* labels) {
  FILE* fp = fopen("harness_h5.txt", "w");
  fprintf(fp, "file %s\n", filename);
  fprintf(fp, "meta %s|%d|%.17g|%d|%.17g|%d|%.17g|%.17g\n", meta->title, meta->nharmonics,
          meta->period, meta->nharmonics_y, meta->period_y, meta->nlayer, meta->theta,
          meta->phi);
  fprintf(fp, "dims %d %d\n", npol, nfreq);
  for (int p = 0; p < npol; ++p) fprintf(fp, "label %s\n", labels[p]);
  for (int i = 0; i < npol * nfreq; ++i)
    fprintf(fp, "row %.17g %.17g %.17g %.17g %.17g\n", rows[i].frequency, rows[i].lambda,
            rows[i].R, rows[i].T, rows[i].A);
  fclose(fp);
  return 0;
}

int rcwa_write_hdf5_orders(const char* filename, int npol, int nfreq, int nx, int ny,
                           const double* R, const double* T) {
  FILE* fp = fopen("harness_h5_orders.txt", "w");
  fprintf(fp, "file %s\ndims %d %d %d %d\n", filename, npol, nfreq, nx, ny);
  for (long i = 0; i < (long)npol * nfreq * nx * ny; ++i) fprintf(fp, "%.17g %.17g\n", R[i], T[i]);
  fclose(fp);
  return 0;
}
#endif

int rcwa_run(FILE* fp_log);
}

int main(int argc, char** argv) {
  const int scenario = argc > 1 ? atoi(argv[1]) : 0;
  static orcwa_rcwalayer layers[4];
  static orcwa_rcwapole poles[3];
  static double freqs[4];
  std::memset(layers, 0, sizeof(layers));
  if (scenario == 0 || scenario == 2) {
    NRcwaHarmonics = 3;
    RcwaPeriod = 3.0e-7;
    NRcwaHarmonicsY = 0;
    RcwaPeriodY = 3.0e-7;
    NRcwaLayer = 3;
    layers[0] = {1.0, 1.0, 0.0, 0.0, 0.5, 1.0, 0.0};
    layers[1] = {4.0, 1.0, 0.0, 0.0, 0.5, 1.0, 2.0e-7};
    layers[2] = {2.25, 2.25, 0.0, 0.0, 0.5, 1.0, 0.0};
    RcwaTheta = 0.0;
    RcwaPhi = 0.0;
    NFreq1 = 3;
    freqs[0] = 4.3e14;
    freqs[1] = 6.5e14;
    freqs[2] = scenario == 2 ? 0.0 : 1.2e15;
  } else {
    NRcwaHarmonics = 2;
    RcwaPeriod = 4.0e-7;
    NRcwaHarmonicsY = 1;
    RcwaPeriodY = 5.0e-7;
    NRcwaLayer = 4;
    layers[0] = {1.0, 1.0, 0.0, 0.0, 0.5, 0.5, 0.0};
    layers[1] = {6.0, 1.0, 0.05, 0.0, 0.3, 0.6, 1.5e-7};
    layers[2] = {2.0, 1.0, 0.0, 0.0, 0.7, 0.25, 0.8e-7};
    layers[3] = {2.25, 2.25, 0.0, 0.0, 0.5, 0.5, 0.0};
    // layer 2 side 1 dispersive: two poles (einf from the last row), a
    // zero-strength row adds nothing
    NRcwaPole = 3;
    poles[0] = {2, 1, 1.5, 1.2e15, 0.1e15, 3.0e15};
    poles[1] = {2, 1, 2.0, 0.0, 0.0, 0.0};
    poles[2] = {2, 1, 2.5, 0.8e15, 0.05e15, 5.0e15};
    RcwaPole = poles;
    RcwaTheta = 25.0;
    RcwaPhi = 40.0;
    NFreq1 = 2;
    freqs[0] = 5.0e14;
    freqs[1] = 7.0e14;
  }
  RcwaLayer = layers;
  Freq1 = freqs;
  FILE* log = fopen("harness_log.txt", "w");
  const int rc = rcwa_run(log);
  fclose(log);
  printf("rc %d\n", rc);
  return 0;
}
