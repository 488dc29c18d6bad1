// rcwa_run(FILE*): the .ofd RCWA entry point (sol/rcwa_bridge.cpp:145-406 of
// the reference) on the batched engine.  Host code only: it reads the input
// globals once, turns the frequency sweep into one batch of instances,
// solves it through the batched C ABI (orcwa_gpu_solve_batch), and writes the
// reference's outputs from the results.
//
// Kept identical on purpose (they are the caller-visible contract): the
// global names and record layouts (include/orcwa_rcwa.h:27-70,
// sol/rcwa_bridge.cpp:35-40), the CSV file names, headers and row formats
// (:211-230, :322-334), the log lines (:154-178, :336-357, :372-402), the
// edge-merge tolerance (:52-67), the eps1 rectangle rule (:72-89), the
// Lorentz pole form (:105-123), the TE/TM basis (:187-196), the propagating
// order test (:285-300) and the HDF5 writer calls (:375-402).
//
// Differences: every frequency is one instance of a single GPU batch, so the
// rows are written after the whole sweep is solved; a frequency whose solve
// reports status bits gets a "*** warning" log line naming them (the
// reference logs its eigen-solver recovery to std::clog,
// rcwa/MKLEigenSolver.hpp:160-176).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/openrcwa_gpu.h"

extern "C" {
// Input globals.  The caller's input module defines them
// (sol/input_data.c:13-24, sol/rcwa_bridge.cpp:35-40); these weak
// definitions only let the library load on its own.
__attribute__((weak)) int NRcwaPole = 0;
__attribute__((weak)) orcwa_rcwapole* RcwaPole = nullptr;
__attribute__((weak)) int NRcwaHarmonics = 0;
__attribute__((weak)) double RcwaPeriod = 0.0;
__attribute__((weak)) int NRcwaHarmonicsY = 0;
__attribute__((weak)) double RcwaPeriodY = 0.0;
__attribute__((weak)) int NRcwaLayer = 0;
__attribute__((weak)) orcwa_rcwalayer* RcwaLayer = nullptr;
__attribute__((weak)) double RcwaTheta = 0.0;
__attribute__((weak)) double RcwaPhi = 0.0;
__attribute__((weak)) int IRcwaIncidence = 0;
__attribute__((weak)) int NFreq1 = 0;
__attribute__((weak)) double* Freq1 = nullptr;
__attribute__((weak)) char Title[256] = "";

// one log line (sol/monitor.c:12-16); the host program's own wins
__attribute__((weak)) void monitor1(FILE* fp, const char msg[]) {
  if (!fp) return;
  fprintf(fp, "%s\n", msg);
  fflush(fp);
}

// The host program's HDF5 writers (sol/rcwa_hdf5.c:91-241, declared in
// include/rcwa_hdf5.h); unresolved (null) when it links none.
typedef struct {
  double frequency, lambda, R, T, A;
} orcwa_spectrum_row;
typedef struct {
  const char* title;
  int nharmonics;
  double period;
  int nharmonics_y;
  double period_y;
  int nlayer;
  double theta, phi;
} orcwa_h5meta;
int rcwa_write_hdf5(const char* filename, const orcwa_h5meta* meta, const orcwa_spectrum_row* rows,
                    int npol, int nfreq, const char* const* pol_labels) __attribute__((weak));
int rcwa_write_hdf5_orders(const char* filename, int npol, int nfreq, int nx, int ny,
                           const double* R, const double* T) __attribute__((weak));
}

namespace {

using cplx = std::complex<double>;
constexpr double kC0 = 2.99792458e8;        // m/s
constexpr double kPi = 3.14159265358979323846;

// The permittivity of one side (1 = inside the eps1 rectangle, 2 = outside)
// of one layer: a constant, or einf plus Lorentz poles when the pole list has
// a non-zero-strength pole for it.  einf comes from the last pole row of the
// (layer, side); rows of zero strength add nothing.
struct SideMaterial {
  cplx fixed;
  bool lorentz = false;
  double einf = 0.0;
  std::vector<double> ae, be, ce;

  cplx at(double omega) const {
    if (!lorentz) return fixed;
    cplx e(einf, 0.0);
    for (size_t k = 0; k < ae.size(); ++k)
      e += ae[k] * ae[k] / cplx(ce[k] * ce[k] - omega * omega, -be[k] * omega);
    return e;
  }
  bool lossless() const {
    if (!lorentz) return fixed.imag() == 0.0;
    for (double b : be)
      if (b != 0.0) return false;
    return true;
  }
};

// The run as read from the globals.
struct Run {
  int nx_harm = 0, ny_harm = 0;
  double lx = 0.0, ly = 0.0;
  int nlayer = 0;
  std::vector<SideMaterial> side;            // [layer][2]
  std::vector<double> fill, fill_y, thick;   // per layer
  std::vector<double> freq;                  // the valid frequencies
  bool bad_freq = false;                     // a non-positive frequency ended the list
  int n_lorentz = 0;                         // (layer, side) pairs with poles
  bool all_lossless = true;
  double px_te = 0, py_te = 0, px_tm = 0, py_tm = 0;
  double sin_t = 0, cos_p = 1, sin_p = 0;
};

Run read_globals() {
  Run r;
  r.nx_harm = NRcwaHarmonics;
  r.ny_harm = NRcwaHarmonicsY;
  r.lx = RcwaPeriod;
  r.ly = RcwaPeriodY;
  r.nlayer = NRcwaLayer;
  r.side.resize(2 * (size_t)r.nlayer);
  for (int l = 0; l < r.nlayer; ++l) {
    const orcwa_rcwalayer& g = RcwaLayer[l];
    r.side[2 * l].fixed = cplx(g.eps1, g.eps1i);
    r.side[2 * l + 1].fixed = cplx(g.eps2, g.eps2i);
    r.fill.push_back(g.fill);
    r.fill_y.push_back(g.fill_y);
    r.thick.push_back((l == 0 || l == r.nlayer - 1) ? 0.0 : g.thickness);
  }
  for (int k = 0; k < NRcwaPole; ++k) {
    const orcwa_rcwapole& q = RcwaPole[k];
    if (q.layer < 0 || q.layer >= r.nlayer || (q.side != 1 && q.side != 2)) continue;
    SideMaterial& m = r.side[2 * q.layer + (q.side - 1)];
    m.einf = q.einf;
    if (q.ae != 0.0) {
      m.lorentz = true;
      m.ae.push_back(q.ae);
      m.be.push_back(q.be);
      m.ce.push_back(q.ce);
    }
  }
  for (const SideMaterial& m : r.side) {
    r.n_lorentz += m.lorentz ? 1 : 0;
    r.all_lossless = r.all_lossless && m.lossless();
  }
  for (int f = 0; f < NFreq1; ++f) {
    if (!(Freq1[f] > 0.0)) {
      r.bad_freq = true;
      break;
    }
    r.freq.push_back(Freq1[f]);
  }
  // incident polarisations: unit transverse amplitudes of TE (s) and TM (p);
  // at normal incidence the plane of incidence is undefined and x, y are used
  const double t = RcwaTheta * kPi / 180.0, p = RcwaPhi * kPi / 180.0;
  r.sin_t = std::sin(t);
  r.cos_p = std::cos(p);
  r.sin_p = std::sin(p);
  if (std::fabs(r.sin_t) < 1e-9) {
    r.px_te = 0.0, r.py_te = 1.0;
    r.px_tm = 1.0, r.py_tm = 0.0;
  } else {
    r.px_te = -r.sin_p, r.py_te = r.cos_p;
    r.px_tm = std::cos(t) * r.cos_p, r.py_tm = std::cos(t) * r.sin_p;
  }
  return r;
}

// Cell edges shared by every layer along one axis: 0, each layer's rectangle
// edge fill * period (fill < 1), and the period, sorted, with edges closer
// than 1e-12 of the period merged.
std::vector<double> merged_edges(double period, const std::vector<double>& fills) {
  std::vector<double> cand{0.0, period};
  for (double f : fills)
    if (f < 1.0) cand.push_back(f * period);
  std::sort(cand.begin(), cand.end());
  std::vector<double> out;
  for (double v : cand)
    if (out.empty() || v - out.back() > 1e-12 * period) out.push_back(v);
  return out;
}

// Bit names of orcwa_status for the warning lines.
std::string status_text(int st) {
  static const struct {
    int bit;
    const char* name;
  } names[] = {{ORCWA_STATUS_EIG_RESIDUAL, "eigen residual > tolerance"},
               {ORCWA_STATUS_GRAZING, "grazing order (R/T not computed)"},
               {ORCWA_STATUS_SINGULAR, "singular interface"},
               {ORCWA_STATUS_NOCONV, "eigen iteration limit"},
               {ORCWA_STATUS_EIG_PERTURBED, "perturbed eigen recovery"},
               {ORCWA_STATUS_ILLCOND, "growing mode, ill-conditioned cascade"}};
  std::string s;
  for (const auto& n : names)
    if (st & n.bit) s += (s.empty() ? "" : ", ") + std::string(n.name);
  return s;
}

}  // namespace

extern "C" int rcwa_run(FILE* fp_log) {
  char line[BUFSIZ];
  if (NRcwaLayer < 2 || RcwaPeriod <= 0 || NFreq1 < 1) {
    printf("%s\n", "*** invalid rcwa setup");
    return 1;
  }
  const Run run = read_globals();
  if (run.ny_harm > 0)
    snprintf(line, sizeof(line),
             "RCWA : 2D Nx=%d Ny=%d (orders %dx%d), period=%.4e x %.4e[m], %d layers, %d frequencies",
             run.nx_harm, run.ny_harm, 2 * run.nx_harm + 1, 2 * run.ny_harm + 1, run.lx, run.ly,
             run.nlayer, NFreq1);
  else
    snprintf(line, sizeof(line), "RCWA : N=%d (order %d), period=%.4e[m], %d layers, %d frequencies",
             run.nx_harm, 2 * run.nx_harm + 1, run.lx, run.nlayer, NFreq1);
  monitor1(fp_log, line);
  snprintf(line, sizeof(line), "RCWA : incidence theta=%.4g[deg] phi=%.4g[deg]", RcwaTheta, RcwaPhi);
  monitor1(fp_log, line);
  if (run.n_lorentz > 0) {
    snprintf(line, sizeof(line),
             "RCWA : dispersive (Lorentz) materials : %d (layer, eps1/eps2) pairs", run.n_lorentz);
    monitor1(fp_log, line);
  }

  const std::vector<double> ex = merged_edges(run.lx, run.fill);
  const std::vector<double> ey = merged_edges(run.ly, run.fill_y);
  const int ncx = (int)ex.size(), ncy = (int)ey.size(), ncell = (ncx - 1) * (ncy - 1);
  const int L = run.nlayer, nfx = 2 * run.nx_harm + 1, nfy = 2 * run.ny_harm + 1;
  const int h = nfx * nfy;
  const int nf = (int)run.freq.size();

  FILE* csv = fopen("rcwa_efficiency.csv", "w");
  if (!csv) {
    printf("*** file %s open error.\n", "rcwa_efficiency.csv");
    return 1;
  }
  FILE* ord = fopen("rcwa_orders.csv", "w");
  if (!ord) {
    printf("*** file %s open error.\n", "rcwa_orders.csv");
    fclose(csv);
    return 1;
  }
  fprintf(csv, "frequency[Hz],lambda[m],R_TE,T_TE,R_TM,T_TM\n");
  fprintf(ord, "frequency[Hz],lambda[m],pol,m,n,R,T\n");

  // ---- the sweep as one batch: instance f = frequency f, TE and TM
  std::vector<double> grid((size_t)nf * L * ncell * 2), amps((size_t)nf * 8);
  std::vector<double> kx0(nf), ky0(nf);
  std::vector<orcwa_instance> inst(nf);
  std::vector<double> ref((size_t)nf * 2 * h, 0.0), trn((size_t)nf * 2 * h, 0.0);
  std::vector<orcwa_result> res(nf);
  bool warned_outer = false;
  for (int f = 0; f < nf; ++f) {
    const double omega = 2.0 * kPi * run.freq[f], lambda = kC0 / run.freq[f];
    const cplx e_top = run.side[0].at(omega);
    // Bloch offsets from the incidence medium's index (clamped at 1e-6)
    const double k0n = 2.0 * kPi / lambda * std::max(std::sqrt(e_top).real(), 1e-6);
    kx0[f] = k0n * run.sin_t * run.cos_p;
    ky0[f] = k0n * run.sin_t * run.sin_p;
    if (run.n_lorentz > 0 && !warned_outer) {
      const cplx e_bot = run.side[2 * (L - 1)].at(omega);
      if (e_top.imag() > 1e-6 * std::fabs(e_top.real()) ||
          e_bot.imag() > 1e-6 * std::fabs(e_bot.real())) {
        snprintf(line, sizeof(line),
                 "*** warning : a semi-infinite layer is lossy at f=%.4e[Hz];"
                 " R/T normalization assumes lossless outer media",
                 run.freq[f]);
        monitor1(fp_log, line);
        warned_outer = true;
      }
    }
    double* g = &grid[(size_t)f * L * ncell * 2];
    for (int l = 0; l < L; ++l) {
      const cplx in = run.side[2 * l].at(omega), out = run.side[2 * l + 1].at(omega);
      const double xr = run.fill[l] * ex.back(), yr = run.fill_y[l] * ey.back();
      for (int iy = 0; iy + 1 < ncy; ++iy)
        for (int ix = 0; ix + 1 < ncx; ++ix) {
          // a cell is eps1 when its centre lies in [0, fill Lx) x [0, fill_y Ly)
          const bool inside = 0.5 * (ex[ix] + ex[ix + 1]) < xr && 0.5 * (ey[iy] + ey[iy + 1]) < yr;
          const cplx v = inside ? in : out;
          double* c = g + ((size_t)l * ncell + (size_t)iy * (ncx - 1) + ix) * 2;
          c[0] = v.real();
          c[1] = v.imag();
        }
    }
    double* a = &amps[(size_t)f * 8];
    const double te_tm[8] = {run.px_te, 0, run.py_te, 0, run.px_tm, 0, run.py_tm, 0};
    std::memcpy(a, te_tm, sizeof(te_tm));
    orcwa_instance& in = inst[f];
    std::memset(&in, 0, sizeof(in));
    in.lambda = lambda;
    in.kx0 = kx0[f];
    in.ky0 = ky0[f];
    in.eps = g;
    in.n_pol = 2;
    in.pol = a;
    res[f].ref = &ref[(size_t)f * 2 * h];
    res[f].trn = &trn[(size_t)f * 2 * h];
    res[f].amp = nullptr;
    res[f].status = 0;
  }
  if (nf > 0) {
    orcwa_stack st;
    st.nx_harm = run.nx_harm;
    st.ny_harm = run.ny_harm;
    st.n_cx = ncx;
    st.n_cy = ncy;
    st.coord_x = ex.data();
    st.coord_y = ey.data();
    st.n_layers = L;
    st.thickness = run.thick.data();
    orcwa_gpu_context* ctx = nullptr;
    int rc = orcwa_gpu_create(&ctx, 0);
    if (rc == ORCWA_OK) rc = orcwa_gpu_solve_batch(ctx, &st, inst.data(), nf, res.data());
    if (ctx) orcwa_gpu_destroy(ctx);
    if (rc != ORCWA_OK) {
      printf("*** RCWA GPU engine error %d\n", rc);
      fclose(csv);
      fclose(ord);
      return 1;
    }
  }

  // ---- outputs, frequency by frequency
  // spectrum rows and order maps are [pol][frequency] (HDF5 layout,
  // include/rcwa_hdf5.h); an order counts as propagating on a side when
  // kx^2 + ky^2 < k0^2 Re(eps) of that half-space, evanescent ones are 0
  std::vector<orcwa_spectrum_row> rows(2 * (size_t)nf);
  std::vector<double> mapR(2 * (size_t)nf * h, 0.0), mapT(2 * (size_t)nf * h, 0.0);
  for (int f = 0; f < nf; ++f) {
    const double freq = run.freq[f], lambda = kC0 / freq, omega = 2.0 * kPi * freq;
    const double k0 = 2.0 * kPi / lambda;
    const double lim_r = k0 * k0 * run.side[0].at(omega).real();
    const double lim_t = k0 * k0 * run.side[2 * (L - 1)].at(omega).real();
    double tot_r[2] = {0, 0}, tot_t[2] = {0, 0};
    for (int p = 0; p < 2; ++p)
      for (int k = 0; k < h; ++k) {
        tot_r[p] += ref[((size_t)f * 2 + p) * h + k];
        tot_t[p] += trn[((size_t)f * 2 + p) * h + k];
      }
    for (int p = 0; p < 2; ++p) {
      for (int k = 0; k < h; ++k) {
        const int m = k / nfy - run.nx_harm, n = k % nfy - run.ny_harm;
        const double kx = kx0[f] + 2.0 * kPi * m / run.lx;
        const double ky = ky0[f] + 2.0 * kPi * n / run.ly;
        const double kt2 = kx * kx + ky * ky;
        const bool up = kt2 < lim_r, down = kt2 < lim_t;
        const double r = up ? ref[((size_t)f * 2 + p) * h + k] : 0.0;
        const double t = down ? trn[((size_t)f * 2 + p) * h + k] : 0.0;
        mapR[((size_t)p * nf + f) * h + k] = r;
        mapT[((size_t)p * nf + f) * h + k] = t;
        if (up || down)
          fprintf(ord, "%.8e,%.8e,%s,%d,%d,%.8e,%.8e\n", freq, lambda, p == 0 ? "TE" : "TM", m,
                  n, r, t);
      }
    }
    fprintf(csv, "%.8e,%.8e,%.8e,%.8e,%.8e,%.8e\n", freq, lambda, tot_r[0], tot_t[0], tot_r[1],
            tot_t[1]);
    for (int p = 0; p < 2; ++p)
      rows[(size_t)p * nf + f] = {freq, lambda, tot_r[p], tot_t[p], 1.0 - tot_r[p] - tot_t[p]};
    snprintf(line, sizeof(line),
             "  f=%.4e[Hz] lambda=%.4e[m] R/T(TE)=%.5f/%.5f R/T(TM)=%.5f/%.5f", freq, lambda,
             tot_r[0], tot_t[0], tot_r[1], tot_t[1]);
    monitor1(fp_log, line);
    if (res[f].status != 0) {
      snprintf(line, sizeof(line), "*** warning : f=%.4e[Hz] solver status %d (%s)", freq,
               res[f].status, status_text(res[f].status).c_str());
      monitor1(fp_log, line);
    }
    for (int p = 0; p < 2; ++p) {
      const double sum = tot_r[p] + tot_t[p];
      if (run.all_lossless && std::fabs(sum - 1.0) > 1e-3) {
        snprintf(line, sizeof(line), "*** warning : R+T=%.6f (energy not conserved)", sum);
        monitor1(fp_log, line);
      } else if (!run.all_lossless && sum > 1.0 + 1e-3) {
        snprintf(line, sizeof(line),
                 "*** warning : R+T=%.6f > 1 with lossy layers"
                 " (check the sign of the eps imaginary part)",
                 sum);
        monitor1(fp_log, line);
      }
    }
  }
  if (run.bad_freq) printf("%s\n", "*** invalid frequency1 data");
  fclose(csv);
  fclose(ord);
  if (run.bad_freq) return 1;

  snprintf(line, sizeof(line), "output : %s, %s", "rcwa_efficiency.csv", "rcwa_orders.csv");
  monitor1(fp_log, line);
  const char* h5 = "time_series_data.h5";
  orcwa_h5meta meta;
  meta.title = Title;
  meta.nharmonics = run.nx_harm;
  meta.period = run.lx;
  meta.nharmonics_y = run.ny_harm;
  meta.period_y = run.ny_harm > 0 ? run.ly : 0.0;
  meta.nlayer = run.nlayer;
  meta.theta = RcwaTheta;
  meta.phi = RcwaPhi;
  const char* labels[2] = {"TE", "TM"};
  const bool written = rcwa_write_hdf5 && rcwa_write_hdf5_orders &&
                       rcwa_write_hdf5(h5, &meta, rows.data(), 2, nf, labels) == 0 &&
                       rcwa_write_hdf5_orders(h5, 2, nf, nfx, nfy, mapR.data(), mapT.data()) == 0;
  if (written)
    snprintf(line, sizeof(line), "output : %s", h5);
  else
    snprintf(line, sizeof(line), "*** warning : %s write failed (CSV is still valid)", h5);
  monitor1(fp_log, line);
  return 0;
}
