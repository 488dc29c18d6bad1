// rcwa_run(FILE*): the reference's C entry point of the .ofd RCWA path
// (sol/rcwa_bridge.cpp:145-406), on the batched GPU engine.  Same globals,
// same rcwa_efficiency.csv / rcwa_orders.csv bytes, same log lines; every
// frequency of the sweep is one instance of a single GPU batch instead of
// one RCWASolver per frequency.
//
// The input globals are defined by the caller's input module
// (sol/input_data.c:13-24 and src/sol_Main.c in the reference); the weak
// definitions below only let the library load on its own.  HDF5 output goes
// through rcwa_write_hdf5 / rcwa_write_hdf5_orders when the host program
// provides them (sol/rcwa_hdf5.c), and is reported as not written otherwise,
// which the reference also treats as non-fatal (sol/rcwa_bridge.cpp:394-402).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/openrcwa_gpu.h"

extern "C" {
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

// message line to the log file (sol/monitor.c:12-16)
__attribute__((weak)) void monitor1(FILE* fp, const char msg[]) {
  if (!fp) return;
  fprintf(fp, "%s\n", msg);
  fflush(fp);
}

// HDF5 writers of the host program, when linked (sol/rcwa_hdf5.c)
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

typedef std::complex<double> cplx;
const double C0 = 2.99792458e8;  // m/s

std::vector<double> common_edges(double period, bool use_y) {
  std::vector<double> e{0.0};
  for (int n = 0; n < NRcwaLayer; n++) {
    const double f = use_y ? RcwaLayer[n].fill_y : RcwaLayer[n].fill;
    if (f < 1.0) e.push_back(f * period);
  }
  e.push_back(period);
  std::sort(e.begin(), e.end());
  std::vector<double> u;
  for (double v : e)
    if (u.empty() || (v - u.back()) > 1e-12 * period) u.push_back(v);
  return u;
}

bool has_dispersion(int n, int side) {
  for (int p = 0; p < NRcwaPole; p++) {
    const orcwa_rcwapole& q = RcwaPole[p];
    if (q.layer == n && q.side == side && q.ae != 0.0) return true;
  }
  return false;
}

cplx layer_eps(int n, int side, double omega) {
  const orcwa_rcwalayer& l = RcwaLayer[n];
  if (!has_dispersion(n, side)) return side == 1 ? cplx(l.eps1, l.eps1i) : cplx(l.eps2, l.eps2i);
  double einf = 0.0;
  for (int p = 0; p < NRcwaPole; p++)
    if (RcwaPole[p].layer == n && RcwaPole[p].side == side) einf = RcwaPole[p].einf;
  cplx e(einf, 0.0);
  for (int p = 0; p < NRcwaPole; p++) {
    const orcwa_rcwapole& q = RcwaPole[p];
    if (q.layer != n || q.side != side || q.ae == 0.0) continue;
    e += cplx(q.ae * q.ae, 0.0) / cplx(q.ce * q.ce - omega * omega, -q.be * omega);
  }
  return e;
}

bool is_lossless(int n) {
  for (int side = 1; side <= 2; side++) {
    if (has_dispersion(n, side)) {
      for (int p = 0; p < NRcwaPole; p++) {
        const orcwa_rcwapole& q = RcwaPole[p];
        if (q.layer == n && q.side == side && q.ae != 0.0 && q.be != 0.0) return false;
      }
    } else {
      const double im = side == 1 ? RcwaLayer[n].eps1i : RcwaLayer[n].eps2i;
      if (im != 0.0) return false;
    }
  }
  return true;
}

}  // namespace

extern "C" int rcwa_run(FILE* fp_log) {
  char str[BUFSIZ];
  if (NRcwaLayer < 2 || RcwaPeriod <= 0 || NFreq1 < 1) {
    printf("%s\n", "*** invalid rcwa setup");
    return 1;
  }
  if (NRcwaHarmonicsY > 0) {
    snprintf(str, sizeof(str),
             "RCWA : 2D Nx=%d Ny=%d (orders %dx%d), period=%.4e x %.4e[m], %d layers, %d frequencies",
             NRcwaHarmonics, NRcwaHarmonicsY, 2 * NRcwaHarmonics + 1, 2 * NRcwaHarmonicsY + 1,
             RcwaPeriod, RcwaPeriodY, NRcwaLayer, NFreq1);
  } else {
    snprintf(str, sizeof(str), "RCWA : N=%d (order %d), period=%.4e[m], %d layers, %d frequencies",
             NRcwaHarmonics, 2 * NRcwaHarmonics + 1, RcwaPeriod, NRcwaLayer, NFreq1);
  }
  monitor1(fp_log, str);
  snprintf(str, sizeof(str), "RCWA : incidence theta=%.4g[deg] phi=%.4g[deg]", RcwaTheta, RcwaPhi);
  monitor1(fp_log, str);

  bool lossless = true;
  int ndisp = 0;
  for (int n = 0; n < NRcwaLayer; n++) {
    if (!is_lossless(n)) lossless = false;
    ndisp += has_dispersion(n, 1) + has_dispersion(n, 2);
  }
  if (ndisp > 0) {
    snprintf(str, sizeof(str), "RCWA : dispersive (Lorentz) materials : %d (layer, eps1/eps2) pairs",
             ndisp);
    monitor1(fp_log, str);
  }
  const double thetaRad = RcwaTheta * M_PI / 180.0;
  const double phiRad = RcwaPhi * M_PI / 180.0;
  double tmX, tmY, teX, teY;
  if (std::fabs(std::sin(thetaRad)) < 1e-9) {
    tmX = 1.0; tmY = 0.0;
    teX = 0.0; teY = 1.0;
  } else {
    tmX = std::cos(thetaRad) * std::cos(phiRad);
    tmY = std::cos(thetaRad) * std::sin(phiRad);
    teX = -std::sin(phiRad);
    teY = std::cos(phiRad);
  }
  const std::vector<double> xe = common_edges(RcwaPeriod, false);
  const std::vector<double> ye = common_edges(RcwaPeriodY, true);
  const int L = NRcwaLayer;
  std::vector<double> thickness(L);
  for (int n = 0; n < L; n++) thickness[n] = (n == 0 || n == L - 1) ? 0.0 : RcwaLayer[n].thickness;

  const char csvname[] = "rcwa_efficiency.csv";
  FILE* csv = fopen(csvname, "w");
  if (!csv) {
    printf("*** file %s open error.\n", csvname);
    return 1;
  }
  fprintf(csv, "frequency[Hz],lambda[m],R_TE,T_TE,R_TM,T_TM\n");
  const char ordname[] = "rcwa_orders.csv";
  FILE* ord = fopen(ordname, "w");
  if (!ord) {
    printf("*** file %s open error.\n", ordname);
    fclose(csv);
    return 1;
  }
  fprintf(ord, "frequency[Hz],lambda[m],pol,m,n,R,T\n");
  const int nOrdX = 2 * NRcwaHarmonics + 1, nOrdY = 2 * NRcwaHarmonicsY + 1;
  const size_t nOrd = (size_t)nOrdX * nOrdY;
  std::vector<double> ordR(2 * (size_t)NFreq1 * nOrd, 0.0), ordT(2 * (size_t)NFreq1 * nOrd, 0.0);
  std::vector<orcwa_spectrum_row> spectrum(2 * (size_t)NFreq1);

  // frequencies up to the first invalid one (the reference stops there)
  int nvalid = 0;
  while (nvalid < NFreq1 && Freq1[nvalid] > 0) ++nvalid;
  const int ierr = (nvalid < NFreq1) ? 1 : 0;

  // ---- one batch for the whole sweep
  const int ncx = (int)xe.size(), ncy = (int)ye.size();
  const int ncell = (ncx - 1) * (ncy - 1);
  std::vector<double> eps((size_t)nvalid * L * ncell * 2);
  std::vector<double> pols((size_t)nvalid * 8);
  std::vector<orcwa_instance> inst(nvalid);
  std::vector<double> ref((size_t)nvalid * 2 * nOrd), trn((size_t)nvalid * 2 * nOrd);
  std::vector<orcwa_result> res(nvalid);
  std::vector<double> kx0v(nvalid), ky0v(nvalid);
  static bool warned = false;
  for (int f = 0; f < nvalid; f++) {
    const double freq = Freq1[f], lambda = C0 / freq, k0 = 2.0 * M_PI / lambda;
    const double omega = 2.0 * M_PI * freq;
    const cplx eps_inc = layer_eps(0, 1, omega);
    const double n_inc = std::max(std::sqrt(eps_inc).real(), 1e-6);
    if (ndisp > 0 && !warned) {
      const cplx eps_trn = layer_eps(L - 1, 1, omega);
      if (eps_inc.imag() > 1e-6 * std::fabs(eps_inc.real()) ||
          eps_trn.imag() > 1e-6 * std::fabs(eps_trn.real())) {
        snprintf(str, sizeof(str),
                 "*** warning : a semi-infinite layer is lossy at f=%.4e[Hz];"
                 " R/T normalization assumes lossless outer media",
                 freq);
        monitor1(fp_log, str);
        warned = true;
      }
    }
    for (int n = 0; n < L; n++) {
      const orcwa_rcwalayer& l = RcwaLayer[n];
      const cplx e1 = layer_eps(n, 1, omega), e2 = layer_eps(n, 2, omega);
      for (int j = 0; j < ncy - 1; j++) {
        const double yc = 0.5 * (ye[j] + ye[j + 1]);
        for (int i = 0; i < ncx - 1; i++) {
          const double xc = 0.5 * (xe[i] + xe[i + 1]);
          const bool in1 = (xc < l.fill * xe.back()) && (yc < l.fill_y * ye.back());
          const cplx v = in1 ? e1 : e2;
          double* dst = &eps[(((size_t)f * L + n) * ncell + (size_t)j * (ncx - 1) + i) * 2];
          dst[0] = v.real();
          dst[1] = v.imag();
        }
      }
    }
    kx0v[f] = k0 * n_inc * std::sin(thetaRad) * std::cos(phiRad);
    ky0v[f] = k0 * n_inc * std::sin(thetaRad) * std::sin(phiRad);
    double* pp = &pols[(size_t)f * 8];
    pp[0] = teX; pp[1] = 0; pp[2] = teY; pp[3] = 0;
    pp[4] = tmX; pp[5] = 0; pp[6] = tmY; pp[7] = 0;
    inst[f].lambda = lambda;
    inst[f].kx0 = kx0v[f];
    inst[f].ky0 = ky0v[f];
    inst[f].eps = &eps[(size_t)f * L * ncell * 2];
    inst[f].n_pol = 2;
    inst[f].pol = pp;
    inst[f].mu = nullptr;
    inst[f].coord_x = nullptr;
    inst[f].coord_y = nullptr;
    inst[f].thickness = nullptr;
    res[f].ref = &ref[(size_t)f * 2 * nOrd];
    res[f].trn = &trn[(size_t)f * 2 * nOrd];
    res[f].amp = nullptr;
    res[f].status = 0;
  }
  if (nvalid > 0) {
    orcwa_stack st;
    st.nx_harm = NRcwaHarmonics;
    st.ny_harm = NRcwaHarmonicsY;
    st.n_cx = ncx;
    st.n_cy = ncy;
    st.coord_x = xe.data();
    st.coord_y = ye.data();
    st.n_layers = L;
    st.thickness = thickness.data();
    orcwa_gpu_context* ctx = nullptr;
    int rc = orcwa_gpu_create(&ctx, 0);
    if (rc == ORCWA_OK) rc = orcwa_gpu_solve_batch(ctx, &st, inst.data(), nvalid, res.data());
    orcwa_gpu_destroy(ctx);
    if (rc != ORCWA_OK) {
      printf("*** RCWA GPU engine error %d\n", rc);
      fclose(csv);
      fclose(ord);
      return 1;
    }
  }

  for (int f = 0; f < nvalid; f++) {
    const double freq = Freq1[f], lambda = C0 / freq, k0 = 2.0 * M_PI / lambda;
    const double omega = 2.0 * M_PI * freq;
    const cplx eps_inc = layer_eps(0, 1, omega);
    const double kr2 = k0 * k0 * eps_inc.real();
    const double kt2 = k0 * k0 * layer_eps(L - 1, 1, omega).real();
    std::vector<char> propR(nOrd), propT(nOrd);
    for (size_t k = 0; k < nOrd; k++) {
      const int m = (int)(k / (size_t)nOrdY) - NRcwaHarmonics;
      const int n = (int)(k % (size_t)nOrdY) - NRcwaHarmonicsY;
      const double kx = kx0v[f] + 2.0 * M_PI * m / RcwaPeriod;
      const double ky = ky0v[f] + 2.0 * M_PI * n / RcwaPeriodY;
      const double kt = kx * kx + ky * ky;
      propR[k] = (kt < kr2);
      propT[k] = (kt < kt2);
    }
    double R[2], T[2];
    for (int pol = 0; pol < 2; pol++) {
      const double* r = &ref[((size_t)f * 2 + pol) * nOrd];
      const double* t = &trn[((size_t)f * 2 + pol) * nOrd];
      double sr = 0.0, stt = 0.0;
      for (size_t k = 0; k < nOrd; k++) {
        sr += r[k];
        stt += t[k];
      }
      R[pol] = sr;
      T[pol] = stt;
      const size_t base = ((size_t)pol * NFreq1 + f) * nOrd;
      for (size_t k = 0; k < nOrd; k++) {
        ordR[base + k] = propR[k] ? r[k] : 0.0;
        ordT[base + k] = propT[k] ? t[k] : 0.0;
      }
    }
    for (int pol = 0; pol < 2; pol++) {
      const size_t base = ((size_t)pol * NFreq1 + f) * nOrd;
      for (size_t k = 0; k < nOrd; k++) {
        if (!propR[k] && !propT[k]) continue;
        const int m = (int)(k / (size_t)nOrdY) - NRcwaHarmonics;
        const int n = (int)(k % (size_t)nOrdY) - NRcwaHarmonicsY;
        fprintf(ord, "%.8e,%.8e,%s,%d,%d,%.8e,%.8e\n", freq, lambda, pol == 0 ? "TE" : "TM", m, n,
                ordR[base + k], ordT[base + k]);
      }
    }
    fprintf(csv, "%.8e,%.8e,%.8e,%.8e,%.8e,%.8e\n", freq, lambda, R[0], T[0], R[1], T[1]);
    for (int pol = 0; pol < 2; pol++) {
      orcwa_spectrum_row& row = spectrum[(size_t)pol * NFreq1 + f];
      row.frequency = freq;
      row.lambda = lambda;
      row.R = R[pol];
      row.T = T[pol];
      row.A = 1.0 - R[pol] - T[pol];
    }
    snprintf(str, sizeof(str), "  f=%.4e[Hz] lambda=%.4e[m] R/T(TE)=%.5f/%.5f R/T(TM)=%.5f/%.5f",
             freq, lambda, R[0], T[0], R[1], T[1]);
    monitor1(fp_log, str);
    for (int pol = 0; pol < 2; pol++) {
      const double sum = R[pol] + T[pol];
      if (lossless) {
        if (std::fabs(sum - 1.0) > 1e-3) {
          snprintf(str, sizeof(str), "*** warning : R+T=%.6f (energy not conserved)", sum);
          monitor1(fp_log, str);
        }
      } else if (sum > 1.0 + 1e-3) {
        snprintf(str, sizeof(str),
                 "*** warning : R+T=%.6f > 1 with lossy layers"
                 " (check the sign of the eps imaginary part)",
                 sum);
        monitor1(fp_log, str);
      }
    }
  }
  if (ierr) printf("%s\n", "*** invalid frequency1 data");
  fclose(csv);
  fclose(ord);

  if (!ierr) {
    snprintf(str, sizeof(str), "output : %s, %s", csvname, ordname);
    monitor1(fp_log, str);
    const char h5name[] = "time_series_data.h5";
    orcwa_h5meta meta;
    meta.title = Title;
    meta.nharmonics = NRcwaHarmonics;
    meta.period = RcwaPeriod;
    meta.nharmonics_y = NRcwaHarmonicsY;
    meta.period_y = (NRcwaHarmonicsY > 0) ? RcwaPeriodY : 0.0;
    meta.nlayer = NRcwaLayer;
    meta.theta = RcwaTheta;
    meta.phi = RcwaPhi;
    const char* pol_labels[2] = {"TE", "TM"};
    if (rcwa_write_hdf5 && rcwa_write_hdf5_orders &&
        rcwa_write_hdf5(h5name, &meta, spectrum.data(), 2, nvalid, pol_labels) == 0 &&
        rcwa_write_hdf5_orders(h5name, 2, nvalid, nOrdX, nOrdY, ordR.data(), ordT.data()) == 0) {
      snprintf(str, sizeof(str), "output : %s", h5name);
    } else {
      snprintf(str, sizeof(str), "*** warning : %s write failed (CSV is still valid)", h5name);
    }
    monitor1(fp_log, str);
  }
  return ierr;
}
