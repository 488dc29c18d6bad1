/*
 * openrcwa_gpu.hpp -- header-only C++17 facade of the MI355X engine with the
 * reference's C++ API, over the C ABI of openrcwa_gpu.h.
 *
 *  - openrcwa::Layer / openrcwa::RCWASolver: the per-wavelength solver of
 *    rcwa/Layer.h:29-37 and rcwa/RCWASolver.h:67-185 (addLayer, disablePML,
 *    setBlochWavevector, solve, generateHorizontalPlaneWave,
 *    scatterPlaneWave), each solve one GPU batch of one instance.
 *  - openrcwa::RCWAProblem / RCWAResult / RCWAFieldRequest / runRCWA: the
 *    .orcwa driver of rcwa/RCWADriver.h:16-46 (problem fields as
 *    rcwa/RCWAInput.h:25-110), every wavelength of a problem solved as one GPU
 *    batch.
 *
 * Link with libopenrcwa_gpu.so.  Lengths and wavelengths share one unit, as
 * in the reference (micrometres on the .orcwa path).  Errors are reported as
 * the reference reports them: messages on std::cerr and early returns
 * (RCWASolver), or an error string (runRCWA).
 */
#ifndef OPENRCWA_GPU_HPP
#define OPENRCWA_GPU_HPP

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <iostream>
#include <set>
#include <string>
#include <vector>

#include "openrcwa_gpu.h"

namespace openrcwa {

using scalar = double;
using scalex = std::complex<double>;

namespace detail {

constexpr scalar kPi = 3.1415926535897932;   // rcwa/defns.h:25

// One shared engine context per process and device (created on first use).
inline orcwa_gpu_context* context(int device) {
  static std::vector<orcwa_gpu_context*> ctxs;
  if ((int)ctxs.size() <= device) ctxs.resize(device + 1, nullptr);
  if (!ctxs[device] && orcwa_gpu_create(&ctxs[device], device) != ORCWA_OK) ctxs[device] = nullptr;
  return ctxs[device];
}

// Merged sorted edges of several layers (all share the first/last edges).
inline std::vector<scalar> merge_edges(const std::vector<const std::vector<scalar>*>& lists) {
  std::set<scalar> s;
  for (const auto* l : lists) s.insert(l->begin(), l->end());
  return std::vector<scalar>(s.begin(), s.end());
}

// Index of the cell of `edges` containing x (cells [e_i, e_{i+1}]).
inline int cell_of(const std::vector<scalar>& edges, scalar x) {
  int i = (int)(std::upper_bound(edges.begin(), edges.end(), x) - edges.begin()) - 1;
  return std::max(0, std::min((int)edges.size() - 2, i));
}

// Per-order R/T of RCWASolver::scatterPlaneWave (rcwa/RCWASolver.cpp:1792-1908)
// from the harmonic amplitudes (rx, ry) reflected and (tx, ty) transmitted.
// Returns false (REF/TRN left empty) where the reference returns silently.
inline bool efficiencies(int nx, int ny, scalar Lx, scalar Ly, scalar kx0, scalar ky0, scalar k0,
                         scalex eps_ref, scalex mu_ref, scalex eps_trn, scalex mu_trn,
                         const std::vector<scalex>& r, const std::vector<scalex>& t,
                         std::vector<scalar>& REF, std::vector<scalar>& TRN) {
  const int h = nx * ny, Nx = (nx - 1) / 2, Ny = (ny - 1) / 2;
  std::vector<scalar> kzr(h), kzt(h), rr(h), tt(h);
  const scalex kref2 = k0 * k0 * eps_ref * mu_ref, ktrn2 = k0 * k0 * eps_trn * mu_trn;
  for (int k = 0; k < h; ++k) {
    const int m = k / ny - Nx, n = k % ny - Ny;
    const scalar kx = 2 * kPi * m / Lx + kx0, ky = 2 * kPi * n / Ly + ky0;
    const scalex zr = -std::sqrt(kref2 - kx * kx - ky * ky);
    const scalex zt = std::sqrt(ktrn2 - kx * kx - ky * ky);
    if (zr == scalex(0.0) || zt == scalex(0.0)) {
      std::cerr << "One plane wave component moves traversely!\n";
      return false;
    }
    const scalex rz = (kx * r[k] + ky * r[h + k]) / zr;
    const scalex tz = (kx * t[k] + ky * t[h + k]) / zt;
    rr[k] = std::norm(r[k]) + std::norm(r[h + k]) + std::norm(rz);
    tt[k] = std::norm(t[k]) + std::norm(t[h + k]) + std::norm(tz);
    kzr[k] = std::abs(zr.real());
    kzt[k] = std::abs(zt.real());
  }
  const int inc = ny * (nx - 1) / 2 + (ny - 1) / 2;
  const scalar kzinc = kzr[inc];
  if (kzinc < 1e-12 * k0) {
    std::cerr << "Warning: incident kz is near zero (grazing incidence?)\n";
    return false;
  }
  const scalar mu_ratio = std::abs(mu_ref / mu_trn);
  REF.assign(h, 0.0);
  TRN.assign(h, 0.0);
  for (int k = 0; k < h; ++k) {
    REF[k] = kzr[k] * rr[k] / kzinc;
    TRN[k] = mu_ratio * kzt[k] * tt[k] / kzinc;
  }
  return true;
}

}  // namespace detail

// ---------------------------------------------------------------------------
// Layer(coordX, coordY, eps[, mu]) (rcwa/Layer.h:29-37): eps and mu are
// (nCy - 1) x (nCx - 1) cell values, row = y cell, stored row-major.
// ---------------------------------------------------------------------------
class Layer {
 public:
  Layer(const std::vector<scalar>& coordX, const std::vector<scalar>& coordY,
        const std::vector<scalex>& eps, const std::vector<scalex>& mu = {})
      : coordX_(coordX), coordY_(coordY), eps_(eps), mu_(mu) {
    if (mu_.empty()) mu_.assign(eps_.size(), scalex(1.0));
  }
  int nCx() const { return (int)coordX_.size(); }
  int nCy() const { return (int)coordY_.size(); }
  scalar Lx() const { return coordX_.back() - coordX_.front(); }
  scalar Ly() const { return coordY_.back() - coordY_.front(); }
  const std::vector<scalar>& coordX() const { return coordX_; }
  const std::vector<scalar>& coordY() const { return coordY_; }
  const std::vector<scalex>& eps() const { return eps_; }
  const std::vector<scalex>& mu() const { return mu_; }
  // value of the cell containing (x, y)
  scalex eps(scalar x, scalar y) const { return at(eps_, x, y); }
  scalex mu(scalar x, scalar y) const { return at(mu_, x, y); }
  bool isMagnetic() const {
    for (const scalex& m : mu_)
      if (m != scalex(1.0)) return true;
    return false;
  }

 private:
  scalex at(const std::vector<scalex>& g, scalar x, scalar y) const {
    const int ix = detail::cell_of(coordX_, x), iy = detail::cell_of(coordY_, y);
    return g[(size_t)iy * (nCx() - 1) + ix];
  }
  std::vector<scalar> coordX_, coordY_;
  std::vector<scalex> eps_, mu_;
};

// ---------------------------------------------------------------------------
// RCWASolver (rcwa/RCWASolver.h:67-185).  solve() runs the layer-stack solve
// on the GPU and keeps the plane-wave response of the incidence-side zeroth
// order for the two transverse unit excitations; scatterPlaneWave combines
// them for the (px, py) that generateHorizontalPlaneWave put into c.  The
// mode basis of a homogeneous layer is the identity here, so c holds the
// incident harmonic amplitudes [px e_0; py e_0] and only such c are accepted.
// Layers with different interior cell edges are solved on the union of their
// edges (the same piecewise-constant functions, so the same Fourier series).
// ---------------------------------------------------------------------------
class RCWASolver {
 public:
  RCWASolver(int nX, int nY, int device = 0)
      : nX_(nX), nY_(nY), nx_(2 * nX + 1), ny_(2 * nY + 1), device_(device) {}

  void addLayer(const Layer& layer) {
    if (!layers_.empty()) {
      const Layer& f = layers_.front();
      // the reference compares the period and the first and last two edges
      // (rcwa/RCWASolver.cpp:106-133) and drops a mismatching layer
      auto ends = [](const std::vector<scalar>& c) {
        return std::vector<scalar>{c[0], c[1], c[c.size() - 2], c[c.size() - 1]};
      };
      if (f.Lx() != layer.Lx() || f.Ly() != layer.Ly() ||
          ends(f.coordX()) != ends(layer.coordX()) || ends(f.coordY()) != ends(layer.coordY())) {
        std::cerr << "The new layer does not match layers before!\n";
        return;
      }
    }
    layers_.push_back(layer);
  }
  int nLayerTypes() const { return (int)layers_.size(); }
  const Layer& getLayer(int index) const { return layers_[index]; }
  // The GPU engine solves the periodic problem (the PML-off path of the
  // reference's plain solve, sol/rcwa_bridge.cpp:244-282).
  void enablePML() { pml_ = true; }
  void disablePML() { pml_ = false; }
  void setBlochWavevector(scalar kx0, scalar ky0) {
    kx0_ = kx0;
    ky0_ = ky0;
  }
  int nx() const { return nx_; }
  int ny() const { return ny_; }
  bool isScatterMatrixSolved() const { return solved_; }
  // orcwa_status bits of the last solve
  int status() const { return status_; }

  void solve(scalar lambda, const std::vector<int>& layerStack,
             const std::vector<scalar>& thickness) {
    solved_ = false;
    if (pml_) {
      std::cerr << "RCWASolver (GPU): PML boundaries are not supported; call disablePML()\n";
      return;
    }
    if (layerStack.size() < 2 || thickness.size() != layerStack.size()) {
      std::cerr << "RCWASolver (GPU): the stack needs at least two layers and one thickness "
                   "per layer\n";
      return;
    }
    std::vector<const std::vector<scalar>*> xs, ys;
    for (int t : layerStack) {
      if (t < 0 || t >= (int)layers_.size()) {
        std::cerr << "RCWASolver (GPU): layer type out of range\n";
        return;
      }
      xs.push_back(&layers_[t].coordX());
      ys.push_back(&layers_[t].coordY());
    }
    cx_ = detail::merge_edges(xs);
    cy_ = detail::merge_edges(ys);
    const int L = (int)layerStack.size(), ncx = (int)cx_.size(), ncy = (int)cy_.size();
    const size_t cells = (size_t)(ncx - 1) * (ncy - 1);
    std::vector<scalex> eps(cells * L), mu(cells * L);
    bool magnetic = false;
    for (int l = 0; l < L; ++l) {
      const Layer& ly = layers_[layerStack[l]];
      magnetic = magnetic || ly.isMagnetic();
      for (int iy = 0; iy + 1 < ncy; ++iy)
        for (int ix = 0; ix + 1 < ncx; ++ix) {
          const scalar xc = 0.5 * (cx_[ix] + cx_[ix + 1]), yc = 0.5 * (cy_[iy] + cy_[iy + 1]);
          eps[l * cells + iy * (ncx - 1) + ix] = ly.eps(xc, yc);
          mu[l * cells + iy * (ncx - 1) + ix] = ly.mu(xc, yc);
        }
    }
    orcwa_stack st{nX_, nY_, ncx, ncy, cx_.data(), cy_.data(), L, thickness.data()};
    const double basis[8] = {1, 0, 0, 0, 0, 0, 1, 0};   // s_x = (1, 0), s_y = (0, 1)
    orcwa_instance in{};
    in.lambda = lambda;
    in.kx0 = kx0_;
    in.ky0 = ky0_;
    in.eps = reinterpret_cast<const double*>(eps.data());
    in.n_pol = 2;
    in.pol = basis;
    in.mu = magnetic ? reinterpret_cast<const double*>(mu.data()) : nullptr;
    const int h = nx_ * ny_;
    amp_.assign((size_t)2 * 4 * h, scalex(0.0));
    std::vector<double> ref(2 * h), trn(2 * h);
    orcwa_result res{ref.data(), trn.data(), reinterpret_cast<double*>(amp_.data()), 0};
    orcwa_gpu_context* ctx = detail::context(device_);
    if (!ctx) {
      std::cerr << "RCWASolver (GPU): no usable GPU\n";
      return;
    }
    const int rc = orcwa_gpu_solve_batch(ctx, &st, &in, 1, &res);
    status_ = res.status;
    if (rc != ORCWA_OK) {
      std::cerr << "RCWASolver (GPU): solve failed (error " << rc << ")\n";
      return;
    }
    lambda_ = lambda;
    stack_ = layerStack;
    solved_ = true;
  }

  // c = [px e_0; py e_0], e_0 the zeroth order (rcwa/RCWASolver.cpp:1778-1789)
  void generateHorizontalPlaneWave(scalex px, scalex py, int /*layerType*/,
                                   std::vector<scalex>& c) const {
    const int h = nx_ * ny_, i0 = ny_ * (nx_ - 1) / 2 + (ny_ - 1) / 2;
    c.assign((size_t)2 * h, scalex(0.0));
    c[i0] = px;
    c[h + i0] = py;
  }

  void scatterPlaneWave(const std::vector<scalex>& cInc, int refLayerType, int trnLayerType,
                        scalar k0, std::vector<scalar>& REF, std::vector<scalar>& TRN) const {
    if (!solved_) {
      std::cerr << "The scatter matrix has not been solved!\n";
      return;
    }
    const int h = nx_ * ny_, i0 = ny_ * (nx_ - 1) / 2 + (ny_ - 1) / 2;
    for (int k = 0; k < 2 * h; ++k)
      if (k != i0 && k != h + i0 && cInc[k] != scalex(0.0)) {
        std::cerr << "RCWASolver (GPU): only plane-wave excitations "
                     "(generateHorizontalPlaneWave) are supported\n";
        return;
      }
    const scalex px = cInc[i0], py = cInc[h + i0];
    // amp_[e][q][k]: excitation e, q = rx, ry, tx, ty
    std::vector<scalex> r(2 * h), t(2 * h);
    auto A = [&](int e, int q, int k) { return amp_[((size_t)e * 4 + q) * h + k]; };
    for (int k = 0; k < h; ++k) {
      r[k] = px * A(0, 0, k) + py * A(1, 0, k);
      r[h + k] = px * A(0, 1, k) + py * A(1, 1, k);
      t[k] = px * A(0, 2, k) + py * A(1, 2, k);
      t[h + k] = px * A(0, 3, k) + py * A(1, 3, k);
    }
    const Layer& lr = layers_[refLayerType];
    const Layer& lt = layers_[trnLayerType];
    detail::efficiencies(nx_, ny_, cx_.back() - cx_.front(), cy_.back() - cy_.front(), kx0_, ky0_,
                         k0, lr.eps(lr.coordX()[0], lr.coordY()[0]),
                         lr.mu(lr.coordX()[0], lr.coordY()[0]),
                         lt.eps(lt.coordX()[0], lt.coordY()[0]),
                         lt.mu(lt.coordX()[0], lt.coordY()[0]), r, t, REF, TRN);
  }

 private:
  int nX_, nY_, nx_, ny_, device_;
  bool pml_ = true;                 // the reference's default (rcwa/RCWASolver.h:67-78)
  scalar kx0_ = 0.0, ky0_ = 0.0, lambda_ = 0.0;
  std::vector<Layer> layers_;
  std::vector<int> stack_;
  std::vector<scalar> cx_, cy_;
  std::vector<scalex> amp_;
  bool solved_ = false;
  int status_ = 0;
};

// ---------------------------------------------------------------------------
// The .orcwa driver: problem fields of rcwa/RCWAInput.h:25-110, results of
// rcwa/RCWADriver.h:16-28, runRCWA of rcwa/RCWADriver.cpp:195-378.
// ---------------------------------------------------------------------------
enum RCWAShape {
  RCWA_SHAPE_BOX = 1,
  RCWA_SHAPE_SPHERE = 2,
  RCWA_SHAPE_CYL_Z = 11,
  RCWA_SHAPE_CYL_X = 12,
  RCWA_SHAPE_CYL_Y = 13
};

struct RCWABox {
  int material = 0;
  int shape = RCWA_SHAPE_BOX;
  scalar x0 = 0, x1 = 0, y0 = 0, y1 = 0, z0 = 0, z1 = 0;

  // cross-section at height zc (rcwa/RCWAInput.cpp:49-89)
  bool containsXYZ(scalar xc, scalar yc, scalar zc) const {
    if (zc < z0 || zc > z1) return false;
    const scalar cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1), cz = 0.5 * (z0 + z1);
    const scalar rx = 0.5 * (x1 - x0), ry = 0.5 * (y1 - y0), rz = 0.5 * (z1 - z0);
    switch (shape) {
      case RCWA_SHAPE_BOX:
        return x0 <= xc && xc <= x1 && y0 <= yc && yc <= y1;
      case RCWA_SHAPE_SPHERE: {
        if (rx <= 0 || ry <= 0 || rz <= 0) return false;
        const scalar dx = (xc - cx) / rx, dy = (yc - cy) / ry, dz = (zc - cz) / rz;
        return dx * dx + dy * dy + dz * dz <= 1.0;
      }
      case RCWA_SHAPE_CYL_Z: {
        if (rx <= 0 || ry <= 0) return false;
        const scalar dx = (xc - cx) / rx, dy = (yc - cy) / ry;
        return dx * dx + dy * dy <= 1.0;
      }
      case RCWA_SHAPE_CYL_X: {
        if (ry <= 0 || rz <= 0 || xc < x0 || xc > x1) return false;
        const scalar dy = (yc - cy) / ry, dz = (zc - cz) / rz;
        return dy * dy + dz * dz <= 1.0;
      }
      case RCWA_SHAPE_CYL_Y: {
        if (rx <= 0 || rz <= 0 || yc < y0 || yc > y1) return false;
        const scalar dx = (xc - cx) / rx, dz = (zc - cz) / rz;
        return dx * dx + dz * dz <= 1.0;
      }
      default:
        return false;
    }
  }
};

struct RCWAProblem {
  std::string title;
  scalar xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0;   // um
  scalar Lx() const { return xmax - xmin; }
  scalar Ly() const { return ymax - ymin; }
  int nHx = 10, nHy = 0;
  std::vector<scalex> materialEps;        // 0 air, 1 PEC, 2.. user materials
  std::vector<scalex> materialMu;
  std::vector<scalar> materialSigma;      // S/m
  std::vector<scalar> materialMagSigma;   // Ohm/m
  struct LorentzPole {
    scalar ae = 0, be = 0, ce = 0;
  };
  struct MaterialDispersion {
    scalar einf = 0;
    std::vector<LorentzPole> poles;
    bool active() const { return !poles.empty(); }
  };
  std::vector<MaterialDispersion> materialDispersion;
  std::vector<RCWABox> boxes;
  scalar theta = 0, phi = 0;   // deg
  int pol = 1;                 // 1 TM, 2 TE, 3 linear at psi, 4 right, 5 left circular
  scalar psi = 0;
  scalex backgroundEps = scalex(1.0, 0.0);
  std::vector<scalar> lambdas;  // um
  bool periodicX = true, periodicY = true;
};

struct RCWAResult {
  scalar lambda;                 // um
  scalar R, T, A;
  std::vector<scalar> REF_orders;   // k = (2 nHy + 1) i + j for order (i - nHx, j - nHy)
  std::vector<scalar> TRN_orders;
  int status = 0;                // orcwa_status bits (an addition of this engine)
};

enum FieldComponent { Ex = 0, Ey = 1, Ez = 2, Hx = 3, Hy = 4, Hz = 5 };   // rcwa/defns.h:27-34
enum SliceType { sliceXY = 0, sliceXZ = 1, sliceYZ = 2 };
enum SaveOption { modulation = 0, realPart = 1, imagPart = 2 };

struct RCWAFieldRequest {
  FieldComponent component = Ez;
  SliceType slice = sliceXZ;
  scalar coord = 0.0;
  SaveOption opt = modulation;
  std::string path;         // field image CSV (first wavelength)
  std::string devicePath;   // permittivity map CSV (real part)
  bool wantsAnything() const { return !path.empty() || !devicePath.empty(); }
};

namespace detail {

constexpr scalar kC0um = 2.99792458e14;   // um/s
constexpr scalar kEps0 = 8.854187817e-12;
constexpr scalar kMu0 = 1.25663706212e-6;
constexpr int kCurvedCells = 8;           // in-plane staircase of curved shapes
constexpr int kCurvedSlices = 8;          // z slices of shapes varying with z

inline std::vector<scalar> z_edges(const RCWAProblem& p) {   // rcwa/RCWAInput.cpp:413-453
  std::set<scalar> e{p.zmin, p.zmax};
  for (const RCWABox& b : p.boxes) {
    if (p.zmin <= b.z0 && b.z0 <= p.zmax) e.insert(b.z0);
    if (p.zmin <= b.z1 && b.z1 <= p.zmax) e.insert(b.z1);
    if (b.shape == RCWA_SHAPE_SPHERE || b.shape == RCWA_SHAPE_CYL_X ||
        b.shape == RCWA_SHAPE_CYL_Y)
      for (int i = 1; i < kCurvedSlices; ++i) {
        const scalar z = b.z0 + (b.z1 - b.z0) * i / (scalar)kCurvedSlices;
        if (p.zmin <= z && z <= p.zmax) e.insert(z);
      }
  }
  return std::vector<scalar>(e.begin(), e.end());
}

// cell edges shared by every slab (rcwa/RCWADriver.cpp:52-83)
inline std::vector<scalar> plane_grid(const RCWAProblem& p, bool xside) {
  const scalar lo = xside ? p.xmin : p.ymin, hi = xside ? p.xmax : p.ymax;
  std::set<scalar> e{lo, hi};
  for (const RCWABox& b : p.boxes) {
    const scalar a0 = std::max(xside ? b.x0 : b.y0, lo), a1 = std::min(xside ? b.x1 : b.y1, hi);
    if (lo < a0 && a0 < hi) e.insert(a0);
    if (lo < a1 && a1 < hi) e.insert(a1);
    const bool curved = b.shape == RCWA_SHAPE_SPHERE || b.shape == RCWA_SHAPE_CYL_Z ||
                        b.shape == (xside ? RCWA_SHAPE_CYL_Y : RCWA_SHAPE_CYL_X);
    if (curved && a0 < a1)
      for (int i = 1; i < kCurvedCells; ++i) {
        const scalar v = a0 + (a1 - a0) * i / (scalar)kCurvedCells;
        if (lo < v && v < hi) e.insert(v);
      }
  }
  return std::vector<scalar>(e.begin(), e.end());
}

// permittivity at a point (rcwa/RCWADriver.cpp:87-126): the last containing box
// wins; Lorentz sum for a dispersive material; +i sigma / (omega eps0)
inline scalex eps_at(const RCWAProblem& p, scalar x, scalar y, scalar z, scalar lam) {
  scalex e = p.backgroundEps;
  const scalar omega = 2.0 * kPi * kC0um / lam;
  for (const RCWABox& b : p.boxes) {
    if (!b.containsXYZ(x, y, z)) continue;
    const int m = b.material;
    if (m < 0 || m >= (int)p.materialEps.size()) continue;
    e = p.materialEps[m];
    if (m < (int)p.materialDispersion.size() && p.materialDispersion[m].active()) {
      const auto& d = p.materialDispersion[m];
      e = scalex(d.einf, 0.0);
      for (const auto& q : d.poles)
        e += scalex(q.ae * q.ae, 0.0) / scalex(q.ce * q.ce - omega * omega, -q.be * omega);
    }
    if (m < (int)p.materialSigma.size() && p.materialSigma[m] != 0.0)
      e += scalex(0.0, p.materialSigma[m] / (omega * kEps0));
  }
  return e;
}

// permeability at a point (rcwa/RCWADriver.cpp:129-149)
inline scalex mu_at(const RCWAProblem& p, scalar x, scalar y, scalar z, scalar lam) {
  scalex m = 1.0;
  for (const RCWABox& b : p.boxes) {
    if (!b.containsXYZ(x, y, z)) continue;
    const int mi = b.material;
    if (mi < 0 || mi >= (int)p.materialMu.size()) continue;
    m = p.materialMu[mi];
    if (mi < (int)p.materialMagSigma.size() && p.materialMagSigma[mi] != 0.0) {
      const scalar omega = 2.0 * kPi * kC0um / lam;
      m += scalex(0.0, p.materialMagSigma[mi] / (omega * kMu0));
    }
  }
  return m;
}

inline bool magnetic(const RCWAProblem& p) {   // rcwa/RCWADriver.cpp:152-160
  for (const scalex& m : p.materialMu)
    if (std::abs(m - 1.0) > 1e-12) return true;
  for (scalar s : p.materialMagSigma)
    if (s != 0.0) return true;
  return false;
}

// transverse (px, py) of the unit incident field (rcwa/RCWADriver.cpp:232-276)
inline void polarisation(const RCWAProblem& p, scalex& px, scalex& py) {
  const scalar th = p.theta * kPi / 180.0, ph = p.phi * kPi / 180.0;
  scalar tm[2], te[2];
  if (std::abs(std::sin(th)) < 1e-9) {
    tm[0] = 1.0; tm[1] = 0.0; te[0] = 0.0; te[1] = 1.0;
  } else {
    tm[0] = std::cos(th) * std::cos(ph);
    tm[1] = std::cos(th) * std::sin(ph);
    te[0] = -std::sin(ph);
    te[1] = std::cos(ph);
  }
  const scalar s2 = 1.0 / std::sqrt(2.0);
  scalex atm = 1.0, ate = 0.0;
  if (p.pol == 2) {
    atm = 0.0; ate = 1.0;
  } else if (p.pol == 3) {
    atm = std::cos(p.psi * kPi / 180.0);
    ate = std::sin(p.psi * kPi / 180.0);
  } else if (p.pol == 4) {
    atm = s2; ate = scalex(0.0, s2);
  } else if (p.pol == 5) {
    atm = s2; ate = scalex(0.0, -s2);
  }
  px = atm * tm[0] + ate * te[0];
  py = atm * tm[1] + ate * te[1];
}

inline void write_csv(const std::string& path, int rows, int cols, const std::vector<scalex>& f,
                      SaveOption opt) {   // writeFieldCSV, rcwa/RCWASolver.cpp:25-50
  FILE* fp = std::fopen(path.c_str(), "w");
  if (!fp) return;
  for (int r = 0; r < rows; ++r) {
    for (int c = 0; c < cols; ++c) {
      const scalex v = f[(size_t)r * cols + c];
      const scalar x = opt == realPart ? v.real() : opt == imagPart ? v.imag() : std::abs(v);
      std::fprintf(fp, c ? ", %.8g" : "%.8g", x);
    }
    std::fprintf(fp, "\n");
  }
  std::fclose(fp);
}

}  // namespace detail

// Solve every wavelength of `prob` (one GPU batch).  Slabs run from the top
// (largest z, incidence side) down; the top and bottom slabs must be
// homogeneous.  On error the partial result (empty here: the batch is solved
// at once) is returned and err is set.  A grazing wavelength gets empty
// per-order vectors and R = T = 0, as the reference's silent return leaves them.
inline std::vector<RCWAResult> runRCWA(const RCWAProblem& prob, std::string& err,
                                       const RCWAFieldRequest* fieldReq = nullptr,
                                       int device = 0) {
  using namespace detail;
  std::vector<RCWAResult> out;
  err.clear();
  const std::vector<scalar> ze = z_edges(prob);
  if (ze.size() < 2) {
    err = "cannot form z slabs (check zmesh / geometry)";
    return out;
  }
  if (fieldReq && fieldReq->wantsAnything() && ze.size() - 1 < 3) {
    err = "field images need an interior layer (split zmesh / geometry into 3 or more slabs)";
    return out;
  }
  if (prob.lambdas.empty()) {
    err = "no wavelengths";
    return out;
  }
  const bool mag = magnetic(prob);
  const std::vector<scalar> xg = plane_grid(prob, true), yg = plane_grid(prob, false);
  std::vector<scalar> zc, thick;
  for (size_t i = ze.size() - 1; i > 0; --i) {
    zc.push_back(0.5 * (ze[i] + ze[i - 1]));
    thick.push_back(ze[i] - ze[i - 1]);
  }
  if (zc.size() == 1) {   // one slab: solved as two copies of itself (R = 0, T = I)
    zc.push_back(zc[0]);
    thick.push_back(thick[0]);
  }
  const int L = (int)zc.size(), ncx = (int)xg.size(), ncy = (int)yg.size();
  const size_t cells = (size_t)(ncx - 1) * (ncy - 1);
  const int B = (int)prob.lambdas.size();
  std::vector<scalex> eps((size_t)B * L * cells), mu(mag ? (size_t)B * L * cells : 0);
  for (int b = 0; b < B; ++b)
    for (int s = 0; s < L; ++s)
      for (int iy = 0; iy + 1 < ncy; ++iy)
        for (int ix = 0; ix + 1 < ncx; ++ix) {
          const scalar x = 0.5 * (xg[ix] + xg[ix + 1]), y = 0.5 * (yg[iy] + yg[iy + 1]);
          const size_t o = ((size_t)b * L + s) * cells + (size_t)iy * (ncx - 1) + ix;
          eps[o] = eps_at(prob, x, y, zc[s], prob.lambdas[b]);
          if (mag) mu[o] = mu_at(prob, x, y, zc[s], prob.lambdas[b]);
        }
  for (int b = 0; b < B; ++b)
    for (int s : {0, L - 1}) {
      const size_t o = ((size_t)b * L + s) * cells;
      for (size_t c = 1; c < cells; ++c)
        if (eps[o + c] != eps[o] || (mag && mu[o + c] != mu[o])) {
          err = "the top and bottom slabs must be homogeneous";
          return out;
        }
    }
  scalex px, py;
  polarisation(prob, px, py);
  const double pol[4] = {px.real(), px.imag(), py.real(), py.imag()};
  const scalar th = prob.theta * kPi / 180.0, ph = prob.phi * kPi / 180.0;
  const scalar xm = 0.5 * (prob.xmin + prob.xmax), ym = 0.5 * (prob.ymin + prob.ymax);
  std::vector<scalar> tk = thick;
  tk.front() = tk.back() = 0.0;   // the half-spaces' thickness is never used
  orcwa_stack st{prob.nHx, prob.nHy, ncx, ncy, xg.data(), yg.data(), L, tk.data()};
  std::vector<orcwa_instance> inst(B);
  for (int b = 0; b < B; ++b) {
    const scalar lam = prob.lambdas[b], k0 = 2.0 * kPi / lam;
    const scalex ei = eps_at(prob, xm, ym, zc[0], lam);
    const scalex mi = mag ? mu_at(prob, xm, ym, zc[0], lam) : scalex(1.0);
    // n = sqrt(eps mu) of the incidence medium (rcwa/RCWADriver.cpp:289-296)
    const scalar ninc = std::max(std::sqrt(ei * mi).real(), 1e-6);
    orcwa_instance& in = inst[b];
    in = orcwa_instance{};
    in.lambda = lam;
    in.kx0 = k0 * ninc * std::sin(th) * std::cos(ph);
    in.ky0 = k0 * ninc * std::sin(th) * std::sin(ph);
    in.eps = reinterpret_cast<const double*>(eps.data() + (size_t)b * L * cells);
    in.mu = mag ? reinterpret_cast<const double*>(mu.data() + (size_t)b * L * cells) : nullptr;
    in.n_pol = 1;
    in.pol = pol;
  }
  const int h = (2 * prob.nHx + 1) * (2 * prob.nHy + 1);
  std::vector<double> ref((size_t)B * h), trn((size_t)B * h);
  std::vector<orcwa_result> res(B);
  for (int b = 0; b < B; ++b)
    res[b] = orcwa_result{ref.data() + (size_t)b * h, trn.data() + (size_t)b * h, nullptr, 0};
  orcwa_gpu_context* ctx = context(device);
  if (!ctx) {
    err = "no usable GPU";
    return out;
  }
  if (orcwa_gpu_solve_batch(ctx, &st, inst.data(), B, res.data()) != ORCWA_OK) {
    err = "GPU solve failed";
    return out;
  }
  for (int b = 0; b < B; ++b) {
    RCWAResult r;
    r.lambda = prob.lambdas[b];
    r.status = res[b].status;
    if (!(r.status & ORCWA_STATUS_GRAZING)) {
      r.REF_orders.assign(ref.begin() + (size_t)b * h, ref.begin() + (size_t)(b + 1) * h);
      r.TRN_orders.assign(trn.begin() + (size_t)b * h, trn.begin() + (size_t)(b + 1) * h);
    }
    r.R = r.T = 0.0;
    for (scalar v : r.REF_orders) r.R += v;
    for (scalar v : r.TRN_orders) r.T += v;
    r.A = 1.0 - r.R - r.T;
    out.push_back(std::move(r));
  }
  if (fieldReq && !fieldReq->path.empty()) {
    orcwa_field_request fr{(int)fieldReq->component, (int)fieldReq->slice, fieldReq->coord, 0, 0};
    int rows = 0, cols = 0, fst = 0;
    if (orcwa_gpu_field_image(ctx, &st, &inst[0], &fr, &rows, &cols, nullptr, nullptr) ==
        ORCWA_OK) {
      std::vector<scalex> f((size_t)rows * cols);
      if (orcwa_gpu_field_image(ctx, &st, &inst[0], &fr, &rows, &cols,
                                reinterpret_cast<double*>(f.data()), &fst) == ORCWA_OK)
        write_csv(fieldReq->path, rows, cols, f, fieldReq->opt);
      else
        err = "field image failed";
    } else {
      err = "field image failed";
    }
  }
  if (fieldReq && !fieldReq->devicePath.empty()) {
    // saveDeviceImage (rcwa/RCWASolver.cpp:1322-1487): 2000 cell-centred
    // samples per row; XZ/YZ rows run down the interior slabs
    const scalar Lx = xg.back() - xg.front(), Ly = yg.back() - yg.front();
    scalar zint = 0.0;
    for (int s = 1; s + 1 < L; ++s) zint += thick[s];
    const int n1 = 2000;
    scalar s1, l1, s2, l2;
    if (fieldReq->slice == sliceXY) {
      s1 = xg.front(); l1 = Lx; s2 = yg.front(); l2 = Ly;
    } else if (fieldReq->slice == sliceYZ) {
      s1 = yg.front(); l1 = Ly; s2 = 0.0; l2 = zint;
    } else {
      s1 = xg.front(); l1 = Lx; s2 = 0.0; l2 = zint;
    }
    const int n2 = std::max(1, (int)(n1 * l2 / l1));
    auto slab_at = [&](scalar z) {
      scalar z0 = 0.0;
      for (int j = 1; j + 1 < L; ++j) {
        if (z0 <= z && z < z0 + thick[j]) return j;
        z0 += thick[j];
      }
      return -1;
    };
    auto lookup = [&](int s, scalar x, scalar y) -> scalex {
      while (x < xg.front()) x += Lx;
      while (x > xg.back()) x -= Lx;
      if (y < yg.front() || y > yg.back()) return 0.0;
      const int ix = std::min(ncx - 2, (int)(std::lower_bound(xg.begin() + 1, xg.end(), x) -
                                             (xg.begin() + 1)));
      const int iy = std::min(ncy - 2, (int)(std::lower_bound(yg.begin() + 1, yg.end(), y) -
                                             (yg.begin() + 1)));
      return eps[(size_t)s * cells + (size_t)iy * (ncx - 1) + ix];
    };
    std::vector<scalex> img((size_t)n1 * n2);
    bool ok = true;
    for (int i = 0; i < n2 && ok; ++i) {
      const scalar a2 = s2 + l2 / n2 * (i + 0.5);
      const int s = fieldReq->slice == sliceXY ? slab_at(fieldReq->coord) : slab_at(a2);
      if (s < 0) {
        ok = false;
        break;
      }
      for (int j = 0; j < n1; ++j) {
        const scalar a1 = s1 + l1 / n1 * (j + 0.5);
        img[(size_t)i * n1 + j] = fieldReq->slice == sliceXY   ? lookup(s, a1, a2)
                                  : fieldReq->slice == sliceXZ ? lookup(s, a1, fieldReq->coord)
                                                               : lookup(s, fieldReq->coord, a1);
      }
    }
    if (ok)
      write_csv(fieldReq->devicePath, n2, n1, img, realPart);
    else
      err = "Device image requires at least one finite-thickness interior layer!";
  }
  return out;
}

}  // namespace openrcwa

#endif /* OPENRCWA_GPU_HPP */
