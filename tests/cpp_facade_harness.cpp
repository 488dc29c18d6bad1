// A C++ caller of the header-only facade include/openrcwa_gpu.hpp, linked
// against libopenrcwa_gpu.so (built with the library by
// openrcwa_amd/build.py).  tests/test_gpu_cpp_facade.py runs it and checks its
// output against the FP64 oracle and the Python .orcwa driver.
//
//   cpp_facade_harness 0   RCWASolver: the 1D grating of data/sample/grating.ofd
//                          (N = 5, 461 nm), TE and TM per order
//   cpp_facade_harness 1   RCWASolver: a 2D stack whose layers have different
//                          interior cell edges, oblique and conical incidence
//   cpp_facade_harness 2   runRCWA: a cylinder and a box in a magnetic,
//                          conductive slab, 3 wavelengths, field and device CSVs
//   cpp_facade_harness 3   RCWASolver: a layer with mismatching edges is dropped
//
// Output lines: "<tag> <index> <values...>" with %.17g numbers.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../include/openrcwa_gpu.hpp"

using namespace openrcwa;

static void print_vec(const char* tag, int idx, const std::vector<scalar>& v) {
  std::printf("%s %d", tag, idx);
  for (scalar x : v) std::printf(" %.17g", x);
  std::printf("\n");
}

static int grating1d() {
  // um (the reference's .ofd is in m: scale-free); f = 6.5e14 Hz
  const scalar P = 0.3, lam = 2.99792458e8 / 6.5e14 * 1e6;
  RCWASolver s(5, 0);
  // every layer on the same edges (addLayer compares the first and last two)
  s.addLayer(Layer({0.0, 0.5 * P, P}, {0.0, 1.0}, {scalex(1.0), scalex(1.0)}));
  s.addLayer(Layer({0.0, 0.5 * P, P}, {0.0, 1.0}, {scalex(4.0), scalex(1.0)}));
  s.addLayer(Layer({0.0, 0.5 * P, P}, {0.0, 1.0}, {scalex(2.25), scalex(2.25)}));
  s.disablePML();
  s.setBlochWavevector(0.0, 0.0);
  s.solve(lam, {0, 1, 2}, {0.0, 0.2, 0.0});
  if (!s.isScatterMatrixSolved()) return 1;
  const scalar k0 = 2 * 3.1415926535897932 / lam;
  const scalex pols[2][2] = {{0.0, 1.0}, {1.0, 0.0}};   // TE (y), TM (x) at normal incidence
  for (int p = 0; p < 2; ++p) {
    std::vector<scalex> c;
    std::vector<scalar> REF, TRN;
    s.generateHorizontalPlaneWave(pols[p][0], pols[p][1], 0, c);
    s.scatterPlaneWave(c, 0, 2, k0, REF, TRN);
    print_vec("REF", p, REF);
    print_vec("TRN", p, TRN);
  }
  std::printf("STATUS 0 %d\n", s.status());
  return 0;
}

static int stack2d() {
  const scalar Px = 0.4, Py = 0.35, lam = 0.55, th = 20.0 * 3.1415926535897932 / 180.0,
               ph = 30.0 * 3.1415926535897932 / 180.0;
  RCWASolver s(3, 2);
  s.addLayer(Layer({0.0, 0.1, 0.3, Px}, {0.0, 0.05, 0.3, Py}, std::vector<scalex>(9, 1.0)));
  // rectangle [0.1, 0.22] x [0.05, 0.2] of eps 6 + 0.1i
  std::vector<scalex> e1(16, 1.0);
  const std::vector<scalar> x1 = {0.0, 0.1, 0.22, 0.3, Px}, y1 = {0.0, 0.05, 0.2, 0.3, Py};
  e1[1 * 4 + 1] = scalex(6.0, 0.1);
  s.addLayer(Layer(x1, y1, e1));
  // L shape of eps 3 on other interior edges
  std::vector<scalex> e2(9, 2.0);
  const std::vector<scalar> x2 = {0.0, 0.1, 0.3, Px}, y2 = {0.0, 0.05, 0.3, Py};
  e2[0] = scalex(3.0);
  e2[1] = scalex(3.0);
  e2[3] = scalex(3.0);
  s.addLayer(Layer(x2, y2, e2));
  s.addLayer(Layer({0.0, 0.1, 0.3, Px}, {0.0, 0.05, 0.3, Py}, std::vector<scalex>(9, 2.25)));
  s.disablePML();
  const scalar k0 = 2 * 3.1415926535897932 / lam;
  s.setBlochWavevector(k0 * std::sin(th) * std::cos(ph), k0 * std::sin(th) * std::sin(ph));
  s.solve(lam, {0, 1, 2, 1, 3}, {0.0, 0.12, 0.08, 0.05, 0.0});
  if (!s.isScatterMatrixSolved()) return 1;
  const scalex pols[2][2] = {{-std::sin(ph), std::cos(ph)},
                             {std::cos(th) * std::cos(ph), std::cos(th) * std::sin(ph)}};
  for (int p = 0; p < 2; ++p) {
    std::vector<scalex> c;
    std::vector<scalar> REF, TRN;
    s.generateHorizontalPlaneWave(pols[p][0], pols[p][1], 0, c);
    s.scatterPlaneWave(c, 0, 3, k0, REF, TRN);
    print_vec("REF", p, REF);
    print_vec("TRN", p, TRN);
  }
  std::printf("STATUS 0 %d\n", s.status());
  return 0;
}

static int driver(const char* fieldcsv, const char* devicecsv) {
  RCWAProblem p;
  p.title = "facade";
  p.xmin = 0.0; p.xmax = 0.5; p.ymin = 0.0; p.ymax = 0.5; p.zmin = 0.0; p.zmax = 0.6;
  p.nHx = 3; p.nHy = 3;
  p.materialEps = {scalex(1.0), scalex(1.0), scalex(2.25), scalex(5.0, 0.05), scalex(3.0)};
  p.materialMu = {scalex(1.0), scalex(1.0), scalex(1.0), scalex(1.0), scalex(1.5, 0.02)};
  p.materialSigma = {0.0, 0.0, 0.0, 0.0, 2.0e4};
  p.materialMagSigma = {0.0, 0.0, 0.0, 0.0, 0.0};
  p.materialDispersion.resize(5);
  RCWABox sub;   // substrate slab z in [0, 0.2]
  sub.material = 2; sub.x0 = 0.0; sub.x1 = 0.5; sub.y0 = 0.0; sub.y1 = 0.5; sub.z0 = 0.0; sub.z1 = 0.2;
  RCWABox mag;   // magnetic, conductive slab z in [0.2, 0.35]
  mag.material = 4; mag.x0 = 0.0; mag.x1 = 0.5; mag.y0 = 0.0; mag.y1 = 0.5; mag.z0 = 0.2; mag.z1 = 0.35;
  RCWABox cyl;   // cylinder in the magnetic slab
  cyl.material = 3; cyl.shape = RCWA_SHAPE_CYL_Z;
  cyl.x0 = 0.1; cyl.x1 = 0.4; cyl.y0 = 0.05; cyl.y1 = 0.35; cyl.z0 = 0.2; cyl.z1 = 0.35;
  RCWABox box;   // box above it, z in [0.35, 0.45]
  box.material = 3; box.x0 = 0.0; box.x1 = 0.2; box.y0 = 0.1; box.y1 = 0.5; box.z0 = 0.35; box.z1 = 0.45;
  p.boxes = {sub, mag, cyl, box};
  p.theta = 15.0; p.phi = 25.0; p.pol = 4;
  p.lambdas = {0.45, 0.6, 0.75};
  RCWAFieldRequest fr;
  fr.component = Ex;
  fr.slice = sliceXZ;
  fr.coord = 0.2;
  fr.opt = realPart;
  fr.path = fieldcsv;
  fr.devicePath = devicecsv;
  std::string err;
  std::vector<RCWAResult> res = runRCWA(p, err, &fr);
  if (!err.empty()) {
    std::printf("ERROR %s\n", err.c_str());
    return 1;
  }
  for (size_t b = 0; b < res.size(); ++b) {
    std::printf("TOT %d %.17g %.17g %.17g %.17g\n", (int)b, res[b].lambda, res[b].R, res[b].T,
                res[b].A);
    print_vec("REF", (int)b, res[b].REF_orders);
    print_vec("TRN", (int)b, res[b].TRN_orders);
    std::printf("STATUS %d %d\n", (int)b, res[b].status);
  }
  return 0;
}

static int mismatch() {
  RCWASolver s(1, 0);
  s.addLayer(Layer({0.0, 0.1, 0.3}, {0.0, 1.0}, {scalex(1.0), scalex(1.0)}));
  s.addLayer(Layer({0.0, 0.15, 0.3}, {0.0, 1.0}, {scalex(2.0), scalex(1.0)}));   // l1 differs
  s.addLayer(Layer({0.0, 0.1, 0.3}, {0.0, 1.0}, {scalex(3.0), scalex(1.0)}));   // matches
  std::printf("NTYPES 0 %d\n", s.nLayerTypes());
  return 0;
}

int main(int argc, char** argv) {
  const int sc = argc > 1 ? std::atoi(argv[1]) : 0;
  switch (sc) {
    case 0: return grating1d();
    case 1: return stack2d();
    case 2: return driver(argc > 2 ? argv[2] : "", argc > 3 ? argv[3] : "");
    case 3: return mismatch();
    default: return 2;
  }
}
