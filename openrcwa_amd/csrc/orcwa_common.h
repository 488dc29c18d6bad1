// Common device/host helpers for the OpenRCWA MI355X engine.
//
// Complex numbers are interleaved double pairs (re, im), the layout of
// std::complex<double> in the reference (rcwa/defns.h:7-8).  All matrices are
// column-major unless a kernel states otherwise.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

struct cd {
  double re, im;
};

__host__ __device__ inline cd mk(double r, double i) { cd z; z.re = r; z.im = i; return z; }
__host__ __device__ inline cd operator+(cd a, cd b) { return mk(a.re + b.re, a.im + b.im); }
__host__ __device__ inline cd operator-(cd a, cd b) { return mk(a.re - b.re, a.im - b.im); }
__host__ __device__ inline cd operator-(cd a) { return mk(-a.re, -a.im); }
__host__ __device__ inline cd operator*(cd a, cd b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
__host__ __device__ inline cd operator*(double s, cd a) { return mk(s * a.re, s * a.im); }
__host__ __device__ inline cd operator*(cd a, double s) { return mk(s * a.re, s * a.im); }
__host__ __device__ inline cd& operator+=(cd& a, cd b) { a.re += b.re; a.im += b.im; return a; }
__host__ __device__ inline cd& operator-=(cd& a, cd b) { a.re -= b.re; a.im -= b.im; return a; }
__host__ __device__ inline cd conj(cd a) { return mk(a.re, -a.im); }
__host__ __device__ inline double norm2(cd a) { return a.re * a.re + a.im * a.im; }
__host__ __device__ inline double cabs1(cd a) { return fabs(a.re) + fabs(a.im); }
__host__ __device__ inline double cabs(cd a) { return hypot(a.re, a.im); }
// a * conj(b)
__host__ __device__ inline cd mulc(cd a, cd b) {
  return mk(a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im);
}
// conj(a) * b
__host__ __device__ inline cd cmul(cd a, cd b) {
  return mk(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}
__host__ __device__ inline cd cdiv(cd a, cd b) {
  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return mk((a.re + a.im * r) / d, (a.im - a.re * r) / d);
  } else {
    double r = b.re / b.im, d = b.re * r + b.im;
    return mk((a.re * r + a.im) / d, (a.im * r - a.re) / d);
  }
}
__host__ __device__ inline cd cinv(cd b) { return cdiv(mk(1.0, 0.0), b); }

// Principal complex square root (branch cut on the negative real axis,
// Re >= 0), as std::sqrt(std::complex<double>).
__host__ __device__ inline cd csqrt_principal(cd z) {
  if (z.re == 0.0 && z.im == 0.0) return mk(0.0, z.im);
  double t = sqrt(0.5 * (fabs(z.re) + hypot(z.re, z.im)));
  if (z.re >= 0.0) return mk(t, z.im / (2.0 * t));
  return mk(fabs(z.im) / (2.0 * t), copysign(t, z.im));
}

// GammaSqrtOp (rcwa/WaveEquationSolver.cpp:8-22): principal root, negated
// if Im < 0, negated again if |Im| < |Re| and Re < 0.
__host__ __device__ inline cd gamma_sqrt(cd g) {
  cd v = csqrt_principal(g);
  if (v.im < 0) v = -v;
  if (fabs(v.im) < fabs(v.re) && v.re < 0) v = -v;
  return v;
}

__host__ __device__ inline cd cexp(cd z) {
  double e = exp(z.re);
  double s, c;
  sincos(z.im, &s, &c);
  return mk(e * c, e * s);
}

// wavefront (64-lane) helpers
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}
// orders LDS accesses between the lanes of one wavefront
__device__ inline void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

#define ORCWA_HIP_CHECK(expr)                                                       \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      fprintf(stderr, "[openrcwa] HIP error %s at %s:%d: %s\n", hipGetErrorName(_e), \
              __FILE__, __LINE__, #expr);                                           \
      return ORCWA_ERR_HIP;                                                         \
    }                                                                               \
  } while (0)
