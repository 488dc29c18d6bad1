// Definition of the opaque context of include/openrcwa_gpu.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <map>
#include <string>
#include <vector>

struct orcwa_gpu_context {
  int device = 0;
  hipStream_t stream = nullptr;
  // Device buffers of destroyed plans, kept for later plans on this context:
  // hipFree waits for the whole device, so releasing a plan's workspace while
  // other contexts' streams run would stall them all.
  struct Cached {
    void* ptr;
    size_t bytes;
  };
  std::vector<Cached> cache;
  size_t cached_bytes = 0;
  // workspace budget of this context's plans in bytes (0: 70 % of the free
  // device memory); set by orcwa_gpu_set_workspace_limit
  size_t workspace_limit = 0;
  // plans created on this context and not yet destroyed: the context refuses
  // to be destroyed while any is live
  int live_plans = 0;
  // residual acceptance of the eigen solves (orcwa_dbg_set_eig_tolerance;
  // 1e-8 as rcwa/MKLEigenSolver.hpp:75-79)
  double eig_tol = 1e-8;
  // precision of the cascade / star-product GEMMs of later plans and star
  // products (orcwa_gpu_set_gemm_precision; GemmPrec values)
  int gemm_prec = 0;
  // autotuned kernel variants per layer-matrix size n (orcwa_plan_autotune,
  // orcwa_gpu_set_tuning_file): QrVariant and GemmTile values
  struct Tune {
    int qr = 0, tile = 0;
    double ms = 0.0;
  };
  std::map<int, Tune> tuning;
  std::string tuning_file;       // appended to by orcwa_plan_autotune when writable
  bool tuning_append = false;
};

namespace orcwa {

// A device buffer of at least `bytes` for a plan on this context: the
// smallest cached buffer that fits, else a new hipMalloc (after the cache is
// freed, if the first attempt fails).  *cap receives the buffer's size.
hipError_t ctx_alloc(orcwa_gpu_context* c, size_t bytes, void** out, size_t* cap);
// Give a buffer back to the context (it stays allocated until the context is
// destroyed, or the cache is trimmed).
void ctx_release(orcwa_gpu_context* c, void* ptr, size_t cap);
// Free every cached buffer.
void ctx_trim(orcwa_gpu_context* c);

}  // namespace orcwa
