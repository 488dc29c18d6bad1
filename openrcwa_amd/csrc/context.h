// Definition of the opaque context of include/openrcwa_gpu.h.
#pragma once

#include <hip/hip_runtime.h>

struct orcwa_gpu_context {
  int device = 0;
  hipStream_t stream = nullptr;
};
