"""MI355X-native RCWA layer-stack engine (HIP kernels for gfx950 behind a C ABI).

Modules:
  engine     contexts, plans and batched solves over include/openrcwa_gpu.h
  ofd        batched host logic of the .ofd RCWA path
  orcwa      the .orcwa input path (parser, driver, orcwa_rcwa CSV command)
  surrogate  the learned surrogate (MLP ensemble trained on the GPU solver's
             output, with plain-solver fallback) over include/openrcwa_surrogate.h
  workloads  seeded synthetic batches
  build      hipcc build of libopenrcwa_gpu.so

The reference's C++ API (RCWASolver, Layer, runRCWA) is the header-only
include/openrcwa_gpu.hpp over the same library.
"""
__all__ = ["engine", "ofd", "orcwa", "surrogate", "workloads"]
