"""MI355X-native RCWA layer-stack engine (HIP kernels for gfx950 behind a C ABI).

Modules:
  engine     contexts, plans and batched solves over include/openrcwa_gpu.h
  ofd        batched host logic of the .ofd RCWA path
  orcwa      the .orcwa input path (parser, driver, orcwa_rcwa CSV command)
  workloads  seeded synthetic batches
  build      hipcc build of libopenrcwa_gpu.so
"""
__all__ = ["engine", "ofd", "orcwa", "workloads"]
