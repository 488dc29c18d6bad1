"""MI355X-native RCWA layer-stack engine (HIP kernels for gfx950 behind a C ABI).

Modules:
  engine     contexts, plans and batched solves over include/openrcwa_gpu.h
  solver     RCWASolver / Layer facade with the reference's method names
  ofd        batched host logic of the .ofd RCWA path
  workloads  seeded synthetic batches
  build      hipcc build of libopenrcwa_gpu.so
"""
__all__ = ["engine", "solver", "ofd", "workloads"]
