"""ctypes binding of ``include/openrcwa_gpu.h`` (the engine's C ABI).

There is no CPU fallback: if ``libopenrcwa_gpu.so`` is missing or no GPU is
present, the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_void_p

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopenrcwa_gpu.so")

ORCWA_OK = 0
ERRORS = {1: "invalid argument", 2: "HIP runtime error / no GPU", 3: "out of device memory",
          4: "unsupported input"}

STATUS_EIG_RESIDUAL = 1
STATUS_GRAZING = 2
STATUS_SINGULAR = 4
STATUS_NOCONV = 8
STATUS_EIG_PERTURBED = 16
STATUS_ILLCOND = 32

# orcwa_gemm_precision
GEMM_PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16x6": 3}


class OrcwaStack(ctypes.Structure):
    _fields_ = [("nx_harm", c_int), ("ny_harm", c_int), ("n_cx", c_int), ("n_cy", c_int),
                ("coord_x", POINTER(c_double)), ("coord_y", POINTER(c_double)),
                ("n_layers", c_int), ("thickness", POINTER(c_double))]


class OrcwaInstance(ctypes.Structure):
    _fields_ = [("lam", c_double), ("kx0", c_double), ("ky0", c_double),
                ("eps", POINTER(c_double)), ("n_pol", c_int), ("pol", POINTER(c_double)),
                ("mu", POINTER(c_double)), ("coord_x", POINTER(c_double)),
                ("coord_y", POINTER(c_double)), ("thickness", POINTER(c_double))]


class OrcwaResult(ctypes.Structure):
    _fields_ = [("ref", POINTER(c_double)), ("trn", POINTER(c_double)),
                ("amp", POINTER(c_double)), ("status", c_int)]


class OrcwaFieldRequest(ctypes.Structure):
    _fields_ = [("component", c_int), ("slice", c_int), ("coord", c_double),
                ("nres1", c_int), ("pol", c_int)]


class OrcwaZStack(ctypes.Structure):
    _fields_ = [("nx_harm", c_int), ("ny_harm", c_int), ("n_cx", c_int), ("n_cy", c_int),
                ("n_slices", c_int), ("lam", c_double), ("kx0", c_double), ("ky0", c_double),
                ("z0", c_double), ("z1", c_double), ("coord_x", POINTER(c_double)),
                ("coord_y", POINTER(c_double)), ("eps", POINTER(c_double))]


class OrcwaZResult(ctypes.Structure):
    _fields_ = [(k, POINTER(c_double)) for k in ("Rud", "Tuu", "Rdu", "Tdd", "W0", "V0", "W1",
                                                 "V1", "gam0", "gam1")] + [("status", c_int)]


SAMPLER_FN = ctypes.CFUNCTYPE(c_int, c_void_p, c_double, POINTER(c_double), POINTER(c_double),
                              POINTER(c_double))


class OrcwaZSampler(ctypes.Structure):
    _fields_ = [("nx_harm", c_int), ("ny_harm", c_int), ("n_cx", c_int), ("n_cy", c_int),
                ("lam", c_double), ("kx0", c_double), ("ky0", c_double), ("z0", c_double),
                ("z1", c_double), ("sample", SAMPLER_FN), ("user", c_void_p)]


FIELD_COMPONENTS = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5}
SLICES = {"xy": 0, "xz": 1, "yz": 2}


# every symbol include/openrcwa_gpu.h declares
EXPORTED = ["orcwa_gpu_create", "orcwa_gpu_destroy", "orcwa_gpu_set_workspace_limit",
            "orcwa_gpu_set_gemm_precision",
            "orcwa_gpu_solve_batch",
            "orcwa_plan_create", "orcwa_plan_execute", "orcwa_plan_sync",
            "orcwa_plan_results", "orcwa_plan_destroy", "orcwa_plan_phase_ms",
            "orcwa_gpu_slab_smatrix", "orcwa_gpu_redheffer_star", "orcwa_gpu_field_image",
            "orcwa_gpu_rcwa_integrate", "orcwa_gpu_redheffer_integrate", "rcwa_run",
            "orcwa_graph_create", "orcwa_graph_launch", "orcwa_graph_sync", "orcwa_graph_reruns",
            "orcwa_graph_destroy", "orcwa_gpu_set_tuning_file", "orcwa_plan_autotune"]

# every symbol include/openrcwa_surrogate.h declares
EXPORTED_SURROGATE = ["orcwa_mlp_create", "orcwa_mlp_destroy", "orcwa_mlp_param_count",
                      "orcwa_mlp_get_params", "orcwa_mlp_set_params", "orcwa_mlp_set_data",
                      "orcwa_mlp_train", "orcwa_mlp_gradient", "orcwa_mlp_set_precision",
                      "orcwa_mlp_set_error_scale",
                      "orcwa_mlp_forward", "orcwa_mlp_predict", "orcwa_mlp_predict_device",
                      "orcwa_mlp_last_ms"]

MLP_FP32 = 0
MLP_BF16 = 1


class OrcwaMlpTrainOpts(ctypes.Structure):
    _fields_ = [("epochs", c_int), ("batch", c_int), ("lr_max", ctypes.c_float),
                ("lr_min", ctypes.c_float), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float), ("seed", ctypes.c_ulonglong), ("use_graph", c_int)]

_lib = None


class EngineError(RuntimeError):
    pass


def load(path: str | None = None):
    """Load the native library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise EngineError(f"{p} is missing: build it with `python -m openrcwa_amd.build`")
    lib = ctypes.CDLL(p)
    lib.orcwa_gpu_create.argtypes = [POINTER(c_void_p), c_int]
    lib.orcwa_gpu_destroy.argtypes = [c_void_p]
    lib.orcwa_gpu_set_workspace_limit.argtypes = [c_void_p, ctypes.c_size_t]
    lib.orcwa_gpu_set_gemm_precision.argtypes = [c_void_p, c_int]
    lib.orcwa_gpu_solve_batch.argtypes = [c_void_p, POINTER(OrcwaStack), POINTER(OrcwaInstance),
                                          c_int, POINTER(OrcwaResult)]
    lib.orcwa_plan_create.argtypes = [c_void_p, POINTER(OrcwaStack), POINTER(OrcwaInstance),
                                      c_int, POINTER(c_void_p)]
    lib.orcwa_plan_execute.argtypes = [c_void_p]
    lib.orcwa_plan_sync.argtypes = [c_void_p]
    lib.orcwa_plan_results.argtypes = [c_void_p, POINTER(OrcwaResult)]
    lib.orcwa_plan_destroy.argtypes = [c_void_p]
    lib.orcwa_plan_phase_ms.argtypes = [c_void_p, c_int]
    lib.orcwa_plan_phase_ms.restype = c_double
    lib.orcwa_gpu_set_tuning_file.argtypes = [c_void_p, ctypes.c_char_p, c_int]
    lib.orcwa_plan_autotune.argtypes = [c_void_p, POINTER(c_int), POINTER(c_int)]
    lib.orcwa_graph_create.argtypes = [c_void_p, POINTER(c_void_p), c_int, c_int,
                                       POINTER(c_void_p)]
    lib.orcwa_graph_launch.argtypes = [c_void_p]
    lib.orcwa_graph_sync.argtypes = [c_void_p]
    lib.orcwa_graph_reruns.argtypes = [c_void_p]
    lib.orcwa_graph_destroy.argtypes = [c_void_p]
    lib.orcwa_gpu_rcwa_integrate.argtypes = [c_void_p, POINTER(OrcwaZStack),
                                             POINTER(OrcwaZResult)]
    lib.orcwa_gpu_redheffer_integrate.argtypes = [c_void_p, POINTER(OrcwaZSampler), c_double,
                                                  c_int, POINTER(OrcwaZResult), POINTER(c_int)]
    lib.orcwa_gpu_field_image.argtypes = [c_void_p, POINTER(OrcwaStack), POINTER(OrcwaInstance),
                                          POINTER(OrcwaFieldRequest), POINTER(c_int),
                                          POINTER(c_int), POINTER(c_double), POINTER(c_int)]
    c_fp = POINTER(ctypes.c_float)
    lib.orcwa_mlp_create.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int,
                                     ctypes.c_ulonglong, POINTER(c_void_p)]
    lib.orcwa_mlp_destroy.argtypes = [c_void_p]
    lib.orcwa_mlp_param_count.argtypes = [c_void_p]
    lib.orcwa_mlp_param_count.restype = ctypes.c_longlong
    lib.orcwa_mlp_get_params.argtypes = [c_void_p, c_fp]
    lib.orcwa_mlp_set_params.argtypes = [c_void_p, c_fp]
    lib.orcwa_mlp_set_data.argtypes = [c_void_p, c_fp, c_fp, ctypes.c_longlong]
    lib.orcwa_mlp_train.argtypes = [c_void_p, POINTER(OrcwaMlpTrainOpts), c_fp]
    lib.orcwa_mlp_gradient.argtypes = [c_void_p, c_fp, c_fp, c_int, c_fp, c_fp]
    lib.orcwa_mlp_set_precision.argtypes = [c_void_p, c_int]
    lib.orcwa_mlp_set_error_scale.argtypes = [c_void_p, c_fp]
    lib.orcwa_mlp_forward.argtypes = [c_void_p, c_fp, ctypes.c_longlong, c_fp]
    lib.orcwa_mlp_predict.argtypes = [c_void_p, c_fp, ctypes.c_longlong, c_fp, c_fp]
    lib.orcwa_mlp_predict_device.argtypes = [c_void_p, c_void_p, ctypes.c_longlong, c_void_p,
                                             c_void_p]
    lib.orcwa_mlp_last_ms.argtypes = [c_void_p]
    lib.orcwa_mlp_last_ms.restype = c_double
    _lib = lib
    return lib


def check(rc: int, what: str):
    if rc != ORCWA_OK:
        raise EngineError(f"{what} failed: {ERRORS.get(rc, rc)}")
