"""The surrogate's allocation-free inference entry point
(orcwa_mlp_predict_device_noalloc: activations in a static device buffer)
against orcwa_mlp_predict_device (activations in buffers the network
allocates), through the C ABI on device buffers, on two calls in a row (the
second smaller, so the buffer holds the first call's data)."""
import ctypes
from ctypes import POINTER, c_float, c_longlong, c_void_p

import numpy as np
import pytest

from openrcwa_amd import _lib, surrogate as sg

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [POINTER(c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [c_void_p]
        self.h.hipMemcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []

    def malloc(self, nbytes):
        p = c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.h.hipMemcpy(dst, a.ctypes.data_as(c_void_p), a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(a.ctypes.data_as(c_void_p), src, a.nbytes, 2) == 0


def test_noalloc_inference_equals_allocating_path(gpu_ctx):
    lib = _lib.load()
    fn = lib.orcwa_mlp_predict_device_noalloc
    fn.argtypes = [c_void_p, c_void_p, c_longlong, c_void_p, c_void_p]
    hip = Hip()
    net = sg.MLP(gpu_ctx, sg.N_FEAT, 512, 4, sg.N_OUT, 4, seed=9)
    bufs = []
    try:
        net.set_error_scale(np.linspace(0.5, 2.0, sg.N_OUT))
        rng = np.random.default_rng(3)
        big = 9000                          # more rows than one pass of the static buffer
        x = rng.uniform(-1, 1, size=(big, sg.N_FEAT)).astype(np.float32)
        dx = hip.malloc(x.nbytes)
        bufs.append(dx)
        hip.h2d(dx, x)
        outs = {}
        for name in ("ref", "noalloc"):
            dm = hip.malloc(big * sg.N_OUT * 4)
            de = hip.malloc(big * 4)
            bufs += [dm, de]
            outs[name] = (dm, de)
        for count in (big, 1500):           # the second call is smaller
            _lib.check(lib.orcwa_mlp_predict_device(net._h, ctypes.cast(dx, POINTER(c_float)),
                                                    count, outs["ref"][0], outs["ref"][1]),
                       "orcwa_mlp_predict_device")
            _lib.check(fn(net._h, dx, count, outs["noalloc"][0], outs["noalloc"][1]),
                       "orcwa_mlp_predict_device_noalloc")
            got = {}
            for name, (dm, de) in outs.items():
                m = np.zeros((count, sg.N_OUT), dtype=np.float32)
                e = np.zeros(count, dtype=np.float32)
                hip.d2h(m, dm)
                hip.d2h(e, de)
                got[name] = (m, e)
            # the same kernels on the same rows: identical results
            assert np.array_equal(got["ref"][0], got["noalloc"][0])
            assert np.array_equal(got["ref"][1], got["noalloc"][1])
            assert np.abs(got["ref"][0]).max() > 0
        # bad arguments are refused before any device work
        assert fn(net._h, dx, 0, outs["noalloc"][0], outs["noalloc"][1]) == 1
        assert fn(net._h, None, 10, outs["noalloc"][0], outs["noalloc"][1]) == 1
    finally:
        for b in bufs:
            hip.h.hipFree(b)
        net.close()
