"""ctypes wrappers of include/openrcwa_gpu_debug.h for the GPU component tests."""
import ctypes
from ctypes import POINTER, c_double, c_int

import numpy as np

from openrcwa_amd import _lib


def _lib_dbg():
    lib = _lib.load()
    lib.orcwa_dbg_zgemm.argtypes = [ctypes.c_void_p, c_int, c_int, c_int, POINTER(c_double),
                                    POINTER(c_double), c_int, c_int, c_int, POINTER(c_double),
                                    c_int, c_int, c_int, POINTER(c_double), POINTER(c_double),
                                    c_int, c_int]
    lib.orcwa_dbg_zgesv.argtypes = [ctypes.c_void_p, c_int, c_int, POINTER(c_double),
                                    POINTER(c_double), POINTER(c_double), POINTER(c_int), c_int]
    lib.orcwa_dbg_zgeev.argtypes = [ctypes.c_void_p, c_int, POINTER(c_double), POINTER(c_double),
                                    POINTER(c_double), POINTER(c_double), POINTER(c_int), c_int]
    lib.orcwa_dbg_set_eig_tolerance.argtypes = [ctypes.c_void_p, c_double]
    lib.orcwa_dbg_convolution.argtypes = [ctypes.c_void_p, c_int, c_int, c_int, c_int,
                                          POINTER(c_double), POINTER(c_double), POINTER(c_double),
                                          POINTER(c_double), POINTER(c_double), POINTER(c_double)]
    return lib


def _p(a):
    return a.ctypes.data_as(POINTER(c_double))


def fortran_batch(mats):
    """(batch, rows, cols) complex -> contiguous column-major per matrix."""
    return np.array(np.transpose(np.asarray(mats, dtype=complex), (0, 2, 1)), order="C", copy=True)


def from_fortran(buf, rows, cols):
    return np.transpose(buf.reshape(-1, cols, rows), (0, 2, 1))


def zgemm(ctx, A, B, C, alpha, beta, tA=0, cA=0, tB=0, cB=0):
    lib = _lib_dbg()
    A, B, C = (np.asarray(x, dtype=complex) for x in (A, B, C))
    batch = A.shape[0]
    ldA, ldB, ldC = A.shape[1], B.shape[1], C.shape[1]
    M, N = C.shape[1], C.shape[2]
    K = A.shape[1] if tA else A.shape[2]
    a = fortran_batch(A)
    b = fortran_batch(B)
    c = fortran_batch(C)
    al = np.array([alpha.real, alpha.imag])
    be = np.array([beta.real, beta.imag])
    rc = lib.orcwa_dbg_zgemm(ctx.handle, M, N, K, _p(al), _p(a.view(np.float64)), ldA, tA, cA,
                             _p(b.view(np.float64)), ldB, tB, cB, _p(be), _p(c.view(np.float64)),
                             ldC, batch)
    _lib.check(rc, "orcwa_dbg_zgemm")
    return from_fortran(c, M, N)


def zgesv(ctx, A, B):
    lib = _lib_dbg()
    A, B = np.asarray(A, dtype=complex), np.asarray(B, dtype=complex)
    batch, n, nrhs = B.shape
    a, b = fortran_batch(A), fortran_batch(B)
    x = np.zeros_like(b)
    info = np.zeros(batch, dtype=np.int32)
    rc = lib.orcwa_dbg_zgesv(ctx.handle, n, nrhs, _p(a.view(np.float64)), _p(b.view(np.float64)),
                             _p(x.view(np.float64)), info.ctypes.data_as(POINTER(c_int)), batch)
    _lib.check(rc, "orcwa_dbg_zgesv")
    return from_fortran(x, n, nrhs), info


def zgeev(ctx, A):
    lib = _lib_dbg()
    A = np.asarray(A, dtype=complex)
    batch, n, _ = A.shape
    a = fortran_batch(A)
    lam = np.zeros((batch, n), dtype=complex)
    w = np.zeros_like(a)
    resid = np.zeros(batch)
    info = np.zeros(batch, dtype=np.int32)
    rc = lib.orcwa_dbg_zgeev(ctx.handle, n, _p(a.view(np.float64)), _p(lam.view(np.float64)),
                             _p(w.view(np.float64)), _p(resid), info.ctypes.data_as(POINTER(c_int)),
                             batch)
    _lib.check(rc, "orcwa_dbg_zgeev")
    return lam, from_fortran(w, n, n), resid, info


def convolution(ctx, nX, nY, coord_x, coord_y, eps):
    lib = _lib_dbg()
    cx = np.ascontiguousarray(coord_x, dtype=np.float64)
    cy = np.ascontiguousarray(coord_y, dtype=np.float64)
    e = np.ascontiguousarray(eps, dtype=complex)
    h = (2 * nX + 1) * (2 * nY + 1)
    outs = [np.zeros((h, h), dtype=complex, order="F") for _ in range(3)]
    # Fortran-ordered outputs receive the column-major matrices directly
    rc = lib.orcwa_dbg_convolution(ctx.handle, nX, nY, cx.size, cy.size, _p(cx), _p(cy),
                                   _p(e.view(np.float64)), *[_p(o) for o in outs])
    _lib.check(rc, "orcwa_dbg_convolution")
    return outs


def set_eig_tolerance(ctx, tol):
    """Residual acceptance of later eigen solves on ctx (default 1e-8; a
    negative value rejects every solve)."""
    _lib.check(_lib_dbg().orcwa_dbg_set_eig_tolerance(ctx.handle, float(tol)),
               "orcwa_dbg_set_eig_tolerance")
