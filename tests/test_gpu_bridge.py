"""The .ofd entry point rcwa_run(FILE*) (sol/rcwa_bridge.cpp:145) through the
library's exported globals: the CSV files it writes are compared with the
oracle's .ofd path (per-frequency totals and per-order values at 1e-9) and
with the reference's byte formats (headers, "%.8e" fields, TE block then TM
block, only orders propagating on either side; sol/rcwa_bridge.cpp:211-230,
322-334)."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import _lib

pytestmark = pytest.mark.gpu
C0 = 2.99792458e8


class RcwaLayer(ctypes.Structure):          # include/orcwa_rcwa.h:27-35
    _fields_ = [(k, ctypes.c_double) for k in
                ("eps1", "eps2", "eps1i", "eps2i", "fill", "fill_y", "thickness")]


class RcwaPole(ctypes.Structure):           # include/orcwa_rcwa.h:45-52
    _fields_ = [("layer", ctypes.c_int), ("side", ctypes.c_int), ("einf", ctypes.c_double),
                ("ae", ctypes.c_double), ("be", ctypes.c_double), ("ce", ctypes.c_double)]


FMT = r"[-+]?\d\.\d{8}e[-+]\d{2}"


def run_bridge(tmp_path, monkeypatch, nX, period, layers, freqs, theta=0.0, phi=0.0, nY=0,
               period_y=0.0, poles=()):
    lib = _lib.load()
    keep = []
    ctypes.c_int.in_dll(lib, "NRcwaHarmonics").value = nX
    ctypes.c_double.in_dll(lib, "RcwaPeriod").value = period
    ctypes.c_int.in_dll(lib, "NRcwaHarmonicsY").value = nY
    # the input module sets the y period of a 1D grating to the x period
    # (sol/input_data.c:580-584)
    ctypes.c_double.in_dll(lib, "RcwaPeriodY").value = period_y if nY > 0 else period
    arr = (RcwaLayer * len(layers))()
    for i, l in enumerate(layers):
        arr[i] = RcwaLayer(l["eps1"], l.get("eps2", l["eps1"]), l.get("eps1i", 0.0),
                           l.get("eps2i", 0.0), l.get("fill", 0.5), l.get("fill_y", 1.0),
                           l.get("thickness", 0.0))
    keep.append(arr)
    ctypes.c_int.in_dll(lib, "NRcwaLayer").value = len(layers)
    ctypes.c_void_p.in_dll(lib, "RcwaLayer").value = ctypes.addressof(arr)
    parr = (RcwaPole * max(1, len(poles)))()
    for i, p in enumerate(poles):
        parr[i] = RcwaPole(*p)
    keep.append(parr)
    ctypes.c_int.in_dll(lib, "NRcwaPole").value = len(poles)
    ctypes.c_void_p.in_dll(lib, "RcwaPole").value = ctypes.addressof(parr)
    ctypes.c_double.in_dll(lib, "RcwaTheta").value = theta
    ctypes.c_double.in_dll(lib, "RcwaPhi").value = phi
    farr = (ctypes.c_double * len(freqs))(*freqs)
    keep.append(farr)
    ctypes.c_int.in_dll(lib, "NFreq1").value = len(freqs)
    ctypes.c_void_p.in_dll(lib, "Freq1").value = ctypes.addressof(farr)
    (ctypes.c_char * 256).in_dll(lib, "Title").value = b"bridge test"
    monkeypatch.chdir(tmp_path)
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    fp = libc.fopen(str(tmp_path / "log.txt").encode(), b"w")
    lib.rcwa_run.argtypes = [ctypes.c_void_p]
    try:
        rc = lib.rcwa_run(fp)
    finally:
        libc.fclose(fp)
    eff = (tmp_path / "rcwa_efficiency.csv").read_text().splitlines()
    ords = (tmp_path / "rcwa_orders.csv").read_text().splitlines()
    log = (tmp_path / "log.txt").read_text()
    return rc, eff, ords, log


def oracle_poles(poles):
    out = {}
    for layer, side, einf, ae, be, ce in poles:
        out.setdefault((layer, side), []).append((einf, ae, be, ce))
    return out or None


def check(tmp_path, monkeypatch, nX, period, layers, freqs, theta=0.0, phi=0.0, nY=0,
          period_y=0.0, poles=()):
    rc, eff, ords, log = run_bridge(tmp_path, monkeypatch, nX, period, layers, freqs, theta, phi,
                                    nY, period_y, poles)
    assert rc == 0, log
    assert eff[0] == "frequency[Hz],lambda[m],R_TE,T_TE,R_TM,T_TM"
    assert ords[0] == "frequency[Hz],lambda[m],pol,m,n,R,T"
    assert len(eff) == 1 + len(freqs)
    row_re = re.compile(",".join([FMT] * 6) + "$")
    ord_re = re.compile(f"{FMT},{FMT},(TE|TM),-?\\d+,-?\\d+,{FMT},{FMT}$")
    orows = [r.split(",") for r in ords[1:]]
    assert all(ord_re.match(r) for r in ords[1:])
    py = period_y if nY > 0 else None
    worst = 0.0
    for f, freq in enumerate(freqs):
        assert row_re.match(eff[f + 1]), eff[f + 1]
        vals = [float(x) for x in eff[f + 1].split(",")]
        ref = o.ofd_solve_frequency(nX, period, layers, freq, theta, phi, nY, py,
                                    oracle_poles(poles))
        assert vals[0] == float(f"{freq:.8e}")
        assert abs(vals[1] - C0 / freq) <= 1e-8 * C0 / freq
        for p in range(2):
            worst = max(worst, abs(vals[2 + 2 * p] - ref["R"][p]), abs(vals[3 + 2 * p] - ref["T"][p]))
        # per-order block of this frequency: TE rows, then TM rows
        mine = [r for r in orows if float(r[0]) == vals[0]]
        pols = [r[2] for r in mine]
        assert pols == sorted(pols, key=lambda s: s != "TE")
        k0 = 2 * math.pi * freq / C0
        n_inc = max(math.sqrt(layers[0]["eps1"]), 1e-6)
        kx0 = k0 * n_inc * math.sin(math.radians(theta)) * math.cos(math.radians(phi))
        ky0 = k0 * n_inc * math.sin(math.radians(theta)) * math.sin(math.radians(phi))
        nyo = 2 * nY + 1
        expect = []
        for p, lab in enumerate(("TE", "TM")):
            for k in range((2 * nX + 1) * nyo):
                m, n = k // nyo - nX, k % nyo - nY
                kx = kx0 + 2 * math.pi * m / period
                ky = ky0 + (2 * math.pi * n / period_y if nY > 0 else 0.0)
                kt = kx * kx + ky * ky
                pr = kt < k0 * k0 * layers[0]["eps1"]
                pt = kt < k0 * k0 * layers[-1]["eps1"]
                if pr or pt:
                    expect.append((lab, m, n, ref["REF"][p][k] if pr else 0.0,
                                   ref["TRN"][p][k] if pt else 0.0))
        assert [(r[2], int(r[3]), int(r[4])) for r in mine] == [e[:3] for e in expect]
        for r, e in zip(mine, expect):
            worst = max(worst, abs(float(r[5]) - e[3]), abs(float(r[6]) - e[4]))
    assert worst < 1e-9, worst
    return log


def test_bridge_grating_sample(tmp_path, monkeypatch):
    """data/sample/grating.ofd with rcwa = 5 3.0e-7 and 5 frequencies."""
    layers = [dict(eps1=1.0, eps2=1.0, fill=0.5, thickness=0.0),
              dict(eps1=4.0, eps2=1.0, fill=0.5, thickness=2.0e-7),
              dict(eps1=2.25, eps2=2.25, fill=0.5, thickness=0.0)]
    log = check(tmp_path, monkeypatch, 5, 3.0e-7, layers, list(np.linspace(4.3e14, 7.5e14, 5)))
    assert "RCWA : N=5 (order 11), period=3.0000e-07[m], 3 layers, 5 frequencies" in log
    assert "output : rcwa_efficiency.csv, rcwa_orders.csv" in log


def test_bridge_2d_oblique_dispersive(tmp_path, monkeypatch):
    layers = [dict(eps1=1.0, eps2=1.0, fill=0.5, fill_y=0.5),
              dict(eps1=2.25, eps2=1.0, eps1i=0.02, fill=0.6, fill_y=0.4, thickness=1.5e-7),
              dict(eps1=1.0, eps2=1.0, fill=0.3, fill_y=0.7, thickness=5e-8),
              dict(eps1=2.25, eps2=2.25, fill=0.5, fill_y=0.5)]
    poles = [(2, 1, 1.0, 1.36e16, 3.0e13, 0.0)]          # Drude-like eps1 of layer 2
    log = check(tmp_path, monkeypatch, 2, 4e-7, layers, [4.5e14, 5.5e14, 6.5e14], theta=20.0,
                phi=30.0, nY=2, period_y=4.5e-7, poles=poles)
    assert "RCWA : 2D Nx=2 Ny=2" in log
    assert "dispersive (Lorentz) materials : 1" in log
