"""Host logic of the .ofd RCWA path, batched: every frequency of a sweep is
one instance of a single GPU batch instead of one RCWASolver per frequency
(sol/rcwa_bridge.cpp:244-366).

Layers follow ``rcwalayer_t`` (include/orcwa_rcwa.h:27-35): dicts with keys
eps1, eps2, eps1i, eps2i, fill, fill_y, thickness.  Poles follow
``rcwapole_t`` (include/orcwa_rcwa.h:45-52): ``{(layer, side): [(einf, ae, be, ce), ...]}``.
"""
from __future__ import annotations

import math

import numpy as np

from .engine import BatchResult, Context, solve_batch

C0 = 2.99792458e8  # m/s (sol/rcwa_bridge.cpp:44)


def common_edges(period: float, fills) -> list:
    """Common cell edges of all layers (sol/rcwa_bridge.cpp:52-67)."""
    e = sorted([0.0] + [f * period for f in fills if f < 1.0] + [period])
    u = []
    for v in e:
        if not u or (v - u[-1]) > 1e-12 * period:
            u.append(v)
    return u


def layer_grid(fill, fill_y, e1, e2, xe, ye) -> np.ndarray:
    """eps1 on [0, fill Lx] x [0, fill_y Ly], eps2 elsewhere; rows = y cells
    (sol/rcwa_bridge.cpp:72-89)."""
    px, py = xe[-1], ye[-1]
    xc = 0.5 * (np.asarray(xe[:-1]) + np.asarray(xe[1:]))
    yc = 0.5 * (np.asarray(ye[:-1]) + np.asarray(ye[1:]))
    in1 = (xc[None, :] < fill * px) & (yc[:, None] < fill_y * py)
    return np.where(in1, complex(e1), complex(e2))


def has_dispersion(poles, n, side) -> bool:
    return any(p[1] != 0.0 for p in poles.get((n, side), []))


def layer_eps(layer, poles, n, side, omega) -> complex:
    """Static or Lorentz permittivity (sol/rcwa_bridge.cpp:97-123)."""
    if not has_dispersion(poles, n, side):
        return complex(layer["eps1"], layer.get("eps1i", 0.0)) if side == 1 else \
            complex(layer["eps2"], layer.get("eps2i", 0.0))
    rows = poles[(n, side)]
    e = complex(rows[-1][0], 0.0)
    for (_, ae, be, ce) in rows:
        if ae == 0.0:
            continue
        e += complex(ae * ae, 0.0) / complex(ce * ce - omega * omega, -be * omega)
    return e


def polarisation_basis(theta_deg: float, phi_deg: float):
    """((teX, teY), (tmX, tmY)) of sol/rcwa_bridge.cpp:187-196."""
    th = theta_deg * math.pi / 180.0
    ph = phi_deg * math.pi / 180.0
    if abs(math.sin(th)) < 1e-9:
        return (0.0, 1.0), (1.0, 0.0)
    return ((-math.sin(ph), math.cos(ph)),
            (math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph)))


def solve_ofd(ctx: Context, nX: int, period: float, layers, freqs, theta=0.0, phi=0.0,
              nY: int = 0, period_y: float | None = None, poles=None):
    """All frequencies of one .ofd RCWA input as one batch.  Returns a dict
    with R, T (F, 2) [TE, TM], per-order REF/TRN masked as rcwa_orders.csv
    (F, 2, h), the raw result and the propagation masks."""
    poles = poles or {}
    if period_y is None or nY == 0:
        period_y = period
    freqs = np.atleast_1d(np.asarray(freqs, dtype=float))
    F = freqs.size
    L = len(layers)
    xe = common_edges(period, [l["fill"] for l in layers])
    ye = common_edges(period_y, [l.get("fill_y", 1.0) for l in layers])
    lam = C0 / freqs
    k0 = 2.0 * math.pi / lam
    omega = 2.0 * math.pi * freqs
    eps = np.empty((F, L, len(ye) - 1, len(xe) - 1), dtype=complex)
    for f in range(F):
        for n, l in enumerate(layers):
            eps[f, n] = layer_grid(l["fill"], l.get("fill_y", 1.0),
                                   layer_eps(l, poles, n, 1, omega[f]),
                                   layer_eps(l, poles, n, 2, omega[f]), xe, ye)
    eps_inc = np.array([layer_eps(layers[0], poles, 0, 1, w) for w in omega])
    eps_trn = np.array([layer_eps(layers[-1], poles, L - 1, 1, w) for w in omega])
    n_inc = np.maximum(np.sqrt(eps_inc.astype(complex)).real, 1e-6)
    th, ph = theta * math.pi / 180.0, phi * math.pi / 180.0
    kx0 = k0 * n_inc * math.sin(th) * math.cos(ph)
    ky0 = k0 * n_inc * math.sin(th) * math.sin(ph)
    thick = np.array([0.0 if n in (0, L - 1) else layers[n]["thickness"] for n in range(L)])
    te, tm = polarisation_basis(theta, phi)
    pols = np.array([[te[0], te[1]], [tm[0], tm[1]]], dtype=complex)
    res: BatchResult = solve_batch(ctx, nX, nY, xe, ye, thick, lam, kx0, ky0, eps, pols)
    nOX, nOY = 2 * nX + 1, 2 * nY + 1
    k = np.arange(nOX * nOY)
    m = k // nOY - nX
    nn = k % nOY - nY
    kx = kx0[:, None] + 2.0 * math.pi * m[None, :] / period
    ky = ky0[:, None] + 2.0 * math.pi * nn[None, :] / period_y
    kt = kx * kx + ky * ky
    propR = kt < (k0 * k0 * eps_inc.real)[:, None]
    propT = kt < (k0 * k0 * eps_trn.real)[:, None]
    REF = np.where(propR[:, None, :], res.ref, 0.0)
    TRN = np.where(propT[:, None, :], res.trn, 0.0)
    return dict(freq=freqs, lam=lam, R=res.ref.sum(-1), T=res.trn.sum(-1), REF=REF, TRN=TRN,
                propR=propR, propT=propT, m=m, n=nn, status=res.status, raw=res)
