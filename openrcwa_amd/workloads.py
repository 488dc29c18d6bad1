"""Seeded synthetic workloads of the layer-stack solve.

``pillar2d`` is the 2D rectangular-pillar batch: Nx = Ny = 7 (15 x 15 = 225
orders, 450 x 450 layer matrices), periods 400 x 400 nm, superstrate eps = 1,
four patterned layers, substrate eps = 2.25.  Each patterned layer holds one
eps1 rectangle [0, fx Lx] x [0, fy Ly] with fx, fy ~ U(0.2, 0.8),
eps1 = U(2, 12) + i U(0, 0.1), eps2 = 1, thickness ~ U(50, 300) nm.  The batch
is 256 frequencies (4.3e14 .. 7.5e14 Hz, linear) x 8 incidence directions
(theta in {0, 35/3, 70/3, 35} deg x phi in {0, 45} deg), TE and TM.

``grating1d`` is data/sample/grating.ofd with N = 5: period 300 nm, air /
eps 4.0|1.0 at fill 0.5, 200 nm / eps 2.25, normal incidence, 16 frequencies
4.3e14 .. 7.5e14 Hz.

``uneven_stacks`` is the uneven batch: per stack N ~ U{3..12} ((2N+1)^2
orders), 1 to 10 finite layers between air and eps 2.25, period 400 nm.  A
layer is a rectangle of eps U(2, 12) + i U(0, 0.1) in air, a rectangle of a
Lorentz material (eps_inf U(1, 4), 1 to 3 poles ae^2 / (ce^2 - w^2 - i be w)
with ae, be, ce in rad/s around 1e15), or a homogeneous lossy metal film
(Re eps U(-30, -5), Im eps U(0.5, 3)); thickness U(20, 300) nm; wavelength
U(400, 800) nm, theta U(0, 40) deg, phi U(0, 90) deg.

``large2d`` is the large 2D grating: Nx = Ny = 15 (31 x 31 = 961 orders,
1922 x 1922 layer matrices), periods 400 x 400 nm, air / six patterned layers /
eps 2.25.  Each layer holds one inclusion of eps U(2, 12) + i U(0, 0.05) in
air, a rectangle or a circle; a circle is the 8 x 8 staircase of cells of
its bounding box whose centres lie inside it (the curved-shape rule of the
.orcwa input path).  Thickness U(50, 300) nm; normal incidence, 64
frequencies 4.3e14 .. 7.5e14 Hz.

Random numbers: splitmix64, 53-bit uniforms in [0, 1).
"""
from __future__ import annotations

import math

import numpy as np

from .ofd import C0, common_edges, layer_grid, polarisation_basis

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def uniforms(self, count: int) -> np.ndarray:
        """The next `count` uniform() values in [0, 1) at once (the same
        sequence as count calls)."""
        k = np.arange(1, count + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + k * np.uint64(0x9E3779B97F4A7C15)
            self.state = int(z[-1]) if count else self.state
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


PILLAR_SEED = 20261202


def pillar2d(n_freq=256, thetas=(0.0, 35.0 / 3.0, 70.0 / 3.0, 35.0), phis=(0.0, 45.0),
             nX=7, nY=7, seed=PILLAR_SEED, n_layers=4, period=400e-9):
    """The 2D pillar workload as batch arrays (instance order: frequency-major,
    then theta, then phi).  Lengths in metres."""
    rng = SplitMix64(seed)
    lay = []
    for _ in range(n_layers):
        fx = rng.uniform(0.2, 0.8)
        fy = rng.uniform(0.2, 0.8)
        e1 = complex(rng.uniform(2.0, 12.0), rng.uniform(0.0, 0.1))
        d = rng.uniform(50e-9, 300e-9)
        lay.append(dict(eps1=e1.real, eps1i=e1.imag, eps2=1.0, eps2i=0.0, fill=fx, fill_y=fy,
                        thickness=d))
    layers = ([dict(eps1=1.0, eps2=1.0, fill=0.5, fill_y=0.5, thickness=0.0)] + lay +
              [dict(eps1=2.25, eps2=2.25, fill=0.5, fill_y=0.5, thickness=0.0)])
    L = len(layers)
    xe = common_edges(period, [l["fill"] for l in layers])
    ye = common_edges(period, [l["fill_y"] for l in layers])
    eps = np.stack([layer_grid(l["fill"], l["fill_y"], complex(l["eps1"], l.get("eps1i", 0.0)),
                               complex(l["eps2"], l.get("eps2i", 0.0)), xe, ye) for l in layers])
    thick = np.array([0.0 if i in (0, L - 1) else layers[i]["thickness"] for i in range(L)])
    freqs = np.linspace(4.3e14, 7.5e14, n_freq)
    lam, kx0, ky0, pols, ang = [], [], [], [], []
    for f in freqs:
        lm = C0 / f
        k0 = 2.0 * math.pi / lm
        for th in thetas:
            for ph in phis:
                t, p = math.radians(th), math.radians(ph)
                lam.append(lm)
                kx0.append(k0 * math.sin(t) * math.cos(p))   # n_inc = 1
                ky0.append(k0 * math.sin(t) * math.sin(p))
                te, tm = polarisation_basis(th, ph)
                pols.append([[te[0], te[1]], [tm[0], tm[1]]])
                ang.append((th, ph))
    lam = np.array(lam)
    kx0 = np.array(kx0)
    ky0 = np.array(ky0)
    pols = np.array(pols, dtype=complex)
    # Rayleigh anomaly flags: |kz| < 1e-9 k0 for some order in either half-space
    k0v = 2.0 * math.pi / lam
    m = np.repeat(np.arange(-nX, nX + 1), 2 * nY + 1)
    n = np.tile(np.arange(-nY, nY + 1), 2 * nX + 1)
    kx = kx0[:, None] + 2 * math.pi * m[None, :] / period
    ky = ky0[:, None] + 2 * math.pi * n[None, :] / period
    kt = kx * kx + ky * ky
    rayleigh = np.zeros(lam.size, dtype=bool)
    for e in (1.0, 2.25):
        rayleigh |= (np.sqrt(np.abs(e * k0v[:, None] ** 2 - kt)) < 1e-9 * k0v[:, None]).any(1)
    return dict(nX=nX, nY=nY, coord_x=np.array(xe), coord_y=np.array(ye), thickness=thick,
                eps=eps, lam=lam, kx0=kx0, ky0=ky0, pols=pols, angles=ang, layers=layers,
                period=period, rayleigh=rayleigh, freqs=freqs)


UNEVEN_SEED = 20261204


def lorentz(einf, poles, omega):
    """eps(w) = einf + sum ae^2 / (ce^2 - w^2 - i be w) (the pole form of
    sol/rcwa_bridge.cpp:105-123 and rcwa/RCWADriver.cpp:96-123)."""
    e = complex(einf)
    for ae, be, ce in poles:
        e += ae * ae / (ce * ce - omega * omega - 1j * be * omega)
    return e


def uneven_stacks(count=1024, seed=UNEVEN_SEED, n_range=(3, 12), layer_range=(1, 10),
                  period=400e-9):
    """A list of independent stacks (dicts with nX, nY, coord_x, coord_y, eps
    (L, ncy-1, ncx-1), thickness, lam, kx0, ky0, theta, phi, pol (TE, TM)).
    Lengths in metres."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        N = n_range[0] + int(rng.uniform() * (n_range[1] - n_range[0] + 1))
        nl = layer_range[0] + int(rng.uniform() * (layer_range[1] - layer_range[0] + 1))
        lam = rng.uniform(400e-9, 800e-9)
        omega = 2.0 * math.pi * C0 / lam
        lay = []
        for _ in range(nl):
            kind = int(rng.uniform() * 3)
            d = rng.uniform(20e-9, 300e-9)
            if kind == 2:        # homogeneous metal film
                e = complex(rng.uniform(-30.0, -5.0), rng.uniform(0.5, 3.0))
                lay.append(dict(rect=None, eps1=e, thickness=d))
                continue
            w, h = rng.uniform(0.2, 0.8) * period, rng.uniform(0.2, 0.8) * period
            x0, y0 = rng.uniform(0.0, period - w), rng.uniform(0.0, period - h)
            if kind == 0:
                e = complex(rng.uniform(2.0, 12.0), rng.uniform(0.0, 0.1))
            else:
                einf = rng.uniform(1.0, 4.0)
                npole = 1 + int(rng.uniform() * 3)
                poles = [(rng.uniform(0.5e15, 2e15), rng.uniform(0.01e15, 0.3e15),
                          rng.uniform(1e15, 6e15)) for _ in range(npole)]
                e = lorentz(einf, poles, omega)
            lay.append(dict(rect=(x0, x0 + w, y0, y0 + h), eps1=e, thickness=d))
        xe = sorted({0.0, period, *[v for l in lay if l["rect"] for v in l["rect"][:2]]})
        ye = sorted({0.0, period, *[v for l in lay if l["rect"] for v in l["rect"][2:]]})
        xc = np.array([0.5 * (xe[i] + xe[i + 1]) for i in range(len(xe) - 1)])
        yc = np.array([0.5 * (ye[j] + ye[j + 1]) for j in range(len(ye) - 1)])
        grids = [np.ones((yc.size, xc.size), dtype=complex)]
        for l in lay:
            if l["rect"] is None:
                grids.append(np.full((yc.size, xc.size), l["eps1"], dtype=complex))
            else:
                x0, x1, y0, y1 = l["rect"]
                g = np.ones((yc.size, xc.size), dtype=complex)
                g[np.ix_((yc > y0) & (yc < y1), (xc > x0) & (xc < x1))] = l["eps1"]
                grids.append(g)
        grids.append(np.full((yc.size, xc.size), 2.25 + 0j))
        th, ph = rng.uniform(0.0, 40.0), rng.uniform(0.0, 90.0)
        k0 = 2.0 * math.pi / lam
        t, p = math.radians(th), math.radians(ph)
        te, tm = polarisation_basis(th, ph)
        out.append(dict(nX=N, nY=N, period=period, coord_x=np.array(xe), coord_y=np.array(ye),
                        eps=np.stack(grids),
                        thickness=np.array([0.0] + [l["thickness"] for l in lay] + [0.0]),
                        lam=lam, kx0=k0 * math.sin(t) * math.cos(p),
                        ky0=k0 * math.sin(t) * math.sin(p), theta=th, phi=ph,
                        pol=np.array([[te[0], te[1]], [tm[0], tm[1]]], dtype=complex)))
    return out


LARGE_SEED = 20261203


def large2d(n_freq=64, n_layers=6, nX=15, nY=15, seed=LARGE_SEED, period=400e-9):
    """The large 2D grating as batch arrays (one instance per frequency, normal
    incidence, TE and TM).  Lengths in metres."""
    rng = SplitMix64(seed)
    shapes = []
    for _ in range(n_layers):
        circle = rng.uniform() < 0.5
        e1 = complex(rng.uniform(2.0, 12.0), rng.uniform(0.0, 0.05))
        d = rng.uniform(50e-9, 300e-9)
        if circle:
            r = rng.uniform(0.15, 0.4) * period
            cx = cy = 0.5 * period
            ex = list(np.linspace(cx - r, cx + r, 9))
            ey = list(np.linspace(cy - r, cy + r, 9))
        else:
            w, h = rng.uniform(0.2, 0.8) * period, rng.uniform(0.2, 0.8) * period
            x0, y0 = rng.uniform(0.0, period - w), rng.uniform(0.0, period - h)
            ex, ey, r, cx, cy = [x0, x0 + w], [y0, y0 + h], 0.0, 0.0, 0.0
        shapes.append(dict(circle=circle, eps1=e1, thickness=d, ex=ex, ey=ey, r=r, cx=cx, cy=cy))
    xe = sorted({0.0, period, *[x for sh in shapes for x in sh["ex"]]})
    ye = sorted({0.0, period, *[y for sh in shapes for y in sh["ey"]]})
    xc = [0.5 * (xe[i] + xe[i + 1]) for i in range(len(xe) - 1)]
    yc = [0.5 * (ye[j] + ye[j + 1]) for j in range(len(ye) - 1)]

    def grid(sh):
        g = np.ones((len(yc), len(xc)), dtype=complex)
        for j, y in enumerate(yc):
            for i, x in enumerate(xc):
                inside = sh["ex"][0] < x < sh["ex"][-1] and sh["ey"][0] < y < sh["ey"][-1]
                if inside and sh["circle"]:
                    # the staircase cell holding (x, y) is in when its centre is
                    ix = int(np.searchsorted(sh["ex"], x)) - 1
                    iy = int(np.searchsorted(sh["ey"], y)) - 1
                    mx = 0.5 * (sh["ex"][ix] + sh["ex"][ix + 1])
                    my = 0.5 * (sh["ey"][iy] + sh["ey"][iy + 1])
                    inside = (mx - sh["cx"]) ** 2 + (my - sh["cy"]) ** 2 <= sh["r"] ** 2
                if inside:
                    g[j, i] = sh["eps1"]
        return g
    eps = np.stack([np.ones((len(yc), len(xc)), dtype=complex)] + [grid(sh) for sh in shapes] +
                   [np.full((len(yc), len(xc)), 2.25 + 0j)])
    thick = np.array([0.0] + [sh["thickness"] for sh in shapes] + [0.0])
    freqs = np.linspace(4.3e14, 7.5e14, n_freq)
    lam = C0 / freqs
    te, tm = polarisation_basis(0.0, 0.0)
    pols = np.array([[[te[0], te[1]], [tm[0], tm[1]]]] * n_freq, dtype=complex)
    return dict(nX=nX, nY=nY, coord_x=np.array(xe), coord_y=np.array(ye), thickness=thick,
                eps=eps, lam=lam, kx0=np.zeros(n_freq), ky0=np.zeros(n_freq), pols=pols,
                shapes=shapes, period=period, freqs=freqs)


def grating1d(nX=5, n_freq=16):
    """data/sample/grating.ofd with rcwa = nX 3.0e-7 and n_freq frequencies."""
    layers = [dict(eps1=1.0, eps2=1.0, fill=0.5, thickness=0.0),
              dict(eps1=4.0, eps2=1.0, fill=0.5, thickness=2.0e-7),
              dict(eps1=2.25, eps2=2.25, fill=0.5, thickness=0.0)]
    freqs = np.linspace(4.3e14, 7.5e14, n_freq)
    return dict(nX=nX, period=3.0e-7, layers=layers, freqs=freqs)
