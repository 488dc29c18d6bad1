"""The `.orcwa` input path on the GPU engine.

* `parse_rcwa_input` follows the reference parser (rcwa/RCWAInput.cpp:134-409):
  the same keywords, defaults, unit conversion (metres in the file, micrometres
  inside), warnings and errors.  `derive_z_edges` follows
  rcwa/RCWAInput.cpp:413-453, `RCWABox.contains_xy/contains_xyz` follow
  rcwa/RCWAInput.cpp:12-89.
* `build_batch` turns a parsed problem into one batch of layer-stack solves
  (one instance per wavelength) with the layer construction of the reference
  driver (rcwa/RCWADriver.cpp:53-192 and 203-315: common staircase grids, slab
  centres, eps(lambda) with Lorentz poles and the conductivity term,
  polarisation 1-5, Bloch vector from the incidence slab); `run_rcwa` solves
  it on the GPU in one call and returns the reference's `RCWAResult` list
  (rcwa/RCWADriver.h:16-28).
* `FieldRequest` and the field / device images of the first wavelength
  (rcwa/RCWADriver.h:30-41, rcwa/RCWADriver.cpp:211-216, 342-359): the field
  on the GPU (`engine.field_image`, RCWASolver::saveFieldImage), the
  permittivity map from the layer grids (RCWASolver::saveDeviceImage,
  rcwa/RCWASolver.cpp:1322-1487), both written as writeFieldCSV does
  (rcwa/RCWASolver.cpp:25-50).
* `main` is the `orcwa_rcwa` command (src/rcwa_Main.cpp:54-241): the same
  options and the same CSV files.

Magnetic materials (mu != 1, a magnetic conductivity) pass their
permeability grids to the engine (rcwa/RCWADriver.cpp:129-149, 181-190).
Not supported on this path (reported as an error, not approximated): a
patterned top or bottom slab (the engine's plane-wave kernel takes the
half-spaces as homogeneous, which the reference driver assumes,
rcwa/RCWADriver.h:11-13).
"""
from __future__ import annotations

import math
import re
import sys
from dataclasses import dataclass, field

import numpy as np

C0 = 2.99792458e8            # m/s
C0_UM = 2.99792458e14        # um/s
EPS0 = 8.854187817e-12       # F/m
MU0 = 1.25663706212e-6       # H/m
M_TO_UM = 1.0e6

SHAPE_BOX, SHAPE_SPHERE, SHAPE_CYL_Z, SHAPE_CYL_X, SHAPE_CYL_Y = 1, 2, 11, 12, 13
_SHAPES = (SHAPE_BOX, SHAPE_SPHERE, SHAPE_CYL_Z, SHAPE_CYL_X, SHAPE_CYL_Y)
N_CURVED_CELLS = 8           # in-plane staircase of curved cross-sections
N_CURVED_SLICES = 8          # z slices of shapes whose cross-section varies with z

_FDTD_PREFIXES = ("plot", "far1d", "far2d", "near1d", "near2d", "near3d")
_FDTD_ONLY = {"abc", "feed", "rfeed", "point", "load", "inductor", "source", "solver",
              "timestep", "pulsewidth", "freqdiv", "matchingloss", "name"}


class OrcwaInputError(ValueError):
    """Invalid input (the reference's parse/driver error string)."""


@dataclass
class RCWABox:
    material: int = 0
    shape: int = SHAPE_BOX
    x0: float = 0.0
    x1: float = 0.0
    y0: float = 0.0
    y1: float = 0.0
    z0: float = 0.0
    z1: float = 0.0

    def contains_xy(self, xc: float, yc: float) -> bool:
        """z-independent cross-section test (rcwa/RCWAInput.cpp:12-44)."""
        s = self.shape
        if s == SHAPE_BOX:
            return self.x0 <= xc <= self.x1 and self.y0 <= yc <= self.y1
        if s in (SHAPE_SPHERE, SHAPE_CYL_Z):
            cx, cy = 0.5 * (self.x0 + self.x1), 0.5 * (self.y0 + self.y1)
            rx, ry = 0.5 * (self.x1 - self.x0), 0.5 * (self.y1 - self.y0)
            if rx <= 0 or ry <= 0:
                return False
            dx, dy = (xc - cx) / rx, (yc - cy) / ry
            return dx * dx + dy * dy <= 1.0
        if s == SHAPE_CYL_X:
            return self.y0 <= yc <= self.y1
        if s == SHAPE_CYL_Y:
            return self.x0 <= xc <= self.x1
        return False

    def contains_xyz(self, xc: float, yc: float, zc: float) -> bool:
        """Cross-section at height zc (rcwa/RCWAInput.cpp:49-89)."""
        if zc < self.z0 or zc > self.z1:
            return False
        cx, cy, cz = 0.5 * (self.x0 + self.x1), 0.5 * (self.y0 + self.y1), 0.5 * (self.z0 + self.z1)
        rx, ry, rz = 0.5 * (self.x1 - self.x0), 0.5 * (self.y1 - self.y0), 0.5 * (self.z1 - self.z0)
        s = self.shape
        if s == SHAPE_BOX:
            return self.x0 <= xc <= self.x1 and self.y0 <= yc <= self.y1
        if s == SHAPE_SPHERE:
            if rx <= 0 or ry <= 0 or rz <= 0:
                return False
            dx, dy, dz = (xc - cx) / rx, (yc - cy) / ry, (zc - cz) / rz
            return dx * dx + dy * dy + dz * dz <= 1.0
        if s == SHAPE_CYL_Z:
            if rx <= 0 or ry <= 0:
                return False
            dx, dy = (xc - cx) / rx, (yc - cy) / ry
            return dx * dx + dy * dy <= 1.0
        if s == SHAPE_CYL_X:
            if ry <= 0 or rz <= 0 or xc < self.x0 or xc > self.x1:
                return False
            dy, dz = (yc - cy) / ry, (zc - cz) / rz
            return dy * dy + dz * dz <= 1.0
        if s == SHAPE_CYL_Y:
            if rx <= 0 or rz <= 0 or yc < self.y0 or yc > self.y1:
                return False
            dx, dz = (xc - cx) / rx, (zc - cz) / rz
            return dx * dx + dz * dz <= 1.0
        return False


@dataclass
class LorentzPole:
    ae: float = 0.0
    be: float = 0.0
    ce: float = 0.0


@dataclass
class MaterialDispersion:
    einf: float = 0.0
    poles: list = field(default_factory=list)

    def active(self) -> bool:
        return bool(self.poles)


@dataclass
class RCWAProblem:
    """rcwa/RCWAInput.h:46-110 (lengths and wavelengths in um)."""
    title: str = ""
    xmin: float = 0.0
    xmax: float = 0.0
    ymin: float = 0.0
    ymax: float = 0.0
    zmin: float = 0.0
    zmax: float = 0.0
    nHx: int = 10
    nHy: int = 0
    material_eps: list = field(default_factory=list)
    material_mu: list = field(default_factory=list)
    material_sigma: list = field(default_factory=list)
    material_mag_sigma: list = field(default_factory=list)
    material_dispersion: list = field(default_factory=list)
    boxes: list = field(default_factory=list)
    theta: float = 0.0
    phi: float = 0.0
    pol: int = 1
    psi: float = 0.0
    background_eps: complex = 1.0 + 0.0j
    lambdas: list = field(default_factory=list)
    periodic_x: bool = True
    periodic_y: bool = True

    @property
    def Lx(self) -> float:
        return self.xmax - self.xmin

    @property
    def Ly(self) -> float:
        return self.ymax - self.ymin


@dataclass
class RCWAResult:
    """rcwa/RCWADriver.h:16-28: per-wavelength totals and per-order
    efficiencies (k = (2 nHy + 1) i + j for order (i - nHx, j - nHy))."""
    lam: float
    R: float
    T: float
    A: float
    REF_orders: np.ndarray
    TRN_orders: np.ndarray
    status: int = 0


# number conversions with the prefix semantics of std::stod / std::stoi
_FLOAT_PREFIX = re.compile(
    r"\s*[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.IGNORECASE)
_INT_PREFIX = re.compile(r"\s*[+-]?\d+")


def _stod(s: str) -> float:
    m = _FLOAT_PREFIX.match(s)
    if not m:
        raise ValueError("stod")
    return float(m.group(0))


def _stoi(s: str) -> int:
    m = _INT_PREFIX.match(s)
    if not m:
        raise ValueError("stoi")
    return int(m.group(0))


def _is_fdtd_only(key: str) -> bool:
    return key.startswith(_FDTD_PREFIXES) or key in _FDTD_ONLY


def _warn(msg: str):
    print(f"*** warning: {msg}", file=sys.stderr)


def parse_rcwa_input(path: str) -> RCWAProblem:
    """Parse a `.orcwa` file (rcwa/RCWAInput.cpp:134-409).  Raises
    OrcwaInputError where the reference returns false with an error."""
    try:
        fin = open(path, "r", encoding="utf-8", errors="replace", newline="\n")
    except OSError:
        raise OrcwaInputError(f"cannot open file: {path}")
    prob = RCWAProblem()
    # material 0 = air, 1 = PEC placeholder (loss as a large positive Im eps)
    prob.material_eps = [1.0 + 0.0j, 1.0 + 1.0e8j]
    prob.material_mu = [1.0 + 0.0j, 1.0 + 0.0j]
    prob.material_sigma = [0.0, 0.0]
    prob.material_mag_sigma = [0.0, 0.0]
    prob.material_dispersion = [MaterialDispersion(), MaterialDispersion()]
    have_f1 = have_f2 = False
    nline = 0
    with fin:
        for raw in fin:
            line = raw[:-1] if raw.endswith("\n") else raw
            if not line:
                continue
            tok = line.split()
            if not tok:
                continue
            if nline == 0:
                if tok[0] not in ("OpenRCWA", "OpenTHFD"):
                    raise OrcwaInputError("not an OpenRCWA file")
                nline += 1
                continue
            nline += 1
            key = tok[0]
            if key == "end":
                break
            if key == "title":
                pos = line.find("=")
                if pos >= 0 and pos + 2 <= len(line):
                    prob.title = line[pos + 1:]
                prob.title = prob.title.lstrip(" \t")
                continue
            if len(tok) < 3 or tok[1] != "=":
                continue
            V = tok[2:]
            nv = len(V)
            try:
                if key == "xmesh":
                    prob.xmin = _stod(V[0]) * M_TO_UM
                    prob.xmax = _stod(V[nv - 1]) * M_TO_UM
                elif key == "ymesh":
                    prob.ymin = _stod(V[0]) * M_TO_UM
                    prob.ymax = _stod(V[nv - 1]) * M_TO_UM
                elif key == "zmesh":
                    prob.zmin = _stod(V[0]) * M_TO_UM
                    prob.zmax = _stod(V[nv - 1]) * M_TO_UM
                elif key == "material":
                    if nv >= 2:
                        epsr = _stod(V[1])
                        esgm = _stod(V[2]) if nv >= 3 else 0.0
                        amur = _stod(V[3]) if nv >= 4 else 1.0
                        msgm = _stod(V[4]) if nv >= 5 else 0.0
                        prob.material_eps.append(complex(epsr, 0.0))
                        prob.material_mu.append(complex(amur, 0.0))
                        prob.material_sigma.append(esgm)
                        prob.material_mag_sigma.append(msgm)
                        prob.material_dispersion.append(MaterialDispersion())
                elif key in ("material_eps", "material_index"):
                    if nv >= 2:
                        m = _stoi(V[0])
                        if 0 <= m < len(prob.material_eps):
                            a = _stod(V[1])
                            b = _stod(V[2]) if nv >= 3 else 0.0
                            z = complex(a, b)
                            prob.material_eps[m] = z * z if key == "material_index" else z
                        else:
                            _warn(f"{key}: material {m} is not defined (ignored)")
                elif key == "material_mu":
                    if nv >= 2:
                        m = _stoi(V[0])
                        if 0 <= m < len(prob.material_mu):
                            a = _stod(V[1])
                            b = _stod(V[2]) if nv >= 3 else 0.0
                            prob.material_mu[m] = complex(a, b)
                        else:
                            _warn(f"material_mu: material {m} is not defined (ignored)")
                elif key == "material_dispersion":
                    if nv >= 5:
                        m = _stoi(V[0])
                        if 0 <= m < len(prob.material_dispersion):
                            d = prob.material_dispersion[m]
                            d.einf = _stod(V[1])
                            p = LorentzPole(_stod(V[2]), _stod(V[3]), _stod(V[4]))
                            if p.ae != 0.0:
                                d.poles.append(p)
                        else:
                            _warn(f"material_dispersion: material {m} is not defined (ignored)")
                elif key == "geometry":
                    if nv >= 8:
                        m = _stoi(V[0])
                        shape = _stoi(V[1])
                        box = RCWABox(m, shape, *(_stod(V[i]) * M_TO_UM for i in range(2, 8)))
                        if shape in _SHAPES:
                            prob.boxes.append(box)
                        else:
                            _warn(f"unsupported geometry shape={shape} (ignored)")
                elif key == "planewave":
                    if nv >= 3:
                        prob.theta = _stod(V[0])
                        prob.phi = _stod(V[1])
                        prob.pol = _stoi(V[2])
                        if nv >= 4:
                            prob.psi = _stod(V[3])
                        if prob.pol < 1 or prob.pol > 5:
                            _warn(f"unsupported polarisation pol={prob.pol} "
                                  "(1=TM, 2=TE, 3=linear[psi], 4=right, 5=left); using TM")
                            prob.pol = 1
                elif key == "pbc":
                    if nv >= 2:
                        prob.periodic_x = _stoi(V[0]) != 0
                        prob.periodic_y = _stoi(V[1]) != 0
                        if not prob.periodic_x or not prob.periodic_y:
                            _warn("RCWA boundaries are periodic (pbc = 0 is ignored)")
                elif key == "rcwaorder":
                    if nv >= 1:
                        prob.nHx = _stoi(V[0])
                    if nv >= 2:
                        prob.nHy = _stoi(V[1])
                elif key in ("frequency1", "frequency2", "frequency"):
                    first = key != "frequency2"
                    already = have_f1 if first else have_f2
                    if not already and nv >= 3:
                        f0, f1, ndiv = _stod(V[0]), _stod(V[1]), max(_stoi(V[2]), 0)
                        for i in range(ndiv + 1):
                            f = f0 if ndiv == 0 else f0 + (f1 - f0) * i / float(ndiv)
                            if f > 0:
                                prob.lambdas.append((C0 / f) * M_TO_UM)
                        if first:
                            have_f1 = True
                        else:
                            have_f2 = True
                elif key in ("wavelength", "lambda"):
                    if nv >= 3:
                        l0, l1, ndiv = _stod(V[0]), _stod(V[1]), max(_stoi(V[2]), 0)
                        for i in range(ndiv + 1):
                            lam = l0 if ndiv == 0 else l0 + (l1 - l0) * i / float(ndiv)
                            if lam > 0:
                                prob.lambdas.append(lam)
                    elif nv >= 1:
                        lam = _stod(V[0])
                        if lam > 0:
                            prob.lambdas.append(lam)
                elif key == "background":
                    if nv >= 1:
                        prob.background_eps = complex(_stod(V[0]), _stod(V[1]) if nv >= 2 else 0.0)
                elif _is_fdtd_only(key):
                    pass
                else:
                    _warn(f'unknown keyword "{key}" (ignored)')
            except (ValueError, IndexError) as e:
                raise OrcwaInputError(f"parse error ({key}): {e}")
    if not prob.lambdas:
        raise OrcwaInputError("no wavelength given (use frequency1 / frequency2 or wavelength)")
    prob.lambdas = sorted(set(prob.lambdas))
    if prob.Lx <= 0:
        raise OrcwaInputError("invalid x period (check xmesh)")
    return prob


def derive_z_edges(prob: RCWAProblem) -> list:
    """Ascending z edges of the slabs (rcwa/RCWAInput.cpp:413-453)."""
    edges = {prob.zmin, prob.zmax}
    for b in prob.boxes:
        if prob.zmin <= b.z0 <= prob.zmax:
            edges.add(b.z0)
        if prob.zmin <= b.z1 <= prob.zmax:
            edges.add(b.z1)
        if b.shape in (SHAPE_SPHERE, SHAPE_CYL_X, SHAPE_CYL_Y):
            for i in range(1, N_CURVED_SLICES):
                z = b.z0 + (b.z1 - b.z0) * i / float(N_CURVED_SLICES)
                if prob.zmin <= z <= prob.zmax:
                    edges.add(z)
    return sorted(edges)


def _curved_in_x(b: RCWABox) -> bool:
    return b.shape in (SHAPE_SPHERE, SHAPE_CYL_Z, SHAPE_CYL_Y)


def _curved_in_y(b: RCWABox) -> bool:
    return b.shape in (SHAPE_SPHERE, SHAPE_CYL_Z, SHAPE_CYL_X)


def _common_grid(lo: float, hi: float, boxes, side: str) -> list:
    edges = {lo, hi}
    for b in boxes:
        a0, a1 = (b.x0, b.x1) if side == "x" else (b.y0, b.y1)
        a0, a1 = max(a0, lo), min(a1, hi)
        if lo < a0 < hi:
            edges.add(a0)
        if lo < a1 < hi:
            edges.add(a1)
        curved = _curved_in_x(b) if side == "x" else _curved_in_y(b)
        if curved and a0 < a1:
            for i in range(1, N_CURVED_CELLS):
                v = a0 + (a1 - a0) * i / float(N_CURVED_CELLS)
                if lo < v < hi:
                    edges.add(v)
    return sorted(edges)


def common_x_grid(prob: RCWAProblem) -> list:
    """x edges shared by every slab (rcwa/RCWADriver.cpp:52-66)."""
    return _common_grid(prob.xmin, prob.xmax, prob.boxes, "x")


def common_y_grid(prob: RCWAProblem) -> list:
    """y edges shared by every slab (rcwa/RCWADriver.cpp:69-83)."""
    return _common_grid(prob.ymin, prob.ymax, prob.boxes, "y")


def eps_at(prob: RCWAProblem, xc: float, yc: float, zc: float, lam: float) -> complex:
    """Permittivity at a point (rcwa/RCWADriver.cpp:87-126): the last box that
    contains the point wins; a dispersive material uses its Lorentz sum, and
    the conductivity adds +i sigma / (omega eps0) (exp(-i omega t))."""
    e = complex(prob.background_eps)
    omega = 2.0 * math.pi * C0_UM / lam
    for b in prob.boxes:
        if not b.contains_xyz(xc, yc, zc):
            continue
        m = b.material
        if m < 0 or m >= len(prob.material_eps):
            continue
        e = complex(prob.material_eps[m])
        if m < len(prob.material_dispersion):
            d = prob.material_dispersion[m]
            if d.active():
                e = complex(d.einf, 0.0)
                for p in d.poles:
                    e += complex(p.ae * p.ae, 0.0) / complex(p.ce * p.ce - omega * omega, -p.be * omega)
        if m < len(prob.material_sigma) and prob.material_sigma[m] != 0.0:
            e += complex(0.0, prob.material_sigma[m] / (omega * EPS0))
    return e


def mu_at(prob: RCWAProblem, xc: float, yc: float, zc: float, lam: float) -> complex:
    """Relative permeability at a point (rcwa/RCWADriver.cpp:129-149): 1 in
    the background, the last containing box's material otherwise, plus
    +i msigma / (omega mu0) for a magnetic conductivity."""
    m = 1.0 + 0.0j
    for b in prob.boxes:
        if not b.contains_xyz(xc, yc, zc):
            continue
        mi = b.material
        if mi < 0 or mi >= len(prob.material_mu):
            continue
        m = complex(prob.material_mu[mi])
        if mi < len(prob.material_mag_sigma) and prob.material_mag_sigma[mi] != 0.0:
            omega = 2.0 * math.pi * C0_UM / lam
            m += complex(0.0, prob.material_mag_sigma[mi] / (omega * MU0))
    return m


def has_magnetic_material(prob: RCWAProblem) -> bool:
    """rcwa/RCWADriver.cpp:152-160."""
    return (any(abs(complex(m) - 1.0) > 1e-12 for m in prob.material_mu)
            or any(s != 0.0 for s in prob.material_mag_sigma))


def incident_polarisation(prob: RCWAProblem) -> tuple:
    """Transverse (px, py) of the unit incident field for pol 1-5
    (rcwa/RCWADriver.cpp:232-276)."""
    th = prob.theta * math.pi / 180.0
    ph = prob.phi * math.pi / 180.0
    if abs(math.sin(th)) < 1e-9:
        tm, te = (1.0, 0.0), (0.0, 1.0)
    else:
        tm = (math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph))
        te = (-math.sin(ph), math.cos(ph))
    s2 = 1.0 / math.sqrt(2.0)
    if prob.pol == 2:
        a_tm, a_te = 0.0, 1.0
    elif prob.pol == 3:
        psi = prob.psi * math.pi / 180.0
        a_tm, a_te = math.cos(psi), math.sin(psi)
    elif prob.pol == 4:
        a_tm, a_te = complex(s2), complex(0.0, s2)
    elif prob.pol == 5:
        a_tm, a_te = complex(s2), complex(0.0, -s2)
    else:
        a_tm, a_te = 1.0, 0.0
    return complex(a_tm * tm[0] + a_te * te[0]), complex(a_tm * tm[1] + a_te * te[1])


def build_batch(prob: RCWAProblem) -> dict:
    """One layer-stack instance per wavelength, slabs from the top (largest
    z, incidence side) down (rcwa/RCWADriver.cpp:203-315)."""
    zedges = derive_z_edges(prob)
    if len(zedges) < 2:
        raise OrcwaInputError("cannot form z slabs (check zmesh / geometry)")
    magnetic = has_magnetic_material(prob)
    xg = common_x_grid(prob)
    yg = common_y_grid(prob)
    centres, thick = [], []
    for i in range(len(zedges) - 1, 0, -1):
        centres.append(0.5 * (zedges[i] + zedges[i - 1]))
        thick.append(zedges[i] - zedges[i - 1])
    if len(centres) == 1:
        # one slab: incidence and exit are the same layer, which the reference
        # solves as a one-layer stack (R = 0, T = I); the engine takes it as two
        # copies of the slab joined by an identity interface
        centres.append(centres[0])
        thick.append(thick[0])
    L = len(centres)
    lams = np.asarray(prob.lambdas, dtype=float)
    B = lams.size
    xc = [0.5 * (xg[i] + xg[i + 1]) for i in range(len(xg) - 1)]
    yc = [0.5 * (yg[j] + yg[j + 1]) for j in range(len(yg) - 1)]
    eps = np.empty((B, L, len(yc), len(xc)), dtype=complex)
    mu = np.ones((B, L, len(yc), len(xc)), dtype=complex) if magnetic else None
    for b, lam in enumerate(lams):
        for s, zc in enumerate(centres):
            for j, y in enumerate(yc):
                for i, x in enumerate(xc):
                    eps[b, s, j, i] = eps_at(prob, x, y, zc, lam)
                    if magnetic:
                        mu[b, s, j, i] = mu_at(prob, x, y, zc, lam)
    for s in (0, L - 1):
        for grid in (eps, mu):
            if grid is None:
                continue
            g = grid[:, s]
            if not np.all(g == g[:, :1, :1]):
                raise OrcwaInputError("the top and bottom slabs must be homogeneous "
                                      f"(slab {s} is patterned)")
    th = prob.theta * math.pi / 180.0
    ph = prob.phi * math.pi / 180.0
    k0 = 2.0 * math.pi / lams
    xm, ym = 0.5 * (prob.xmin + prob.xmax), 0.5 * (prob.ymin + prob.ymax)
    eps_inc = np.array([eps_at(prob, xm, ym, centres[0], lam) for lam in lams])
    mu_inc = (np.array([mu_at(prob, xm, ym, centres[0], lam) for lam in lams]) if magnetic
              else np.ones_like(eps_inc))
    # n = sqrt(eps mu) of the incidence medium (rcwa/RCWADriver.cpp:289-296)
    n_inc = np.maximum(np.sqrt(eps_inc * mu_inc).real, 1e-6)
    if np.any(np.abs(eps_inc.imag) > 1e-6 * np.abs(eps_inc.real)):
        _warn("the incidence medium is lossy (R/T assume a lossless incidence medium)")
    kx0 = k0 * n_inc * math.sin(th) * math.cos(ph)
    ky0 = k0 * n_inc * math.sin(th) * math.sin(ph)
    px, py = incident_polarisation(prob)
    return dict(xgrid=np.asarray(xg), ygrid=np.asarray(yg), thickness=np.asarray(thick),
                centres=np.asarray(centres), lam=lams, kx0=kx0, ky0=ky0, eps=eps, mu=mu,
                pol=np.array([[px, py]], dtype=complex), nX=prob.nHx, nY=prob.nHy)


@dataclass
class FieldRequest:
    """RCWAFieldRequest (rcwa/RCWADriver.h:30-41): the component, slice,
    coordinate [um] and save option of the field image, and the two output
    paths (empty = not written)."""
    component: str = "Ez"
    slice: str = "xz"
    coord: float = 0.0
    opt: str = "mod"
    path: str = ""
    device_path: str = ""

    def wants_anything(self) -> bool:
        return bool(self.path) or bool(self.device_path)


def write_field_csv(path: str, field: np.ndarray, opt: str):
    """writeFieldCSV (rcwa/RCWASolver.cpp:25-50): |f|, Re f or Im f, one row
    per line, ", " between values, 8 significant digits."""
    if opt == "re":
        vals = field.real
    elif opt == "im":
        vals = field.imag
    else:
        vals = np.abs(field)
    with open(path, "w") as f:
        for row in vals:
            f.write(", ".join(format(float(v), ".8g") for v in row) + "\n")


def _image_samples(start: float, length: float, count: int) -> np.ndarray:
    step = length / count
    return start + step * (np.arange(count) + 0.5)


def device_image(bt: dict, slice_type: str, coord: float, b: int = 0) -> np.ndarray:
    """RCWASolver::saveDeviceImage (rcwa/RCWASolver.cpp:1322-1487) for instance
    b: the layer permittivity at 2000 cell-centred samples per row; rows run
    down the interior slabs (XZ, YZ; a row belongs to the slab with
    z0 <= z < z1) or along y (XY, the slab at depth coord).  Layer::eps
    wraps x into the period and gives 0 for y outside it
    (rcwa/Layer.cpp:383-428)."""
    xg, yg, thick = bt["xgrid"], bt["ygrid"], bt["thickness"]
    L = thick.size
    eps = bt["eps"][b]
    Lx, Ly = xg[-1] - xg[0], yg[-1] - yg[0]
    zint = float(sum(thick[1:L - 1]))
    nres1 = 2000
    if slice_type == "xy":
        s1, l1, s2, l2 = xg[0], Lx, yg[0], Ly
    elif slice_type == "yz":
        s1, l1, s2, l2 = yg[0], Ly, 0.0, zint
    else:
        s1, l1, s2, l2 = xg[0], Lx, 0.0, zint
    nres2 = max(1, int(nres1 * l2 / l1))
    a1 = _image_samples(s1, l1, nres1)
    a2 = _image_samples(s2, l2, nres2)

    def slab_at(z):
        z0 = 0.0
        for j in range(1, L - 1):
            z1 = z0 + thick[j]
            if z0 <= z < z1:
                return j
            z0 = z1
        raise OrcwaInputError("Device image requires at least one finite-thickness "
                              "interior layer!")

    def wrap_x(x):
        while x < xg[0]:
            x += Lx
        while x > xg[-1]:
            x -= Lx
        return x

    def lookup(grid, xs, ys):
        xs = np.array([wrap_x(float(x)) for x in np.atleast_1d(xs)])
        ys = np.atleast_1d(np.asarray(ys, dtype=float))
        ix = np.searchsorted(xg[1:], xs, side="left")
        iy = np.clip(np.searchsorted(yg[1:], ys, side="left"), 0, len(yg) - 2)
        out = grid[iy, ix]
        return np.where((ys < yg[0]) | (ys > yg[-1]), 0.0, out)

    out = np.zeros((nres2, nres1), dtype=complex)
    if slice_type == "xy":
        g = eps[slab_at(coord)]
        for i, y in enumerate(a2):
            out[i] = lookup(g, a1, np.full(nres1, y))
    else:
        for i, z in enumerate(a2):
            g = eps[slab_at(z)]
            out[i] = (lookup(g, a1, np.full(nres1, coord)) if slice_type == "xz"
                      else lookup(g, np.full(nres1, coord), a1))
    return out


def field_image(bt: dict, ctx, component: str, slice_type: str, coord: float,
                b: int = 0) -> np.ndarray:
    """The field image of instance b (saveFieldImage with the driver's
    plane-wave coefficients, rcwa/RCWADriver.cpp:342-352), on the GPU."""
    from . import engine
    thick = bt["thickness"].copy()
    thick[0] = thick[-1] = 0.0
    mu = None if bt["mu"] is None else bt["mu"][b]
    img, _ = engine.field_image(ctx, bt["nX"], bt["nY"], bt["xgrid"], bt["ygrid"], thick,
                                bt["lam"][b], bt["kx0"][b], bt["ky0"][b], bt["eps"][b],
                                bt["pol"][0], component, slice_type, coord, 0, mu)
    return img


def run_rcwa(prob: RCWAProblem, ctx=None, field_req: FieldRequest | None = None) -> list:
    """Solve every wavelength of `prob` as one GPU batch (runRCWA,
    rcwa/RCWADriver.cpp:195-378).  A grazing instance returns empty
    per-order arrays and R = T = 0, as the reference's silent return from
    scatterPlaneWave leaves REF/TRN empty.  With a field request the images
    of the first wavelength are written; they need an interior slab
    (rcwa/RCWADriver.cpp:209-216)."""
    from . import engine
    from ._lib import STATUS_GRAZING
    if field_req is not None and field_req.wants_anything() and len(derive_z_edges(prob)) - 1 < 3:
        raise OrcwaInputError("field images need an interior layer "
                              "(split zmesh / geometry into 3 or more slabs)")
    bt = build_batch(prob)
    own = ctx is None
    if own:
        ctx = engine.Context(0)
    try:
        # the thickness of the two half-spaces is never used by the solve
        thick = bt["thickness"].copy()
        thick[0] = thick[-1] = 0.0
        res = engine.solve_batch(ctx, bt["nX"], bt["nY"], bt["xgrid"], bt["ygrid"], thick,
                                 bt["lam"], bt["kx0"], bt["ky0"], bt["eps"], bt["pol"],
                                 bt["mu"])
        if field_req is not None and field_req.path:
            write_field_csv(field_req.path, field_image(bt, ctx, field_req.component,
                                                        field_req.slice, field_req.coord),
                            field_req.opt)
        if field_req is not None and field_req.device_path:
            write_field_csv(field_req.device_path,
                            device_image(bt, field_req.slice, field_req.coord), "re")
    finally:
        if own:
            ctx.close()
    out = []
    for b, lam in enumerate(bt["lam"]):
        st = int(res.status[b])
        if st & STATUS_GRAZING:
            ref = np.zeros(0)
            trn = np.zeros(0)
        else:
            ref = np.array(res.ref[b, 0])
            trn = np.array(res.trn[b, 0])
        R, T = float(ref.sum()), float(trn.sum())
        out.append(RCWAResult(float(lam), R, T, 1.0 - R - T, ref, trn, st))
    return out


def _g8(v: float) -> str:
    """A double as a C++ stream with precision 8 writes it."""
    return format(v, ".8g")


def write_csv(path: str, results, verbose: bool, n_orders: int):
    """The R/T CSV of src/rcwa_Main.cpp:160-187."""
    with open(path, "w") as f:
        head = "# lambda[um], R, T, A, R+T"
        if verbose:
            head += "".join(f", R_order_{i}" for i in range(n_orders))
            head += "".join(f", T_order_{i}" for i in range(n_orders))
        f.write(head + "\n")
        for r in results:
            row = f"{_g8(r.lam)}, {_g8(r.R)}, {_g8(r.T)}, {_g8(r.A)}, {_g8(r.R + r.T)}"
            if verbose:
                row += "".join(f", {_g8(v)}" for v in r.REF_orders)
                row += "".join(f", {_g8(v)}" for v in r.TRN_orders)
            f.write(row + "\n")


_USAGE = ("Usage : python -m openrcwa_amd.orcwa [-o <out.csv>] [-v] [field options] "
          "<datafile.orcwa>\n"
          "  -o <out.csv>  : output CSV file (default: <input>_rcwa.csv)\n"
          "  -v            : verbose - also write per-order efficiencies\n"
          "  field image of the first wavelength:\n"
          "  -field <c>    : Ex|Ey|Ez|Hx|Hy|Hz\n"
          "  -device       : also write the permittivity map (real part)\n"
          "  -slice <p>    : xy|xz|yz (default xz)\n"
          "  -scoord <v>   : slice position [um] (default 0)\n"
          "  -sopt <o>     : mod|re|im (default mod)\n"
          "  -fieldout <f> : field CSV file (default: <input>_field.csv)\n"
          "  -deviceout <f>: device CSV file (default: <input>_device.csv)\n")


def main(argv=None) -> int:
    """The orcwa_rcwa command (src/rcwa_Main.cpp:54-241) for R/T output."""
    argv = sys.argv[1:] if argv is None else list(argv)
    infile, outfile, verbose = "", "", False
    req = FieldRequest()
    want_field = want_device = False
    field_out = device_out = ""
    i = 0
    while i < len(argv):
        a = argv[i]
        more = i + 1 < len(argv)
        if a == "-o" and more:
            outfile = argv[i + 1]
            i += 1
        elif a in ("-v", "--verbose"):
            verbose = True
        elif a == "-field" and more:
            i += 1
            if argv[i] not in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
                print(f"*** invalid -field argument: {argv[i]}", file=sys.stderr)
                return 1
            req.component = argv[i]
            want_field = True
        elif a == "-slice" and more:
            i += 1
            if argv[i] not in ("xy", "xz", "yz"):
                print(f"*** invalid -slice argument: {argv[i]}", file=sys.stderr)
                return 1
            req.slice = argv[i]
        elif a == "-scoord" and more:
            i += 1
            try:
                req.coord = _stod(argv[i])
            except ValueError:
                print(f"*** invalid -scoord argument: {argv[i]}", file=sys.stderr)
                return 1
        elif a == "-sopt" and more:
            i += 1
            if argv[i] not in ("mod", "re", "im"):
                print(f"*** invalid -sopt argument: {argv[i]}", file=sys.stderr)
                return 1
            req.opt = argv[i]
        elif a == "-fieldout" and more:
            i += 1
            field_out = argv[i]
        elif a == "-device":
            want_device = True
        elif a == "-deviceout" and more:
            i += 1
            device_out = argv[i]
            want_device = True
        elif a in ("--help", "-h"):
            print(_USAGE, end="")
            return 0
        else:
            infile = a
        i += 1
    if not infile:
        print(_USAGE, end="")
        return 1
    stem = infile[:infile.rfind(".")] if "." in infile else infile
    if not outfile:
        outfile = stem + "_rcwa.csv"
    if want_field:
        req.path = field_out or stem + "_field.csv"
    if want_device:
        req.device_path = device_out or stem + "_device.csv"
    try:
        prob = parse_rcwa_input(infile)
    except OrcwaInputError as e:
        print(f"*** input error: {e}", file=sys.stderr)
        return 1
    print("<<< OpenRCWA (RCWA solver, MI355X) >>>")
    print(f"Title    = {prob.title}")
    print(f"Period   = {prob.Lx:g} x {prob.Ly:g} [um]")
    print(f"Harmonics= {2 * prob.nHx + 1} x {2 * prob.nHy + 1}")
    print(f"Incidence= theta {prob.theta:g}, phi {prob.phi:g}, pol {prob.pol}")
    print(f"Layers   = {len(derive_z_edges(prob)) - 1}")
    print(f"Wavelens = {len(prob.lambdas)} ({prob.lambdas[-1]:g} - {prob.lambdas[0]:g} um)")
    print("=== solving ===")
    try:
        results = run_rcwa(prob, field_req=req if req.wants_anything() else None)
    except OrcwaInputError as e:
        print(f"*** solve error: {e}", file=sys.stderr)
        return 1
    n_ord = (2 * prob.nHx + 1) * (2 * prob.nHy + 1)
    write_csv(outfile, results, verbose, n_ord)
    for r in results:
        print(f"  lambda = {r.lam:.6g} um   R = {r.R:.6g}   T = {r.T:.6g}   A = {r.A:.6g}"
              f"   R+T = {r.R + r.T:.6g}")
    print("*** warning: the HDF5 spectrum file is not written on this path (the CSV is complete)",
          file=sys.stderr)
    print(f"=== output -> {outfile} ===")
    if want_field:
        print(f"=== field  -> {req.path} ===")
    if want_device:
        print(f"=== device -> {req.device_path} ===")
    print("=== normal end ===")
    return 0


if __name__ == "__main__":
    sys.exit(main())
