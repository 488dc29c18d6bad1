"""Learned surrogate of a two-layer 2D grating, trained on the GPU solver's output.

Design space (12 parameters): two patterned layers between air and a substrate
of eps 2.25, period 300 nm in x and y, Nx = Ny = 5 (11 x 11 = 121 orders,
242 x 242 layer matrices).  Layer j (j = 1, 2, top to bottom) holds a square
inclusion [0, f_j P] x [0, f_j P] of material j in air.  Two sets of ranges
(SPACES): "sweep", the default, a design sweep around a nominal grating, and
"wide":

                                          sweep          wide
  f1, f2        fill fractions            U(0.3, 0.7)    U(0.2, 0.8)
  d1, d2        thicknesses (nm)          U(60, 140)     U(50, 200)
  eps1, eps2    Re eps of each material   U(2, 3.5)      U(2, 6)
                Im eps                    U(0, 0.05)     U(0, 0.1)
  lambda        wavelength (nm)           U(400, 800)    U(400, 800)
  theta, phi    incidence (deg)           theta U(0, 20) theta U(0, 30), phi U(0, 45)
  pol           TE (-sin phi, cos phi) or TM (cos theta cos phi, cos theta sin phi)
                (the plane-wave amplitudes of sol/rcwa_bridge.cpp:187-196, without
                its theta = 0 special case so that the map stays continuous)

Outputs (18): R and T of the orders (m, n) in {-1, 0, 1}^2, the per-order
efficiencies of RCWASolver::scatterPlaneWave (rcwa/RCWASolver.cpp:1792-1908).
With P = 300 nm and lambda >= 400 nm no other order propagates in air or in
the substrate (|m| >= 2 needs |kx| >= 2 lambda / P - sin theta > 1.5 k0), so the
18 values carry all the reflected and transmitted power.

The network (include/openrcwa_surrogate.h) is an ensemble of independently
initialised 4 x 512 GELU MLPs trained by the engine's HIP kernels; the spread
of the members (max - min over members, the largest over outputs), scaled by a
factor calibrated on held-out solver data, is the error estimate.  A query
whose estimate exceeds the tolerance (1e-3 by default), or whose prediction
leaves the physical range (an efficiency outside [0, 1], sum R + T > 1), is
answered by the plain GPU solver instead.
"""
from __future__ import annotations

import ctypes
import math
import time
from ctypes import POINTER, c_float, c_void_p

import numpy as np

from . import _lib
from .engine import Context, Plan
from .workloads import SplitMix64

C0 = 2.99792458e8
PERIOD = 300e-9
NX = NY = 5
EPS_SUB = 2.25
ORDERS = [(m, n) for m in (-1, 0, 1) for n in (-1, 0, 1)]
N_OUT = 2 * len(ORDERS)
PARAMS = ["fill1", "fill2", "d1", "d2", "eps1_re", "eps1_im", "eps2_re", "eps2_im", "lambda",
          "theta", "phi", "pol"]
# Parameter ranges of the two design spaces (see the module docstring).
SPACES = {
    "sweep": {"fill1": (0.3, 0.7), "fill2": (0.3, 0.7), "d1": (60e-9, 140e-9),
              "d2": (60e-9, 140e-9), "eps1_re": (2.0, 3.5), "eps1_im": (0.0, 0.05),
              "eps2_re": (2.0, 3.5), "eps2_im": (0.0, 0.05), "lambda": (400e-9, 800e-9),
              "theta": (0.0, 20.0), "phi": (0.0, 45.0), "pol": (0.0, 1.0)},
    "wide": {"fill1": (0.2, 0.8), "fill2": (0.2, 0.8), "d1": (50e-9, 200e-9),
             "d2": (50e-9, 200e-9), "eps1_re": (2.0, 6.0), "eps1_im": (0.0, 0.1),
             "eps2_re": (2.0, 6.0), "eps2_im": (0.0, 0.1), "lambda": (400e-9, 800e-9),
             "theta": (0.0, 30.0), "phi": (0.0, 45.0), "pol": (0.0, 1.0)},
}
DEFAULT_SPACE = "sweep"
RANGES = SPACES[DEFAULT_SPACE]


def ranges(space=DEFAULT_SPACE) -> dict:
    """The parameter ranges of a design space (a name of SPACES or a dict)."""
    return SPACES[space] if isinstance(space, str) else space
N_IN = len(PARAMS)
SEED_TRAIN = 20261203
SEED_HELDOUT = 20261204
SEED_CALIB = 20261205


def order_index(m, n, nX=NX, nY=NY):
    """Harmonic index k = (2Ny+1)(m+Nx) + (n+Ny) (rcwa/RCWASolver.cpp:255-264)."""
    return (2 * nY + 1) * (m + nX) + (n + nY)


ORDER_IDX = np.array([order_index(m, n) for m, n in ORDERS])


def sample_geometries(count: int, seed: int, space=DEFAULT_SPACE) -> np.ndarray:
    """count x 11 parameter rows (all but pol) from splitmix64 uniforms."""
    rg = ranges(space)
    u = SplitMix64(seed).uniforms(count * (N_IN - 1)).reshape(count, N_IN - 1)
    lo = np.array([rg[k][0] for k in PARAMS[:-1]])
    hi = np.array([rg[k][1] for k in PARAMS[:-1]])
    return lo + (hi - lo) * u


def sample_queries(count: int, seed: int, space=DEFAULT_SPACE) -> np.ndarray:
    """count x 12 parameter rows with a random polarisation (0 = TE, 1 = TM)."""
    g = sample_geometries(count, seed, space)
    pol = (SplitMix64(seed ^ 0x5DEECE66D).uniforms(count) >= 0.5).astype(np.float64)
    return np.concatenate([g, pol[:, None]], axis=1)


def scale_inputs(p: np.ndarray, space=DEFAULT_SPACE) -> np.ndarray:
    """Parameters -> network inputs in [-1, 1] over the space's ranges (float32)."""
    p = np.asarray(p, dtype=np.float64)
    rg = ranges(space)
    lo = np.array([rg[k][0] for k in PARAMS])
    hi = np.array([rg[k][1] for k in PARAMS])
    return (2.0 * (p - lo) / (hi - lo) - 1.0).astype(np.float32)


def features(p: np.ndarray, space=DEFAULT_SPACE) -> np.ndarray:
    """Network inputs of parameter rows: the 12 scaled parameters, then for each
    of the 8 non-zero orders of ORDERS in air and in the substrate the
    normalised z wavenumber kz / k0 = sqrt(n^2 - (sin t cos f + m lam / P)^2 -
    (sin t sin f + n lam / P)^2), as its real (propagating) and imaginary
    (evanescent) parts.  kz is where the efficiencies are not smooth: an
    order's power switches on where its kz crosses zero (the Rayleigh
    anomalies), so giving it to the network keeps the map it learns smooth."""
    p = np.asarray(p, dtype=np.float64)
    x = scale_inputs(p, space)
    r = p[:, 8] / PERIOD
    th, ph = np.deg2rad(p[:, 9]), np.deg2rad(p[:, 10])
    kx, ky = np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph)
    cols = []
    for nmed in (1.0, math.sqrt(EPS_SUB)):
        for m, n in ORDERS:
            if m == 0 and n == 0:
                continue
            q = nmed * nmed - (kx + m * r) ** 2 - (ky + n * r) ** 2
            cols.append(np.sqrt(np.maximum(q, 0.0)))
            cols.append(np.sqrt(np.maximum(-q, 0.0)) / 3.0)
    return np.concatenate([x, np.stack(cols, 1).astype(np.float32)], 1)


N_FEAT = N_IN + 32


def stack_arrays(geoms: np.ndarray):
    """Batch arrays of the solver for geometry rows (count x >= 11): per-instance
    cell edges (the common edges of both layers, sorted, a repeated edge being an
    empty cell), eps grids (count, 4, 3, 3), thicknesses, wavelength, Bloch
    offsets and the TE / TM amplitudes."""
    g = np.asarray(geoms, dtype=np.float64)
    B = g.shape[0]
    f1, f2, d1, d2 = g[:, 0], g[:, 1], g[:, 2], g[:, 3]
    e1 = g[:, 4] + 1j * g[:, 5]
    e2 = g[:, 6] + 1j * g[:, 7]
    lam = g[:, 8]
    th = np.deg2rad(g[:, 9])
    ph = np.deg2rad(g[:, 10])
    edges = np.sort(np.stack([np.zeros(B), f1 * PERIOD, f2 * PERIOD, np.full(B, PERIOD)], 1), 1)
    right = edges[:, 1:]                                   # right edge of each cell
    tol = 1e-9 * PERIOD
    in1 = right <= (f1 * PERIOD)[:, None] + tol            # cells inside [0, f1 P]
    in2 = right <= (f2 * PERIOD)[:, None] + tol
    eps = np.ones((B, 4, 3, 3), dtype=np.complex128)
    eps[:, 1] = np.where(in1[:, :, None] & in1[:, None, :], e1[:, None, None], 1.0)
    eps[:, 2] = np.where(in2[:, :, None] & in2[:, None, :], e2[:, None, None], 1.0)
    eps[:, 3] = EPS_SUB
    thick = np.stack([np.zeros(B), d1, d2, np.zeros(B)], 1)
    k0 = 2.0 * math.pi / lam
    kx0 = k0 * np.sin(th) * np.cos(ph)
    ky0 = k0 * np.sin(th) * np.sin(ph)
    pols = np.zeros((B, 2, 2), dtype=np.complex128)
    pols[:, 0, 0], pols[:, 0, 1] = -np.sin(ph), np.cos(ph)
    pols[:, 1, 0], pols[:, 1, 1] = np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph)
    return dict(edges=edges, eps=eps, thickness=thick, lam=lam, kx0=kx0, ky0=ky0, pols=pols)


def solve_geometries(ctx: Context, geoms: np.ndarray, chunk: int = 4096, progress=None):
    """Plain GPU solve of every geometry row: (count, 2, 18) efficiencies (TE, TM;
    R of ORDERS then T of ORDERS) and the per-instance status bits.
    progress(done, total) is called after each chunk."""
    g = np.asarray(geoms, dtype=np.float64)
    out = np.zeros((g.shape[0], 2, N_OUT))
    status = np.zeros(g.shape[0], dtype=np.int32)
    for i0 in range(0, g.shape[0], chunk):
        a = stack_arrays(g[i0:i0 + chunk])
        B = a["lam"].size
        plan = Plan(ctx, NX, NY, a["edges"][0], a["edges"][0], a["thickness"][0], a["lam"],
                    a["kx0"], a["ky0"], a["eps"], a["pols"],
                    inst_coords=(a["edges"], a["edges"]), inst_thickness=a["thickness"])
        try:
            plan.execute()
            plan.sync()
            r = plan.results()
        finally:
            plan.close()
        out[i0:i0 + B, :, :len(ORDERS)] = r.ref[:, :, ORDER_IDX]
        out[i0:i0 + B, :, len(ORDERS):] = r.trn[:, :, ORDER_IDX]
        status[i0:i0 + B] = r.status
        if progress is not None:
            progress(i0 + B, g.shape[0])
    return out, status


def solve_queries(ctx: Context, queries: np.ndarray, chunk: int = 4096):
    """Plain GPU answer to each 12-parameter query: (count, 18)."""
    q = np.asarray(queries, dtype=np.float64)
    res, status = solve_geometries(ctx, q[:, :N_IN - 1], chunk)
    pol = (q[:, N_IN - 1] >= 0.5).astype(int)
    return res[np.arange(q.shape[0]), pol], status


def make_dataset(ctx: Context, n_solves: int, seed: int = SEED_TRAIN, chunk: int = 4096,
                 progress=None, space=DEFAULT_SPACE):
    """GPU-generated training rows: every solved geometry gives a TE and a TM
    row.  Returns (params (2 n, 12), targets (2 n, 18) float64, status (n,))."""
    g = sample_geometries(n_solves, seed, space)
    res, status = solve_geometries(ctx, g, chunk, progress)
    params = np.concatenate([np.concatenate([g, np.zeros((n_solves, 1))], 1),
                             np.concatenate([g, np.ones((n_solves, 1))], 1)])
    targets = np.concatenate([res[:, 0], res[:, 1]])
    return params, targets, status


def _fp(a):
    return a.ctypes.data_as(POINTER(c_float))


class MLP:
    """orcwa_mlp_*: an ensemble of MLPs on the GPU (include/openrcwa_surrogate.h)."""

    def __init__(self, ctx: Context, n_in=N_IN, width=512, depth=4, n_out=N_OUT, members=4,
                 seed=1):
        self._lib = _lib.load()
        self.ctx = ctx
        self.n_in, self.width, self.depth, self.n_out, self.members = \
            int(n_in), int(width), int(depth), int(n_out), int(members)
        h = c_void_p()
        _lib.check(self._lib.orcwa_mlp_create(ctx.handle, self.n_in, self.width, self.depth,
                                              self.n_out, self.members, int(seed),
                                              ctypes.byref(h)), "orcwa_mlp_create")
        self._h = h
        self.P = int(self._lib.orcwa_mlp_param_count(h))
        ctx._plans.add(self)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orcwa_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        p = np.zeros((self.members, self.P), dtype=np.float32)
        _lib.check(self._lib.orcwa_mlp_get_params(self._h, _fp(p)), "orcwa_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(np.broadcast_to(np.asarray(p, dtype=np.float32),
                                                 (self.members, self.P)))
        _lib.check(self._lib.orcwa_mlp_set_params(self._h, _fp(p)), "orcwa_mlp_set_params")

    def set_data(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        if x.shape[1] != self.n_in or t.shape[1] != self.n_out or x.shape[0] != t.shape[0]:
            raise ValueError("training rows must be (count, n_in) and (count, n_out)")
        _lib.check(self._lib.orcwa_mlp_set_data(self._h, _fp(x), _fp(t), x.shape[0]),
                   "orcwa_mlp_set_data")

    def train(self, epochs, batch=4096, lr_max=1e-3, lr_min=1e-5, seed=7, use_graph=True):
        o = _lib.OrcwaMlpTrainOpts(int(epochs), int(batch), float(lr_max), float(lr_min), 0.0,
                                   0.0, 0.0, int(seed), 1 if use_graph else 0)
        loss = np.zeros(int(epochs), dtype=np.float32)
        _lib.check(self._lib.orcwa_mlp_train(self._h, ctypes.byref(o), _fp(loss)),
                   "orcwa_mlp_train")
        return loss

    def gradient(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros((self.members, self.P), dtype=np.float32)
        loss = np.zeros(self.members, dtype=np.float32)
        _lib.check(self._lib.orcwa_mlp_gradient(self._h, _fp(x), _fp(t), x.shape[0], _fp(g),
                                                _fp(loss)), "orcwa_mlp_gradient")
        return loss, g

    def set_precision(self, bf16: bool):
        _lib.check(self._lib.orcwa_mlp_set_precision(self._h, _lib.MLP_BF16 if bf16 else
                                                     _lib.MLP_FP32), "orcwa_mlp_set_precision")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((self.members, x.shape[0], self.n_out), dtype=np.float32)
        _lib.check(self._lib.orcwa_mlp_forward(self._h, _fp(x), x.shape[0], _fp(y)),
                   "orcwa_mlp_forward")
        return y

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        mean = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        err = np.zeros(x.shape[0], dtype=np.float32)
        _lib.check(self._lib.orcwa_mlp_predict(self._h, _fp(x), x.shape[0], _fp(mean), _fp(err)),
                   "orcwa_mlp_predict")
        return mean, err

    def set_error_scale(self, scale):
        sc = np.ascontiguousarray(np.broadcast_to(np.asarray(scale, dtype=np.float32),
                                                  (self.n_out,)))
        _lib.check(self._lib.orcwa_mlp_set_error_scale(self._h, _fp(sc)),
                   "orcwa_mlp_set_error_scale")

    def last_ms(self) -> float:
        return float(self._lib.orcwa_mlp_last_ms(self._h))


class Surrogate:
    """Network + calibrated error estimate + plain-solver fallback."""

    def __init__(self, ctx: Context, members=4, width=512, depth=4, seed=1, tol=1e-3,
                 space=DEFAULT_SPACE):
        self.ctx = ctx
        self.space = space
        self.net = MLP(ctx, N_FEAT, width, depth, N_OUT, members, seed)
        self.tol = float(tol)
        self.kappa = None
        self.tau = None

    def fit(self, params, targets, epochs, batch=4096, lr_max=1e-3, lr_min=1e-5, seed=7,
            use_graph=True):
        self.net.set_data(features(params, self.space), targets)
        return self.net.train(epochs, batch, lr_max, lr_min, seed, use_graph)

    def calibrate(self, params, targets, margin=1.5, tol=None):
        """Threshold of the error estimate for tolerance tol (default: the
        surrogate's).  The estimate of a query is its largest member spread,
        max_c (max - min over members of output c); tau is the smallest
        estimate among the calibration queries whose answer is off by more than
        tol in some output, divided by margin, so every such calibration query
        is flagged.  (One scale for all outputs flags far fewer queries than a
        scale per output calibrated the same way: a query off in one output is
        often caught by the spread of another.)  The scale tol / tau goes to the
        network (orcwa_mlp_set_error_scale), whose estimate is then compared
        with tol.  Returns (tau, fallback fraction on this set)."""
        if tol is not None:
            self.tol = float(tol)
        x = features(params, self.space)
        y = self.net.forward(x).astype(np.float64)             # (members, rows, n_out)
        mean = y.mean(0)
        spread = (y.max(0) - y.min(0)).max(1)
        bad = (np.abs(mean - targets).max(1) > self.tol) & ~self._phys(mean)
        tau = float(spread[bad].min()) / margin if bad.any() else float(spread.max())
        self.tau = max(tau, 1e-30)
        self.kappa = np.full(N_OUT, self.tol / self.tau)
        self.net.set_error_scale(self.kappa)
        m, est = self.net.predict(x)
        flagged = self._flags(m, est)
        return self.tau, float(flagged.mean())

    def _phys(self, mean):
        n = len(ORDERS)
        tot = mean[:, :n].sum(1) + mean[:, n:].sum(1)
        return ((mean.min(1) < -self.tol) | (mean.max(1) > 1.0 + self.tol) |
                (tot > 1.0 + self.tol))

    def _flags(self, mean, est):
        # est: the scaled estimate (tol / tau) max_c spread_c of predict()
        return (est > self.tol) | self._phys(mean)

    def query(self, params):
        """Answer 12-parameter queries: network prediction, with the flagged ones
        re-solved by the plain GPU solver.  Returns (answers (count, 18), flagged
        mask, timings dict in seconds)."""
        if self.kappa is None:
            raise RuntimeError("calibrate() the error estimate before querying")
        p = np.asarray(params, dtype=np.float64)
        t0 = time.perf_counter()
        mean, spread = self.net.predict(features(p, self.space))
        t1 = time.perf_counter()
        flagged = self._flags(mean, spread)
        out = mean.astype(np.float64)
        if flagged.any():
            ans, _ = solve_queries(self.ctx, p[flagged])
            out[flagged] = ans
        t2 = time.perf_counter()
        return out, flagged, {"network_s": t1 - t0, "fallback_s": t2 - t1, "total_s": t2 - t0,
                              "network_gpu_ms": self.net.last_ms()}
