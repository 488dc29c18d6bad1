"""Host-side API over the C ABI: contexts, plans and batched solves.

A batch is many independent instances of one stack layout (same harmonics,
cell edges and layer count): wavelength, Bloch vector, per-layer permittivity
grid and incident polarisations vary per instance.  Lengths and wavelengths
share one unit, as in the reference (metres on the .ofd path,
sol/rcwa_bridge.cpp:15-16; micrometres on the .orcwa path).
"""
from __future__ import annotations

import ctypes
import os
import weakref
from ctypes import POINTER, c_double, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _lib

# autotuned kernel variants per layer-matrix size (orcwa_plan_autotune output)
TUNING_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.txt")


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


class Context:
    """One device + one stream (orcwa_gpu_create)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        self.device = int(device)
        self._plans = weakref.WeakSet()
        h = c_void_p()
        _lib.check(self._lib.orcwa_gpu_create(ctypes.byref(h), int(device)), "orcwa_gpu_create")
        self._h = h
        if os.path.exists(TUNING_FILE):
            # the committed table of autotuned kernel variants (read only)
            self.set_tuning_file(TUNING_FILE)

    @property
    def handle(self):
        if not getattr(self, "_h", None):
            raise _lib.EngineError("the context is closed")
        return self._h

    def set_workspace_limit(self, nbytes: int):
        """Device-memory budget of the plans later created on this context
        (orcwa_gpu_set_workspace_limit; 0 restores the default)."""
        _lib.check(self._lib.orcwa_gpu_set_workspace_limit(self.handle, int(nbytes)),
                   "orcwa_gpu_set_workspace_limit")

    def set_tuning_file(self, path, append=False):
        """Read autotuned kernel variants from `path` (and append later tuning
        results to it when append) -- orcwa_gpu_set_tuning_file."""
        _lib.check(self._lib.orcwa_gpu_set_tuning_file(self.handle, str(path).encode(),
                                                       1 if append else 0),
                   "orcwa_gpu_set_tuning_file")

    def set_gemm_precision(self, precision):
        """Precision of the star-product GEMMs of later plans and star products:
        "fp64", "fp32", "bf16x3" or "bf16x6" (orcwa_gpu_set_gemm_precision)."""
        prec = _lib.GEMM_PRECISIONS.get(precision, precision)
        _lib.check(self._lib.orcwa_gpu_set_gemm_precision(self.handle, int(prec)),
                   "orcwa_gpu_set_gemm_precision")

    def close(self):
        """Destroy the context; its outstanding plans are closed first (the C
        ABI refuses to destroy a context with live plans)."""
        if getattr(self, "_h", None):
            # graphs first: a plan used by a live graph refuses destruction
            for p in sorted(list(self._plans), key=lambda x: not isinstance(x, Graph)):
                p.close()
            _lib.check(self._lib.orcwa_gpu_destroy(self._h), "orcwa_gpu_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


@dataclass
class BatchResult:
    ref: np.ndarray      # (B, npol, h) per-order reflection efficiencies
    trn: np.ndarray      # (B, npol, h)
    amp: np.ndarray      # (B, 2, 4, h) complex basis responses (rx, ry, tx, ty)
    status: np.ndarray   # (B,) orcwa_status bits

    @property
    def R(self):
        return self.ref.sum(axis=-1)

    @property
    def T(self):
        return self.trn.sum(axis=-1)


class Plan:
    """orcwa_plan_*: inputs copied to device memory once; execute() runs the
    whole batch on the GPU."""

    def __init__(self, ctx: Context, nX: int, nY: int, coord_x, coord_y, thickness,
                 lam, kx0, ky0, eps, pols, mu=None, inst_coords=None, inst_thickness=None):
        self._lib = _lib.load()
        self.ctx = ctx
        self.coord_x = np.ascontiguousarray(coord_x, dtype=np.float64)
        self.coord_y = np.ascontiguousarray(coord_y, dtype=np.float64)
        self.thickness = np.ascontiguousarray(thickness, dtype=np.float64)
        L = self.thickness.size
        ncx, ncy = self.coord_x.size, self.coord_y.size
        lam = np.atleast_1d(np.asarray(lam, dtype=np.float64))
        B = lam.size
        kx0 = np.broadcast_to(np.asarray(kx0, dtype=np.float64), (B,))
        ky0 = np.broadcast_to(np.asarray(ky0, dtype=np.float64), (B,))
        eps = np.asarray(eps, dtype=np.complex128)
        shared = eps.ndim == 3
        if shared:
            eps = eps.reshape(1, L, ncy - 1, ncx - 1)
        if eps.shape[1:] != (L, ncy - 1, ncx - 1) or eps.shape[0] not in (1, B):
            raise ValueError(f"eps must be (B|1, L, ncy-1, ncx-1); got {eps.shape}")
        self.eps = np.ascontiguousarray(eps)
        self.mu = None
        if mu is not None:
            mu = np.asarray(mu, dtype=np.complex128)
            if mu.ndim == 3:
                mu = mu.reshape(1, L, ncy - 1, ncx - 1)
            if mu.shape[1:] != (L, ncy - 1, ncx - 1) or mu.shape[0] not in (1, B):
                raise ValueError(f"mu must be (B|1, L, ncy-1, ncx-1); got {mu.shape}")
            self.mu = np.ascontiguousarray(mu)
        pols = np.asarray(pols, dtype=np.complex128)
        if pols.ndim == 2:
            pols = np.broadcast_to(pols, (B,) + pols.shape)
        if pols.shape[0] != B or pols.shape[2] != 2 or not 1 <= pols.shape[1] <= 8:
            raise ValueError("pols must be (B, npol<=8, 2) complex (px, py)")
        self.npol = pols.shape[1]
        self.pols = np.ascontiguousarray(pols)
        self.B = B
        self.nX, self.nY = int(nX), int(nY)
        self.h = (2 * self.nX + 1) * (2 * self.nY + 1)

        st = _lib.OrcwaStack()
        st.nx_harm, st.ny_harm = self.nX, self.nY
        st.n_cx, st.n_cy = ncx, ncy
        st.coord_x, st.coord_y = _dptr(self.coord_x), _dptr(self.coord_y)
        st.n_layers = L
        st.thickness = _dptr(self.thickness)
        self._stack = st
        inst = (_lib.OrcwaInstance * B)()
        eps_flat = self.eps.view(np.float64).reshape(self.eps.shape[0], -1)
        pol_flat = self.pols.view(np.float64).reshape(B, -1)
        mu_flat = None if self.mu is None else self.mu.view(np.float64).reshape(self.mu.shape[0], -1)
        self._eps_flat, self._pol_flat, self._mu_flat = eps_flat, pol_flat, mu_flat
        for i in range(B):
            inst[i].lam = float(lam[i])
            inst[i].kx0 = float(kx0[i])
            inst[i].ky0 = float(ky0[i])
            inst[i].eps = _dptr(eps_flat[0 if shared or self.eps.shape[0] == 1 else i])
            inst[i].n_pol = self.npol
            inst[i].pol = _dptr(pol_flat[i])
            if mu_flat is not None:
                inst[i].mu = _dptr(mu_flat[0 if mu_flat.shape[0] == 1 else i])
        if inst_coords is not None:
            # per-instance cell edges (B, ncx) and (B, ncy)
            self._icx = np.ascontiguousarray(inst_coords[0], dtype=np.float64)
            self._icy = np.ascontiguousarray(inst_coords[1], dtype=np.float64)
            for i in range(B):
                inst[i].coord_x = _dptr(self._icx[i])
                inst[i].coord_y = _dptr(self._icy[i])
        if inst_thickness is not None:
            self._ith = np.ascontiguousarray(inst_thickness, dtype=np.float64)
            for i in range(B):
                inst[i].thickness = _dptr(self._ith[i])
        self._inst = inst
        h = c_void_p()
        _lib.check(self._lib.orcwa_plan_create(ctx.handle, ctypes.byref(st), inst, B,
                                               ctypes.byref(h)), "orcwa_plan_create")
        self._h = h
        ctx._plans.add(self)

    def execute(self):
        _lib.check(self._lib.orcwa_plan_execute(self._h), "orcwa_plan_execute")

    def sync(self):
        _lib.check(self._lib.orcwa_plan_sync(self._h), "orcwa_plan_sync")

    def autotune(self):
        """Time the plan with each kernel variant and keep the fastest
        (orcwa_plan_autotune).  Returns (qr_variant, gemm_tile)."""
        q, t = ctypes.c_int(), ctypes.c_int()
        _lib.check(self._lib.orcwa_plan_autotune(self._h, ctypes.byref(q), ctypes.byref(t)),
                   "orcwa_plan_autotune")
        return q.value, t.value

    def phase_ms(self, phase: int) -> float:
        return float(self._lib.orcwa_plan_phase_ms(self._h, int(phase)))

    def results(self) -> BatchResult:
        B, npol, h = self.B, self.npol, self.h
        ref = np.zeros((B, npol, h))
        trn = np.zeros((B, npol, h))
        amp = np.zeros((B, 2, 4, h), dtype=np.complex128)
        res = (_lib.OrcwaResult * B)()
        for i in range(B):
            res[i].ref = _dptr(ref[i])
            res[i].trn = _dptr(trn[i])
            res[i].amp = amp[i].ctypes.data_as(POINTER(c_double))
        _lib.check(self._lib.orcwa_plan_results(self._h, res), "orcwa_plan_results")
        status = np.array([res[i].status for i in range(B)], dtype=np.int32)
        return BatchResult(ref, trn, amp, status)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orcwa_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_batch(ctx: Context, nX, nY, coord_x, coord_y, thickness, lam, kx0, ky0, eps,
                pols, mu=None) -> BatchResult:
    """Plan + execute + results (orcwa_gpu_solve_batch semantics)."""
    p = Plan(ctx, nX, nY, coord_x, coord_y, thickness, lam, kx0, ky0, eps, pols, mu)
    try:
        p.execute()
        p.sync()
        return p.results()
    finally:
        p.close()


def _uneven_groups(stacks, pad_layers=True):
    """Group independent stacks by (Nx, Ny, periods) — and layer count unless
    pad_layers — into batches whose instances carry their own edges (padded to
    the group's edge count with empty cells: a repeated edge) and thicknesses.
    With pad_layers, a stack with fewer layers than its group gets
    zero-thickness copies of its (homogeneous) exit half-space inserted above
    it: an interface between equal media has E = H = I and a zero-thickness
    layer a unit phase, so each such cascade step passes R and Tdd through
    unchanged, and one batch per order count keeps the GPU full.  Largest
    groups (by n^3 x layers x stacks) first."""
    groups = {}
    for i, s_ in enumerate(stacks):
        if s_.get("mu") is not None and not np.all(np.asarray(s_["mu"]) == 1.0):
            raise ValueError(f"stack {i}: magnetic layers (mu != 1) are not supported by "
                             "solve_uneven; solve them with a Plan that takes mu")
        cx, cy = np.asarray(s_["coord_x"]), np.asarray(s_["coord_y"])
        key = (int(s_["nX"]), int(s_["nY"]), 0 if pad_layers else len(s_["thickness"]),
               float(cx[-1] - cx[0]), float(cy[-1] - cy[0]))
        groups.setdefault(key, []).append(i)
    nlay = {k: max(len(stacks[i]["thickness"]) for i in idx) for k, idx in groups.items()}
    order = sorted(groups.items(), key=lambda kv: -((2 * kv[0][0] + 1) * (2 * kv[0][1] + 1)) ** 3
                   * nlay[kv[0]] * len(kv[1]))
    for key, idx in order:
        nX, nY, L = key[0], key[1], nlay[key]
        mx = max(len(stacks[i]["coord_x"]) for i in idx)
        my = max(len(stacks[i]["coord_y"]) for i in idx)
        cxs, cys, eps, thick = [], [], [], []
        for i in idx:
            s_ = stacks[i]
            cx, cy = np.asarray(s_["coord_x"], float), np.asarray(s_["coord_y"], float)
            e = np.asarray(s_["eps"], complex)
            t = np.asarray(s_["thickness"], float)
            # empty cells at the far edges: repeat the last edge and the last
            # column / row of every grid
            px, py = mx - cx.size, my - cy.size
            cxs.append(np.concatenate([cx, np.full(px, cx[-1])]))
            cys.append(np.concatenate([cy, np.full(py, cy[-1])]))
            e = np.concatenate([e, np.repeat(e[:, :, -1:], px, axis=2)], axis=2)
            e = np.concatenate([e, np.repeat(e[:, -1:, :], py, axis=1)], axis=1)
            pl = L - t.size
            if pl > 0:
                if not np.all(e[-1] == e[-1].flat[0]):
                    raise ValueError("layer padding needs a homogeneous exit half-space")
                e = np.concatenate([e[:-1], np.repeat(e[-1:], pl, axis=0), e[-1:]], axis=0)
                t = np.concatenate([t[:-1], np.zeros(pl), t[-1:]])
            eps.append(e)
            thick.append(t)
        thick = np.array(thick)
        yield dict(nX=nX, nY=nY, coord_x=cxs[0], coord_y=cys[0], thickness=thick[0],
                   lam=[stacks[i]["lam"] for i in idx], kx0=[stacks[i]["kx0"] for i in idx],
                   ky0=[stacks[i]["ky0"] for i in idx], eps=np.stack(eps),
                   pol=np.stack([stacks[i]["pol"] for i in idx]),
                   inst_coords=(np.stack(cxs), np.stack(cys)), inst_thickness=thick), idx


def _device_free_bytes(device: int) -> int:
    """Free device memory (hipMemGetInfo), or 0 if it cannot be read."""
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
        if hip.hipSetDevice(int(device)) != 0:
            return 0
        if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
            return 0
        return int(free.value)
    except OSError:
        return 0


def _group_plan(ctx: Context, g):
    return Plan(ctx, g["nX"], g["nY"], g["coord_x"], g["coord_y"], g["thickness"], g["lam"],
                g["kx0"], g["ky0"], g["eps"], g["pol"], inst_coords=g["inst_coords"],
                inst_thickness=g["inst_thickness"])


def uneven_plans(ctx: Context, stacks, pad_layers=True):
    """Plans for independent stacks of different order counts, layer counts,
    cell layouts and thicknesses (dicts as workloads.uneven_stacks makes
    them), one per (Nx, Ny, periods) group (and layer count unless
    pad_layers).  Yields (plan, indices) one group at a time (a group's
    workspace is released when its plan is closed)."""
    for g, idx in _uneven_groups(stacks, pad_layers):
        yield _group_plan(ctx, g), idx


def solve_uneven(ctx: Context, stacks, streams: int = 1, pad_layers: bool = True):
    """Solve independent stacks of different sizes (grouped as
    _uneven_groups does).  Returns a list of (ref (npol, h), trn (npol, h),
    status) in the input order.  With streams > 1 the groups are spread over
    that many contexts (one stream each, on the device of ctx) driven from as
    many host threads, so small groups share the GPU instead of running one
    after another; their plans are then built one at a time, each context's
    workspace limited to an equal share of the device memory free at the
    start (unless ORCWA_WORKSPACE_GB is set)."""
    import os
    import queue
    import threading

    out = [None] * len(stacks)
    create_lock = threading.Lock()

    def run(c, g, idx):
        with create_lock:
            p = _group_plan(c, g)
        try:
            p.execute()
            p.sync()
            r = p.results()
        finally:
            p.close()
        for k, i in enumerate(idx):
            out[i] = (r.ref[k], r.trn[k], int(r.status[k]))

    if streams <= 1:
        for g, idx in _uneven_groups(stacks, pad_layers):
            run(ctx, g, idx)
        return out

    work = queue.Queue()
    for item in _uneven_groups(stacks, pad_layers):
        work.put(item)
    share = None
    if "ORCWA_WORKSPACE_GB" not in os.environ:
        free = _device_free_bytes(ctx.device)
        if free:
            share = 0.8 * free / streams / 1e9
    ctxs = [ctx] + [Context(ctx.device) for _ in range(streams - 1)]
    if share is not None:
        for c in ctxs:
            c.set_workspace_limit(int(share * 1e9))
    errors = []

    def worker(c):
        while not errors:
            try:
                g, idx = work.get_nowait()
            except queue.Empty:
                return
            try:
                run(c, g, idx)
            except Exception as e:      # first failure stops every worker
                errors.append(e)

    threads = [threading.Thread(target=worker, args=(c,)) for c in ctxs]
    try:
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    finally:
        if share is not None:
            ctx.set_workspace_limit(0)
        for c in ctxs[1:]:
            c.close()
    if errors:
        raise errors[0]
    return out


class Graph:
    """orcwa_graph_*: the executes of several plans captured once into one HIP
    graph and replayed with one launch."""

    def __init__(self, ctx: Context, plans, streams: int = 4):
        self._lib = _lib.load()
        self.ctx = ctx
        self.plans = list(plans)
        arr = (c_void_p * len(self.plans))(*[p_._h.value for p_ in self.plans])
        h = c_void_p()
        _lib.check(self._lib.orcwa_graph_create(ctx.handle, arr, len(self.plans), int(streams),
                                                ctypes.byref(h)), "orcwa_graph_create")
        self._h = h
        ctx._plans.add(self)

    def launch(self):
        _lib.check(self._lib.orcwa_graph_launch(self._h), "orcwa_graph_launch")

    def sync(self):
        _lib.check(self._lib.orcwa_graph_sync(self._h), "orcwa_graph_sync")

    @property
    def reruns(self) -> int:
        return int(self._lib.orcwa_graph_reruns(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.orcwa_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _group_bytes(g) -> float:
    """Device bytes an unchunked plan of a group would take (per instance: the
    patterned layers' seven n x n slots, the cascade's L + 3; per distinct grid
    ten h x h products)."""
    h = (2 * g["nX"] + 1) * (2 * g["nY"] + 1)
    n = 2 * h
    eps = g["eps"]
    L = eps.shape[1]
    flat = eps.reshape(eps.shape[0], L, -1)
    patterned = ~np.all(flat == flat[:, :, :1], axis=2)
    lp = int(patterned.sum(axis=1).max())
    slot = 16.0 * n * n
    return eps.shape[0] * (7 * lp + (L - lp) + L + 3) * slot + \
        int(patterned.sum()) * 10.0 * 16.0 * h * h


class UnevenBatch:
    """An uneven batch (stacks of different order counts, layer counts, cell
    layouts and thicknesses, grouped as _uneven_groups does) with every group's
    plan resident at once and their executes captured in one HIP graph: a
    solve is then one graph launch, the groups overlapping on `streams` forks.
    Each plan's workspace is a share of `frac` of the free device memory in
    proportion to what the group would take unchunked."""

    def __init__(self, ctx: Context, stacks, streams: int = 4, pad_layers: bool = True,
                 frac: float = 0.8):
        self.ctx = ctx
        self.count = len(stacks)
        groups = list(_uneven_groups(stacks, pad_layers))
        need = [_group_bytes(g) for g, _ in groups]
        free = _device_free_bytes(ctx.device)
        budget = frac * free if free else sum(need)
        total = sum(need)
        self.plans, self.index = [], []
        try:
            for (g, idx), b in zip(groups, need):
                ctx.set_workspace_limit(int(b if total <= budget else budget * b / total) + (1 << 20))
                self.plans.append(_group_plan(ctx, g))
                self.index.append(idx)
        finally:
            ctx.set_workspace_limit(0)
        self.graph = Graph(ctx, self.plans, streams)

    def run(self):
        """One solve of the whole batch: a graph launch, then the results as
        solve_uneven returns them."""
        self.graph.launch()
        self.graph.sync()
        out = [None] * self.count
        for p_, idx in zip(self.plans, self.index):
            r = p_.results()
            for k, i in enumerate(idx):
                out[i] = (r.ref[k], r.trn[k], int(r.status[k]))
        return out

    def close(self):
        self.graph.close()
        for p_ in self.plans:
            p_.close()
        self.plans = []


def field_image(ctx: Context, nX, nY, coord_x, coord_y, thickness, lam, kx0, ky0, eps, pol,
                component, slice_type, coord, nres1=0, mu=None):
    """Complex field of one instance on a slice (orcwa_gpu_field_image; the
    inputCoeffs form of RCWASolver::saveFieldImage, rcwa/RCWASolver.cpp:707-1011).
    eps (and mu): (L, ncy-1, ncx-1); pol = (px, py); component in "Ex".."Hz" or
    0..5; slice_type in "xy", "xz", "yz" or 0..2.  Returns (image (rows, cols)
    complex, status)."""
    comp = _lib.FIELD_COMPONENTS.get(component, component)
    sl = _lib.SLICES.get(slice_type, slice_type)
    lib = _lib.load()
    coord_x = np.ascontiguousarray(coord_x, dtype=np.float64)
    coord_y = np.ascontiguousarray(coord_y, dtype=np.float64)
    thickness = np.ascontiguousarray(thickness, dtype=np.float64)
    L = thickness.size
    eps = np.ascontiguousarray(np.asarray(eps, dtype=np.complex128).reshape(
        L, coord_y.size - 1, coord_x.size - 1))
    pols = np.ascontiguousarray(np.asarray(pol, dtype=np.complex128).reshape(1, 2))
    st = _lib.OrcwaStack()
    st.nx_harm, st.ny_harm = int(nX), int(nY)
    st.n_cx, st.n_cy = coord_x.size, coord_y.size
    st.coord_x, st.coord_y = _dptr(coord_x), _dptr(coord_y)
    st.n_layers = L
    st.thickness = _dptr(thickness)
    inst = _lib.OrcwaInstance()
    inst.lam, inst.kx0, inst.ky0 = float(lam), float(kx0), float(ky0)
    eps_flat = eps.view(np.float64).ravel()
    pol_flat = pols.view(np.float64).ravel()
    inst.eps = _dptr(eps_flat)
    inst.n_pol = 1
    inst.pol = _dptr(pol_flat)
    mu_flat = None
    if mu is not None:
        mu_flat = np.ascontiguousarray(np.asarray(mu, dtype=np.complex128).reshape(
            L, coord_y.size - 1, coord_x.size - 1)).view(np.float64).ravel()
        inst.mu = _dptr(mu_flat)
    req = _lib.OrcwaFieldRequest(int(comp), int(sl), float(coord), int(nres1), 0)
    rows, cols, status = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.check(lib.orcwa_gpu_field_image(ctx.handle, ctypes.byref(st), ctypes.byref(inst),
                                         ctypes.byref(req), ctypes.byref(rows),
                                         ctypes.byref(cols), None, None),
               "orcwa_gpu_field_image")
    img = np.zeros((rows.value, cols.value), dtype=np.complex128)
    _lib.check(lib.orcwa_gpu_field_image(ctx.handle, ctypes.byref(st), ctypes.byref(inst),
                                         ctypes.byref(req), ctypes.byref(rows),
                                         ctypes.byref(cols), _dptr(img.view(np.float64)),
                                         ctypes.byref(status)),
               "orcwa_gpu_field_image")
    return img, int(status.value)


def integrator_points(z0, z1, N):
    """The z of the samples the integrator takes, in its order: z1, the slab
    midpoints z0 + (i + 1/2) dz, z0 (core/RCWAIntegrator.cpp:22-75)."""
    dz = (z1 - z0) / float(N)
    return [z1] + [z0 + (i + 0.5) * dz for i in range(N)] + [z0]


def rcwa_integrate(ctx: Context, nX, nY, sample, lam, z0, z1, N, kx0=0.0, ky0=0.0):
    """A z-varying structure by the midpoint staircase (orcwa_gpu_rcwa_integrate;
    RCWAIntegrator::compute(N), core/RCWAIntegrator.cpp:20-86).  sample(z)
    returns (coord_x, coord_y, eps) of the cross-section at z (the
    VarLayerSampler of core/VarLayerSampler.cpp:52-68 with periodic K); every
    sample has the same number of edges and the same periods.  Returns a dict
    with Rud, Tuu, Rdu, Tdd, W0, V0, W1, V1 (n x n) and gam0, gam1 (n), and
    the status bits."""
    lib = _lib.load()
    pts = integrator_points(z0, z1, N)
    smp = [sample(z) for z in pts]
    cx = np.ascontiguousarray([np.asarray(s_[0], dtype=np.float64) for s_ in smp])
    cy = np.ascontiguousarray([np.asarray(s_[1], dtype=np.float64) for s_ in smp])
    eps = np.ascontiguousarray([np.asarray(s_[2], dtype=np.complex128).reshape(
        cy.shape[1] - 1, cx.shape[1] - 1) for s_ in smp])
    zs = _lib.OrcwaZStack(int(nX), int(nY), cx.shape[1], cy.shape[1], int(N), float(lam),
                          float(kx0), float(ky0), float(z0), float(z1), _dptr(cx), _dptr(cy),
                          _dptr(eps.view(np.float64)))
    n = 2 * (2 * int(nX) + 1) * (2 * int(nY) + 1)
    mats = {k: np.zeros((n, n), dtype=np.complex128)
            for k in ("Rud", "Tuu", "Rdu", "Tdd", "W0", "V0", "W1", "V1")}
    vecs = {k: np.zeros(n, dtype=np.complex128) for k in ("gam0", "gam1")}
    res = _lib.OrcwaZResult()
    for k, a in list(mats.items()) + list(vecs.items()):
        setattr(res, k, _dptr(a.view(np.float64)))
    _lib.check(lib.orcwa_gpu_rcwa_integrate(ctx.handle, ctypes.byref(zs), ctypes.byref(res)),
               "orcwa_gpu_rcwa_integrate")
    out = {k: a.T.copy() for k, a in mats.items()}    # column-major on the device
    out.update(vecs)
    out["status"] = int(res.status)
    return out


def _zresult(n):
    mats = {k: np.zeros((n, n), dtype=np.complex128)
            for k in ("Rud", "Tuu", "Rdu", "Tdd", "W0", "V0", "W1", "V1")}
    vecs = {k: np.zeros(n, dtype=np.complex128) for k in ("gam0", "gam1")}
    res = _lib.OrcwaZResult()
    for k, a in list(mats.items()) + list(vecs.items()):
        setattr(res, k, _dptr(a.view(np.float64)))
    return res, mats, vecs


def _zout(res, mats, vecs):
    out = {k: a.T.copy() for k, a in mats.items()}    # column-major on the device
    out.update(vecs)
    out["status"] = int(res.status)
    return out


def redheffer_integrate(ctx: Context, nX, nY, sample, lam, z0, z1, pe, kx0=0.0, ky0=0.0,
                        max_segments=0):
    """The adaptive first-order perturbation integrator
    (orcwa_gpu_redheffer_integrate; RedhefferIntegrator::compute,
    core/RedhefferIntegrator.cpp:173-399).  sample(z) -> (coord_x, coord_y, eps)
    as for rcwa_integrate, called at the points the integrator chooses.
    Returns the dict of rcwa_integrate plus N, the number of segments."""
    lib = _lib.load()
    cx0, cy0, _ = sample(z0)
    ncx, ncy = len(cx0), len(cy0)
    failure = []

    def cb(user, z, px, py, pe_):
        try:
            cx, cy, eps = sample(z)
            np.ctypeslib.as_array(px, shape=(ncx,))[:] = np.asarray(cx, dtype=np.float64)
            np.ctypeslib.as_array(py, shape=(ncy,))[:] = np.asarray(cy, dtype=np.float64)
            e = np.asarray(eps, dtype=np.complex128).reshape(ncy - 1, ncx - 1)
            np.ctypeslib.as_array(pe_, shape=(2 * e.size,))[:] = e.view(np.float64).ravel()
            return 0
        except Exception as exc:           # reported after the call returns
            failure.append(exc)
            return 1
    fn = _lib.SAMPLER_FN(cb)
    zs = _lib.OrcwaZSampler(int(nX), int(nY), ncx, ncy, float(lam), float(kx0), float(ky0),
                            float(z0), float(z1), fn, None)
    n = 2 * (2 * int(nX) + 1) * (2 * int(nY) + 1)
    res, mats, vecs = _zresult(n)
    nseg = ctypes.c_int(0)
    rc = lib.orcwa_gpu_redheffer_integrate(ctx.handle, ctypes.byref(zs), float(pe),
                                           int(max_segments), ctypes.byref(res),
                                           ctypes.byref(nseg))
    if failure:
        raise failure[0]
    _lib.check(rc, "orcwa_gpu_redheffer_integrate")
    out = _zout(res, mats, vecs)
    out["N"] = int(nseg.value)
    return out


def _cplx(a, shape):
    a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.complex128), shape))
    return a


def slab_smatrix(ctx: Context, W, V, gamma, W0, V0, lam, wz):
    """Slab S-matrices (Rud, Tuu) of a batch of slabs in reference bases
    (orcwa_gpu_slab_smatrix; dtmm::RCWAScatterMatrix::compute_v2,
    rcwa/RCWAScatterMatrix.cpp:88-147).  W, V, W0, V0: (B, n, n); gamma: (B, n);
    lam, wz: (B,) or scalars.  Returns (Rud, Tuu, info)."""
    lib = _lib.load()
    W = np.asarray(W, dtype=np.complex128)
    B, n, _ = W.shape
    # column-major per matrix: store the transpose of each row-major matrix
    cm = [np.ascontiguousarray(np.swapaxes(_cplx(x, (B, n, n)), 1, 2)) for x in (W, V, W0, V0)]
    g = _cplx(gamma, (B, n))
    lam = np.ascontiguousarray(np.broadcast_to(np.asarray(lam, dtype=np.float64), (B,)))
    wz = np.ascontiguousarray(np.broadcast_to(np.asarray(wz, dtype=np.float64), (B,)))
    Rud = np.zeros((B, n, n), dtype=np.complex128)
    Tuu = np.zeros((B, n, n), dtype=np.complex128)
    info = np.zeros(B, dtype=np.int32)
    _lib.check(lib.orcwa_gpu_slab_smatrix(
        ctx.handle, n, B, _dptr(lam), _dptr(wz), _dptr(cm[0].view(np.float64)),
        _dptr(cm[1].view(np.float64)), _dptr(g.view(np.float64)), _dptr(cm[2].view(np.float64)),
        _dptr(cm[3].view(np.float64)), _dptr(Rud.view(np.float64)), _dptr(Tuu.view(np.float64)),
        info.ctypes.data_as(POINTER(ctypes.c_int))), "orcwa_gpu_slab_smatrix")
    return np.swapaxes(Rud, 1, 2), np.swapaxes(Tuu, 1, 2), info


def redheffer_star(ctx: Context, top, bottom):
    """Star product top * bottom of batches of S-matrices given as
    (rud, tuu, rdu, tdd) arrays of shape (B, n, n) each (orcwa_gpu_redheffer_star;
    core/RCWAIntegrator.cpp:57-67).  Returns ((rud, tuu, rdu, tdd), info)."""
    lib = _lib.load()
    t = np.stack([np.asarray(x, dtype=np.complex128) for x in top], axis=1)
    b = np.stack([np.asarray(x, dtype=np.complex128) for x in bottom], axis=1)
    B, _, n, _ = t.shape
    tc = np.ascontiguousarray(np.swapaxes(t, 2, 3))
    bc = np.ascontiguousarray(np.swapaxes(b, 2, 3))
    out = np.zeros_like(tc)
    info = np.zeros(B, dtype=np.int32)
    _lib.check(lib.orcwa_gpu_redheffer_star(
        ctx.handle, n, B, _dptr(tc.view(np.float64)), _dptr(bc.view(np.float64)),
        _dptr(out.view(np.float64)), info.ctypes.data_as(POINTER(ctypes.c_int))),
        "orcwa_gpu_redheffer_star")
    o = np.swapaxes(out, 2, 3)
    return (o[:, 0], o[:, 1], o[:, 2], o[:, 3]), info
