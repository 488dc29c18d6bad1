"""CPU tests of the .ofd front end rcwa_run (openrcwa_amd/csrc/ofd_bridge.cpp,
the reference's sol/rcwa_bridge.cpp:145-406) without a GPU: the front end is
host code over the batched C ABI, so it is compiled here with g++ against a
stand-in engine and stand-in HDF5 writers (tests/ofd_bridge_harness.cpp).

Checked: the batch handed to the engine (merged cell edges, eps1 rectangle
cells, the Lorentz pole form with einf from the last row, Bloch offsets, the
TE/TM basis), the CSV bytes (headers, %.8e rows, totals over all orders,
TE-then-TM order rows only for orders propagating on either side), the
warning line for a frequency with status bits, the invalid-frequency exit,
and the HDF5 hand-off (sol/rcwa_hdf5.c:91-241, include/rcwa_hdf5.h): the
meta record, rows [npol][nfreq], per-order maps [npol][nfreq][nx][ny] with
evanescent orders zeroed, and the log line when no writer is linked."""
import cmath
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
C0 = 2.99792458e8

SCEN = {
    0: dict(nx=3, ny=0, lx=3.0e-7, ly=3.0e-7, theta=0.0, phi=0.0,
            layers=[(1.0, 1.0, 0, 0, 0.5, 1.0, 0.0), (4.0, 1.0, 0, 0, 0.5, 1.0, 2e-7),
                    (2.25, 2.25, 0, 0, 0.5, 1.0, 0.0)],
            poles=[], freqs=[4.3e14, 6.5e14, 1.2e15]),
    1: dict(nx=2, ny=1, lx=4.0e-7, ly=5.0e-7, theta=25.0, phi=40.0,
            layers=[(1.0, 1.0, 0, 0, 0.5, 0.5, 0.0), (6.0, 1.0, 0.05, 0, 0.3, 0.6, 1.5e-7),
                    (2.0, 1.0, 0, 0, 0.7, 0.25, 0.8e-7), (2.25, 2.25, 0, 0, 0.5, 0.5, 0.0)],
            poles=[(2, 1, 1.5, 1.2e15, 0.1e15, 3.0e15), (2, 1, 2.0, 0.0, 0.0, 0.0),
                   (2, 1, 2.5, 0.8e15, 0.05e15, 5.0e15)],
            freqs=[5.0e14, 7.0e14]),
}


def fake(f, p, k, trn):
    return 1e-4 * (1 + k) + 1e-2 * p + 1e-1 * f + 5e-5 * trn


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("ofd")
    srcs = [os.path.join(ROOT, "openrcwa_amd", "csrc", "ofd_bridge.cpp"),
            os.path.join(ROOT, "tests", "ofd_bridge_harness.cpp")]
    exe = {}
    for name, extra in (("h5", []), ("noh5", ["-DNO_H5"])):
        out = str(d / f"harness_{name}")
        subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", *extra, *srcs, "-o", out],
                       check=True, capture_output=True, text=True)
        exe[name] = out
    return exe


def run(exe, scenario, tmp_path):
    r = subprocess.run([exe, str(scenario)], cwd=tmp_path, capture_output=True, text=True,
                       check=True)
    files = {p.name: p.read_text() for p in tmp_path.iterdir() if p.is_file()}
    return r.stdout, files


def edges(period, fills):
    cand = sorted([0.0, period] + [f * period for f in fills if f < 1.0])
    out = []
    for v in cand:
        if not out or v - out[-1] > 1e-12 * period:
            out.append(v)
    return out


def side_eps(sc, l, side, omega):
    lay = sc["layers"][l]
    rows = [q for q in sc["poles"] if q[0] == l and q[1] == side]
    if any(q[3] != 0.0 for q in rows):
        e = complex(rows[-1][2])
        for q in rows:
            if q[3] != 0.0:
                e += q[3] ** 2 / (q[5] ** 2 - omega ** 2 - 1j * q[4] * omega)
        return e
    return complex(lay[0], lay[2]) if side == 1 else complex(lay[1], lay[3])


def parse_batch(text):
    inst, cur = [], None
    for line in text.splitlines():
        w = line.split()
        if w[0] == "inst":
            cur = dict(lam=float(w[2]), kx0=float(w[3]), ky0=float(w[4]), pol=[], eps=[])
            inst.append(cur)
        elif w[0] == "pol":
            cur["pol"].append(float(w[1]))
        elif w[0] == "eps":
            cur["eps"].append(complex(float(w[1]), float(w[2])))
    return inst


@pytest.mark.parametrize("scenario", [0, 1])
def test_rcwa_run_batch_csv_and_hdf5(harness, tmp_path, scenario):
    sc = SCEN[scenario]
    out, files = run(harness["h5"], scenario, tmp_path)
    assert "rc 0" in out
    nfx, nfy = 2 * sc["nx"] + 1, 2 * sc["ny"] + 1
    h, nf = nfx * nfy, len(sc["freqs"])
    xe = edges(sc["lx"], [l[4] for l in sc["layers"]])
    ye = edges(sc["ly"], [l[5] for l in sc["layers"]])
    th, ph = math.radians(sc["theta"]), math.radians(sc["phi"])
    if abs(math.sin(th)) < 1e-9:
        te, tm = (0.0, 1.0), (1.0, 0.0)
    else:
        te = (-math.sin(ph), math.cos(ph))
        tm = (math.cos(th) * math.cos(ph), math.cos(th) * math.sin(ph))

    # ---- the batch handed to the engine
    inst = parse_batch(files["harness_batch.txt"])
    assert len(inst) == nf
    for f, fr in enumerate(sc["freqs"]):
        omega, lam = 2 * math.pi * fr, C0 / fr
        k0 = 2 * math.pi / lam
        assert abs(inst[f]["lam"] - lam) < 1e-15 * lam
        n_inc = max(cmath.sqrt(side_eps(sc, 0, 1, omega)).real, 1e-6)
        assert abs(inst[f]["kx0"] - k0 * n_inc * math.sin(th) * math.cos(ph)) < 1e-9 * k0
        assert abs(inst[f]["ky0"] - k0 * n_inc * math.sin(th) * math.sin(ph)) < 1e-9 * k0
        assert np.allclose(inst[f]["pol"], [te[0], 0, te[1], 0, tm[0], 0, tm[1], 0], atol=1e-15)
        eps = np.array(inst[f]["eps"]).reshape(len(sc["layers"]), len(ye) - 1, len(xe) - 1)
        for l, lay in enumerate(sc["layers"]):
            for iy in range(len(ye) - 1):
                for ix in range(len(xe) - 1):
                    xc, yc = 0.5 * (xe[ix] + xe[ix + 1]), 0.5 * (ye[iy] + ye[iy + 1])
                    inside = xc < lay[4] * xe[-1] and yc < lay[5] * ye[-1]
                    want = side_eps(sc, l, 1 if inside else 2, omega)
                    assert abs(eps[l, iy, ix] - want) < 1e-12 * abs(want)

    # ---- CSV outputs and the HDF5 hand-off
    csv = files["rcwa_efficiency.csv"].splitlines()
    assert csv[0] == "frequency[Hz],lambda[m],R_TE,T_TE,R_TM,T_TM"
    ords = files["rcwa_orders.csv"].splitlines()
    assert ords[0] == "frequency[Hz],lambda[m],pol,m,n,R,T"
    h5 = files["harness_h5.txt"].splitlines()
    assert h5[0] == "file time_series_data.h5"
    meta = h5[1][5:].split("|")
    assert meta[0] == "harness title" and int(meta[1]) == sc["nx"] and int(meta[3]) == sc["ny"]
    assert float(meta[2]) == sc["lx"] and float(meta[4]) == (sc["ly"] if sc["ny"] > 0 else 0.0)
    assert int(meta[5]) == len(sc["layers"])
    assert float(meta[6]) == sc["theta"] and float(meta[7]) == sc["phi"]
    assert h5[2] == f"dims 2 {nf}" and h5[3] == "label TE" and h5[4] == "label TM"
    rows = [list(map(float, r.split()[1:])) for r in h5[5:]]
    omap = files["harness_h5_orders.txt"].splitlines()
    assert omap[1] == f"dims 2 {nf} {nfx} {nfy}"
    maps = np.array([list(map(float, r.split())) for r in omap[2:]]).reshape(2, nf, nfx, nfy, 2)
    want_rows, line = [], 1
    for f, fr in enumerate(sc["freqs"]):
        omega, lam = 2 * math.pi * fr, C0 / fr
        k0 = 2 * math.pi / lam
        kx0, ky0 = inst[f]["kx0"], inst[f]["ky0"]
        tot = [[sum(fake(f, p, k, t) for k in range(h)) for t in (0, 1)] for p in (0, 1)]
        assert csv[1 + f] == ",".join("%.8e" % v for v in
                                      (fr, lam, tot[0][0], tot[0][1], tot[1][0], tot[1][1]))
        lim_r = k0 ** 2 * side_eps(sc, 0, 1, omega).real
        lim_t = k0 ** 2 * side_eps(sc, len(sc["layers"]) - 1, 1, omega).real
        for p, name in enumerate(("TE", "TM")):
            for k in range(h):
                m, n = k // nfy - sc["nx"], k % nfy - sc["ny"]
                kt2 = (kx0 + 2 * math.pi * m / sc["lx"]) ** 2 + (ky0 + 2 * math.pi * n / sc["ly"]) ** 2
                up, down = kt2 < lim_r, kt2 < lim_t
                r = fake(f, p, k, 0) if up else 0.0
                t = fake(f, p, k, 1) if down else 0.0
                assert maps[p, f, k // nfy, k % nfy, 0] == pytest.approx(r, abs=1e-16)
                assert maps[p, f, k // nfy, k % nfy, 1] == pytest.approx(t, abs=1e-16)
                if up or down:
                    assert ords[line] == "%.8e,%.8e,%s,%d,%d,%.8e,%.8e" % (fr, lam, name, m, n,
                                                                          r, t)
                    line += 1
        for p in (0, 1):
            want_rows.append((p, f, [fr, lam, tot[p][0], tot[p][1], 1 - tot[p][0] - tot[p][1]]))
    assert line == len(ords)
    for p, f, v in want_rows:
        assert np.allclose(rows[p * nf + f], v, rtol=1e-15, atol=1e-15)

    log = files["harness_log.txt"]
    assert "output : rcwa_efficiency.csv, rcwa_orders.csv" in log
    assert "output : time_series_data.h5" in log
    # the stand-in engine reports status bits for frequency 1
    assert ("*** warning : f=%.4e[Hz] solver status 48 (perturbed eigen recovery, growing mode, "
            "ill-conditioned cascade)" % sc["freqs"][1]) in log
    assert log.count("solver status") == 1
    if sc["poles"]:
        assert "RCWA : dispersive (Lorentz) materials : 1 (layer, eps1/eps2) pairs" in log


def test_rcwa_run_invalid_frequency(harness, tmp_path):
    out, files = run(harness["h5"], 2, tmp_path)
    assert "*** invalid frequency1 data" in out and "rc 1" in out
    assert "harness_h5.txt" not in files
    assert len(files["rcwa_efficiency.csv"].splitlines()) == 3      # header + 2 valid rows


def test_rcwa_run_without_hdf5_writer(harness, tmp_path):
    out, files = run(harness["noh5"], 0, tmp_path)
    assert "rc 0" in out
    assert "*** warning : time_series_data.h5 write failed (CSV is still valid)" in \
        files["harness_log.txt"]
