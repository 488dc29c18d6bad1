"""Host side of uneven batches (no GPU): how engine._uneven_groups splits
independent stacks into batches, and that its padding — of cell edges (a
repeated edge = an empty cell) and of layers (zero-thickness copies of the
exit half-space) — leaves the physics unchanged, checked with the oracle on
the padded and unpadded stacks (RCWASolver::addLayer takes each stack's own
edges, rcwa/RCWASolver.cpp:108-137; the cascade over the layers,
rcwa/RCWASolver.cpp:378-492)."""
import numpy as np
import pytest

from oracle import rcwa_oracle as o
from openrcwa_amd import engine, workloads


def _stacks():
    return (workloads.uneven_stacks(30, n_range=(2, 4), layer_range=(1, 3)) +
            workloads.uneven_stacks(5, seed=9, n_range=(3, 3), layer_range=(2, 2)))


@pytest.mark.parametrize("pad_layers", [False, True])
def test_groups_cover_every_stack_once(pad_layers):
    stacks = _stacks()
    seen = []
    for g, idx in engine._uneven_groups(stacks, pad_layers):
        seen += list(idx)
        L = len(g["thickness"])
        mx, my = g["inst_coords"][0].shape[1], g["inst_coords"][1].shape[1]
        assert g["eps"].shape == (len(idx), L, my - 1, mx - 1)
        assert g["inst_thickness"].shape == (len(idx), L)
        for k, i in enumerate(idx):
            s_ = stacks[i]
            Ls = len(s_["thickness"])
            assert (s_["nX"], s_["nY"]) == (g["nX"], g["nY"])
            assert Ls == L if not pad_layers else Ls <= L
            cx, cy = g["inst_coords"][0][k], g["inst_coords"][1][k]
            assert np.all(np.diff(cx) >= 0) and np.all(np.diff(cy) >= 0)
            # the stack's own edges first, then repeats of its last edge
            assert np.array_equal(cx[:len(s_["coord_x"])], s_["coord_x"])
            assert np.all(cx[len(s_["coord_x"]):] == s_["coord_x"][-1])
            th = g["inst_thickness"][k]
            assert np.array_equal(th[:Ls - 1], s_["thickness"][:-1]) and th[-1] == 0.0
            assert np.all(th[Ls - 1:-1] == 0.0)
            ny0, nx0 = s_["eps"].shape[1:]
            e = g["eps"][k][:, :ny0, :nx0]
            assert np.array_equal(e[:Ls - 1], s_["eps"][:-1])
            assert np.all(e[Ls - 1:] == s_["eps"][-1])
    assert sorted(seen) == list(range(len(stacks)))


def test_groups_largest_first():
    stacks = _stacks()
    costs = [((2 * g["nX"] + 1) * (2 * g["nY"] + 1)) ** 3 * len(g["thickness"]) * len(idx)
             for g, idx in engine._uneven_groups(stacks)]
    assert costs == sorted(costs, reverse=True)


def test_padding_with_empty_cells_and_layers_is_exact():
    # stacks padded with empty cells and with zero-thickness exit-medium
    # layers: solving a padded instance gives the unpadded stack's R and T
    # order by order
    stacks = _stacks()
    checked = 0
    for g, idx in engine._uneven_groups(stacks):
        for k, i in enumerate(idx):
            s_ = stacks[i]
            cx, cy = g["inst_coords"][0][k], g["inst_coords"][1][k]
            if (cx.size == len(s_["coord_x"]) or cy.size == len(s_["coord_y"]) or
                    len(g["thickness"]) < len(s_["thickness"]) + 2):
                continue
            pols = [tuple(p) for p in s_["pol"]]
            args = (s_["lam"], s_["theta"], s_["phi"])
            R0, T0, _ = o.solve_stack(s_["nX"], s_["nY"], s_["period"], s_["period"],
                                      s_["coord_x"], s_["coord_y"], list(s_["eps"]),
                                      list(s_["thickness"]), *args, pols=pols)
            R1, T1, _ = o.solve_stack(g["nX"], g["nY"], s_["period"], s_["period"], cx, cy,
                                      list(g["eps"][k]), list(g["inst_thickness"][k]), *args,
                                      pols=pols)
            for q in range(2):
                assert np.abs(np.asarray(R1[q]) - R0[q]).max() < 1e-12
                assert np.abs(np.asarray(T1[q]) - T0[q]).max() < 1e-12
            checked += 1
            if checked >= 3:
                return
    assert checked > 0
