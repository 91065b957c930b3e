"""The oracle against the known answers of the reference's own unit tests.

Each case names the reference test it restates (tests/golden fixture file).
"""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle_lib import (OrTraversalCall, Patch, darr, iarr, invalid_patch, lib)

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "golden_reference_tests.json")))


def _coarse_fine(case):
    c, f = case["coarse"], case["fine"]
    coarse = Patch(1, 0, c["n"], c["gw"], c["pos"], c["size"], c["level"],
                   c["time"], c["dt"], c["min_neighbor_time"])
    fine = Patch(1, 0, f["n"], f["gw"], f["pos"], f["size"], f["level"],
                 f["time"], f["dt"], f["min_neighbor_time"])
    # GhostLayerCompositorTest.cpp:180-196: uNew = 1, uOld = -1 (with ghost
    # layer), then the four interior uOld values.
    coarse.unew[:] = c["unew_all"]
    coarse.uold[:] = c["uold_all"]
    for k, v in c["uold_interior"].items():
        x, y = map(int, k.split(","))
        coarse.set_uold(0, x, y, v)
    return coarse, fine


@pytest.mark.parametrize("name", ["interpolation_left_ghostlayer",
                                  "interpolation_right_ghostlayer"])
def test_interpolation_golden(name):
    case = GOLD[name]
    coarse, fine = _coarse_fine(case)
    inv = invalid_patch()
    slots = [inv, inv, inv, inv]
    slots[case["slots"]["fine"]] = fine
    for s in case["slots"]["coarse"]:
        slots[s] = coarse
    for i, p in enumerate(slots):
        if p is inv and i not in case["slots"]["coarse"] and i != case["slots"]["fine"]:
            slots[i] = invalid_patch()
    arr = (ctypes.POINTER(type(fine.s)) * 4)(*[ctypes.pointer(p.s) for p in slots])
    lib().or_fill_ghostlayers_and_update_neighbor_times(arr, case["level"], -1, 0)
    for k, (num, den) in case["expected_fine_uold"].items():
        x, y = map(int, k.split(","))
        assert fine.uold_at(0, x, y) == pytest.approx(num / den, rel=1e-14, abs=1e-14)


def test_overlapping_region_golden():
    for c in GOLD["overlapping_region"]["cases"]:
        v = lib().or_overlapping_region(darr(*c["p1"]), darr(*c["s1"]),
                                        darr(*c["p2"]), darr(*c["s2"]))
        assert v == pytest.approx(c["expected"], abs=1e-15)


@pytest.mark.parametrize("name", ["restriction_regions_overlapping_bounds",
                                  "restriction_regions_partial",
                                  "restriction_regions_infinite_lower"])
def test_restriction_regions_golden(name):
    c = GOLD[name]
    off = (ctypes.c_int * 8)()
    size = (ctypes.c_int * 8)()
    lib().or_regions_overlapped_by_ghostlayers(
        darr(*c["lower"]), darr(*c["upper"]), darr(*c["src_pos"]),
        darr(*c["src_size"]), darr(*c["src_subcell"]), iarr(*c["src_n"]),
        off, size)
    for r, (eo, es) in enumerate(c["expected"]):
        assert [off[2 * r], off[2 * r + 1]] == eo, (name, r)
        assert [size[2 * r], size[2 * r + 1]] == es, (name, r)


def test_traversal_orders_golden():
    f = (OrTraversalCall * 8)()
    e = (OrTraversalCall * 4)()
    assert lib().or_face_traversal_order(f) == 8
    assert lib().or_edge_traversal_order(e) == 4
    got_f = [[c.src, c.dst, c.dir[0], c.dir[1]] for c in f]
    got_e = [[c.src, c.dst, c.dir[0], c.dir[1]] for c in e]
    assert got_f == GOLD["face_traversal_order"]["expected"]
    assert got_e == GOLD["edge_traversal_order"]["expected"]


@pytest.mark.parametrize("veto", [0, 1, 2, 3])
def test_timestepping_veto_golden(veto):
    c = GOLD["timestepping_veto"]
    pc = c["patch"]
    patches = []
    for i in range(4):
        t = c["veto_time"] if i == veto else pc["time"]
        p = Patch(pc["nq"], 0, pc["n"], pc["gw"], pc["pos"], pc["size"],
                  pc["level"], t, pc["dt"], pc["min_neighbor_time"])
        lib().or_reset_min_neighbor_time_constraint(p.ref)
        patches.append(p)
    arr = (ctypes.POINTER(type(patches[0].s)) * 4)(
        *[ctypes.pointer(p.s) for p in patches])
    lib().or_fill_ghostlayers_and_update_neighbor_times(arr, 0, -1, 0)
    for i, p in enumerate(patches):
        assert p.s.min_neighbor_time_constraint == pytest.approx(
            c["expected_min_neighbor_time_constraint"])
        blocked = lib().or_is_blocked_by_neighbors(p.ref)
        assert bool(blocked) == (i == veto)


def test_coarse_patch_time_interval_golden():
    c = GOLD["coarse_patch_time_interval"]
    t = np.array(c["fine_times"])
    dt = np.array(c["fine_dts"])
    out = (ctypes.c_double * 2)()
    lib().or_minimal_fine_grid_interval(
        t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        dt.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(t), out)
    assert out[0] == pytest.approx(c["expected_time"])
    assert out[1] == pytest.approx(c["expected_dt"])
