"""Pins of the restriction's [0, 1] time basis in the oracle.

DefaultRestriction restricts a fine leaf into a virtual coarse patch at the
fixed times 0 and 1 (DefaultRestriction.cpph:35-52: tfOld = (0 - tOld)/dt,
tfNew = (1 - tOld)/dt), and DefaultTransfer reads a virtual source at
time factor dst.timeUNew / 1 (DefaultTransfer.cpph:30-41).  Composed, the
two reproduce a source that is linear in time at the destination's time:
ghost = u(T) = f + (T - t0) g for uOld = f, uNew = f + dt g.  A swapped
pair of time factors, or the destination's own interval in place of
[0, 1], gives another time (1 - T, or T - t0 + ...) and fails these checks.

The reference's own known-answer test of this chain
(GridLevelTransferTest::testRestrictionToVirtualPatch) is replayed as well:
its geometry (which fine cells overlap which coarse ghost cell, by how
much) is pinned by its closed forms, and the time weights it expects are
recorded next to the ones the present code produces.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from oracle_lib import Patch, darr, iarr, invalid_patch, lib

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "golden_reference_tests.json")))


def _patch(c, nq, gw, virtual=False):
    p = Patch(nq, 0, c["n"], gw, c["pos"], c["size"], c["level"], c["time"], c["dt"],
              c["min_neighbor_time"])
    if virtual:
        p.s.is_virtual = 1
        p.s.is_leaf = 0
        p.s.min_neighbor_time = c["min_neighbor_time"]
        p.s.max_neighbor_dt = c["max_neighbor_dt"]
    return p


def _weights(case):
    """Area weight of every fine cell in every coarse ghost cell of the
    neighbour (the virtual cells (0..1, 0..1) it mirrors), from the
    oracle's overlap function (pinned by GridLevelTransferTest's overlap
    areas)."""
    f, v = case["fine"], case["virtual"]
    fs = f["size"] / f["n"]
    vs = v["size"] / v["n"]
    w = {}
    for gx, vx in ((-2, 0), (-1, 1)):
        for gy in (0, 1):
            cells = []
            for x in range(f["n"]):
                for y in range(f["n"]):
                    a = lib().or_overlapping_region(
                        darr(v["pos"][0] + vx * vs, v["pos"][1] + gy * vs), darr(vs, vs),
                        darr(f["pos"][0] + x * fs, f["pos"][1] + y * fs), darr(fs, fs))
                    if a > 0:
                        cells.append((x, y, a / (vs * vs)))
            w[f"{gx},{gy}"] = cells
    return w


def _at_weight(case, cells, wnew):
    un = case["fine"]["unew_xy"]
    uo = case["fine"]["uold_xy"]
    return sum(a * (wnew * un[x][y] + (1.0 - wnew) * uo[x][y]) for x, y, a in cells)


def test_restriction_to_virtual_patch_geometry_matches_reference_closed_forms():
    case = GOLD["restriction_to_virtual_patch"]
    w = _weights(case)
    # the reference's four values, with the time weight each of them uses:
    # the first weighs uNew by 2 (2 uNew - uOld), the other three by 4/3
    used = {"-2,0": 2.0, "-2,1": 4.0 / 3.0, "-1,0": 4.0 / 3.0, "-1,1": 4.0 / 3.0}
    for k, cells in w.items():
        assert len(cells) == 4
        assert sorted(a for _, _, a in cells) == pytest.approx(
            [case["region_fraction"] * f for f in (0.25, 0.5, 0.5, 1.0)], rel=1e-14)
        assert _at_weight(case, cells, used[k]) == pytest.approx(
            case["expected_neighbour_uold"][k], rel=1e-13, abs=1e-15)


def test_restriction_to_virtual_patch_chain():
    """stepUp's restriction of the fine leaf (GridLevelTransfer.cpp:140-198)
    and finalizeVirtualSubgrid's ghost fill at the virtual patch's vertices
    (:222-236) on the reference test's patches: the neighbour's ghost cells
    receive the fine data at the neighbour's time 2 = the fine uNew (time
    weight 1), not the weights the reference test lists (2 and 4/3, which
    no single time factor of the present code yields)."""
    case = GOLD["restriction_to_virtual_patch"]
    gw = case["gw"]
    vp = _patch(case["virtual"], 1, gw, virtual=True)
    nb = _patch(case["neighbour"], 1, gw)
    fp = _patch(case["fine"], 1, gw)
    fp.unew[0] = np.array(case["fine"]["unew_xy"])
    fp.uold[0, gw:-gw, gw:-gw] = np.array(case["fine"]["uold_xy"])
    # the neighbour's two-cell ghost layer (x in [0, 1]) covers the whole
    # fine patch, so the overlapped-regions restriction is the full one
    lib().or_restrict_default(fp.ref, vp.ref, 0)
    # the virtual patch holds the fine data at times 0 and 1 (linear in time)
    un = np.array(case["fine"]["unew_xy"])
    uo = np.array(case["fine"]["uold_xy"])
    t0, dt = case["fine"]["time"], case["fine"]["dt"]
    w = _weights(case)
    for k, cells in w.items():
        vx = 0 if k.startswith("-2") else 1
        vy = int(k.split(",")[1])
        old = sum(a * (uo[x, y] + (0.0 - t0) / dt * (un[x, y] - uo[x, y])) for x, y, a in cells)
        new = sum(a * (uo[x, y] + (1.0 - t0) / dt * (un[x, y] - uo[x, y])) for x, y, a in cells)
        assert vp.uold_at(0, vx, vy) == pytest.approx(old, rel=1e-13, abs=1e-14)
        assert vp.unew[0, vx, vy] == pytest.approx(new, rel=1e-13, abs=1e-14)
    # ghost fill at the virtual patch's vertices 1 and 3 (patches "1 0 / 3 2")
    for vertex in ("1", "3"):
        slots = [invalid_patch() for _ in range(4)]
        for s, which in case["vertex_patches"][vertex].items():
            slots[int(s)] = {"neighbour": nb, "virtual": vp}[which]
        arr = (ctypes.POINTER(type(vp.s)) * 4)(*[ctypes.pointer(p.s) for p in slots])
        lib().or_fill_ghostlayers_and_update_neighbor_times(arr, 0, -1, 0)
    for k, cells in w.items():
        gx, gy = map(int, k.split(","))
        got = nb.uold_at(0, gx, gy)
        assert got == pytest.approx(_at_weight(case, cells, 1.0), rel=1e-13)
        assert got != pytest.approx(case["expected_neighbour_uold"][k], rel=1e-3)


@pytest.mark.parametrize("t0,dt", [(0.37, 0.05), (2.5, 0.2), (0.0, 0.125)])
def test_restriction_time_basis_closed_form(t0, dt):
    """uOld = f, uNew = f + dt g on a 9x9 fine leaf over [t0, t0 + dt];
    restrict into a virtual 3x3 patch, then transfer from it into a
    neighbour's ghost column at the neighbour's time T: the ghost holds
    the area average of f + (T - t0) g, for T inside, before and after the
    source interval."""
    r = np.random.default_rng(7)
    f = r.uniform(0.5, 2.0, (9, 9))
    g = r.uniform(-3.0, 3.0, (9, 9))
    fine = Patch(1, 0, 9, 1, 1.0 / 3, 1.0 / 3, 1, t0, dt)
    fine.uold[0, 1:-1, 1:-1] = f
    fine.unew[0] = f + dt * g
    virt = Patch(1, 0, 3, 1, 0.0, 1.0, 0, 0.0, 0.0)
    virt.s.is_virtual = 1
    virt.s.is_leaf = 0
    lib().or_restrict_default(fine.ref, virt.ref, 0)
    # the fine leaf covers the virtual cell (1, 1) exactly
    avg_f = f.mean()
    avg_g = g.mean()
    assert virt.uold[0, 2, 2] == pytest.approx(avg_f + (0.0 - t0) * avg_g, rel=1e-12, abs=1e-12)
    assert virt.unew[0, 1, 1] == pytest.approx(avg_f + (1.0 - t0) * avg_g, rel=1e-12, abs=1e-12)
    for T in (t0 + 0.5 * dt, t0 + dt, t0 - 0.3, t0 + 4.0 * dt):
        dst_dt = 0.01
        nb = Patch(1, 0, 3, 1, 1.0, 1.0, 0, T - dst_dt, dst_dt)
        # virtual cell column x = 1 -> the neighbour's lower ghost column
        lib().or_transfer_ghostlayer(iarr(1, 3), iarr(1, 0), iarr(-1, 0), virt.ref, nb.ref)
        want = avg_f + (T - t0) * avg_g
        tol = 1e-12 * (1.0 + abs(T) + abs(t0)) / max(dt, 1e-3)
        assert nb.uold_at(0, -1, 1) == pytest.approx(want, abs=tol)


def test_restriction_zero_timestep_uses_unit_factors():
    """dt = 0 leaves both time factors at 1 (DefaultRestriction.cpph:45-50):
    the virtual patch receives uNew in both arrays."""
    fine = Patch(1, 0, 3, 1, 1.0 / 3, 1.0 / 3, 1, 0.7, 0.0)
    fine.uold[0, 1:-1, 1:-1] = 5.0
    fine.unew[0] = 2.0
    virt = Patch(1, 0, 3, 1, 0.0, 1.0, 0, 0.0, 0.0)
    virt.s.is_virtual = 1
    virt.s.is_leaf = 0
    lib().or_restrict_default(fine.ref, virt.ref, 0)
    assert virt.unew[0, 1, 1] == pytest.approx(2.0, rel=1e-15)
    assert virt.uold[0, 2, 2] == pytest.approx(2.0, rel=1e-15)


@pytest.mark.parametrize("t1,dt1", [(0.41, 0.02), (0.37, 0.0), (3.0, 0.5)])
def test_switch_to_leaf_evaluates_the_restricted_basis_at_the_patch_times(t1, dt1):
    """Patch::switchToLeaf (Patch.cpp:793-830): a virtual patch restricted
    from a linear-in-time leaf over [t0, t0 + dt] becomes a leaf whose uOld
    and uNew are the leaf data's average at t1 and t1 + dt1."""
    t0, dt = 0.37, 0.05
    r = np.random.default_rng(11)
    f = r.uniform(0.5, 2.0, (9, 9))
    g = r.uniform(-3.0, 3.0, (9, 9))
    fine = Patch(1, 0, 9, 1, 1.0 / 3, 1.0 / 3, 1, t0, dt)
    fine.uold[0, 1:-1, 1:-1] = f
    fine.unew[0] = f + dt * g
    virt = Patch(1, 0, 3, 1, 0.0, 1.0, 0, 0.0, 0.0)
    virt.s.is_virtual = 1
    virt.s.is_leaf = 0
    lib().or_restrict_default(fine.ref, virt.ref, 0)
    before_old, before_new = np.array(virt.uold), np.array(virt.unew)
    virt.s.time, virt.s.dt = t1, dt1
    lib().or_switch_to_leaf(virt.ref)
    assert virt.s.is_leaf == 1 and virt.s.is_virtual == 0
    tol = 1e-12 * (1.0 + t1 + t0) / dt
    assert virt.uold[0, 2, 2] == pytest.approx(f.mean() + (t1 - t0) * g.mean(), abs=tol)
    assert virt.unew[0, 1, 1] == pytest.approx(f.mean() + (t1 + dt1 - t0) * g.mean(), abs=tol)
    # the ghost ring of uOld is left alone
    assert np.array_equal(virt.uold[0, 0], before_old[0, 0])


@pytest.mark.parametrize("sides", [("left",), ("right", "top"), ("left", "bottom", "top")])
def test_only_overlapped_restriction_equals_full_restriction_where_read(sides):
    """GridLevelTransfer restricts only the regions of a leaf that coarse
    neighbours' ghost layers overlap (restrictSolution(onlyOverlapped),
    DefaultRestriction.cpph:108-150, Region.cpp:244-315); the batched step
    restricts whole leaves.  With refinement factor 3 the regions align with
    coarse cells, so every coarse cell a neighbour's ghost layer reads holds
    bitwise the same value either way."""
    r = np.random.default_rng(len(sides))
    n = 6
    sub = 1.0 / n
    # bounds of the region overlapped by ghost layers (Patch.cpp:916-941):
    # `upper` from the lower sides, `lower` from the upper sides
    BIG = np.finfo(np.float64).max
    upper, lower = [-BIG, -BIG], [BIG, BIG]
    read = np.zeros((n, n), dtype=bool)
    if "left" in sides:
        upper[0] = sub
        read[0, :] = True
    if "bottom" in sides:
        upper[1] = sub
        read[:, 0] = True
    if "right" in sides:
        lower[0] = 1.0 - sub
        read[-1, :] = True
    if "top" in sides:
        lower[1] = 1.0 - sub
        read[:, -1] = True
    results = []
    for only in (1, 0):
        virt = Patch(1, 0, n, 1, 0.0, 1.0, 0, 0.0, 0.0)
        virt.s.is_virtual = 1
        virt.s.is_leaf = 0
        for d in range(2):
            virt.s.lower_ghost_bounds[d] = lower[d]
            virt.s.upper_ghost_bounds[d] = upper[d]
        rr = np.random.default_rng(5)
        for a in range(3):
            for c in range(3):
                leaf = Patch(1, 0, n, 1, (a / 3.0, c / 3.0), 1.0 / 3.0, 1, 0.25, 0.05)
                leaf.unew[0] = rr.uniform(0.5, 2.0, (n, n))
                leaf.uold[0, 1:-1, 1:-1] = rr.uniform(0.5, 2.0, (n, n))
                lib().or_restrict_default(leaf.ref, virt.ref, only)
        results.append((np.array(virt.unew[0]), np.array(virt.uold[0, 1:-1, 1:-1])))
    (pn, po), (fn, fo) = results
    assert read.any()
    assert np.array_equal(pn[read].view(np.uint64), fn[read].view(np.uint64))
    assert np.array_equal(po[read].view(np.uint64), fo[read].view(np.uint64))
    assert (pn[~read] == 0.0).all()  # the rest is left alone with onlyOverlapped
