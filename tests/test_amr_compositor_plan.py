"""The adaptive-grid driver's face plan against the reference's per-vertex
ghost-layer compositor, on the oracle.

peanoclaw_amd.amr.AmrGrid fills every leaf's face ghosts from a plan it
derives per face (same-level transfer, or interpolation from the coarser
leaf covering the region; amr.AmrGrid._face / _ghost_plan).  The reference
instead visits every vertex of every level with the four patches around it
(GhostLayerCompositor::fillGhostLayersAndUpdateNeighborTimes,
GhostLayerCompositor.cpp:280-320, and the face functor,
GhostlayerCompositorFunctors.cpp:23-77), where a quadrant without a patch
of that level holds the coarser patch covering it (Peano's hanging-vertex
adjacency).  Here the stencils are built from the geometry alone and run
through the oracle's compositor (faces, and corners by the edge functor,
:87-141); the whole ghost ring of every leaf must equal the driver's
bitwise.  Grids: the C1 and C3 refinement structures
and random three-level trees, with random data, b = 0 (so the SWE h <-> h+b
round trips are exact and the order of the calls cannot matter).
"""
import ctypes

import numpy as np
import pytest

from amr_oracle import OracleBackend
from oracle_lib import OrPatch, invalid_patch, lib


def _random_state(g, B, seed):
    rng = np.random.default_rng(seed)
    for pts in g.levels.values():
        for pt in pts.values():
            P = B.p[pt.handle]
            n = pt.n
            P.unew[0] = rng.uniform(0.5, 2.0, (n, n))
            P.unew[1:] = rng.uniform(-0.3, 0.3, (2, n, n))
            P.uold[:] = rng.uniform(0.5, 2.0, P.uold.shape)
            P.pwg[:] = 0.0
            P.s.time, P.s.dt = 0.3, 0.01


def _covering(g, lev, q):
    """The patch of level `lev` at index q, else the coarser patch covering
    that region; None outside the domain."""
    k = 3 ** (lev - 1)
    if not (0 <= q[0] < g.px * k and 0 <= q[1] < g.py * k):
        return None
    while lev >= 1:
        pt = g.levels.get(lev, {}).get(q)
        if pt is not None:
            return pt
        lev -= 1
        q = (q[0] // 3, q[1] // 3)
    return None


def _compositor_fill(g, B):
    """Restriction and uNew -> uOld as the driver does them, then the face
    (and corner) ghosts by the vertex sweep, then the domain walls."""
    for v in g.virtuals:
        B.clear(v.handle)
    for leaf, ancestors in g._restrictions:
        for v in ancestors:
            B.restrict(leaf.handle, v.handle)
    for v in g.virtuals:
        B.post_restrict(v.handle)
    for p in g.leaves:
        B.copy_unew_to_uold(p.handle)
    inv = invalid_patch()
    P4 = ctypes.POINTER(OrPatch) * 4
    calls = 0
    for lev in sorted(g.levels):
        verts = set()
        for (i, j) in g.levels[lev]:
            verts.update(((i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1)))
        for a, b in sorted(verts):
            # "1 0 / 3 2" around the vertex (GhostLayerCompositor.h:52-56)
            quad = ((a, b), (a - 1, b), (a, b - 1), (a - 1, b - 1))
            ps = []
            for q in quad:
                pt = _covering(g, lev, q)
                ps.append(B.p[pt.handle] if pt is not None else inv)
            lib().or_fill_ghostlayers_and_update_neighbor_times(
                P4(*[ctypes.pointer(p.s) for p in ps]), lev, -1, 1)
            calls += 1
    # SolveTimestep::fillBoundaryLayers: domain sides, left, top, right, bottom
    for p in g.leaves:
        k = 3 ** (p.level - 1)
        i, j = p.ij
        for dim, upper in ((0, 0), (1, 1), (0, 1), (1, 0)):
            idx = (i, j)[dim]
            last = (g.px, g.py)[dim] * k - 1
            if (idx == 0 and not upper) or (idx == last and upper):
                B.fill_boundary(p.handle, dim, upper)
    return calls


def _faces(u):
    """The ghost ring: both ghost rows and columns, corners included."""
    return (u[:, 0, :], u[:, -1, :], u[:, 1:-1, 0], u[:, 1:-1, -1])


def _check(build, seed):
    A, B = OracleBackend(), OracleBackend()
    ga, gb = build(A), build(B)
    _random_state(ga, A, seed)
    _random_state(gb, B, seed)
    before = {p.handle: A.uold(p.handle) for p in ga.leaves}
    # the library's rows of the same plan (pc_amr_face_plan, used by the
    # batched driver) equal the driver's
    for x, y in zip(ga.packed_plan_native(), ga.packed_plan()):
        assert np.array_equal(x, y)
    ga.fill_ghosts()
    calls = _compositor_fill(gb, B)
    n_int = n_interp = 0
    for kind, src, dst, *_ in ga._plan[0]:
        n_int += kind == "transfer"
        n_interp += kind != "transfer"
    for p in ga.leaves:
        ua, ub = A.uold(p.handle), B.uold(p.handle)
        for fa, fb, f0 in zip(_faces(ua), _faces(ub), _faces(before[p.handle])):
            assert np.array_equal(fa.view(np.uint64), fb.view(np.uint64)), (p.level, p.ij)
            assert (fa != f0).all(), (p.level, p.ij)  # every face ghost was written
    return len(ga.leaves), n_int, n_interp, calls


def test_c1_structure_face_plan_equals_compositor():
    from peanoclaw_amd.amr import TwoLevelGrid, dam_break_fields
    rng = np.random.default_rng(1)
    centre = tuple(0.5 + rng.uniform(-0.05, 0.05, 2))
    fn = dam_break_fields(centre=centre, radius=0.25)
    # the coarse patches the BreakingDam criterion refines on the initial
    # data (threshold 0.05, one level finer / the coarse width)
    refined = set()
    probe = OracleBackend()
    size, n = 1.0 / 27, 16
    for i in range(27):
        for j in range(27):
            s = probe.new_patch((i * size, j * size), size, n, 1, False)
            ii = np.arange(-1, n + 1)
            X = i * size + (ii[:, None] + 0.5) * (size / n)
            Y = j * size + (ii[None, :] + 0.5) * (size / n)
            X, Y = np.broadcast_arrays(X, Y)
            h, bb = fn(X, Y)
            probe.set_state(s, np.ascontiguousarray(h[1:-1, 1:-1]), bb)
            if probe.flags([s], 0.05, size / n / 3.0, size / n)[0] == 1:
                refined.add((i, j))
    assert 40 <= len(refined) < 400  # the initial ring (6.9% of the coarse patches)
    leaves, nt, ni, calls = _check(lambda B: TwoLevelGrid(B, 27, 27, n, refined, fn), 3)
    print(f"C1 structure: {leaves} leaves, {nt} transfer + {ni} interpolation faces, "
          f"{calls} vertex stencils")
    assert ni > 0


def test_c3_structure_face_plan_equals_compositor():
    from peanoclaw_amd.amr import AmrGrid, front_refinement, wet_dry_fields
    fn = wet_dry_fields()
    refine = front_refinement(9, 9, 27, fn, 3)
    leaves, nt, ni, calls = _check(lambda B: AmrGrid(B, 9, 9, 27, refine, fn, max_level=3), 4)
    print(f"C3 structure: {leaves} leaves, {nt} transfer + {ni} interpolation faces, "
          f"{calls} vertex stencils")
    assert ni > 0


@pytest.mark.parametrize("seed", range(6))
def test_random_three_level_trees_face_plan_equals_compositor(seed):
    from peanoclaw_amd.amr import AmrGrid, dam_break_fields
    rng = np.random.default_rng(100 + seed)
    px, py = int(rng.integers(2, 5)), int(rng.integers(2, 5))
    pick = {}

    def refine(lev, ij):
        key = (lev, ij)
        if key not in pick:
            pick[key] = bool(rng.random() < (0.45 if lev == 1 else 0.3))
        return pick[key]

    def build(B):
        return AmrGrid(B, px, py, 4, refine, dam_break_fields(), max_level=3)

    leaves, nt, ni, calls = _check(build, seed)
    assert leaves > 0
