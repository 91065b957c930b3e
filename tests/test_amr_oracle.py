"""The two-level AMR driver on the oracle backend (CPU).

With every coarse patch refined and flat bathymetry the driver's sequence
(transfers between fine patches, walls, global CFL step, solveTimestep per
patch) must reproduce the uniform fine grid of the pool driver bit for bit.
"""
import ctypes
import math

import numpy as np

from amr_oracle import OracleBackend
from oracle_lib import OrPool, dptr, lib
import pytest

from peanoclaw_amd.amr import AmrGrid, TwoLevelGrid, dam_break_fields, wet_dry_fields


@pytest.mark.parametrize("levels,px,n", [(2, 2, 8), (3, 1, 4)])
def test_fully_refined_grid_equals_uniform_fine_grid(levels, px, n):
    fn = dam_break_fields(centre=(0.47, 0.52), radius=0.25)
    grid = AmrGrid(OracleBackend(), px, px, n, lambda lev, ij: True, fn, max_level=levels)
    finest = grid.levels[levels]
    assert len(grid.leaves) == len(finest) == (px * 3 ** (levels - 1)) ** 2
    # the same fine grid as one pool, ghost fill by or_grid_fill
    P = px * 3 ** (levels - 1)
    s = n + 2
    H = [np.zeros(P * P * s * s), np.zeros(P * P * s * s)]
    HU = [np.zeros_like(H[0]), np.zeros_like(H[0])]
    HV = [np.zeros_like(H[0]), np.zeros_like(H[0])]
    B = np.zeros_like(H[0])
    slot_of = np.arange(P * P, dtype=np.int32)
    for (I, J), fp in finest.items():
        k = I * P + J
        u = grid.b.state(fp.handle)
        H[0][k * s * s:(k + 1) * s * s].reshape(s, s)[1:-1, 1:-1] = u[0]
    off = np.array([k * s * s for k in range(P * P)], dtype=np.int64)
    nx = np.full(P * P, n, dtype=np.int32)
    dx = np.full(P * P, 1.0 / (P * n))
    pl = OrPool()
    pl.np = P * P
    pl.nx = pl.ny = nx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    pl.off = off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
    pl.dx = pl.dy = dx.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for k in range(2):
        pl.h[k], pl.hu[k], pl.hv[k] = dptr(H[k]), dptr(HU[k]), dptr(HV[k])
    pl.b = dptr(B)
    pl.cur = 0
    for _ in range(12):
        c = pl.cur
        lib().or_grid_fill(P, P, n, slot_of.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                           dptr(H[c]), dptr(HU[c]), dptr(HV[c]))
        d_pool = lib().or_pool_step(ctypes.byref(pl), math.inf, 1)
        d_amr = grid.step()
        assert d_amr == d_pool
    c = pl.cur
    for (I, J), fp in finest.items():
        k = I * P + J
        u = grid.b.state(fp.handle)
        for q, arr in enumerate((H[c], HU[c], HV[c])):
            ref = arr[k * s * s:(k + 1) * s * s].reshape(s, s)[1:-1, 1:-1]
            np.testing.assert_array_equal(u[q].view(np.uint64), ref.view(np.uint64))


def test_ghost_fill_reproduces_a_linear_field():
    """Interpolated fine ghosts (coarse -> fine) and transferred coarse ghosts
    (from restricted virtual patches) of a linear field equal the field at
    the ghost cell centres: the region and weight arithmetic is placed
    correctly on both levels."""
    ax, ay = 0.7, -0.4
    hills = [(0.4, 0.6, 0.15, 0.3)]

    def fn(X, Y):
        b = np.zeros(X.shape)
        for x, y, s, a in hills:
            b = b + a * np.exp(-((X - x) ** 2 + (Y - y) ** 2) / (s * s))
        return 2.0 + ax * X + ay * Y - b, b

    be = OracleBackend()
    grid = TwoLevelGrid(be, 4, 4, 6, {(1, 1), (2, 1), (1, 2)}, fn)
    _check_linear_ghosts(grid, be, ax, ay)


def test_three_level_ghost_fill_reproduces_a_linear_field():
    """As above with three levels, including level-3 patches whose face
    neighbour is a level-1 leaf (ghosts interpolated across two levels) and
    level-2 virtual patches restricted from level-3 children."""
    ax, ay = -0.3, 0.55
    hills = [(0.5, 0.4, 0.2, 0.25)]

    def fn(X, Y):
        b = np.zeros(X.shape)
        for x, y, s, a in hills:
            b = b + a * np.exp(-((X - x) ** 2 + (Y - y) ** 2) / (s * s))
        return 2.0 + ax * X + ay * Y - b, b

    refine = {(1, (1, 1)), (1, (2, 1)), (2, (3, 3)), (2, (5, 4)), (2, (6, 3))}
    be = OracleBackend()
    grid = AmrGrid(be, 4, 4, 6, lambda lev, ij: (lev, ij) in refine, fn, max_level=3)
    kinds = {(p.level, grid._face(p, d, u)[1].level if grid._face(p, d, u)[1] else 0)
             for p in grid.leaves for d, u in ((0, 0), (0, 1), (1, 0), (1, 1))}
    assert (3, 1) in kinds and (3, 2) in kinds and (2, 1) in kinds
    _check_linear_ghosts(grid, be, ax, ay)


def _check_linear_ghosts(grid, be, ax, ay):
    grid.fill_ghosts()
    checked = 0
    for p in grid.leaves:
        u = be.uold(p.handle)
        b = be.p[p.handle].pwg[0]
        sub = p.size / p.n
        for dim, upper in ((0, 0), (0, 1), (1, 0), (1, 1)):
            kind, _ = grid._face(p, dim, upper)
            if kind == "wall":
                continue
            for k in range(p.n):
                x, y = ((p.n + 1 if upper else 0), k + 1) if dim == 0 else \
                       (k + 1, (p.n + 1 if upper else 0))
                X = p.pos[0] + (x - 1 + 0.5) * sub
                Y = p.pos[1] + (y - 1 + 0.5) * sub
                surface = u[0, x, y] + b[x, y]
                assert abs(surface - (2.0 + ax * X + ay * Y)) < 1e-12, (kind, p.level, p.ij)
                assert u[1, x, y] == 0.0 and u[2, x, y] == 0.0
                checked += 1
    assert checked > 100


def test_three_level_wet_dry_grid_runs_and_stays_finite():
    """Wet/dry basin on three levels: finite, non-negative depths and
    a positive step while the column collapses onto dry land."""
    fn = wet_dry_fields()
    refine = {(1, (0, 1)), (1, (1, 1)), (1, (2, 2)), (2, (2, 4)), (2, (3, 4)), (2, (7, 7))}
    grid = AmrGrid(OracleBackend(), 3, 3, 6, lambda lev, ij: (lev, ij) in refine, fn,
                   max_level=3)
    assert {p.level for p in grid.leaves} == {1, 2, 3}
    for _ in range(8):
        assert grid.step() > 0
    st = grid.leaf_states()
    assert all(np.isfinite(v).all() for v in st.values())
    assert all((v[0] >= 0.0).all() for v in st.values())
    assert any((v[0] == 0.0).any() for v in st.values())


def test_two_level_grid_runs_and_stays_finite():
    fn = dam_break_fields(centre=(0.5, 0.5), radius=0.25, hills=[(0.3, 0.6, 0.1, 0.2)])
    grid = TwoLevelGrid(OracleBackend(), 4, 4, 8, {(1, 1), (1, 2), (2, 1)}, fn)
    m = []
    for _ in range(10):
        assert grid.step() > 0
        st = grid.leaf_states()
        assert all(np.isfinite(v).all() for v in st.values())
        m.append(sum(float(v[0].sum()) * ((1 / 32) ** 2 if k[0] == 1 else (1 / 96) ** 2)
                     for k, v in st.items()))
    # no flux correction (off by default in the reference): mass moves by the
    # coarse/fine flux mismatch where the dam-break front crosses an interface
    assert abs(m[-1] - m[0]) / m[0] < 5e-3


@pytest.mark.parametrize("levels", [2, 3])
def test_pure_level_transfers_track_the_reference_call_sequence(levels):
    """The batched semantics (level transfers read h + b, sources left
    unchanged) against the reference's in-place transform round trips:
    within 1e-12 relative per step and 1e-10 after 150 steps."""
    fn = wet_dry_fields() if levels == 3 else dam_break_fields(
        centre=(0.47, 0.53), radius=0.25, hills=[(0.3, 0.6, 0.1, 0.25)])
    refine = ({(1, (1, 1)), (1, (2, 1)), (2, (4, 4)), (2, (5, 3)), (2, (3, 5))} if levels == 3
              else {(1, (1, 1)), (1, (1, 2)), (1, (2, 1)), (1, (4, 4))})
    px, n = (3, 9) if levels == 3 else (6, 8)
    kw = dict(max_level=levels)
    ref = AmrGrid(OracleBackend(), px, px, n, lambda l, ij: (l, ij) in refine, fn, **kw)
    pure = AmrGrid(OracleBackend(), px, px, n, lambda l, ij: (l, ij) in refine, fn, pure=True,
                   **kw)
    worst = 0.0
    for it in range(150):
        d0, d1 = ref.step(), pure.step()
        assert abs(d1 - d0) <= 1e-12 * d0
        if it in (0, 1, 4) or it == 149:
            a, b = ref.leaf_states(), pure.leaf_states()
            rel = max(float((np.abs(a[k] - b[k]) / np.maximum(np.abs(a[k]), 1.0)).max())
                      for k in a)
            worst = max(worst, rel)
            assert rel <= (1e-12 if it < 5 else 1e-10), (it, rel)
    print(f"{levels} levels: pure vs reference call sequence, max rel {worst:.2e}")


def _front_grid(pure=False):
    fn = dam_break_fields(centre=(0.47, 0.53), radius=0.25, hills=[(0.3, 0.6, 0.1, 0.2)])
    return AmrGrid(OracleBackend(), 6, 6, 9, lambda lev, ij: False, fn, max_level=2, pure=pure)


def test_regrid_follows_the_front_and_conserves_mass_roughly():
    grid = _front_grid()
    refined, coarsened = grid.regrid()
    assert refined > 0 and coarsened == 0
    first = grid.refined_fraction()
    assert 0.0 < first < 1.0
    m0 = sum(float(v[0].sum()) * (1.0 / (54 * 3 ** (k[0] - 1))) ** 2
             for k, v in grid.leaf_states().items())
    fr = grid.run(60, regrid_every=10)
    st = grid.leaf_states()
    assert all(np.isfinite(v).all() for v in st.values())
    m1 = sum(float(v[0].sum()) * (1.0 / (54 * 3 ** (k[0] - 1))) ** 2 for k, v in st.items())
    assert abs(m1 - m0) / m0 < 5e-3
    assert len(fr) == 6 and all(0.0 < f < 1.0 for f in fr)
    print("refined fractions:", [round(f, 3) for f in [first] + fr])


def test_native_face_plan_follows_regrids():
    """The library's face rows (pc_amr_face_plan, what the batched driver
    uploads) equal the driver's incrementally updated plan after every
    regrid of a front that refines and coarsens."""
    grid = _front_grid()
    grid.regrid()
    for k in range(6):
        grid.run(10)
        grid.regrid()
        for x, y in zip(grid.packed_plan_native(), grid.packed_plan()):
            assert np.array_equal(x, y), k
    # coarsening as well: a lake at rest, fully refined, coarsens back
    def fn(X, Y):
        b = -1.0 + 0.004 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / 0.01)
        return 0.0 - b, b
    lake = AmrGrid(OracleBackend(), 3, 3, 6, lambda lev, ij: True, fn, max_level=2)
    lake.run(5)
    assert lake.regrid() == (0, 9)
    for x, y in zip(lake.packed_plan_native(), lake.packed_plan()):
        assert np.array_equal(x, y)


def test_refined_patches_reproduce_a_linear_field():
    """New patches are interpolated from their parent: a linear surface
    (Gaussian b) is reproduced at every new cell centre."""
    ax, ay = 0.6, -0.3

    def fn(X, Y):
        b = 0.3 * np.exp(-((X - 0.4) ** 2 + (Y - 0.6) ** 2) / 0.05)
        return 2.0 + ax * X + ay * Y - b, b
    be = OracleBackend()
    grid = AmrGrid(be, 3, 3, 6, lambda lev, ij: False, fn, max_level=2)
    # every leaf flagged for refinement: demanded width below the coarse width
    assert grid.regrid(threshold=-1.0) == (9, 0)
    assert len(grid.leaves) == 81
    for p in grid.leaves:
        u = be.state(p.handle)
        uo = be.uold(p.handle)[:, 1:-1, 1:-1]
        b = be.p[p.handle].pwg[0][1:-1, 1:-1]
        c = (np.arange(6) + 0.5) * (p.size / 6)
        X, Y = np.meshgrid(p.pos[0] + c, p.pos[1] + c, indexing="ij")
        for arr in (u, uo):
            np.testing.assert_allclose(arr[0] + b, 2.0 + ax * X + ay * Y, rtol=0, atol=1e-12)
            assert np.abs(arr[1:]).max() == 0.0


def test_coarsening_a_lake_at_rest_keeps_it():
    """A fully refined lake at rest (flat surface over a low bump, depth
    gradient below the threshold): every fine leaf allows coarsening; after
    coarsening the coarse leaves hold the same still surface and it stays
    at rest."""
    def fn(X, Y):
        # negligible at the walls, whose ghost b is not reflected
        b = -1.0 + 0.004 * np.exp(-((X - 0.5) ** 2 + (Y - 0.5) ** 2) / 0.01)
        return 0.0 - b, b
    be = OracleBackend()
    grid = AmrGrid(be, 3, 3, 6, lambda lev, ij: True, fn, max_level=2)
    grid.run(4)
    # fine leaves younger than 5 steps veto the coarsening of their parent
    assert grid.regrid() == (0, 0)
    grid.run(1)
    assert grid.regrid() == (0, 9)
    assert len(grid.leaves) == 9 and all(p.level == 1 for p in grid.leaves)
    grid.run(5)
    for p in grid.leaves:
        u = be.state(p.handle)
        b = be.p[p.handle].pwg[0][1:-1, 1:-1]
        assert np.abs(u[0] + b).max() <= 1e-10
        assert np.abs(u[1:]).max() <= 1e-10
