"""Pins for the oracle where the reference holds no vectors.

The reference's Riemann solver (SWE computeNetUpdates) is absent, so the
f-wave edge solver is pinned by what the mathematics fixes; the block loops
are pinned by decomposition invariance and conservation.  Each check is
chosen so that a plausible slip (dropped source term, wrong sign, swapped
operand, wrong accumulation target) fails it.
"""
import ctypes

import numpy as np
import pytest

from oracle_lib import Patch, dptr, fwave, lib

G = 9.81


def rng():
    return np.random.default_rng(1234)


def test_equal_states_give_zero_updates():
    r = rng()
    for _ in range(200):
        h = r.uniform(0.05, 3.0)
        hu = r.uniform(-1, 1)
        b = r.uniform(-1, 1)
        o = fwave(h, h, hu, hu, b, b)
        assert o[:4] == (0.0, 0.0, 0.0, 0.0)
        u = hu / h
        c = np.sqrt(G * h)
        assert o[4] == pytest.approx(abs(u) + c, rel=1e-14)


def test_lake_at_rest_is_well_balanced():
    r = rng()
    for _ in range(200):
        eta = r.uniform(1.0, 3.0)
        bL, bR = r.uniform(-0.5, 0.5, 2)
        o = fwave(eta - bL, eta - bR, 0.0, 0.0, bL, bR)
        assert max(abs(v) for v in o[:4]) < 1e-12 * G * eta


def test_h_flux_and_momentum_jump_are_conserved_per_edge():
    r = rng()
    for _ in range(500):
        hL, hR = r.uniform(0.05, 3.0, 2)
        huL, huR = r.uniform(-2, 2, 2)
        bL, bR = r.uniform(-0.3, 0.3, 2)
        dhL, dhR, dqL, dqR, s = fwave(hL, hR, huL, huR, bL, bR)
        assert dhL + dhR == pytest.approx(huR - huL, abs=1e-13 * (1 + abs(huR - huL)))
        uL, uR = huL / hL, huR / hR
        d1 = (huR * uR + 0.5 * G * hR * hR) - (huL * uL + 0.5 * G * hL * hL) \
            + 0.5 * G * (hL + hR) * (bR - bL)
        assert dqL + dqR == pytest.approx(d1, abs=1e-12 * (1 + abs(d1)))
        assert s > 0


def test_mirror_symmetry_is_bitwise():
    r = rng()
    for _ in range(500):
        hL, hR = r.uniform(0.0, 3.0, 2)
        huL, huR = r.uniform(-2, 2, 2)
        bL, bR = r.uniform(-0.3, 0.3, 2)
        a = fwave(hL, hR, huL, huR, bL, bR)
        m = fwave(hR, hL, -huR, -huL, bR, bL)
        assert m[0] == a[1] and m[1] == a[0]
        assert m[2] == -a[3] and m[3] == -a[2]
        assert m[4] == a[4]


def test_roe_limit_matches_textbook_decomposition():
    # b = 0, both wet and strongly subsonic with equal depths: the Einfeldt
    # bounds coincide with the Roe speeds, and the f-wave split equals the
    # Roe eigen-decomposition of the flux jump (LeVeque 2002, 15.3-15.5).
    r = rng()
    for _ in range(100):
        h = r.uniform(0.5, 2.0)
        huL, huR = r.uniform(-0.05, 0.05, 2)
        hL = hR = h
        dhL, dhR, dqL, dqR, s = fwave(hL, hR, huL, huR, 0.0, 0.0)
        uL, uR = huL / hL, huR / hR
        sq = np.sqrt(h)
        uhat = (uL * sq + uR * sq) / (2 * sq)
        chat = np.sqrt(G * h)
        lam = np.array([uhat - chat, uhat + chat])
        # Einfeldt bounds with equal depths are exactly these when
        # uL - c >= uhat - c and uR + c <= uhat + c, i.e. only if uL == uR;
        # otherwise compare against the HLLE-bounded speeds directly.
        l1 = min(uL - chat, uhat - chat)
        l2 = max(uR + chat, uhat + chat)
        R = np.array([[1.0, 1.0], [l1, l2]])
        df = np.array([huR - huL,
                       (huR * uR + 0.5 * G * hR ** 2) - (huL * uL + 0.5 * G * hL ** 2)])
        beta = np.linalg.solve(R, df)
        assert l1 < 0 < l2
        assert dhL == pytest.approx(beta[0], rel=1e-12, abs=1e-15)
        assert dhR == pytest.approx(beta[1], rel=1e-12, abs=1e-15)
        assert dqL == pytest.approx(beta[0] * l1, rel=1e-12, abs=1e-15)
        assert dqR == pytest.approx(beta[1] * l2, rel=1e-12, abs=1e-15)
        assert s == pytest.approx(max(abs(l1), abs(l2)), rel=1e-14)
        assert abs(lam[0]) <= s + 1e-12


def test_dry_cases():
    # both dry: nothing moves
    assert fwave(0.0, 0.001, 0.0, 0.0, 0.0, 0.0) == (0.0,) * 5
    # wet | dry: the dry side receives nothing, the wet side sees a wall
    dhL, dhR, dqL, dqR, s = fwave(1.0, 0.0, 0.3, 0.0, 0.0, 0.5)
    assert dhR == 0.0 and dqR == 0.0
    w = fwave(1.0, 1.0, 0.3, -0.3, 0.0, 0.0)
    assert (dhL, dqL, s) == (w[0], w[2], w[4])
    # dry | wet mirrors it
    m = fwave(0.0, 1.0, 0.0, -0.3, 0.5, 0.0)
    assert m[0] == 0.0 and m[2] == 0.0
    assert m[1] == dhL and m[3] == -dqL


def test_wall_edge_carries_no_mass():
    r = rng()
    for _ in range(200):
        h = r.uniform(0.05, 3.0)
        hu = r.uniform(-1, 1)
        dhL, dhR, dqL, dqR, s = fwave(h, h, hu, -hu, 0.3, 0.3)
        # interface mass flux F = huL + dhL vanishes at a reflecting wall
        assert abs(hu + dhL) <= 4e-16 * (abs(hu) + 1)


# ---------------------------------------------------------------- block --

def _dam_tile(n, r0=0.25):
    s = n + 2
    x = (np.arange(s) - 0.5) / n
    X, Y = np.meshgrid(x, x, indexing="ij")
    h = np.where(np.hypot(X - 0.5, Y - 0.5) < r0, 2.0, 1.0)
    return h, np.zeros((s, s)), np.zeros((s, s)), np.zeros((s, s))


def _reflect(h, hu, hv):
    # walls in the order left, top, right, bottom incl. corners
    h[0, :], hu[0, :], hv[0, :] = h[1, :], -hu[1, :], hv[1, :]
    h[:, -1], hu[:, -1], hv[:, -1] = h[:, -2], hu[:, -2], -hv[:, -2]
    h[-1, :], hu[-1, :], hv[-1, :] = h[-2, :], -hu[-2, :], hv[-2, :]
    h[:, 0], hu[:, 0], hv[:, 0] = h[:, 1], hu[:, 1], -hv[:, 1]


def _block_step(n, dx, h, hu, hv, b, maxdt=np.inf):
    net = [np.zeros_like(h) for _ in range(3)]
    smax = lib().or_block_fluxes(n, n, dx, dx, dptr(h), dptr(hu), dptr(hv),
                                 dptr(b), *[dptr(a) for a in net])
    dt = min(lib().or_block_max_timestep(smax, dx, dx), maxdt)
    lib().or_block_update(n, n, dt, dptr(h), dptr(hu), dptr(hv),
                          *[dptr(a) for a in net])
    return dt


def test_block_mass_conservation_with_walls():
    n = 48
    h, hu, hv, b = _dam_tile(n)
    m0 = h[1:-1, 1:-1].sum()
    for _ in range(60):
        _reflect(h, hu, hv)
        _block_step(n, 1.0 / n, h, hu, hv, b)
    assert abs(h[1:-1, 1:-1].sum() - m0) / m0 < 1e-13


def test_block_xy_symmetry_and_mirror_symmetry():
    n = 32
    h, hu, hv, b = _dam_tile(n)
    for _ in range(20):
        _reflect(h, hu, hv)
        _block_step(n, 1.0 / n, h, hu, hv, b)
    I = h[1:-1, 1:-1]
    # the dam is symmetric under x<->y and under both mirrors
    np.testing.assert_allclose(I, I.T, rtol=0, atol=1e-13)
    np.testing.assert_allclose(I, I[::-1, :], rtol=0, atol=1e-13)
    np.testing.assert_allclose(hu[1:-1, 1:-1], hv[1:-1, 1:-1].T, rtol=0, atol=1e-13)


def test_block_accumulation_targets():
    # a single perturbed cell: only it and its 4 face neighbours change, and
    # hu only changes through x-edges, hv only through y-edges
    n = 6
    s = n + 2
    h = np.ones((s, s)); hu = np.zeros((s, s)); hv = np.zeros((s, s)); b = np.zeros((s, s))
    h[3, 3] = 1.5
    h0 = h.copy()
    _reflect(h, hu, hv)
    _block_step(n, 0.1, h, hu, hv, b, maxdt=1e-3)
    changed = np.argwhere(h[1:-1, 1:-1] != h0[1:-1, 1:-1]) + 1
    assert sorted(map(tuple, changed)) == [(2, 3), (3, 2), (3, 3), (3, 4), (4, 3)]
    assert np.all(hu[3, 2] == 0) and np.all(hu[3, 4] == 0)
    assert hu[2, 3] < 0 < hu[4, 3]
    assert hv[3, 2] < 0 < hv[3, 4]
    assert hv[2, 3] == 0 and hv[4, 3] == 0


def test_block_max_timestep_rule():
    lt = lib().or_block_max_timestep
    assert lt(2.0, 0.1, 0.2) == pytest.approx(0.4 * 0.05)
    assert lt(1e-6, 0.1, 0.1) == np.finfo(np.float64).max


def test_solve_timestep_patch_semantics():
    # SWEKernel::solveTimestep on a patch: block on uOld, dt = min(cfl, maxDt),
    # swap uNew/uOld interior, advance time intervals.
    n = 8
    p = Patch(3, 1, n, 1, 0.0, 1.0, 0, 0.5, 0.1)
    x = (np.arange(n) + 0.5) / n
    p.unew[0] = 1.0 + 0.5 * np.exp(-((x[:, None] - 0.4) ** 2 + (x[None, :] - 0.6) ** 2) * 20)
    lib().or_copy_unew_to_uold(p.ref)
    for dim in (0, 1):
        for up in (0, 1):
            lib().or_fill_boundary_layer(p.ref, dim, up)
    before = p.unew.copy()
    lib().or_solve_timestep(p.ref, 1e-3)
    assert p.s.dt == 1e-3 and p.s.time == pytest.approx(0.6)
    assert p.s.est_next > 1e-3
    np.testing.assert_array_equal(p.uold[:, 1:-1, 1:-1], before)
    assert not np.array_equal(p.unew, before)


def test_interpolation_is_exact_on_linear_fields():
    # d-linear interpolation (incl. clamped extrapolation) reproduces linear
    # data exactly; time interpolation is linear between uOld and uNew.
    c = Patch(1, 0, 6, 1, 0.0, 3.0, 0, 0.0, 1.0)
    f = Patch(1, 0, 6, 1, 1.0, 1.0, 1, 0.0, 0.25)
    xs = (np.arange(6) + 0.5) * 0.5
    Z = 0.3 + 1.7 * xs[:, None] - 0.9 * xs[None, :]
    c.unew[0] = Z + 1.0
    c.uold[0, 1:-1, 1:-1] = Z
    size = (ctypes.c_int * 2)(6, 1)
    off = (ctypes.c_int * 2)(0, -1)
    lib().or_interpolate_default(size, off, c.ref, f.ref, 1, 0, 1)
    fx = 1.0 + (np.arange(6) + 0.5) / 6
    fy = 1.0 - 0.5 / 6
    expect = 0.3 + 1.7 * fx - 0.9 * fy + 0.25
    np.testing.assert_allclose(f.uold[0, 1:-1, 0], expect, rtol=0, atol=1e-13)


def test_restriction_conserves_the_area_average():
    src = Patch(1, 0, 9, 1, 1.0 / 3, 1.0 / 3, 1, 0.0, 0.0)
    dst = Patch(1, 0, 3, 1, 0.0, 1.0, 0, 0.0, 0.0)
    dst.s.is_virtual = 1
    dst.s.is_leaf = 0
    r = rng()
    src.unew[0] = r.uniform(0, 1, (9, 9))
    src.uold[0, 1:-1, 1:-1] = src.unew[0]
    lib().or_restrict_default(src.ref, dst.ref, 0)
    # the centre coarse cell is the 9x9 block's mean; others untouched
    assert dst.unew[0, 1, 1] == pytest.approx(src.unew[0].mean(), rel=1e-13)
    assert dst.unew[0].sum() == pytest.approx(src.unew[0].mean(), rel=1e-13)


def test_breaking_dam_refinement_closed_forms():
    p = Patch(3, 1, 16, 1, (0.25, 0.25), 1.0 / 3, 0, 0.0, 0.0)
    lib().or_breaking_dam_init(p.ref, 0.5, 0.5, 0.25, 2.0, 1.0)
    out = (ctypes.c_double * 2)()
    lib().or_demanded_mesh_width(p.ref, 0.05, 1 / 1296, 1 / 432, out)
    assert out[0] == 1 / 1296    # the jump of 1 over one cell is a gradient of 1/dx
    sub = (ctypes.c_double * 2)(*p.s.subcell)
    assert lib().or_refinement_decision(sub, out) == 1
    flat = Patch(3, 1, 16, 1, (0.0, 0.0), 1.0 / 27, 0, 0.0, 0.0)
    lib().or_breaking_dam_init(flat.ref, 0.5, 0.5, 0.25, 2.0, 1.0)
    lib().or_demanded_mesh_width(flat.ref, 0.05, 1 / 1296, 1 / 432, out)
    assert out[0] == 1 / 432
    sub = (ctypes.c_double * 2)(*flat.s.subcell)
    # already at the coarsest width: 3*dx > demanded, so the erase veto holds
    assert lib().or_refinement_decision(sub, out) == 2
    # a fine patch (dx = 1/1296) in flat water may coarsen
    sub = (ctypes.c_double * 2)(1 / 1296, 1 / 1296)
    assert lib().or_refinement_decision(sub, out) == 0


def test_rectangular_grid_ghost_fill_brute_force():
    """or_grid_fill_rect on a 3 x 2 grid of 3 x 5 patches with a hole: every
    face ghost cell holds the facing neighbour's boundary cell (same-level
    transfer), or at a wall (outer border or hole) the patch's own boundary
    cell with the normal momentum negated (SWEKernel.cpp:176-237); checked
    cell by cell in plain Python."""
    gpx, gpy, nx, ny = 3, 2, 3, 5
    grid = [[0, 1], [2, -1], [3, 4]]
    area = (nx + 2) * (ny + 2)
    r = rng()
    H, HU, HV = (r.uniform(1, 2, 5 * area), r.uniform(-1, 1, 5 * area),
                 r.uniform(-1, 1, 5 * area))
    h0, hu0, hv0 = H.copy(), HU.copy(), HV.copy()
    slot_of = np.array(grid, dtype=np.int32).ravel()
    off = np.arange(5, dtype=np.int64) * area
    lib().or_grid_fill_rect(gpx, gpy, nx, ny,
                            slot_of.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                            dptr(H), dptr(HU), dptr(HV))

    def at(i, j):
        return grid[i][j] if 0 <= i < gpx and 0 <= j < gpy else -1

    def idx(s, x, y):
        return s * area + x * (ny + 2) + y

    checked = 0
    for i in range(gpx):
        for j in range(gpy):
            s = at(i, j)
            if s < 0:
                continue
            # interior untouched
            for x in range(1, nx + 1):
                for y in range(1, ny + 1):
                    k = idx(s, x, y)
                    assert (H[k], HU[k], HV[k]) == (h0[k], hu0[k], hv0[k])
            faces = [(0, -1, 0), (nx + 1, 1, 0)]  # x faces: ghost row, di
            for gx, di, _ in faces:
                q = at(i + di, j)
                for y in range(1, ny + 1):
                    k = idx(s, gx, y)
                    if q >= 0:
                        src = idx(q, nx if di < 0 else 1, y)
                        want = (h0[src], hu0[src], hv0[src])
                    else:
                        src = idx(s, 1 if di < 0 else nx, y)
                        want = (h0[src], -hu0[src], hv0[src])
                    assert (H[k], HU[k], HV[k]) == want, (s, gx, y)
                    checked += 1
            for gy, dj in ((0, -1), (ny + 1, 1)):
                q = at(i, j + dj)
                for x in range(1, nx + 1):
                    k = idx(s, x, gy)
                    if q >= 0:
                        src = idx(q, x, ny if dj < 0 else 1)
                        want = (h0[src], hu0[src], hv0[src])
                    else:
                        src = idx(s, x, 1 if dj < 0 else ny)
                        want = (h0[src], hu0[src], -hv0[src])
                    assert (H[k], HU[k], HV[k]) == want, (s, x, gy)
                    checked += 1
    assert checked == 5 * 2 * (nx + ny)
