"""HIP path (through the C ABI) against the CPU oracle on the same inputs.

Bar: bit-exact where the arithmetic is the same sequence of IEEE-754
operations (all of the SWE path), checked as max relative difference
<= 1e-12 per step and <= 1e-10 after many steps, with the exact-equality
count printed for the record.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle_lib import OrPool, Patch as OPatch, dptr as odptr, iarr, lib as olib

pytestmark = pytest.mark.gpu

REL_STEP = 1e-12
REL_LONG = 1e-10


@pytest.fixture(scope="module")
def amd():
    import peanoclaw_amd  # noqa: F401
    from peanoclaw_amd.pool import PatchPool
    pool = PatchPool(0)
    yield pool
    pool.close()


def new_pool():
    from peanoclaw_amd.pool import PatchPool
    return PatchPool(0)


def relmax(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), 1e-300)
    m = np.where(d == 0, 0.0, d / np.maximum(s, 1.0))
    return float(m.max()) if m.size else 0.0


def random_edges(rng, n):
    e = np.empty((n, 6))
    e[:, 0:2] = rng.uniform(0.0, 3.0, (n, 2))
    dry = rng.random((n, 2)) < 0.1
    e[:, 0:2][dry] = rng.uniform(0.0, 0.02, dry.sum())
    e[:, 2:4] = rng.uniform(-2.0, 2.0, (n, 2))
    # signed zeros of the momentum (the sign of a zero quotient must match)
    z = rng.random((n, 2)) < 0.02
    e[:, 2:4][z] = np.where(rng.random(z.sum()) < 0.5, 0.0, -0.0)
    e[:, 4:6] = rng.uniform(-0.5, 0.5, (n, 2))
    return e


def test_edge_solver_bitwise(amd):
    rng = np.random.default_rng(7)
    e = random_edges(rng, 400000)
    got = amd.edge_batch(e)
    ref = np.empty((len(e), 5))
    olib().or_fwave_batch(len(e), odptr(np.ascontiguousarray(e)), odptr(ref))
    same = (got.view(np.uint64) == ref.view(np.uint64)).all(axis=1)
    print("edge solver: bitwise equal", int(same.sum()), "of", len(e))
    np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))


def _oracle_patch_from(unew, uold, b, pos, size, time, dt):
    nx, ny = unew.shape[1:]
    p = OPatch(3, 1, nx, 1, pos, size, 0, time, dt, ny=ny)
    p.unew[:] = unew
    p.uold[:] = uold
    p.pwg[0] = b
    return p


@pytest.mark.parametrize("nx,ny", [(16, 16), (8, 12), (64, 64), (27, 27)])
def test_solve_timestep_single_patch(amd, nx, ny):
    from peanoclaw_amd.pool import make_description
    rng = np.random.default_rng(nx * 100 + ny)
    d = make_description((0.25, 0.5), (0.3, 0.2), nx, ny=ny, time=0.7, timestep_size=0.01)
    s = amd.register(d)
    unew = np.stack([rng.uniform(0.5, 2.0, (nx, ny)), rng.uniform(-0.2, 0.2, (nx, ny)),
                     rng.uniform(-0.2, 0.2, (nx, ny))])
    uold = np.stack([rng.uniform(0.5, 2.0, (nx + 2, ny + 2)),
                     rng.uniform(-0.2, 0.2, (nx + 2, ny + 2)),
                     rng.uniform(-0.2, 0.2, (nx + 2, ny + 2))])
    uold[0, 3, 3] = 0.0  # a dry cell inside
    b = rng.uniform(-0.3, 0.3, (nx + 2, ny + 2))
    amd.upload(s, unew=unew, uold=uold, b=b)
    from peanoclaw_amd.numerics import Patch, SWENumerics
    num = SWENumerics(amd)
    dt, est = num.solveTimestep(Patch(amd, s), 1e-3)
    gn, go, gb = amd.download(s)
    op = _oracle_patch_from(unew, uold, b, (0.25, 0.5), (0.3, 0.2), 0.7, 0.01)
    olib().or_solve_timestep(op.ref, 1e-3)
    assert dt == op.s.dt and est == op.s.est_next
    np.testing.assert_array_equal(gn, op.unew)
    np.testing.assert_array_equal(go, op.uold)
    desc = amd.description(s)
    assert desc.time == op.s.time and desc.timestep_size == op.s.dt
    amd.release(s)


def test_boundary_transfer_copy_swap(amd):
    from peanoclaw_amd.numerics import Patch, SWENumerics
    from peanoclaw_amd.pool import make_description
    rng = np.random.default_rng(3)
    n = 8
    num = SWENumerics(amd)
    sa = amd.register(make_description((0.0, 0.0), 0.5, n, time=0.2, timestep_size=0.1))
    sb = amd.register(make_description((0.5, 0.0), 0.5, n, time=0.25, timestep_size=0.05))
    data = {}
    for s in (sa, sb):
        unew = rng.uniform(0.5, 2, (3, n, n))
        uold = rng.uniform(0.5, 2, (3, n + 2, n + 2))
        b = rng.uniform(0, 0.1, (n + 2, n + 2))
        amd.upload(s, unew=unew, uold=uold, b=b)
        data[s] = (unew, uold, b)
    oa = _oracle_patch_from(*data[sa], (0.0, 0.0), 0.5, 0.2, 0.1)
    ob = _oracle_patch_from(*data[sb], (0.5, 0.0), 0.5, 0.25, 0.05)
    # transfer a's right strip into b's left ghost layer (face functor)
    num.transferGhostlayer((1, n), (n - 1, 0), (-1, 0), Patch(amd, sa), Patch(amd, sb))
    olib().or_transfer_ghostlayer(iarr(1, n), iarr(n - 1, 0), iarr(-1, 0), oa.ref, ob.ref)
    for dim in (0, 1):
        for up in (0, 1):
            num.fillBoundaryLayer(Patch(amd, sa), dim, up)
            olib().or_fill_boundary_layer(oa.ref, dim, up)
    num.copyUNewToUOld(Patch(amd, sb))
    olib().or_copy_unew_to_uold(ob.ref)
    num.swapUNewAndUOld(Patch(amd, sa))
    olib().or_swap_unew_uold(oa.ref)
    for s, o in ((sa, oa), (sb, ob)):
        gn, go, gb = amd.download(s)
        np.testing.assert_array_equal(gn, o.unew)
        np.testing.assert_array_equal(go, o.uold)
    amd.release(sa)
    amd.release(sb)


def test_interpolation_and_restriction_swe(amd):
    from peanoclaw_amd.numerics import Patch, SWENumerics
    from peanoclaw_amd.pool import make_description
    rng = np.random.default_rng(11)
    num = SWENumerics(amd)
    nc, nf = 9, 9
    # coarse leaf at [0,1]^2 and a fine patch (refinement factor 3) inside it
    sc = amd.register(make_description((0.0, 0.0), 1.0, nc, time=0.3, timestep_size=0.02))
    sf = amd.register(make_description((1.0 / 3, 1.0 / 3), 1.0 / 3, nf, level=1, time=0.3,
                                       timestep_size=0.01))
    cu = rng.uniform(0.8, 1.5, (3, nc, nc))
    co = rng.uniform(0.8, 1.5, (3, nc + 2, nc + 2))
    cb = rng.uniform(0.0, 0.3, (nc + 2, nc + 2))
    fu = np.zeros((3, nf, nf))
    fo = rng.uniform(0.8, 1.5, (3, nf + 2, nf + 2))
    fb = rng.uniform(0.0, 0.3, (nf + 2, nf + 2))
    amd.upload(sc, unew=cu, uold=co, b=cb)
    amd.upload(sf, unew=fu, uold=fo, b=fb)
    oc = _oracle_patch_from(cu, co, cb, (0.0, 0.0), 1.0, 0.3, 0.02)
    of = _oracle_patch_from(fu, fo, fb, (1.0 / 3, 1.0 / 3), 1.0 / 3, 0.3, 0.01)
    # fine left ghost layer from the coarse patch (face functor arguments)
    num.interpolateSolution((1, nf), (-1, 0), Patch(amd, sc), Patch(amd, sf), True, False, True)
    assert olib().or_interpolate_swe(iarr(1, nf), iarr(-1, 0), oc.ref, of.ref, 1, 0, 1) == 0
    # whole fine patch into uNew and uOld (Remesh::createCell, Remesh.cpp:399-418)
    num.interpolateSolution((nf, nf), (0, 0), Patch(amd, sc), Patch(amd, sf), False, False,
                            False)
    assert olib().or_interpolate_swe(iarr(nf, nf), iarr(0, 0), oc.ref, of.ref, 0, 0, 0) == 0
    for s, o in ((sc, oc), (sf, of)):
        gn, go, gb = amd.download(s)
        assert relmax(gn, o.unew) == 0.0
        assert relmax(go, o.uold) == 0.0
    # restriction of the fine patch into a virtual coarse patch
    dv = make_description((0.0, 0.0), 1.0, nc, time=0.0, timestep_size=0.0, is_virtual=True)
    sv = amd.register(dv)
    ov = OPatch(3, 1, nc, 1, 0.0, 1.0, 0, 0.0, 0.0)
    ov.s.is_virtual = 1
    ov.s.is_leaf = 0
    vb = rng.uniform(0.0, 0.3, (nc + 2, nc + 2))
    amd.upload(sv, b=vb)
    ov.pwg[0] = vb
    num.restrictSolution(Patch(amd, sf), Patch(amd, sv), False)
    olib().or_restrict_swe(of.ref, ov.ref, 0)
    num.postProcessRestriction(Patch(amd, sv), False)
    olib().or_post_process_restriction_swe(ov.ref, 0)
    gn, go, gb = amd.download(sv)
    assert relmax(gn, ov.unew) == 0.0
    assert relmax(go, ov.uold) == 0.0
    gn, go, gb = amd.download(sf)
    assert relmax(gn, of.unew) == 0.0 and relmax(go, of.uold) == 0.0
    for s in (sc, sf, sv):
        amd.release(s)


class OracleRun:
    """The oracle's pool: the workload's tiles packed in slot order with rows
    of ny + 2 (the oracle's layout), seeded from the device's current state;
    the device tiles have rows of pool.stride(slot)."""

    def __init__(self, pool, wl):
        self.wl = wl
        nslots = max(wl.slots) + 1
        self.off = np.zeros(nslots, dtype=np.int64)
        nx = np.zeros(nslots, dtype=np.int32)
        ny = np.zeros(nslots, dtype=np.int32)
        dx = np.zeros(nslots)
        dy = np.zeros(nslots)
        top = 0
        for s in sorted(wl.slots):
            self.off[s] = top
            top += (pool.dims(s)[0] + 2) * (pool.dims(s)[1] + 2)
        self.top = top
        h, hu, hv, b = self.from_device(pool, with_b=True)
        self.H, self.HU, self.HV = [h, np.zeros(top)], [hu, np.zeros(top)], [hv, np.zeros(top)]
        self.B = b
        for s in wl.slots:
            d = pool.description(s)
            nx[s], ny[s] = d.subdivision_factor[0], d.subdivision_factor[1]
            dx[s] = d.size[0] / d.subdivision_factor[0]
            dy[s] = d.size[1] / d.subdivision_factor[1]
        self.nx, self.ny, self.dx, self.dy = nx, ny, dx, dy
        self.pl = OrPool()
        self.pl.np = nslots
        self.pl.nx = nx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        self.pl.ny = ny.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        self.pl.off = self.off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        self.pl.dx = dx.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self.pl.dy = dy.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        for k in range(2):
            self.pl.h[k] = odptr(self.H[k])
            self.pl.hu[k] = odptr(self.HU[k])
            self.pl.hv[k] = odptr(self.HV[k])
        self.pl.b = odptr(self.B)
        self.pl.cur = 0

    def from_device(self, pool, with_b=False):
        """h, hu, hv (and b) of the device's current state in the oracle's
        layout."""
        slots = sorted(self.wl.slots)
        dev_top = max(pool.offset(s) + (pool.dims(s)[0] + 2) * pool.stride(s) for s in slots)
        raw = [np.zeros(dev_top) for _ in range(4)]
        pool.copy_fields(False, pool.current_state(self.wl.scenario), 0, *raw)
        out = [np.zeros(self.top) for _ in range(4 if with_b else 3)]
        groups = {}
        for s in slots:
            groups.setdefault((pool.dims(s), pool.stride(s)), []).append(s)
        for ((nx, ny), st), ss in groups.items():
            for k0 in range(0, len(ss), 2048):
                part = ss[k0:k0 + 2048]
                doff = np.array([pool.offset(s) for s in part], dtype=np.int64)
                poff = self.off[part]
                x = np.arange(nx + 2)[None, :, None]
                y = np.arange(ny + 2)[None, None, :]
                di = doff[:, None, None] + x * st + y
                pi = poff[:, None, None] + x * (ny + 2) + y
                for o, r in zip(out, raw):
                    o[pi] = r[di]
        return out

    def fill(self):
        c = self.pl.cur
        for L in self.wl.layouts:
            so = np.ascontiguousarray(L.slot_of, dtype=np.int32)
            olib().or_grid_fill_rect(so.shape[0], so.shape[1], L.n, L.ny,
                                     so.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     self.off.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                     odptr(self.H[c]), odptr(self.HU[c]), odptr(self.HV[c]))

    def step(self):
        self.fill()
        return olib().or_pool_step(ctypes.byref(self.pl), math.inf, 8)

    def interior(self, pool, slot):
        nx, ny = self.nx[slot], self.ny[slot]
        o = self.off[slot]
        c = self.pl.cur
        t = lambda a: a[o:o + (nx + 2) * (ny + 2)].reshape(nx + 2, ny + 2)[1:-1, 1:-1]
        return np.stack([t(self.H[c]), t(self.HU[c]), t(self.HV[c])])


def _compare(pool, orun, slots):
    worst = 0.0
    exact = 0
    total = 0
    for s in slots:
        gn, _, _ = pool.download(s)
        ref = orun.interior(pool, s)
        worst = max(worst, relmax(gn, ref))
        exact += int((gn == ref).sum())
        total += gn.size
    return worst, exact, total


MODES = {
    # ghost strips gathered into the tiles, separate CFL pass
    "strips": dict(fused_sides=False, fused=0),
    # ghost values read from neighbours / edge buffers, separate CFL pass
    "sides": dict(fused_sides=True, fused=0),
    # as "sides", next-step CFL speeds computed inside the update kernel
    "sides_fused": dict(fused_sides=True, fused=1),
    # as "sides_fused", exact speeds only where a bound may reach the CFL limit
    "sides_pruned": dict(fused_sides=True, fused=2),
    # pruned with a threshold above every speed: each step takes the exact
    # fallback pass (slack < 1 rejects every pruned minimum)
    "sides_pruned_fallback": dict(fused_sides=True, fused=2, slack=0.5),
}


def _apply_mode(pool, wl, mode):
    m = MODES[mode]
    wl.build_ghost_table(fused_sides=m["fused_sides"])
    pool.launch_config(fused=m["fused"])
    pool.prune_slack(m.get("slack", 1.03125))


@pytest.mark.parametrize("mode", ["strips", "sides", "sides_fused", "sides_pruned",
                                  "sides_pruned_fallback"])
def test_dam_break_grid_200_steps(mode):
    from peanoclaw_amd.scenarios import dam_break_grid
    pool = new_pool()
    wl = dam_break_grid(pool, 9, 9, 16)
    _apply_mode(pool, wl, mode)
    orun = OracleRun(pool, wl)
    m0 = pool.mass()
    dts = []
    for it in range(200):
        pool.step(1)
        dts.append(orun.step())
        if it in (0, 1, 9):
            worst, exact, total = _compare(pool, orun, wl.slots)
            print(f"step {it + 1}: max rel {worst:.3e}, exact {exact}/{total}")
            assert worst <= REL_STEP
    t, dt, est, nsteps = pool.scenario_get(0)
    assert nsteps == 200
    assert dt == dts[-1]
    assert t == pytest.approx(sum(dts[:-1]), rel=1e-14)
    worst, exact, total = _compare(pool, orun, wl.slots)
    print(f"after 200 steps: max rel {worst:.3e}, exact {exact}/{total}")
    assert worst <= REL_LONG
    m1 = pool.mass()
    assert abs(m1 - m0) / m0 < 1e-13
    pool.close()


def test_dam_break_grid_1000_steps_graph():
    """C0 for 1000 steps in the autotuned configuration (pruned speeds,
    row segments, graph replays): bitwise against the oracle after 1000
    steps (the bar is 1e-10), the same dt, mass to 1e-13."""
    from peanoclaw_amd.scenarios import dam_break_grid
    pool = new_pool()
    wl = dam_break_grid(pool, 9, 9, 16)
    _apply_mode(pool, wl, "sides_pruned")
    pool.launch_config(persistent_workgroups=4096, fused=2)
    orun = OracleRun(pool, wl)
    m0 = pool.mass()
    pool.step(1)
    orun.step()
    pool.graph_capture()
    dts = []
    for _ in range(9):
        pool.graph_replay(111)
        pool.synchronize()
        for _ in range(111):
            dts.append(orun.step())
    t, dt, est, nsteps = pool.scenario_get(0)
    assert nsteps == 1000
    assert dt == dts[-1]
    worst, exact, total = _compare(pool, orun, wl.slots)
    m1 = pool.mass()
    print(f"C0 1000 steps (graph): max rel {worst:.3e}, bitwise {exact}/{total}, "
          f"mass drift {(m1 - m0) / m0:.2e}")
    assert worst <= REL_LONG and exact == total
    assert abs(m1 - m0) / m0 < 1e-13
    pool.close()


@pytest.mark.parametrize("mode", ["strips", "sides", "sides_fused", "sides_pruned",
                                  "sides_pruned_fallback"])
def test_uneven_patches_parity_and_forms(mode):
    from peanoclaw_amd.scenarios import uneven_patches
    pool = new_pool()
    wl = uneven_patches(pool, num_patches=240, seed=5, field_seed=6)
    _apply_mode(pool, wl, mode)
    orun = OracleRun(pool, wl)
    for it in range(5):
        pool.step(1)
        orun.step()
        worst, exact, total = _compare(pool, orun, wl.slots)
        print(f"uneven step {it + 1}: max rel {worst:.3e}, exact {exact}/{total}")
        assert worst <= REL_STEP
    # the plain per-patch-launch form continues identically
    pool.step_per_patch(3)
    for _ in range(3):
        orun.step()
    worst, _, _ = _compare(pool, orun, wl.slots)
    assert worst <= REL_STEP
    # hipGraph replays continue identically (replayed several times with
    # eager work and synchronisation between replays)
    pool.graph_capture()
    for _ in range(3):
        pool.graph_replay(2)
        pool.synchronize()
        for _ in range(2):
            orun.step()
        worst, _, _ = _compare(pool, orun, wl.slots)
        assert worst <= REL_STEP
    # persistent launch over the device work list
    pool.launch_config(persistent_workgroups=512, fused=MODES[mode]["fused"])
    pool.step(4)
    for _ in range(4):
        orun.step()
    worst, _, _ = _compare(pool, orun, wl.slots)
    assert worst <= REL_STEP
    pool.close()


def test_uneven_50k_full_size_bitwise():
    """The measured workload at its full size (50,000 uneven patches,
    68,241,728 cells) in the autotuned configuration (pruned speeds, 4096
    persistent workgroups): every interior cell of the pool bitwise equal
    to the oracle after two eager steps and after two more graph replays,
    and the same dt."""
    import time

    from peanoclaw_amd.scenarios import uneven_patches
    t0 = time.perf_counter()
    pool = new_pool()
    wl = uneven_patches(pool, 50000)
    wl.build_ghost_table(fused_sides=True)
    pool.launch_config(persistent_workgroups=4096, fused=2)
    orun = OracleRun(pool, wl)
    top = orun.H[0].size
    mask = np.zeros(top, dtype=bool)
    for s in wl.slots:
        nx, ny, o = int(orun.nx[s]), int(orun.ny[s]), int(orun.off[s])
        mask[o:o + (nx + 2) * (ny + 2)].reshape(nx + 2, ny + 2)[1:-1, 1:-1] = True
    assert int(mask.sum()) == wl.cells_per_step == 68241728
    print(f"50k workload set up in {time.perf_counter() - t0:.1f} s", flush=True)

    def compare():
        got = orun.from_device(pool)
        c = orun.pl.cur
        same = 0
        for g, ref in zip(got, (orun.H[c], orun.HU[c], orun.HV[c])):
            same += int((g[mask].view(np.uint64) == ref[mask].view(np.uint64)).sum())
        return same, 3 * int(mask.sum())

    dts = []
    for _ in range(2):
        pool.step(1)
        dts.append(orun.step())
    same, total = compare()
    print(f"50k uneven, 2 steps: bitwise {same}/{total}", flush=True)
    assert same == total
    pool.graph_capture()
    pool.graph_replay(2)
    pool.synchronize()
    for _ in range(2):
        dts.append(orun.step())
    same, total = compare()
    t, dt, est, n = pool.scenario_get(0)
    print(f"50k uneven, 2 more graph replays: bitwise {same}/{total}; dt {dt:.9e} (oracle "
          f"{dts[-1]:.9e}); {time.perf_counter() - t0:.1f} s in all", flush=True)
    assert same == total
    assert dt == dts[-1] and n == 4
    pool.close()


@pytest.mark.parametrize("fused", [2, 1, 0])
def test_row_segments_do_not_change_results(fused):
    """The update's row segments (whole patches, 2 or 4 segments, automatic)
    switched between steps of one run, in each speed mode: every step
    bitwise against the oracle (a changed split re-derives the per-segment
    interior speed maxima)."""
    from peanoclaw_amd.scenarios import uneven_patches
    pool = new_pool()
    wl = uneven_patches(pool, num_patches=240, seed=7, field_seed=8)
    wl.build_ghost_table(fused_sides=True)
    pool.launch_config(fused=fused)
    orun = OracleRun(pool, wl)
    for it, seg in enumerate([1, 4, 4, 2, 0, 1, 2, 4]):
        pool.row_segments(seg)
        pool.step(1)
        orun.step()
        worst, exact, total = _compare(pool, orun, wl.slots)
        assert exact == total, (it, seg, worst)
    t, dt, est, n = pool.scenario_get(0)
    print(f"row segments 1/4/4/2/auto/1/2/4 (fused {fused}): bitwise {exact}/{total}, dt {dt:.6e}")
    pool.close()


# (nx, ny, grid px, grid py): the smallest and largest sizes, both aspect
# ratios, odd sizes, a lone patch; every wavefront packing (W = 2 ... 64)
RAGGED = [(2, 2, 5, 4), (2, 64, 3, 2), (64, 2, 2, 3), (5, 3, 4, 4), (13, 7, 3, 5),
          (64, 64, 2, 2), (33, 17, 2, 3), (9, 41, 1, 1), (7, 7, 3, 3)]


def _ragged_workload(pool, seed=11):
    """Grids of rectangular patches (dx = dy = 1/64, one scenario, global
    dt); the 7 x 7 grid has a hole.  h from smooth noise in [0.6, 1.8],
    hu, hv ~ U[-0.1, 0.1], b = two Gaussian hills at every cell centre
    including the ghost ring."""
    from peanoclaw_amd.pool import make_description
    from peanoclaw_amd.scenarios import GridLayout, Workload
    rng = np.random.default_rng(seed)
    dx = 1.0 / 64
    layouts, slots, cells = [], [], 0
    x0 = 0.0
    for nx, ny, gpx, gpy in RAGGED:
        slot_of = -np.ones((gpx, gpy), dtype=np.int32)
        hills = [(rng.uniform(0, gpx * nx * dx), rng.uniform(0, gpy * ny * dx),
                  rng.uniform(0.05, 0.2), rng.uniform(0.0, 0.3)) for _ in range(2)]
        for i in range(gpx):
            for j in range(gpy):
                if (nx, ny) == (7, 7) and (i, j) == (1, 1):
                    continue  # hole: its neighbours see walls
                pos = (x0 + i * nx * dx, j * ny * dx)
                s = pool.register(make_description(pos, (nx * dx, ny * dx), nx, ny=ny))
                slot_of[i, j] = s
                X = pos[0] + (np.arange(-1, nx + 1)[:, None] + 0.5) * dx
                Y = pos[1] + (np.arange(-1, ny + 1)[None, :] + 0.5) * dx
                X, Y = np.broadcast_arrays(X, Y)
                b = np.zeros(X.shape)
                for hx_, hy_, sg, a in hills:
                    b = b + a * np.exp(-((X - hx_) ** 2 + (Y - hy_) ** 2) / (sg * sg))
                unew = np.empty((3, nx, ny))
                unew[0] = 1.2 + 0.3 * np.sin(7 * X[1:-1, 1:-1]) * np.cos(5 * Y[1:-1, 1:-1]) \
                    + 0.3 * rng.uniform(-1, 1, (nx, ny))
                unew[1:] = rng.uniform(-0.1, 0.1, (2, nx, ny))
                pool.upload(s, unew=unew, b=b)
                slots.append(s)
                cells += nx * ny
        layouts.append(GridLayout(nx, slot_of, ny=ny))
        x0 += gpx * nx * dx + 0.25
    pool.scenario_set(0, 0.0, 0.0, np.inf)
    return Workload(pool, layouts, 0, cells, slots)


@pytest.mark.parametrize("mode", ["strips", "sides", "sides_fused", "sides_pruned",
                                  "sides_pruned_fallback"])
def test_ragged_patch_grids_bitwise(mode):
    """Rectangular patches of every size class from 2 to 64 cells per side
    (grids, a lone patch and a grid with a hole) in one batched step:
    bitwise against the oracle every step, then the per-patch form and
    graph replays continue bitwise."""
    pool = new_pool()
    wl = _ragged_workload(pool)
    _apply_mode(pool, wl, mode)
    orun = OracleRun(pool, wl)
    for it in range(6):
        pool.step(1)
        orun.step()
        worst, exact, total = _compare(pool, orun, wl.slots)
        if it in (0, 5):
            print(f"ragged step {it + 1} ({mode}): bitwise {exact}/{total}")
        assert exact == total, (it, worst)
    pool.step_per_patch(2)
    pool.graph_capture()
    pool.graph_replay(2)
    pool.synchronize()
    for _ in range(4):
        orun.step()
    worst, exact, total = _compare(pool, orun, wl.slots)
    print(f"ragged after per-patch steps and graph replays ({mode}): bitwise {exact}/{total}")
    assert exact == total, worst
    pool.close()


@pytest.mark.parametrize("mode", ["strips", "sides_fused", "sides_pruned"])
def test_wet_dry_basin_300_steps(mode):
    """Dry land, a sloping basin and a dam column: the dry/wet edge
    branches and the depth/velocity clamps, over 300 steps."""
    from peanoclaw_amd.scenarios import wet_dry_grid
    pool = new_pool()
    wl = wet_dry_grid(pool, 9, 9, 16)
    _apply_mode(pool, wl, mode)
    orun = OracleRun(pool, wl)
    dry0 = sum(int((pool.download(s)[0][0] < 0.01).sum()) for s in wl.slots)
    assert dry0 > 1000
    for it in range(300):
        pool.step(1)
        orun.step()
        if it in (0, 9, 99):
            worst, exact, total = _compare(pool, orun, wl.slots)
            print(f"wet/dry step {it + 1}: max rel {worst:.3e}, exact {exact}/{total}")
            assert worst <= REL_STEP
    worst, exact, total = _compare(pool, orun, wl.slots)
    dry = sum(int((pool.download(s)[0][0] < 0.01).sum()) for s in wl.slots)
    print(f"wet/dry after 300 steps: max rel {worst:.3e}, exact {exact}/{total}, "
          f"dry cells {dry0} -> {dry}")
    assert worst <= REL_LONG
    pool.close()


def test_refinement_flags_bitexact():
    from peanoclaw_amd.scenarios import dam_break_grid
    pool = new_pool()
    wl = dam_break_grid(pool, 27, 27, 16, centre=(0.52, 0.47))
    pool.step(3)
    flags = pool.refinement_flags(0.05, 1.0 / 1296, 1.0 / 432, max(wl.slots) + 1)
    for s in wl.slots:
        gn, _, _ = pool.download(s)
        d = pool.description(s)
        op = OPatch(3, 1, 16, 1, (d.position[0], d.position[1]), d.size[0], 0, 0.0, 0.0)
        op.unew[:] = gn
        out = (ctypes.c_double * 2)()
        olib().or_demanded_mesh_width(op.ref, 0.05, 1.0 / 1296, 1.0 / 432, out)
        sub = (ctypes.c_double * 2)(*op.s.subcell)
        assert flags[s] == olib().or_refinement_decision(sub, out)
        # the per-call form (Numerics::getDemandedMeshWidth) of the same patch
        got = (ctypes.c_double * 2)()
        assert pool.lib.pc_demanded_mesh_width(pool.h, s, 0.05, 1.0 / 1296, 1.0 / 432, got) == 0
        assert list(got) == list(out)
    refined = float((flags[wl.slots] == 1).mean())
    print("refined fraction of coarse patches:", refined)
    assert 0.0 < refined < 1.0
    pool.close()
