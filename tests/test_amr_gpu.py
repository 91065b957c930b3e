"""Two-level AMR through the C ABI (every Numerics call on the GPU) against
the same call sequence on the oracle: bitwise per step."""
import numpy as np
import pytest

from amr_oracle import OracleBackend

pytestmark = pytest.mark.gpu


def _grid_pair(px, n, refined, fn):
    from peanoclaw_amd.amr import DeviceBackend, TwoLevelGrid
    from peanoclaw_amd.pool import PatchPool
    pool = PatchPool(0)
    dev = TwoLevelGrid(DeviceBackend(pool), px, px, n, refined, fn)
    ora = TwoLevelGrid(OracleBackend(), px, px, n, refined, fn)
    return pool, dev, ora


def _bitwise_equal(dev, ora):
    a, b = dev.leaf_states(), ora.leaf_states()
    assert a.keys() == b.keys()
    same = sum(int((a[k].view(np.uint64) == b[k].view(np.uint64)).sum()) for k in a)
    total = sum(v.size for v in a.values())
    return same, total


@pytest.mark.parametrize("hills", [False, True])
def test_two_level_amr_bitwise(hills):
    from peanoclaw_amd.amr import dam_break_fields
    fn = dam_break_fields(centre=(0.47, 0.53), radius=0.25,
                          hills=[(0.3, 0.6, 0.1, 0.25), (0.7, 0.35, 0.08, 0.2)] if hills else ())
    refined = {(1, 1), (1, 2), (2, 1), (4, 4), (3, 2)}
    pool, dev, ora = _grid_pair(6, 8, refined, fn)
    try:
        for it in range(40):
            d1 = dev.step()
            d0 = ora.step()
            assert d1 == d0, f"dt differs at step {it}"
            if it in (0, 1, 9, 39):
                same, total = _bitwise_equal(dev, ora)
                print(f"AMR step {it + 1}: dt {d1:.6e}, bitwise {same}/{total}")
                assert same == total
        assert dev.time == ora.time
    finally:
        pool.close()


def test_two_level_amr_breaking_dam_27x27():
    """The 2-level dam break at its full size: 27 x 27 coarse patches of
    16^2 cells, jittered centre (seed 1), patches refined by factor 3 where
    the BreakingDam criterion (|grad h| > 0.05) asks for it on the initial
    data; a few steps bitwise against the oracle."""
    import time

    from peanoclaw_amd.amr import (DeviceBackend, TwoLevelGrid, dam_break_fields,
                                   refined_set_from_flags)
    from peanoclaw_amd.pool import PatchPool
    rng = np.random.default_rng(1)
    centre = tuple(0.5 + rng.uniform(-0.05, 0.05, 2))
    fn = dam_break_fields(centre=centre, radius=0.25)
    pool = PatchPool(0)
    try:
        refined = refined_set_from_flags(pool, 27, 16, fn)
        frac = len(refined) / 27 ** 2
        dev = TwoLevelGrid(DeviceBackend(pool), 27, 27, 16, refined, fn)
        ora = TwoLevelGrid(OracleBackend(), 27, 27, 16, refined, fn)
        t0 = time.perf_counter()
        for it in range(4):
            assert dev.step() == ora.step()
        el = time.perf_counter() - t0
        same, total = _bitwise_equal(dev, ora)
        print(f"2-level dam break 27x27x16^2: refined fraction {frac:.3f} of coarse patches, "
              f"{len(dev.leaves)} leaves, {dev.cells} leaf cells; bitwise {same}/{total}; "
              f"{el / 4:.3f} s per step for both drivers")
        assert same == total
        assert 0.0 < frac < 1.0
    finally:
        pool.close()


def test_three_level_wet_dry_basin_full_size():
    """The wet/dry stress case at its full size: 9 x 9 coarse patches of 27^2
    cells, three levels (finest dx = 1/2187), sloping basin with dry land
    and a depth-1.5 column; patches refine where dry and wet cells meet or
    the depth jumps (column edge).  Ghosts cross up to two levels.  Steps
    bitwise against the oracle; mass drift and depth clamps reported."""
    import time

    from oracle_lib import clamp_events
    from peanoclaw_amd.amr import AmrGrid, DeviceBackend, front_refinement, wet_dry_fields
    from peanoclaw_amd.pool import PatchPool
    fn = wet_dry_fields()
    px, n = 9, 27
    refine = front_refinement(px, px, n, fn, 3)
    pool = PatchPool(0)
    try:
        dev = AmrGrid(DeviceBackend(pool), px, px, n, refine, fn, max_level=3)
        ora = AmrGrid(OracleBackend(), px, px, n, refine, fn, max_level=3)
        levels = {lev: sum(p.level == lev for p in dev.leaves) for lev in (1, 2, 3)}
        assert all(v > 0 for v in levels.values())

        def mass(grid):
            return sum(float(v[0].sum()) * (1.0 / (px * n * 3 ** (k[0] - 1))) ** 2
                       for k, v in grid.leaf_states().items())
        m0 = mass(dev)
        c0 = clamp_events()
        t0 = time.perf_counter()
        steps = 12
        for it in range(steps):
            assert dev.step() == ora.step(), f"dt differs at step {it}"
        el = time.perf_counter() - t0
        same, total = _bitwise_equal(dev, ora)
        m1 = mass(dev)
        print(f"3-level wet/dry 9x9x27^2: leaves per level {levels}, {dev.cells} leaf cells "
              f"({dev.cells / 2187 ** 2:.3f} of 2187^2); {steps} steps to t={dev.time + dev.dt:.3e}; "
              f"bitwise {same}/{total}; mass drift {(m1 - m0) / m0:.2e}; "
              f"depth clamps {clamp_events() - c0}; {el / steps:.3f} s per step for both drivers")
        assert same == total
    finally:
        pool.close()


@pytest.mark.parametrize("levels", [2, 3])
def test_local_time_stepping_bitwise(levels):
    """Local time stepping (every leaf its own CFL step, blocked by its
    neighbours' time constraints, ghosts interpolated in time): the per-call
    device sequence against the oracle's, bitwise, with the clocks of every
    leaf equal after each grid iteration."""
    from peanoclaw_amd.amr import AmrGrid, DeviceBackend, dam_break_fields
    from peanoclaw_amd.pool import PatchPool
    fn = dam_break_fields(centre=(0.5, 0.5), radius=0.2, hills=[(0.3, 0.6, 0.1, 0.15)])

    def refine(lev, ij):
        return (lev == 1 and abs(ij[0] - 2) <= 1 and abs(ij[1] - 2) <= 1) or (
            lev == 2 and ij == (7, 7))
    pool = PatchPool(0)
    try:
        dev = AmrGrid(DeviceBackend(pool), 5, 5, 8, refine, fn, max_level=levels)
        ora = AmrGrid(OracleBackend(), 5, 5, 8, refine, fn, max_level=levels)
        end = 0.006
        its = steps = 0
        while True:
            k1 = dev.lts_iteration(end)
            k0 = ora.lts_iteration(end)
            assert k1 == k0, f"leaves advanced differ in iteration {its}"
            for p, q in zip(dev.leaves, ora.leaves):
                assert dev.b.clock(p.handle) == ora.b.clock(q.handle)
            its += 1
            steps += k1
            if k1 == 0:
                break
        same, total = _bitwise_equal(dev, ora)
        levels_seen = sorted({p.level for p in dev.leaves})
        print(f"LTS {levels}-level ({levels_seen}), {len(dev.leaves)} leaves: {its} iterations, "
              f"{steps} leaf steps to t={end}; bitwise {same}/{total}")
        assert same == total
    finally:
        pool.close()
