"""Local time stepping of the AMR driver on the oracle backend (CPU).

Every leaf advances with its own CFL step, blocked while its current end
time is ahead of a neighbour's (SolveTimestep::enterCell,
TimeIntervals::isAllowedToAdvanceInTime); ghost values come from the
neighbours' current time intervals.  Checked: a lake at rest over a bump
stays at rest across two and three levels; each leaf reaches the end time;
fine leaves subcycle (about three steps per coarse step at refinement
factor 3); no ghost value is extrapolated in time (every transfer time
factor lies in [0, 1]); the result agrees with the global-step driver to
the order of the time-stepping difference.
"""
import numpy as np
import pytest

from amr_oracle import OracleBackend

from peanoclaw_amd.amr import AmrGrid, dam_break_fields


def _centre_refined(px):
    c = px // 2
    return lambda lev, ij: (lev == 1 and abs(ij[0] - c) <= 1 and abs(ij[1] - c) <= 1) or (
        lev == 2 and ij == (3 * c + 1, 3 * c + 1))


class TimeFactorProbe(OracleBackend):
    """Oracle backend that records, for every transfer from a leaf, where
    the destination's uNew time lies in the source's interval: the time
    factor of DefaultTransfer, (dst uNew time - src uOld time) / src dt, and
    the distances (in time) outside [src uOld time, src uNew time]."""

    def __init__(self):
        super().__init__()
        self.factors = []
        self.outside = []

    def transfer(self, size, soff, doff, src, dst):
        S, D = self.p[src].s, self.p[dst].s
        if S.is_leaf and S.dt > 0.0:
            t = D.time + D.dt
            self.factors.append((t - S.time) / S.dt)
            self.outside.append(max(S.time - t, t - (S.time + S.dt), 0.0))
        super().transfer(size, soff, doff, src, dst)


@pytest.mark.parametrize("levels", [2, 3])
def test_lts_lake_at_rest_stays_at_rest(levels):
    def fn(X, Y):
        # negligible at the walls, whose ghost b is not reflected
        b = -1.0 + 0.3 * np.exp(-((X - 0.45) ** 2 + (Y - 0.55) ** 2) / 0.005)
        return 0.0 - b, b
    grid = AmrGrid(OracleBackend(), 5, 5, 6, _centre_refined(5), fn, max_level=levels)
    assert {p.level for p in grid.leaves} == set(range(1, levels + 1))
    its, steps = grid.run_lts(0.02)
    assert steps > len(grid.leaves)
    for p in grid.leaves:
        t, dt = grid.b.clock(p.handle)
        assert t + dt == pytest.approx(0.02, rel=1e-12, abs=0.0)
        u = grid.b.state(p.handle)
        b = grid.b.p[p.handle].pwg[0][1:-1, 1:-1]
        assert np.abs(u[0] + b).max() <= 1e-12
        assert np.abs(u[1:]).max() <= 1e-12


def test_lts_dam_break_subcycles_without_time_extrapolation():
    fn = dam_break_fields(centre=(0.5, 0.5), radius=0.2, hills=[(0.3, 0.6, 0.1, 0.15)])
    be = TimeFactorProbe()
    grid = AmrGrid(be, 5, 5, 8, _centre_refined(5), fn, max_level=2)
    steps_before = {p.handle: p.age for p in grid.leaves}
    end = 0.015
    its, leaf_steps = grid.run_lts(end)
    taken = {p.handle: p.age - steps_before[p.handle] for p in grid.leaves}
    coarse = [taken[p.handle] for p in grid.leaves if p.level == 1]
    fine = [taken[p.handle] for p in grid.leaves if p.level == 2]
    ratio = np.mean(fine) / np.mean(coarse)
    print(f"LTS 2-level dam break: {its} iterations, {leaf_steps} leaf steps; "
          f"coarse steps {min(coarse)}..{max(coarse)}, fine {min(fine)}..{max(fine)}, "
          f"ratio {ratio:.2f}; time factors in [{min(be.factors):.3f}, {max(be.factors):.3f}]")
    assert 2.0 <= ratio <= 4.0
    # inside the source interval up to the blocking tolerance of 1e-8
    # (tarch::la comparisons), so no ghost value is extrapolated in time
    assert max(be.outside) <= 1e-8
    assert any(0.01 < f < 0.99 for f in be.factors)  # genuinely interpolated
    for p in grid.leaves:
        t, dt = be.clock(p.handle)
        assert t + dt == pytest.approx(end, rel=1e-12, abs=0.0)
        assert np.isfinite(be.state(p.handle)).all()


def test_lts_agrees_with_global_steps_to_first_order():
    fn = dam_break_fields(centre=(0.48, 0.52), radius=0.2)
    end = 0.01

    def run(lts):
        g = AmrGrid(OracleBackend(), 5, 5, 8, _centre_refined(5), fn, max_level=2)
        if lts:
            g.run_lts(end)
        else:
            while g.step(end) > 0.0:
                pass
        return g, g.leaf_states()
    g_lts, a = run(True)
    g_glb, b = run(False)
    for p in g_glb.leaves:
        t, dt = g_glb.b.clock(p.handle)
        assert t + dt == pytest.approx(end, rel=1e-12, abs=0.0)
    diff = max(float(np.abs(a[k][0] - b[k][0]).max()) for k in a)
    l1 = sum(float(np.abs(a[k][0] - b[k][0]).sum()) * (1.0 / (40 * 3 ** (k[0] - 1))) ** 2
             for k in a)
    print(f"LTS vs global steps at t={end}: max |dh| {diff:.3e}, L1 {l1:.3e}")
    # same solution up to the time-stepping difference: far below the dam
    # height difference of 1, not bitwise
    assert 0.0 < diff < 0.05
    assert l1 < 1e-3
