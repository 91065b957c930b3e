"""Augmented Riemann solver (the reference's solverType 2) on the GPU
against the oracle's restatement: bitwise per edge, per patch step, on the
batched wet/dry basin and on the three-level wet/dry AMR grid."""
import numpy as np
import pytest

from oracle_lib import dptr as odptr, lib as olib, wave_propagation_solver
from test_gpu_parity import (OPatch, OracleRun, REL_STEP, _apply_mode, _compare, new_pool,
                             random_edges)

pytestmark = pytest.mark.gpu


def _aug_edges(rng, n):
    """Random edges plus the wet/dry configurations of the solver: exactly
    dry sides, sides just around dryTol, shores above and below the wet
    surface (wall and inundation), fast run-up, lake at rest."""
    e = random_edges(rng, n)
    k = n // 8
    # one dry side with bathymetry above or below the other's surface
    side = rng.random(k) < 0.5
    e[:k, 0] = np.where(side, 0.0, rng.uniform(0.02, 2.0, k))
    e[:k, 1] = np.where(side, rng.uniform(0.02, 2.0, k), 0.0)
    e[:k, 4:6] = rng.uniform(-1.5, 1.5, (k, 2))
    # h close to dryTol
    e[k:2 * k, 0:2] = 0.01 + rng.uniform(-1e-3, 1e-3, (k, 2))
    # lake at rest
    eta = rng.uniform(-0.2, 0.2, k)
    e[2 * k:3 * k, 4:6] = rng.uniform(-2.0, eta[:, None] - 0.02, (k, 2))
    e[2 * k:3 * k, 0] = eta - e[2 * k:3 * k, 4]
    e[2 * k:3 * k, 1] = eta - e[2 * k:3 * k, 5]
    e[2 * k:3 * k, 2:4] = 0.0
    # supersonic flows both ways
    e[3 * k:4 * k, 2:4] = rng.uniform(-15.0, 15.0, (k, 2))
    return np.ascontiguousarray(e)


def test_augrie_edge_solver_bitwise():
    pool = new_pool()
    try:
        pool.wave_propagation_solver(2)
        e = _aug_edges(np.random.default_rng(21), 200000)
        got = pool.edge_batch(e)
        ref = np.empty((len(e), 5))
        with wave_propagation_solver(2):
            olib().or_edge_batch(len(e), odptr(e), odptr(ref))
        same = (got.view(np.uint64) == ref.view(np.uint64)).all(axis=1)
        print("augmented solver: bitwise equal", int(same.sum()), "of", len(e))
        np.testing.assert_array_equal(got.view(np.uint64), ref.view(np.uint64))
        # the switch back to the f-wave takes effect
        pool.wave_propagation_solver(1)
        fw = np.empty((len(e), 5))
        olib().or_fwave_batch(len(e), odptr(e), odptr(fw))
        np.testing.assert_array_equal(pool.edge_batch(e).view(np.uint64), fw.view(np.uint64))
    finally:
        pool.close()


@pytest.mark.parametrize("nx,ny", [(16, 16), (27, 27), (8, 12)])
def test_augrie_solve_timestep_single_patch(nx, ny):
    from peanoclaw_amd.numerics import Patch, SWENumerics
    from peanoclaw_amd.pool import make_description
    pool = new_pool()
    try:
        pool.wave_propagation_solver(2)
        rng = np.random.default_rng(nx * 31 + ny)
        d = make_description((0.25, 0.5), (0.3, 0.2), nx, ny=ny, time=0.7, timestep_size=0.01)
        s = pool.register(d)
        unew = np.stack([rng.uniform(0.5, 2.0, (nx, ny)), rng.uniform(-0.2, 0.2, (nx, ny)),
                         rng.uniform(-0.2, 0.2, (nx, ny))])
        uold = np.stack([rng.uniform(0.0, 2.0, (nx + 2, ny + 2)),
                         rng.uniform(-0.5, 0.5, (nx + 2, ny + 2)),
                         rng.uniform(-0.5, 0.5, (nx + 2, ny + 2))])
        uold[0, rng.random((nx + 2, ny + 2)) < 0.2] = 0.0  # dry cells
        b = rng.uniform(-0.6, 0.6, (nx + 2, ny + 2))
        pool.upload(s, unew=unew, uold=uold, b=b)
        dt, est = SWENumerics(pool).solveTimestep(Patch(pool, s), 1e-3)
        gn, go, _ = pool.download(s)
        op = OPatch(3, 1, nx, 1, (0.25, 0.5), (0.3, 0.2), 0, 0.7, 0.01, ny=ny)
        op.unew[:] = unew
        op.uold[:] = uold
        op.pwg[0] = b
        with wave_propagation_solver(2):
            olib().or_solve_timestep(op.ref, 1e-3)
        assert dt == op.s.dt and est == op.s.est_next
        np.testing.assert_array_equal(gn.view(np.uint64), np.asarray(op.unew).view(np.uint64))
        np.testing.assert_array_equal(go.view(np.uint64), np.asarray(op.uold).view(np.uint64))
    finally:
        pool.close()


@pytest.mark.parametrize("mode", ["strips", "sides", "sides_pruned"])
def test_augrie_wet_dry_basin_batched(mode):
    """The batched step with the augmented solver on the wet/dry basin
    (9 x 9 patches of 16^2): bitwise with the oracle's pool step; the fused
    modes fall back to the exact speed pass."""
    from peanoclaw_amd.scenarios import wet_dry_grid
    pool = new_pool()
    try:
        wl = wet_dry_grid(pool, 9, 9, 16)
        _apply_mode(pool, wl, mode)
        pool.wave_propagation_solver(2)
        orun = OracleRun(pool, wl)
        with wave_propagation_solver(2):
            for it in range(150):
                pool.step(1)
                orun.step()
                if it in (0, 9, 49):
                    worst, exact, total = _compare(pool, orun, wl.slots)
                    print(f"augmented wet/dry step {it + 1}: max rel {worst:.3e}, "
                          f"exact {exact}/{total}")
                    assert worst <= REL_STEP
            worst, exact, total = _compare(pool, orun, wl.slots)
            print(f"augmented wet/dry after 150 steps: max rel {worst:.3e}, exact {exact}/{total}")
            assert exact == total
            # graph replays continue identically
            pool.graph_capture()
            pool.graph_replay(3)
            pool.synchronize()
            for _ in range(3):
                orun.step()
            worst, exact, total = _compare(pool, orun, wl.slots)
            assert exact == total
    finally:
        pool.close()


def test_augrie_three_level_wet_dry_amr():
    """The three-level wet/dry basin (9 x 9 x 27^2, finest dx 1/2187) with
    the augmented solver: every Numerics call on the GPU, bitwise with the
    oracle over a few steps."""
    from amr_oracle import OracleBackend
    from peanoclaw_amd.amr import AmrGrid, DeviceBackend, front_refinement, wet_dry_fields
    from test_amr_gpu import _bitwise_equal
    fn = wet_dry_fields()
    px, n = 9, 27
    refine = front_refinement(px, px, n, fn, 3)
    pool = new_pool()
    try:
        pool.wave_propagation_solver(2)
        dev = AmrGrid(DeviceBackend(pool), px, px, n, refine, fn, max_level=3)
        with wave_propagation_solver(2):
            ora = AmrGrid(OracleBackend(), px, px, n, refine, fn, max_level=3)
            for it in range(6):
                assert dev.step() == ora.step(), f"dt differs at step {it}"
        same, total = _bitwise_equal(dev, ora)
        print(f"augmented 3-level wet/dry: bitwise {same}/{total}")
        assert same == total
    finally:
        pool.close()
