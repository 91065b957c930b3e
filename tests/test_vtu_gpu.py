"""VTU and legacy VTK output straight from the device pool (pc_pool_write_vtu,
pc_pool_write_vtk): the cell
arrays equal the slots' downloaded uNew and b, the times their
descriptors'."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_pool_vtu_matches_download(tmp_path):
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.scenarios import dam_break_grid
    from peanoclaw_amd.vtu import read_vtu, write_pool_vtu
    pool = PatchPool(0)
    try:
        wl = dam_break_grid(pool, 3, 3, 8)
        pool.step(5)
        slots = list(wl.slots)
        f = tmp_path / "pool.vtu"
        write_pool_vtu(pool, f, slots)
        pts, conn, types, cell = read_vtu(f)
        assert len(pts) == (3 * 8 + 1) ** 2 and conn.shape[0] == 9 * 64
        for k, s in enumerate(slots):
            unew, _, b = pool.download(s)
            d = pool.description(s)
            sl = slice(k * 64, (k + 1) * 64)
            for q in range(3):
                assert np.array_equal(cell[f"q{q}"][sl], unew[q].ravel())
            assert np.array_equal(cell["pWG0"][sl], b[1:-1, 1:-1].ravel())
            assert (cell["timeOld"][sl] == d.time).all()
            assert (cell["timeNew"][sl] == d.time + d.timestep_size).all()
        print(f"pool VTU: {len(pts)} points, {conn.shape[0]} cells, t = {cell['timeNew'][0]:.4e}")
    finally:
        pool.close()


def test_pool_vtk_matches_download(tmp_path):
    """Legacy VTK from the device pool (pc_pool_write_vtk): the same cell
    arrays as the downloaded state and descriptors."""
    from peanoclaw_amd.pool import PatchPool
    from peanoclaw_amd.scenarios import dam_break_grid
    from peanoclaw_amd.vtu import read_vtk, write_pool_vtk
    pool = PatchPool(0)
    try:
        wl = dam_break_grid(pool, 3, 3, 8)
        pool.step(3)
        slots = list(wl.slots)
        f = tmp_path / "pool.vtk"
        write_pool_vtk(pool, f, slots)
        pts, conn, types, cell = read_vtk(f)
        assert conn.shape[0] == 9 * 64 and (types == 9).all()
        for k, s in enumerate(slots):
            unew, _, b = pool.download(s)
            d = pool.description(s)
            sl = slice(k * 64, (k + 1) * 64)
            for q in range(3):
                assert np.array_equal(cell[f"q{q}"][sl], unew[q].ravel())
            assert np.array_equal(cell["pWG0"][sl], b[1:-1, 1:-1].ravel())
            assert (cell["timeNew"][sl] == d.time + d.timestep_size).all()
    finally:
        pool.close()
