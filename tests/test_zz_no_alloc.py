"""pc_edge_batch_static, the entry point that allocates nothing, against
pc_edge_batch (which stages through pool memory), through the C ABI: two
calls in a row (the second with fewer edges, so stale staging contents
would show), both edge solvers, bitwise; more edges than its static array
holds fail with PC_E_INVALID and leave it usable."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _edges(rng, n):
    e = np.empty((n, 6))
    e[:, 0:2] = rng.uniform(0.0, 3.0, (n, 2))
    dry = rng.random((n, 2)) < 0.1
    e[:, 0:2][dry] = rng.uniform(0.0, 0.02, dry.sum())
    e[:, 2:4] = rng.uniform(-2.0, 2.0, (n, 2))
    e[:, 4:6] = rng.uniform(-0.5, 0.5, (n, 2))
    return e


def _static(pool, edges):
    from peanoclaw_amd._abi import check
    edges = np.ascontiguousarray(edges, dtype=np.float64)
    out = np.full((edges.shape[0], 5), np.nan)
    dp = ctypes.POINTER(ctypes.c_double)
    check(pool.lib.pc_edge_batch_static(pool.h, edges.shape[0], edges.ctypes.data_as(dp),
                                        out.ctypes.data_as(dp)))
    return out


@pytest.mark.parametrize("solver", [1, 2])
def test_edge_batch_static_equals_edge_batch(solver):
    from peanoclaw_amd._abi import PeanoClawAmdError
    from peanoclaw_amd.pool import PatchPool
    rng = np.random.default_rng(40 + solver)
    with PatchPool(0) as pool:
        pool.wave_propagation_solver(solver)
        for n in (65536, 1000):
            e = _edges(rng, n)
            print(f"solver {solver}: {n} edges through the static staging", flush=True)
            got = _static(pool, e)
            ref = pool.edge_batch(e)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64))
        with pytest.raises(PeanoClawAmdError) as err:
            _static(pool, _edges(rng, 65537))
        assert err.value.code == 1
        e = _edges(rng, 7)
        assert np.array_equal(_static(pool, e).view(np.uint64),
                              pool.edge_batch(e).view(np.uint64))
