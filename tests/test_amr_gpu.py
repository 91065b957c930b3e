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
